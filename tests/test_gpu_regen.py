"""Unweight by regeneration (hpsp_unweight_regenerate) against Unweight by
moving (hpsp_unweight, itself bit-exact with the reference's
Decays::Unweight, see test_gpu_parity): for a container that holds the events
of one Generate, regenerating each event at its destination must give the
same container, index map, weights and accepted count, bit for bit."""
import numpy as np
import pytest
import torch

from gpu_util import DEV, generate, make_decay, unweight, unweight_regenerate
from oracle import ENGINES

pytestmark = pytest.mark.gpu

UW_SEED = 0x180EC6D33CFD0ABA


def bits(t):
    return t.cpu().numpy().view(np.uint64)


def same(dec, gen_seed, first, n, uw_seed):
    cols, _, evw = generate(dec, gen_seed, n, first=first)
    k1, out1, idx1, w1 = unweight(dec, cols, uw_seed, w_in=evw)
    k2, out2, idx2, w2 = unweight_regenerate(dec, gen_seed, first, n, evw, uw_seed)
    assert k1 == k2
    assert np.array_equal(bits(out1[:, :n]), bits(out2[:, :n]))
    assert np.array_equal(idx1.cpu().numpy()[:n], idx2.cpu().numpy()[:n])
    assert np.array_equal(bits(w1[:n]), bits(w2[:n]))
    return k1


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 7, 12])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_regen_equals_move(N, engine):
    rng = np.random.default_rng(300 + 13 * N + ENGINES[engine])
    m = list(rng.uniform(0.1, 1.0, N))
    # room enough above threshold that the weights (products of N-1 momenta)
    # often pass the acceptance test
    M = sum(m) * rng.uniform(1.5, 3.0) + 2.0
    dec = make_decay(M, m, engine=engine)
    seed = int(rng.integers(0, 2**63))
    first = int(rng.integers(0, 2**40))
    k = same(dec, seed, first, 20000 + N, int(rng.integers(0, 2**63)))
    assert 0 <= k <= 20000 + N


def test_regen_moving_mother_and_sizes():
    m = [3.0969, 0.493677, 0.13957]
    dec = make_decay(5.279, m, mother4=[6.0, 0.3, -0.2, 0.5])
    for n in (1, 63, 64, 65, 1023, 1024, 1025, 4096 + 17):
        same(dec, 159753, 7, n, UW_SEED)


@pytest.mark.parametrize("N", [2, 6, 7])
def test_regen_ragged_sizes(N):
    """Partial blocks and tiles at the end, for one staging round (N = 2),
    two (N = 6) and none (N = 7, stored from registers)."""
    rng = np.random.default_rng(40 + N)
    m = list(rng.uniform(0.1, 0.5, N))
    dec = make_decay(sum(m) * 2.0 + 1.0, m)
    for n in (1, 255, 257, 1000, 1025, 3 * 1024 + 200):
        same(dec, 11 + n, 3, n, UW_SEED + n)


def test_regen_headline_size_and_outputs():
    """2^22 events of the benchmark decay; outputs may be skipped."""
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    n = 1 << 22
    k = same(dec, 159753, 0, n, UW_SEED)
    cols, _, evw = generate(dec, 159753, n)
    k2, out, idx, w = unweight_regenerate(dec, 159753, 0, n, evw, UW_SEED, want_idx=False,
                                          want_w=False)
    assert k2 == k and idx is None and w is None
    k3, out3, _, _ = unweight_regenerate(dec, 159753, 0, n, evw, UW_SEED, want_out=False)
    assert k3 == k and out3 is None


def test_regen_empty():
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    w = torch.empty(1, dtype=torch.float64, device=DEV)
    k, _, _, _ = unweight_regenerate(dec, 1, 0, 0, w, UW_SEED)
    assert k == 0


@pytest.mark.parametrize("engine", ["squares3", "philox"])
def test_weights_only_generate(engine):
    """hpsp_generate with no columns writes the same weights; with the
    regenerating Unweight it gives the same container as the eager step."""
    import ctypes
    from hydra_amd import capi
    from hydra_amd.capi import check, lib, ptr, stream_handle
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957], engine=engine)
    n = 50000
    cols, wt, evw = generate(dec, 99, n, first=1234)
    wt2 = torch.empty(n, dtype=torch.float64, device=DEV)
    evw2 = torch.empty(n, dtype=torch.float64, device=DEV)
    check(lib().hpsp_generate(ctypes.byref(dec), 99, 1234, n, capi.FP64, None, n, 0, ptr(wt2),
                              ptr(evw2), stream_handle()), "hpsp_generate")
    torch.cuda.synchronize()
    assert np.array_equal(bits(wt), bits(wt2)) and np.array_equal(bits(evw), bits(evw2))
    k1, out1, idx1, _ = unweight(dec, cols, UW_SEED, w_in=evw)
    k2, out2, idx2, _ = unweight_regenerate(dec, 99, 1234, n, evw2, UW_SEED)
    assert k1 == k2 and np.array_equal(bits(out1), bits(out2))
