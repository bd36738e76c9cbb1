"""Generate followed by Decays::Unweight in one call (hpsp_generate_unweight:
the flags from each event's generation weight where a bound on its distance
from PhaseSpaceWeight settles the test, from PhaseSpaceWeight elsewhere; then
every event generated at its place) against the two calls it replaces
(hpsp_generate of the PhaseSpaceWeight column, then
hpsp_unweight_regenerate, itself bit-exact with the moving Unweight, see
test_gpu_regen): the same container, index map, weights and accepted count,
bit for bit, for every N and engine, ragged sizes, a moving mother, decays
near threshold (where the bound is large) and light daughters beside a heavy
one (where it is largest)."""
import ctypes

import numpy as np
import pytest
import torch

from gpu_util import DEV, check, generate, lib, make_decay, ptr, stream_handle, unweight, \
    unweight_regenerate
from hydra_amd import capi
from oracle import ENGINES

pytestmark = pytest.mark.gpu

UW_SEED = 0x180EC6D33CFD0ABA


def bits(t):
    return t.cpu().numpy().view(np.uint64)


def generate_unweight(dec, gen_seed, first, n, seed, want_idx=True, want_w=True, ld=None):
    ld = ld or max(n, 1)
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty((4 * dec.n, ld), dtype=torch.float64, device=DEV)
    idx = torch.empty(max(n, 1), dtype=torch.int64, device=DEV) if want_idx else None
    w = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if want_w else None
    nacc = torch.full((1,), 12345, dtype=torch.int64, device=DEV)
    check(lib().hpsp_generate_unweight(ctypes.byref(dec), gen_seed, first, n, seed, ptr(out), ld,
                                       ptr(idx), ptr(w), ptr(ws), ws_bytes, ptr(nacc),
                                       stream_handle()), "hpsp_generate_unweight")
    torch.cuda.synchronize()
    return int(nacc.item()), out, idx, w


def same(dec, gen_seed, first, n, uw_seed):
    _, _, evw = generate(dec, gen_seed, n, first=first, want_wt=False)
    k1, out1, idx1, w1 = unweight_regenerate(dec, gen_seed, first, n, evw, uw_seed)
    print(f"N={dec.n} engine={dec.engine} n={n} first={first}", flush=True)
    k2, out2, idx2, w2 = generate_unweight(dec, gen_seed, first, n, uw_seed)
    assert k1 == k2
    assert np.array_equal(bits(out1[:, :n]), bits(out2[:, :n]))
    assert np.array_equal(idx1.cpu().numpy()[:n], idx2.cpu().numpy()[:n])
    assert np.array_equal(bits(w1[:n]), bits(w2[:n]))
    return k1


@pytest.mark.parametrize("N", [2, 3, 4, 5, 6, 7, 12])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_generate_unweight_equals_two_calls(N, engine):
    rng = np.random.default_rng(700 + 13 * N + ENGINES[engine])
    m = list(rng.uniform(0.1, 1.0, N))
    M = sum(m) * rng.uniform(1.5, 3.0) + 2.0
    dec = make_decay(M, m, engine=engine)
    seed = int(rng.integers(0, 2**63))
    first = int(rng.integers(0, 2**40))
    k = same(dec, seed, first, 20000 + N, int(rng.integers(0, 2**63)))
    assert 0 <= k <= 20000 + N


def test_generate_unweight_sizes_and_moving_mother():
    """Partial last blocks, one block, and a boosted (moving) mother."""
    m = [3.0969, 0.493677, 0.13957]
    dec = make_decay(5.279, m, mother4=[6.0, 0.3, -0.2, 0.5])
    for n in (1, 63, 255, 256, 257, 1023, 4096 + 17):
        same(dec, 159753, 7, n, UW_SEED)


def test_generate_unweight_headline_decay_many_blocks():
    """The headline decay over 2^22 events, and the optional outputs left
    out."""
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    n = 1 << 22
    k = same(dec, 159753, 0, n, UW_SEED)
    k2, out2, _, _ = generate_unweight(dec, 159753, 0, n, UW_SEED, want_idx=False, want_w=False)
    assert k2 == k
    k3, out3, _, _ = generate_unweight(dec, 159753, 0, n, UW_SEED)
    assert torch.equal(out2, out3)


@pytest.mark.parametrize("case", ["threshold3", "threshold10", "light", "philox_threshold",
                                  "tiny_first", "tiny_12"])
def test_generate_unweight_ill_conditioned(case):
    """Near threshold and light daughters the bound is large and many events
    take the exact test; with a first daughter of 1 keV the bound is far above
    its limit and every event does."""
    M3 = 3.0969 + 0.493677 + 0.13957
    decays = {
        "threshold3": (M3 + 0.001, [3.0969, 0.493677, 0.13957], "squares3"),
        "threshold10": (10.0, [0.999] * 10, "squares3"),
        "light": (100.0, [0.000511, 0.000511, 90.0], "squares4"),
        "philox_threshold": (M3 + 0.01, [3.0969, 0.493677, 0.13957], "philox"),
        "tiny_first": (5.279, [1e-6, 0.493677, 0.13957], "squares3"),
        "tiny_12": (20.0, [1e-6] + [0.5] * 11, "threefry"),
    }
    M, m, engine = decays[case]
    dec = make_decay(M, m, engine=engine)
    same(dec, 2024, 3, 1 << 20, UW_SEED)


def test_generate_unweight_empty_and_errors():
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    k, _, _, _ = generate_unweight(dec, 1, 0, 0, UW_SEED)
    assert k == 0
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    ws = torch.empty(16, dtype=torch.uint8, device=DEV)
    out = torch.empty((12, 1000), dtype=torch.float64, device=DEV)
    # workspace too small, stride too small, no count
    assert lib().hpsp_generate_unweight(ctypes.byref(dec), 1, 0, 1000, 2, ptr(out), 1000, None,
                                        None, ptr(ws), 16, ptr(nacc), stream_handle()) != 0
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(1000)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    assert lib().hpsp_generate_unweight(ctypes.byref(dec), 1, 0, 1000, 2, ptr(out), 999, None,
                                        None, ptr(ws), ws_bytes, ptr(nacc), stream_handle()) != 0
    assert lib().hpsp_generate_unweight(ctypes.byref(dec), 1, 0, 1000, 2, ptr(out), 1000, None,
                                        None, ptr(ws), ws_bytes, None, stream_handle()) != 0


def doubtful_groups(dec, n, seed=UW_SEED):
    """The groups of 64 events the weight chain leaves to the exact test: the
    counter of hpsp_generate_unweight's list, which follows hpsp_unweight's
    part of the workspace (capi.hip)."""
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    check(lib().hpsp_generate_unweight(ctypes.byref(dec), 159753, 0, n, seed, None, n, None, None,
                                       ptr(ws), ws_bytes, ptr(nacc), stream_handle()),
          "hpsp_generate_unweight")
    torch.cuda.synchronize()
    off = lib().hpsp_unweight_workspace_size(n)
    return int(ws[off:off + 8].cpu().view(torch.int64).item())


@pytest.mark.parametrize("engine,limit", [("squares3", 1e-3), ("squares4", 1e-3),
                                          ("philox", 1e-3), ("threefry_long", 1e-3)])
def test_chain_decides_almost_every_flag(engine, limit):
    """On the headline decay the exact test runs for almost no group (every
    engine: the Random123 ones take their bound's K per event)."""
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957], engine=engine)
    n = 1 << 22
    g = doubtful_groups(dec, n)
    assert g <= limit * (n // 64), (engine, g)


@pytest.mark.parametrize("M,masses", [(5.279, [3.0969, 0.493677, 0.13957]),
                                      (10.0, [0.1, 0.2, 8.0])])
def test_table3_on_device_equals_restatement(oracle, M, masses):
    """The 3-body bound's table the call writes into its workspace (the last
    464 (X, Y) pairs) equals the restatement that tests/test_chain_bound.py
    checks against every event's bound."""
    from test_chain_bound import BIN3, table3
    dec = make_decay(M, masses)
    n = 4096
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(n)
    ws = torch.zeros(ws_bytes, dtype=torch.uint8, device=DEV)
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    check(lib().hpsp_generate_unweight(ctypes.byref(dec), 1, 0, n, UW_SEED, None, n, None, None,
                                       ptr(ws), ws_bytes, ptr(nacc), stream_handle()),
          "hpsp_generate_unweight")
    torch.cuda.synchronize()
    tab_bytes = 2 * BIN3 * 8
    tab = ws[ws_bytes - tab_bytes:].cpu().numpy().view(np.float64).reshape(BIN3, 2)
    X, Y = table3(oracle, M, masses)
    for got, want in ((tab[:, 0], X), (tab[:, 1], Y)):
        fin = np.isfinite(want)
        assert np.array_equal(np.isfinite(got), fin)
        np.testing.assert_allclose(got[fin], want[fin], rtol=1e-13, atol=0)


# ---- the mixed-precision form ------------------------------------------------

def generate_unweight_mixed(dec, gen_seed, first, n, seed, ld=None):
    ld = ld or max(n, 1)
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty((4 * dec.n, ld), dtype=torch.float32, device=DEV)
    idx = torch.empty(max(n, 1), dtype=torch.int64, device=DEV)
    w = torch.empty(max(n, 1), dtype=torch.float64, device=DEV)
    nacc = torch.full((1,), 12345, dtype=torch.int64, device=DEV)
    check(lib().hpsp_generate_unweight_mixed(ctypes.byref(dec), gen_seed, first, n, seed,
                                             ptr(out), ld, ptr(idx), ptr(w), ptr(ws), ws_bytes,
                                             ptr(nacc), stream_handle()),
          "hpsp_generate_unweight_mixed")
    torch.cuda.synchronize()
    return int(nacc.item()), out, idx, w


def same_mixed(dec, gen_seed, first, n, uw_seed):
    """The one call against the mixed Generate (FP32 columns and its weight
    column) followed by the moving Unweight of those columns with that
    column as the weights (Decays.inl:415-455)."""
    cols32, _, evw = generate(dec, gen_seed, n, first=first, prec=capi.MIXED, out_f32=True,
                              want_wt=False)
    k1, out1, idx1, w1 = unweight(dec, cols32[:, :n].double().contiguous(), uw_seed, w_in=evw,
                                  n=n)
    print(f"mixed N={dec.n} engine={dec.engine} n={n} first={first}", flush=True)
    k2, out2, idx2, w2 = generate_unweight_mixed(dec, gen_seed, first, n, uw_seed)
    assert k1 == k2
    assert np.array_equal(bits(out1[:, :n]), bits(out2[:, :n].double()))
    assert np.array_equal(idx1.cpu().numpy()[:n], idx2.cpu().numpy()[:n])
    assert np.array_equal(bits(w1[:n]), bits(w2[:n]))
    return k1


@pytest.mark.parametrize("N", [2, 3, 4, 7, 12])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_generate_unweight_mixed_equals_generate_then_unweight(N, engine):
    rng = np.random.default_rng(900 + 13 * N + ENGINES[engine])
    m = list(rng.uniform(0.1, 1.0, N))
    M = sum(m) * rng.uniform(1.5, 3.0) + 2.0
    dec = make_decay(M, m, engine=engine)
    k = same_mixed(dec, int(rng.integers(0, 2**63)), int(rng.integers(0, 2**40)), 20000 + N,
                   int(rng.integers(0, 2**63)))
    assert 0 <= k <= 20000 + N


def test_generate_unweight_mixed_sizes_and_headline():
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    for n in (1, 63, 257, 1023, 4096 + 17):
        same_mixed(dec, 159753, 7, n, UW_SEED)
    same_mixed(dec, 159753, 0, 1 << 22, UW_SEED)
    near = make_decay(10.0, [0.999] * 10)
    same_mixed(near, 2024, 3, 1 << 18, UW_SEED)
    # where the squared chain's bound is large and lanes take the mixed chain
    M3 = 3.0969 + 0.493677 + 0.13957
    for M, m, eng in ((M3 + 0.001, [3.0969, 0.493677, 0.13957], "squares3"),
                      (100.0, [0.000511, 0.000511, 90.0], "philox"),
                      (5.279, [1e-6, 0.493677, 0.13957], "squares4"),
                      (20.0, [1e-6] + [0.5] * 11, "threefry")):
        same_mixed(make_decay(M, m, engine=eng), 2024, 3, 1 << 18, UW_SEED)


def test_generate_unweight_mixed_empty_and_errors():
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    k, _, _, _ = generate_unweight_mixed(dec, 1, 0, 0, UW_SEED)
    assert k == 0
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    out = torch.empty((12, 1000), dtype=torch.float32, device=DEV)
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(1000)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    assert lib().hpsp_generate_unweight_mixed(ctypes.byref(dec), 1, 0, 1000, 2, ptr(out), 1000,
                                              None, None, ptr(ws), 16, ptr(nacc),
                                              stream_handle()) != 0
    assert lib().hpsp_generate_unweight_mixed(ctypes.byref(dec), 1, 0, 1000, 2, ptr(out), 999,
                                              None, None, ptr(ws), ws_bytes, ptr(nacc),
                                              stream_handle()) != 0


def test_generate_unweight_strides_and_partial_outputs():
    """A column stride above n, and the index map alone (no columns), in both
    forms: the same rows and indices as the full call."""
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957], engine="squares4")
    n = 5000
    k1, out1, idx1, w1 = generate_unweight(dec, 77, 11, n, UW_SEED)
    k2, out2, idx2, w2 = generate_unweight(dec, 77, 11, n, UW_SEED, ld=n + 37)
    assert k1 == k2
    assert np.array_equal(bits(out1[:, :n]), bits(out2[:, :n]))
    assert np.array_equal(idx1.cpu().numpy(), idx2.cpu().numpy())
    # the index map alone
    ws_bytes = lib().hpsp_generate_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    idx3 = torch.empty(n, dtype=torch.int64, device=DEV)
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    check(lib().hpsp_generate_unweight(ctypes.byref(dec), 77, 11, n, UW_SEED, None, n, ptr(idx3),
                                       None, ptr(ws), ws_bytes, ptr(nacc), stream_handle()),
          "hpsp_generate_unweight")
    torch.cuda.synchronize()
    assert int(nacc.item()) == k1
    assert np.array_equal(idx1.cpu().numpy(), idx3.cpu().numpy())
    # the mixed form with a stride above n
    m1 = generate_unweight_mixed(dec, 77, 11, n, UW_SEED)
    m2 = generate_unweight_mixed(dec, 77, 11, n, UW_SEED, ld=n + 37)
    assert m1[0] == m2[0]
    assert np.array_equal(m1[1][:, :n].cpu().numpy(), m2[1][:, :n].cpu().numpy())
    assert np.array_equal(m1[2].cpu().numpy(), m2[2].cpu().numpy())
