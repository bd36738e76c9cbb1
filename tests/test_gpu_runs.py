"""Generation runs: a large launch gives each wave runs of consecutive groups
of 64 events and carries the squares words from group to group; small
launches take one group per wave.  Events are a function of (decay, seed,
index) alone, so one large launch and the same range generated in small
pieces (one group per wave) must agree bit for bit, including a partial last
run, every engine family and each output form (momenta and weights: runs up
to 16 groups; momenta of at most 3 daughters alone: up to 4; FP32 columns of
the mixed form).
"""
import numpy as np
import pytest
import torch

from gpu_util import evaluate, generate, make_decay, make_functor
from hydra_amd import capi

pytestmark = pytest.mark.gpu

PIECE = 1 << 19  # 8192 groups: one group per wave


def bits(t):
    return t.view(torch.int64) if t.dtype == torch.float64 else t.view(torch.int32)


def pieces_match(fulls, piece_fn, n):
    """fulls: (rows, n) tensors of one launch; piece_fn(off, m) -> the same
    tensors for events off .. off + m - 1 generated on their own.  Returns the
    first mismatching offset, or None."""
    for off in range(0, n, PIECE):
        m = min(PIECE, n - off)
        for full, part in zip(fulls, piece_fn(off, m)):
            if not torch.equal(bits(full[:, off:off + m]), bits(part[:, :m])):
                return off
    return None


CASES = [
    # (N, engine, n, want_evw, prec, out_f32): runs of 16, 4, 16, 4, 4, 4 groups
    (3, "squares3", (1 << 23) + 4321, True, capi.FP64, False),
    (3, "squares3", (1 << 21) + 12345, False, capi.FP64, False),
    (7, "squares3", (1 << 23) + 4321, False, capi.FP64, False),
    (3, "philox", (1 << 21) + 12345, True, capi.FP64, False),
    (7, "squares4", (1 << 21) + 12345, True, capi.FP64, False),
    (3, "squares3", (1 << 21) + 12345, True, capi.MIXED, True),
]


@pytest.mark.parametrize("N,engine,n,want_evw,prec,out_f32", CASES)
def test_generate_runs_match_single_groups(N, engine, n, want_evw, prec, out_f32):
    rng = np.random.default_rng(N * 7 + len(engine))
    m = list(rng.uniform(0.1, 0.6, N))
    dec = make_decay(sum(m) * 2.5 + 0.5, m, engine=engine)
    seed, first = 271828, 1 << 33
    cols, _, evw = generate(dec, seed, n, first=first, prec=prec, out_f32=out_f32,
                            want_wt=False, want_evw=want_evw)
    fulls = [cols[:, :n]] + ([evw[:n].view(1, -1)] if want_evw else [])

    def piece(off, k):
        c, _, w = generate(dec, seed, k, first=first + off, prec=prec, out_f32=out_f32,
                           want_wt=False, want_evw=want_evw)
        return [c] + ([w.view(1, -1)] if want_evw else [])

    assert pieces_match(fulls, piece, n) is None


def test_generate_mothers_runs_match_single_groups():
    m = [3.0969, 0.493677, 0.13957]
    dec = make_decay(5.279, m)
    n = (1 << 21) + 12345
    rng = np.random.default_rng(9)
    mt = np.zeros((4, n))
    mt[1:] = rng.normal(0, 2.0, (3, n))
    mt[0] = np.sqrt((5.279 * 1.2) ** 2 + (mt[1:] ** 2).sum(axis=0))
    cols, _, evw = generate(dec, 5, n, mothers=mt, want_wt=False)

    def piece(off, k):
        c, _, w = generate(dec, 5, k, first=off, mothers=mt[:, off:off + k], want_wt=False)
        return [c, w.view(1, -1)]

    assert pieces_match([cols[:, :n], evw[:n].view(1, -1)], piece, n) is None


def test_evaluate_runs_match_single_groups():
    dec = make_decay(5.279, [3.0969, 0.493677, 0.13957])
    F = [make_functor(capi.F_MASS2, (0, 2)), make_functor(capi.F_MASS2, (1, 2)),
         make_functor(capi.F_COS_HELICITY, (1, 2))]
    n = (1 << 21) + 12345
    full = evaluate(dec, F, 77, n, first=12)
    assert pieces_match([full[:, :n]], lambda off, k: [evaluate(dec, F, 77, k, first=12 + off)],
                        n) is None
