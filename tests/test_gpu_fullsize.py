"""Parity at the benchmarked sizes, against the oracle.

* The headline step (3-body decay, 2^27 events: Generate followed by
  Decays::Unweight() in one call, hpsp_generate_unweight, as bench.py times
  it: flags from each event's weight chain, the exact test where in doubt,
  every event generated at its place): the accepted count and an
  order-sensitive digest of the whole stable partition (output row ->
  container index) equal the oracle's streaming Generate + Unweight over the
  same 2^27 events (Decays.inl:415-455), and the momenta of sampled output
  rows equal the oracle's events at their indices within 1e-12 E.
* AverageOn of the three-Breit-Wigner Dalitz functor over 2^30 events in 64
  chunks of 2^24 (PhaseSpace.inl:255-273): mean within 1e-12 relative of the
  oracle's single pass over the same events, sigma within 1e-10.
"""
import ctypes
import math
import os

import numpy as np
import pytest
import torch

from gpu_util import DEV, average_on, bw_functor, make_decay, momentum_error, ptr
from hydra_amd import capi

M3, MASSES3 = 5.279, [3.0969, 0.493677, 0.13957]
SEED, UW_SEED = 159753, 0x180EC6D33CFD0ABA


def threads():
    return max(1, min(16, os.cpu_count() or 1))


@pytest.mark.gpu
def test_headline_step_partition_at_2p27(oracle):
    n = 1 << 27
    dec = make_decay(M3, MASSES3)
    lib = capi.lib()
    ws_bytes = lib.hpsp_generate_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty((12, n), dtype=torch.float64, device=DEV)
    idx = torch.empty(n, dtype=torch.int64, device=DEV)
    nacc_d = torch.zeros(1, dtype=torch.int64, device=DEV)
    capi.check(lib.hpsp_generate_unweight(ctypes.byref(dec), SEED, 0, n, UW_SEED, ptr(out), n,
                                          ptr(idx), None, ptr(ws), ws_bytes, ptr(nacc_d),
                                          capi.stream_handle()), "hpsp_generate_unweight")
    torch.cuda.synchronize()
    nacc = int(nacc_d.item())
    del ws
    idx_h = idx.cpu().numpy().view(np.uint64)
    print(f"gpu: {nacc} accepted of {n}", flush=True)
    onacc, odig = oracle.unweight_digest(M3, MASSES3, SEED, n, UW_SEED, nthreads=threads())
    assert nacc == onacc
    assert oracle.permutation_digest(idx_h) == odig
    # momenta of sampled rows: both ends of the accepted and rejected runs
    # and random rows in between
    rng = np.random.default_rng(2027)
    rows = np.unique(np.concatenate([[0, 1, nacc - 1, nacc, nacc + 1, n - 1],
                                     rng.integers(0, n, 4096)]))
    p_o, _ = oracle.decay_events(M3, MASSES3, SEED, idx_h[rows])
    got = out[:, torch.as_tensor(rows, device=DEV)].cpu().numpy()
    want = np.ascontiguousarray(p_o.transpose(1, 2, 0)).reshape(12, -1)
    err, _ = momentum_error(got, want, 3)
    assert err <= 1e-12, err


@pytest.mark.gpu
def test_average_on_dalitz_2p30_vs_oracle(oracle, golden):
    chunk, nchunks = 1 << 24, 64
    dec = make_decay(M3, MASSES3)
    f = bw_functor(golden["bw3_par"])
    acc = capi.Stats(0.0, 0.0, 0.0)
    for c in range(nchunks):
        s = average_on(dec, SEED, chunk, f=f, first=c * chunk)
        part = capi.Stats(*[float(x) for x in s])
        if c == 0:
            acc = part
        else:
            merged = capi.Stats()
            capi.lib().hpsp_stats_merge(ctypes.byref(acc), ctypes.byref(part),
                                       ctypes.byref(merged))
            acc = merged
    st = oracle.average_on(M3, MASSES3, SEED, chunk * nchunks, fid=1, fpar=golden["bw3_par"],
                           intervals=4 * threads())
    mean, sig = acc.mean, math.sqrt(acc.m2) / acc.w
    omean, osig = st[1], math.sqrt(st[2]) / st[0]
    print(f"gpu mean {mean!r} sigma {sig!r}; oracle {omean!r} {osig!r}", flush=True)
    assert abs(mean - omean) <= 1e-12 * abs(omean)
    assert abs(sig - osig) <= 1e-10 * abs(osig)
    assert acc.w == pytest.approx(st[0], rel=1e-12)
