"""Stored events as particle columns on the GPU: functor values
(hpsp_particles_evaluate) and histogram fills (hpsp_histogram_fill_particles)
over a Decays container or melded containers, and the decay chain of
examples/phase_space/phsp_chain.inl run end to end against the reference's
golden vectors.

Bars: functor values on the same momenta bit for bit with the oracle (built
from + - * / sqrt), the decay-plane angle to 1e-14 (atan2 is the device
library's, the oracle's glibc's); bins exact, weighted sums to rounding.
Against the reference's chain (momenta within 1e-12 E): M^2 to 1e-11 M^2,
cos theta and the angle to 1e-9.
"""
import ctypes

import numpy as np
import pytest
import torch

from gpu_util import DEV, make_decay, momentum_error
from hydra_amd import capi
from hydra_amd.capi import check, lib, make_functor, ptr, stream_handle

pytestmark = pytest.mark.gpu

B0 = (5.27955, [3.0969, 0.493677, 0.13957061])
JPSI = (3.0969, [0.1056583745] * 2)
CHAIN_F = [(capi.F_MASS2, (0, 2)), (capi.F_MASS2, (1, 2)),
           (capi.F_COS_HELICITY, (1, 2, 0, 3)), (capi.F_DECAY_PLANE_ANGLE, (1, 2, 3))]
ORACLE_F = [(3, (0, 2)), (3, (1, 2)), (4, (1, 2, 0, 3)), (6, (1, 2, 3))]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def particles(groups):
    """hpsp_particles over [(tensor (4k, ld), k), ...] in order."""
    p = capi.Particles()
    for t, k in groups:
        ld = t.shape[1]
        for i in range(k):
            p.cols[p.n] = t.data_ptr() + 8 * 4 * i * ld
            p.ld[p.n] = ld
            p.n += 1
    return p


def evaluate(p, functors, n):
    out = torch.empty((len(functors), max(n, 1)), dtype=torch.float64, device=DEV)
    check(lib().hpsp_particles_evaluate(ctypes.byref(p), capi.functor_array(functors),
                                        len(functors), n, ptr(out), max(n, 1), stream_handle()),
          "hpsp_particles_evaluate")
    torch.cuda.synchronize()
    return out.cpu().numpy()[:, :n]


def fill(p, grid, lo, hi, axes, n, w=None):
    h = capi.make_histogram(grid, lo, hi)
    out = torch.empty(int(np.prod(grid)) + 2, dtype=torch.float64, device=DEV)
    check(lib().hpsp_histogram_fill_particles(ctypes.byref(h), ctypes.byref(p),
                                              capi.functor_array(axes), n, ptr(w), ptr(out),
                                              stream_handle()), "hpsp_histogram_fill_particles")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def gpu_chain(golden, n):
    """Generate(B0) then Generate over the J/psi columns in place; returns the
    two column tensors and the melded particle list."""
    s3, s2 = (int(x) for x in golden["chain_seeds"])
    d3 = make_decay(*B0)
    d2 = make_decay(*JPSI)
    c3 = torch.empty((12, n), dtype=torch.float64, device=DEV)
    c2 = torch.empty((8, n), dtype=torch.float64, device=DEV)
    sh = stream_handle()
    check(lib().hpsp_generate(ctypes.byref(d3), s3, 0, n, capi.FP64, ptr(c3), n, 0, None, None,
                              sh), "generate")
    # mothers: the J/psi columns 0..3 of the three-body container, read in place
    check(lib().hpsp_generate_mothers(ctypes.byref(d2), s2, 0, n, ptr(c3), n, capi.FP64, ptr(c2),
                                      n, 0, None, None, sh), "generate_mothers")
    torch.cuda.synchronize()
    return c3, c2, particles([(c3, 3), (c2, 2)])


def rows_of(c3, c2, n):
    a = c3.cpu().numpy().reshape(3, 4, n).transpose(2, 0, 1)
    b = c2.cpu().numpy().reshape(2, 4, n).transpose(2, 0, 1)
    return np.concatenate([a, b], axis=1)


def test_particles_evaluate_vs_oracle(oracle):
    """Functors over particles spread over two allocations with different
    strides, on the same momenta as the oracle."""
    rng = np.random.default_rng(77)
    n = 3000
    P = 7
    pv = rng.normal(0, 1.0, (n, P, 3))
    m = rng.uniform(0.1, 1.0, P)
    E = np.sqrt((pv ** 2).sum(-1) + m ** 2)
    rows = np.concatenate([E[..., None], pv], axis=-1)  # (n, P, 4)
    cols = rows.transpose(1, 2, 0)  # (P, 4, n)
    a = torch.zeros((16, n + 5), dtype=torch.float64, device=DEV)
    b = torch.zeros((12, n + 64), dtype=torch.float64, device=DEV)
    a[:, :n] = torch.as_tensor(cols[:4].reshape(16, n), device=DEV)
    b[:, :n] = torch.as_tensor(cols[4:].reshape(12, n), device=DEV)
    p = particles([(a, 4), (b, 3)])
    bw = [(0.892, 0.05, 1.0, 0.3, 1, 2), (1.43, 0.11, -0.4, 0.8, 0, 5)]
    dev = [make_functor(capi.F_MASS2, (0, 6)), make_functor(capi.F_MASS, (3, 4)),
           make_functor(capi.F_COS_HELICITY, (1, 2, 0, 4)),
           make_functor(capi.F_COS_HELICITY, (5, 6, 4, 3)),
           make_functor(capi.F_BW_SUM, bw), make_functor(capi.F_ONE),
           make_functor(capi.F_DECAY_PLANE_ANGLE, (1, 2, 5)),
           make_functor(capi.F_DECAY_PLANE_ANGLE, (4, 6, 0))]
    ora = [(3, (0, 6)), (5, (3, 4)), (4, (1, 2, 0, 4)), (4, (5, 6, 4, 3)), (1, bw), (0, None),
           (6, (1, 2, 5)), (6, (4, 6, 0))]
    got = evaluate(p, dev, n)
    from oracle import functor_params
    for j, (fid, t) in enumerate(ora):
        exp = np.array([oracle.functor(fid, functor_params(fid, t), rows[k]) for k in range(n)])
        if fid == 6:
            assert np.max(np.abs(got[j] - exp)) <= 1e-14, j
        else:
            assert np.array_equal(bits(got[j]), bits(exp)), j
    # a single Decays container's particles give the same values
    assert evaluate(particles([(a, 4)]), [dev[2]], n).shape == (1, n)


def test_chain_vs_golden(golden):
    n = 512
    c3, c2, p = gpu_chain(golden, n)
    rows = rows_of(c3, c2, n)
    gp = golden["chain_p"]
    assert np.max(np.abs(rows - gp) / np.abs(gp[..., 0:1])) <= 1e-12
    got = evaluate(p, [make_functor(k, t) for k, t in CHAIN_F], n)
    exp = golden["chain_vars"]
    M2 = B0[0] ** 2
    assert np.max(np.abs(got[0] - exp[0])) <= 1e-11 * M2
    assert np.max(np.abs(got[1] - exp[1])) <= 1e-11 * M2
    assert np.max(np.abs(got[2] - exp[2])) <= 1e-9
    d = np.abs(got[3] - exp[3])
    d = np.minimum(d, 2 * np.pi - d)  # the angle is periodic
    assert np.max(d) <= 1e-9


def test_chain_functors_on_gpu_momenta(oracle, golden):
    """On the GPU's own chain the values are the oracle's functors of those
    momenta; a large chain, to cover many waves."""
    n = 200000
    c3, c2, p = gpu_chain(golden, n)
    got = evaluate(p, [make_functor(k, t) for k, t in CHAIN_F], n)
    rows = rows_of(c3, c2, n)
    from oracle import functor_params
    idx = np.random.default_rng(3).choice(n, 3000, replace=False)
    for j, (fid, t) in enumerate(ORACLE_F):
        exp = np.array([oracle.functor(fid, functor_params(fid, t), rows[k]) for k in idx])
        if fid == 6:
            assert np.max(np.abs(got[j][idx] - exp)) <= 1e-14
        else:
            assert np.array_equal(bits(got[j][idx]), bits(exp))
    # the J/psi -> mu mu pairs have the J/psi mass
    mm = evaluate(p, [make_functor(capi.F_MASS, (3, 4))], n)[0]
    assert np.max(np.abs(mm - JPSI[0])) < 1e-12


@pytest.mark.parametrize("weighted", [False, True])
def test_histogram_fill_particles_vs_oracle(oracle, golden, weighted):
    n = 100000
    c3, c2, p = gpu_chain(golden, n)
    rows = rows_of(c3, c2, n)
    grid = [40, 30]
    lo, hi = [-1.0, 0.0], [1.0, 2 * np.pi]
    axes = [make_functor(capi.F_COS_HELICITY, (1, 2, 0, 3)),
            make_functor(capi.F_DECAY_PLANE_ANGLE, (1, 2, 3))]
    w = None
    if weighted:
        d3 = make_decay(*B0)
        wt = torch.empty(n, dtype=torch.float64, device=DEV)
        check(lib().hpsp_event_weight(ctypes.byref(d3), None, ptr(c3), n, n, ptr(wt),
                                      stream_handle()), "event_weight")
        w = wt
    got = fill(p, grid, lo, hi, axes, n, w)
    # axis values as the GPU computes them (the angle's last ulp comes from
    # the device atan2), binned by the oracle
    x = evaluate(p, axes, n)
    wv = None if w is None else w.cpu().numpy()
    exp = oracle.histogram_fill(grid, lo, hi, x, wv)
    if wv is None:
        assert np.array_equal(got, exp) and got.sum() == n
    else:
        scale = oracle.histogram_fill(grid, lo, hi, x, np.abs(wv))
        assert np.all(np.abs(got - exp) <= 1e-13 * scale + 1e-300)
    # and the oracle's own functor values bin the same events (sampled check)
    from oracle import functor_params
    k = np.arange(0, n, 97)
    xo = np.array([[oracle.functor(4, functor_params(4, (1, 2, 0, 3)), rows[i]) for i in k],
                   [oracle.functor(6, functor_params(6, (1, 2, 3)), rows[i]) for i in k]])
    assert np.max(np.abs(xo - x[:, k])) <= 1e-14


def test_particles_edges():
    a = torch.zeros((8, 130), dtype=torch.float64, device=DEV)
    a[0] = 2.0
    a[4] = 3.0
    p = particles([(a, 2)])
    f = [make_functor(capi.F_MASS2, (0, 1))]
    assert evaluate(p, f, 0).shape == (1, 0)
    for n in (1, 65, 130):
        got = evaluate(p, f, n)
        assert np.all(got == 25.0)
    h = fill(p, [5], [0.0], [50.0], f, 130)
    assert h[2] == 130 and h.sum() == 130
