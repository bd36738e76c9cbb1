"""phase_space_range on the GPU (hpsp_range_generate, hpsp_histogram_fill_range)
against the oracle (bit-equal to the reference's range, see test_oracle) and
the reference's golden vectors.

Bars: weights bit for bit (built from + - * / sqrt only); momenta within
1e-12 of each daughter's energy, the bar of Generate; the fused range
histogram puts every element in the bin of the oracle's element (unit-weight
counts exact), weighted sums to rounding (1e-13 of the bin's sum of |w|).
"""
import numpy as np
import pytest

from gpu_util import momentum_error, range_generate, range_histogram
from hydra_amd import capi
from hydra_amd.capi import make_functor

pytestmark = pytest.mark.gpu

MK, MPI = 0.493677, 0.13957061
C1 = [3.0969, 0.493677, 0.13957]
N12 = [0.1 + 0.1 * i for i in range(12)]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def rows(p):
    """AoS (n, N, 4) -> SoA (4N, n)."""
    n, N, _ = p.shape
    return np.ascontiguousarray(p.transpose(1, 2, 0).reshape(4 * N, n))


def random_range(rng, N):
    m = list(rng.uniform(0.1, 1.0, N))
    M = sum(m) * rng.uniform(1.05, 3.0)
    pv = rng.normal(0, 1.5, 3) * rng.integers(0, 2)
    return [float(np.sqrt(M * M + (pv ** 2).sum()))] + list(pv), m


@pytest.mark.parametrize("N", [2, 3, 4, 5, 7, 10, 12])
def test_range_vs_oracle(oracle, N):
    rng = np.random.default_rng(7000 + N)
    m4, m = random_range(rng, N)
    seed = int(rng.integers(0, 2**64, dtype=np.uint64))
    first = int(rng.integers(0, 2**44))
    n = 5000
    dec = capi.make_range(m4, m)
    w, cols = range_generate(dec, seed, n, first=first)
    we, ce = oracle.phase_space_range(m4, m, seed, n, first=first)
    assert np.array_equal(bits(w), bits(we))
    assert momentum_error(cols, ce, N)[0] <= 1e-12


@pytest.mark.parametrize("key,m", [("k3pi", [MPI] * 3), ("c1_moving", C1), ("n12", N12),
                                   ("n2", [1.0, 0.5])])
def test_range_golden(golden, key, m):
    v = golden[f"range_{key}_in"]
    seed, first = (int(x) for x in golden[f"range_{key}_seed"])
    dec = capi.make_range(list(v[:4]), m)
    w, cols = range_generate(dec, seed, 512, first=first)
    assert np.array_equal(bits(w), bits(golden[f"range_{key}_w"]))
    assert momentum_error(cols, rows(golden[f"range_{key}_p"]), len(m))[0] <= 1e-12


def test_range_chunks_and_ld(oracle):
    """Chunks of one range reproduce it; the column stride is honoured."""
    dec = capi.make_range([MK, 0, 0, 0], [MPI] * 3)
    w, cols = range_generate(dec, 321, 3000)
    w2, c2 = range_generate(dec, 321, 1000, first=1700, ld=1536)
    assert np.array_equal(bits(w2), bits(w[1700:2700]))
    assert np.array_equal(bits(c2[:, :1000]), bits(cols[:, 1700:2700]))
    # only the low 32 bits of the seed reach the engine
    w3, _ = range_generate(dec, 2**32 + 321, 3000)
    assert np.array_equal(bits(w3), bits(w))


def test_range_edges(oracle):
    dec = capi.make_range([MK, 0, 0, 0], [MPI] * 3)
    w, _ = range_generate(dec, 5, 0)
    assert w.shape == (0,)
    for n in (1, 63, 65, 257):
        w, cols = range_generate(dec, 5, n, first=2**40 - 3)
        we, ce = oracle.phase_space_range([MK, 0, 0, 0], [MPI] * 3, 5, n, first=2**40 - 3)
        assert np.array_equal(bits(w), bits(we))
        assert momentum_error(cols, ce, 3)[0] <= 1e-12


@pytest.mark.parametrize("wtd", [0, 1])
def test_range_histogram_golden(golden, wtd):
    """The histogram of range_semantics.inl:154-167 as the reference fills it."""
    dec = capi.make_range([MK, 0, 0, 0], [MPI] * 3)
    mode = capi.HIST_WEIGHT_GENERATION if wtd else capi.HIST_WEIGHT_ONE
    got = range_histogram([10], [2 * MPI], [MK - MPI], dec, 321,
                          [make_functor(capi.F_MASS, (0, 1))], 100000, mode=mode)
    exp = golden[f"range_hist_k3pi_w{wtd}"]
    if wtd:
        assert np.max(np.abs(got - exp) / np.maximum(exp, 1.0)) <= 1e-13
    else:
        assert np.array_equal(got, exp)


@pytest.mark.parametrize("N,grid", [(3, [40, 40]), (5, [64]), (12, [7, 8, 9])])
def test_range_histogram_vs_oracle(oracle, N, grid):
    rng = np.random.default_rng(9100 + N)
    m4, m = random_range(rng, N)
    seed = int(rng.integers(0, 2**32))
    first = 12345
    n = 40000
    dec = capi.make_range(m4, m)
    pairs = [(0, 1), (1, N - 1), (0, N - 1)][:len(grid)]
    axes = [make_functor(capi.F_MASS, pr) for pr in pairs]
    # mother.mass() with Vector4R's operation order
    M = float(np.sqrt(((m4[0] * m4[0] - m4[1] * m4[1]) - m4[2] * m4[2]) - m4[3] * m4[3]))
    lo = [m[a] + m[b] for a, b in pairs]
    hi = [M] * len(grid)

    def axes_of(cols):
        p = cols.reshape(N, 4, n).transpose(2, 0, 1)
        return np.array([[oracle.functor(5, pr, p[k]) for k in range(n)] for pr in pairs])

    def check(got, x, wv):
        exp = oracle.histogram_fill(grid, lo, hi, x, wv)
        if wv is None:
            assert np.array_equal(got, exp)
        else:
            scale = oracle.histogram_fill(grid, lo, hi, x, np.abs(wv))
            assert np.all(np.abs(got - exp) <= 1e-13 * scale + 1e-300)

    # unit and generation weights: the oracle's own elements
    we, ce = oracle.phase_space_range(m4, m, seed, n, first=first)
    x = axes_of(ce)
    check(range_histogram(grid, lo, hi, dec, seed, axes, n, first=first), x, None)
    check(range_histogram(grid, lo, hi, dec, seed, axes, n, first=first,
                          mode=capi.HIST_WEIGHT_GENERATION), x, we)
    # PhaseSpaceWeight is recomputed from the momenta, which carry the 1e-12 E
    # tolerance, and near threshold it amplifies their differences: it is
    # checked on the momenta the GPU generates (the fused fill forms the same
    # ones, see test_range_histogram_equals_fill_of_generated)
    _, cg = range_generate(dec, seed, n, first=first)
    check(range_histogram(grid, lo, hi, dec, seed, axes, n, first=first,
                          mode=capi.HIST_WEIGHT_PHSP), axes_of(cg), oracle.event_weight(M, m, cg))


def test_range_histogram_large_grid_uses_device_bins(oracle):
    """A grid beyond the LDS capacity fills the device contents directly."""
    dec = capi.make_range([MK, 0, 0, 0], [MPI] * 3)
    n = 20000
    grid = [200, 200]  # 40002 slots: more than fit one block's LDS
    lo, hi = [2 * MPI] * 2, [MK - MPI] * 2
    axes = [make_functor(capi.F_MASS, (0, 1)), make_functor(capi.F_MASS, (1, 2))]
    got = range_histogram(grid, lo, hi, dec, 99, axes, n)
    we, ce = oracle.phase_space_range([MK, 0, 0, 0], [MPI] * 3, 99, n)
    p = ce.reshape(3, 4, n).transpose(2, 0, 1)
    x = np.array([[oracle.functor(5, pr, p[k]) for k in range(n)] for pr in ((0, 1), (1, 2))])
    assert np.array_equal(got, oracle.histogram_fill(grid, lo, hi, x))


def test_range_histogram_equals_fill_of_generated(oracle):
    """Fusing generation into the fill changes nothing: the same elements,
    generated into columns and filled from them, give the same bins."""
    import ctypes
    import torch
    from gpu_util import DEV
    from hydra_amd.capi import check, lib, ptr, stream_handle
    dec = capi.make_range([6.0, 0.3, -0.2, 0.5], C1)
    n = 300000
    axes = [make_functor(capi.F_MASS2, (0, 1)), make_functor(capi.F_MASS2, (1, 2))]
    grid, lo, hi = [50, 50], [12.0, 0.3], [34.5, 8.5]  # every element inside
    fused = range_histogram(grid, lo, hi, dec, 17, axes, n)
    _, cols = range_generate(dec, 17, n)
    ct = torch.as_tensor(cols, device=DEV)
    h = capi.make_histogram(grid, lo, hi)
    out = torch.empty(2502, dtype=torch.float64, device=DEV)
    check(lib().hpsp_histogram_fill_events(ctypes.byref(h), ctypes.byref(dec),
                                           capi.functor_array(axes), ptr(ct), n, n,
                                           capi.HIST_WEIGHT_ONE, None, ptr(out),
                                           stream_handle()), "fill_events")
    torch.cuda.synchronize()
    assert np.array_equal(fused, out.cpu().numpy())
    assert fused.sum() == n
