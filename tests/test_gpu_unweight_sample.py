"""hydra::unweight (the free function of hydra/Random.h) on the GPU against
the oracle, which reproduces the reference's result bit for bit (its OMP
build's order included, see test_oracle), and the golden vectors.  The kept
set, the kept order and every moved value are compared exactly."""
import numpy as np
import pytest

from gpu_util import unweight_sample
from hydra_amd import capi

pytestmark = pytest.mark.gpu

SEED = capi.UNWEIGHT_SAMPLE_SEED


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_golden(golden):
    x, w = golden["uws_x"], golden["uws_w"]
    for tag, mp, jump in (("auto", -1.0, 0), ("max", 0.5, 10000)):
        k, out, _, _ = unweight_sample(x, w, mp, SEED, jump, order=capi.ORDER_SEQUENTIAL)
        assert k == int(golden[f"uws_{tag}_kept"][0])
        assert np.array_equal(bits(out[0]), bits(golden[f"uws_{tag}_x"]))


@pytest.mark.parametrize("ncols", [1, 3, 12, 13])
@pytest.mark.parametrize("order", [capi.ORDER_STABLE, capi.ORDER_SEQUENTIAL])
def test_vs_oracle(oracle, ncols, order):
    rng = np.random.default_rng(10 * ncols + order)
    n = 300001
    cols = rng.normal(0, 1, (ncols, n))
    w = rng.exponential(1.0, n)
    for mp, jump in ((-1.0, 0), (2.0, 12345)):
        k, out, idx, wo = unweight_sample(cols, w, mp, SEED + ncols, jump, order, want_w=True)
        k2, out2, idx2 = oracle.unweight_sample(cols, w, mp, SEED + ncols, jump, order=order)
        assert k == k2
        assert np.array_equal(idx, idx2.astype(np.int64))
        assert np.array_equal(bits(out), bits(out2))
        assert np.array_equal(bits(wo), bits(w[idx]))


def test_long_swap_chains(oracle):
    """One rejected sample followed by a long run of kept ones: the
    sequential order moves it along the whole run."""
    n = 5000
    w = np.full(n, 1.0)
    w[0] = 0.0    # never kept
    w[3000] = 0.0
    k, _, idx, _ = unweight_sample(np.zeros((1, n)), w, 1.0, SEED, 0, capi.ORDER_SEQUENTIAL,
                                   want_cols=False)
    k2, _, idx2 = oracle.unweight_sample(np.zeros((1, n)), w, 1.0, SEED, 0, order=1)
    assert k == k2 and np.array_equal(idx, idx2.astype(np.int64))


def test_edges(oracle):
    for n in (0, 1, 7, 64, 65, 1024, 1025):
        rng = np.random.default_rng(n)
        w = rng.uniform(0, 1, n)
        for order in (capi.ORDER_STABLE, capi.ORDER_SEQUENTIAL):
            k, out, idx, _ = unweight_sample(np.arange(n, dtype=float)[None, :], w, -1.0, SEED, 0,
                                             order)
            k2, out2, idx2 = oracle.unweight_sample(np.arange(n, dtype=float)[None, :], w, -1.0,
                                                    SEED, 0, order=order)
            assert k == k2 and np.array_equal(out[:, :n], out2)
