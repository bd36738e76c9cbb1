"""Evaluate (PhaseSpace::Evaluate over a mother and over mothers) on the GPU
against the oracle (bit-equal to the reference's Evaluate, see test_oracle) and
the reference's golden vectors.

Bars: the weight column bit for bit (built from + - * / sqrt only); functor
columns within the tolerance their momenta carry (|dp| <= 1e-12 E):
|d(p_a+p_b)^2| <= 1e-11 M^2, |d cos| <= 1e-9, Breit-Wigner sums 1e-8 relative.
"""
import ctypes

import numpy as np
import pytest

from gpu_util import evaluate, generate, make_decay, make_functor
from hydra_amd import capi
from oracle import ENGINES

pytestmark = pytest.mark.gpu

SEED = 159753
C1 = (5.279, [3.0969, 0.493677, 0.13957])
BW3 = [(0.892, 0.05, 0.4967141530, -0.1382643012, 1, 2),
       (1.43, 0.11, 0.6476885381, 1.5230298564, 1, 2),
       (0.775, 0.149, -0.2341533747, -0.2341368981, 0, 2)]


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def random_decay(rng, N):
    m = list(rng.uniform(0.1, 1.0, N))
    return sum(m) * rng.uniform(1.05, 3.0), m


def functor_sets(N, offset):
    """(device functors, oracle functors) seeing P = N + offset four-vectors;
    offset 1 puts the mother first."""
    a, b = offset, N - 1 + offset
    bw = [(0.892, 0.05, 1.0, 0.3, a, b), (1.43, 0.11, -0.4, 0.8, offset, a + (1 if N > 2 else 0))]
    dev = [make_functor(capi.F_MASS2, (a, b)), make_functor(capi.F_COS_HELICITY, (a, b)),
           make_functor(capi.F_BW_SUM, bw), make_functor(capi.F_ONE)]
    ora = [(3, (a, b)), (4, (a, b, offset, N)), (1, bw), (0, None)]
    if N == 2:
        # a pair of two daughters is the whole parent: the helicity angle is
        # undefined (0/0), so that slot holds the constant instead
        dev[1], ora[1] = make_functor(capi.F_ONE), (0, None)
    return dev, ora


def check_columns(got, exp, M):
    assert np.array_equal(bits(got[0]), bits(exp[0])), "weights"
    assert np.max(np.abs(got[1] - exp[1])) <= 1e-11 * M * M
    assert np.max(np.abs(got[2] - exp[2])) <= 1e-9
    assert np.max(np.abs(got[3] - exp[3]) / np.maximum(np.abs(exp[3]), 1e-300)) <= 1e-8
    assert np.all(got[4] == 1.0)


@pytest.mark.parametrize("N", [2, 3, 5, 10, 12])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_evaluate_vs_oracle(oracle, N, engine):
    rng = np.random.default_rng(500 + 10 * N + ENGINES[engine])
    M, m = random_decay(rng, N)
    seed = int(rng.integers(0, 2**63))
    first = int(rng.integers(0, 2**40))
    n = 3000
    dec = make_decay(M, m, engine=engine)
    dev, ora = functor_sets(N, 0)
    got = evaluate(dec, dev, seed, n, first=first).cpu().numpy()
    exp = oracle.evaluate(M, m, seed, n, ora, first=first, engine=engine)
    check_columns(got, exp, M)


@pytest.mark.parametrize("N", [2, 3, 7, 12])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_evaluate_mothers_vs_oracle(oracle, N, engine):
    rng = np.random.default_rng(900 + 10 * N + ENGINES[engine])
    M, m = random_decay(rng, N)
    seed = int(rng.integers(0, 2**63))
    n = 2500
    mt = np.zeros((4, n))
    mt[1:] = rng.normal(0, 2.0, (3, n))
    mt[0] = np.sqrt((M * rng.uniform(1.0, 1.3, n)) ** 2 + (mt[1:] ** 2).sum(axis=0))
    dec = make_decay(M, m, engine=engine)
    dev, ora = functor_sets(N, 1)
    got = evaluate(dec, dev, seed, n, mothers=mt, first=0).cpu().numpy()
    exp = oracle.evaluate_mothers(m, seed, mt, ora, engine=engine)
    check_columns(got, exp, M * 1.3)
    # a chunk of the same mothers with an event offset keys the same engines
    got2 = evaluate(dec, dev, seed, 500, mothers=mt[:, 1000:1500], first=1000).cpu().numpy()
    assert np.array_equal(bits(got2[0]), bits(got[0][1000:1500]))


def test_evaluate_golden(golden):
    M, m = C1
    dec = make_decay(M, m)
    F = [make_functor(capi.F_MASS2, (0, 2)), make_functor(capi.F_MASS2, (1, 2))]
    got = evaluate(dec, F, SEED, 1024).cpu().numpy()
    exp = golden["eval_c1_m2"]
    assert np.array_equal(bits(got[0]), bits(exp[0]))
    assert np.max(np.abs(got[1:] - exp[1:])) <= 1e-11 * M * M
    F = [make_functor(capi.F_COS_HELICITY, (1, 2)), make_functor(capi.F_BW_SUM, BW3)]
    got = evaluate(dec, F, SEED, 1024).cpu().numpy()
    exp = golden["eval_c1_cos_bw"]
    assert np.array_equal(bits(got[0]), bits(exp[0]))
    assert np.max(np.abs(got[1] - exp[1])) <= 1e-9
    assert np.max(np.abs(got[2] - exp[2]) / exp[2]) <= 1e-8
    mothers = golden["genm_mothers"]
    F = [make_functor(capi.F_MASS2, (1, 3)), make_functor(capi.F_MASS2, (0, 1))]
    got = evaluate(dec, F, SEED, mothers.shape[1], mothers=mothers).cpu().numpy()
    exp = golden["evalm_c1_m2"]
    assert np.array_equal(bits(got[0]), bits(exp[0]))
    assert np.max(np.abs(got[1:] - exp[1:])) <= 1e-11 * 40.0
    F = [make_functor(capi.F_COS_HELICITY, (2, 3))]
    got = evaluate(dec, F, SEED, mothers.shape[1], mothers=mothers).cpu().numpy()
    exp = golden["evalm_c1_cos"]
    assert np.array_equal(bits(got[0]), bits(exp[0]))
    assert np.max(np.abs(got[1] - exp[1])) <= 1e-9
    M12 = 20.0
    m12 = [0.1 + 0.1 * i for i in range(12)]
    dec = make_decay(M12, m12)
    F = [make_functor(capi.F_MASS2, (4, 11)), make_functor(capi.F_COS_HELICITY, (3, 7))]
    got = evaluate(dec, F, SEED, 256).cpu().numpy()
    exp = golden["eval_n12_m2"]
    assert np.array_equal(bits(got[0]), bits(exp[0]))
    assert np.max(np.abs(got[1] - exp[1])) <= 1e-11 * M12 * M12
    assert np.max(np.abs(got[2] - exp[2])) <= 1e-9


def test_evaluate_weight_is_generation_weight():
    """The weight column equals Generate's generation weight, bit for bit."""
    M, m = C1
    dec = make_decay(M, m)
    got = evaluate(dec, [make_functor(capi.F_ONE)], 31, 70000, first=123).cpu().numpy()
    _, wt, _ = generate(dec, 31, 70000, first=123, want_evw=False)
    assert np.array_equal(bits(got[0]), bits(wt.cpu().numpy()))


def test_evaluate_all_functor_slots_and_ld(oracle):
    M, m = C1
    dec = make_decay(M, m)
    pairs = [(0, 1), (0, 2), (1, 2), (2, 0), (1, 0), (2, 1), (0, 0), (1, 1)]
    dev = [make_functor(capi.F_MASS2, p) for p in pairs]
    got = evaluate(dec, dev, 5, 1000, ld=1536).cpu().numpy()[:, :1000]
    exp = oracle.evaluate(M, m, 5, 1000, [(3, p) for p in pairs])
    assert np.array_equal(bits(got[0]), bits(exp[0]))
    assert np.max(np.abs(got[1:] - exp[1:])) <= 1e-11 * M * M


def test_evaluate_edges():
    M, m = C1
    dec = make_decay(M, m)
    F = [make_functor(capi.F_ONE)]
    assert evaluate(dec, F, 1, 0).shape == (2, 1)
    got = evaluate(dec, F, 1, 1).cpu().numpy()
    assert got.shape == (2, 1) and got[1, 0] == 1.0
    # 65 events: one full wave and one lane of the next
    got = evaluate(dec, F, 1, 65).cpu().numpy()
    assert np.all(got[0] > 0)
