"""The oracle (oracle/phsp_oracle.c) pinned before it is trusted.

* bit-for-bit against the golden vectors written by the reference itself
  (tests/golden/make_golden.py);
* bit-for-bit against the live reference harness where it is built;
* against published known-answer vectors (SplitMix64, Random123) and
  closed forms of the phase-space algebra that a dropped term, a wrong sign
  or a swapped index would break.
"""
import math

import numpy as np
import pytest

from oracle import ENGINES, cols_to_rows, rows_to_cols

SEED = 159753
UW_SEED = 0x180EC6D33CFD0ABA
UWF_SEED = 0x39ABDC4529B1661C
DECAYS = {
    "c1_3body": (5.279, [3.0969, 0.493677, 0.13957]),
    "c3_10body": (10.0, [0.13957] * 10),
    "c3_10body_threshold": (10.0, [0.999] * 10),
    "n2": (3.0, [1.0, 0.5]),
    "n12": (20.0, [0.1 + 0.1 * i for i in range(12)]),
}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


# ---------------------------------------------------------------- engines
def test_splitmix64_published_sequence(oracle):
    # SplitMix64 (Steele, Lea, Flood 2014) started from state 0.
    x = 0
    got = []
    for _ in range(3):
        r, x = oracle.splitmix64(x)
        got.append(r)
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.mark.parametrize("name,rounds,ctr,key,expect", [
    ("philox2x64", 10, [0, 0], [0], [0xCA00A0459843D731, 0x66C24222C9A845B5]),
    ("philox4x64", 10, [0, 0, 0, 0], [0, 0],
     [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]),
    ("threefry2x64", 20, [0, 0], [0, 0], [0xC2B6E3A8C2C69865, 0x6F81ED42F350084D]),
    ("threefry4x64", 20, [0, 0, 0, 0], [0, 0, 0, 0],
     [0x09218EBDE6C85537, 0x55941F5266D86105, 0x4BD25E16282434DC, 0xEE29EC846BD2E40B]),
])
def test_random123_known_answers(oracle, name, rounds, ctr, key, expect):
    # Random123 known-answer vectors at the library's default round counts;
    # the reference runs the same bijections at 7 (Philox) / 12 (Threefry)
    # rounds, pinned by the golden words below.
    out = oracle.bijection(name, rounds, ctr, key)
    assert [int(v) for v in out] == expect


@pytest.mark.parametrize("engine", list(ENGINES))
@pytest.mark.parametrize("disc", [0, 2**40 + 12345])
def test_engine_words_golden(oracle, golden, engine, disc):
    got = oracle.rng_words(engine, SEED, disc, 64)
    assert np.array_equal(got, golden[f"rng_{engine}_d{disc}"])


def test_discard_equals_draws(oracle):
    # discard(n) is n draws for squares; n blocks for Random123.
    for e, arity in (("squares3", 1), ("squares4", 1), ("philox", 2), ("philox_long", 4),
                     ("threefry", 2), ("threefry_long", 4)):
        full = oracle.rng_words(e, 42, 0, 64 * arity)
        part = oracle.rng_words(e, 42, 8, 16)
        assert np.array_equal(part, full[8 * arity: 8 * arity + 16])


def test_uniform_map(oracle):
    assert oracle.uniform(0, 0) == 0.0
    assert oracle.uniform(0, 2**32 - 1) == (2**32 - 1) / 2**32
    # 64-bit words round in the conversion and can reach 1.0
    assert oracle.uniform(2, 2**64 - 1) == 1.0
    assert oracle.uniform(2, 2**63) == 0.5


# ---------------------------------------------------------------- set-up
@pytest.mark.parametrize("key", list(DECAYS))
def test_params_golden(oracle, golden, key):
    M, m = DECAYS[key]
    rc, ecm, mw = oracle.params(M, m)
    assert rc == 0
    assert np.array_equal(bits(np.array([ecm, mw])), bits(golden[f"params_{key}"]))


def test_params_closed_form(oracle):
    # N = 2: wtmax = PDK(M, m1, m2) = |p*| of the two-body decay.
    M, m1, m2 = 3.0, 1.0, 0.5
    rc, ecm, mw = oracle.params(M, [m1, m2])
    pstar = math.sqrt((M * M - (m1 + m2) ** 2) * (M * M - (m1 - m2) ** 2)) / (2 * M)
    assert abs(1.0 / mw - pstar) <= 4e-16 * pstar
    assert ecm == M - m1 - m2
    assert oracle.params(1.0, [0.6, 0.5])[0] == -2  # not enough energy
    assert oracle.params(0.0, [0.1, 0.1])[0] == -1
    assert oracle.params(1.0, [0.1, -0.1])[0] == -1


# ---------------------------------------------------------------- GENBOD
@pytest.mark.parametrize("key", list(DECAYS))
def test_decay_events_golden_default_engine(oracle, golden, key):
    M, m = DECAYS[key]
    p, wt = oracle.decay_events(M, m, SEED, golden["evts"])
    assert np.array_equal(bits(p), bits(golden[f"decay_{key}_squares3_p"]))
    assert np.array_equal(bits(wt), bits(golden[f"decay_{key}_squares3_wt"]))


@pytest.mark.parametrize("key", ["c1_3body", "c3_10body"])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_decay_events_golden_engines(oracle, golden, key, engine):
    M, m = DECAYS[key]
    p, wt = oracle.decay_events(M, m, SEED, golden["evts"], engine=engine)
    assert np.array_equal(bits(p), bits(golden[f"decay_{key}_{engine}_p"]))
    assert np.array_equal(bits(wt), bits(golden[f"decay_{key}_{engine}_wt"]))


def test_decay_moving_mother_golden(oracle, golden):
    M, m = DECAYS["c1_3body"]
    p, wt = oracle.decay_events(M, m, SEED, golden["evts"][:64], mother4=[6.0, 0.3, -0.2, 0.5])
    assert np.array_equal(bits(p), bits(golden["decay_moving_p"]))
    assert np.array_equal(bits(wt), bits(golden["decay_moving_wt"]))


def test_generate_and_event_weight_golden(oracle, golden):
    M, m = DECAYS["c1_3body"]
    cols, wt = oracle.generate(M, m, SEED, 4096)
    assert np.array_equal(bits(cols), bits(rows_to_cols(golden["gen_c1_p"])))
    w = oracle.event_weight(M, m, cols)
    assert np.array_equal(bits(w), bits(golden["gen_c1_w"]))
    # the weight recomputed from stored momenta (masses of running sums)
    # equals the generation weight only up to cancellation in mass()
    assert np.max(np.abs(w - wt) / wt) < 1e-10
    M, m = DECAYS["c3_10body_threshold"]
    cols, wt = oracle.generate(M, m, SEED, 1024)
    assert np.array_equal(bits(cols), bits(rows_to_cols(golden["gen_thr_p"])))
    assert np.array_equal(bits(oracle.event_weight(M, m, cols)), bits(golden["gen_thr_w"]))


def test_generate_mothers_golden(oracle, golden):
    M, m = DECAYS["c1_3body"]
    cols, wt = oracle.generate_mothers(M, m, SEED, golden["genm_mothers"])
    assert np.array_equal(bits(cols), bits(rows_to_cols(golden["genm_p"])))
    assert np.array_equal(bits(oracle.event_weight(M, m, cols)), bits(golden["genm_w"]))


def test_chunked_generation_is_one_call(oracle):
    # event k depends only on (seed, k): chunks concatenate to one run
    M, m = DECAYS["c1_3body"]
    a, _ = oracle.generate(M, m, SEED, 1000)
    b1, _ = oracle.generate(M, m, SEED, 300, first=0)
    b2, _ = oracle.generate(M, m, SEED, 700, first=300)
    assert np.array_equal(a, np.concatenate([b1, b2], axis=1))


@pytest.mark.parametrize("key", list(DECAYS))
def test_kinematic_invariants(oracle, key):
    M, m = DECAYS[key]
    N = len(m)
    cols, wt = oracle.generate(M, m, 99, 2000)
    p = cols_to_rows(cols, N)
    tot = p.sum(axis=1)
    assert np.max(np.abs(tot[:, 0] - M)) < 1e-12 * M
    assert np.max(np.abs(tot[:, 1:])) < 1e-12 * M
    m2 = p[..., 0] ** 2 - (p[..., 1:] ** 2).sum(axis=-1)
    E2 = p[..., 0] ** 2
    assert np.max(np.abs(m2 - np.array(m) ** 2) / E2) < 1e-12
    assert np.all(wt > 0) and np.all(wt <= 1.0 + 1e-12)


def test_two_body_closed_form(oracle):
    M, m1, m2 = 3.0, 1.0, 0.5
    cols, wt = oracle.generate(M, [m1, m2], 5, 500)
    p = cols_to_rows(cols, 2)
    pstar = math.sqrt((M * M - (m1 + m2) ** 2) * (M * M - (m1 - m2) ** 2)) / (2 * M)
    pm = np.sqrt((p[:, 0, 1:] ** 2).sum(axis=-1))
    assert np.max(np.abs(pm - pstar)) < 1e-14 * M
    assert np.max(np.abs(wt - 1.0)) < 1e-14
    assert np.max(np.abs(p[:, 0, 1:] + p[:, 1, 1:])) < 1e-14 * M


@pytest.mark.parametrize("N", [2, 3, 4, 5])
def test_massless_phase_space_volume(oracle, N):
    # E[prod pd] over GENBOD's sorted uniforms equals the massless N-body
    # phase-space integral in GENBOD's normalisation: M^(N-1) / (2^(2N-3) (N-1)!).
    M = 1.0
    m = [1e-9] * N
    st = oracle.average_on(M, m, 1234, 400000, fid=0, intervals=8)
    rc, ecm, mw = oracle.params(M, m)
    mean_prod = st[0] / 400000 / mw
    expect = M ** (N - 1) / (2 ** (2 * N - 3) * math.factorial(N - 1))
    sd = 3.0 / math.sqrt(400000)  # relative statistical scale (weights are O(mean))
    assert abs(mean_prod / expect - 1.0) < 5 * sd


def test_isotropy(oracle):
    M, m = DECAYS["c1_3body"]
    cols, wt = oracle.generate(M, m, 77, 200000)
    p = cols_to_rows(cols, 3)
    for i in range(3):
        c = p[:, i, 3] / np.sqrt((p[:, i, 1:] ** 2).sum(axis=-1))
        assert abs(c.mean()) < 5 / math.sqrt(3 * len(c))


# ---------------------------------------------------------------- weights / unweighting
def test_unweight_golden(oracle, golden):
    M, m = DECAYS["c1_3body"]
    cols, _ = oracle.generate(M, m, SEED, 4096)
    nacc, out, idx, flags = oracle.unweight(M, m, cols, UW_SEED)
    assert nacc == int(golden["uw_c1_nacc"][0])
    assert np.array_equal(idx, golden["uw_c1_idx"])
    assert np.array_equal(bits(out), bits(rows_to_cols(golden["uw_c1_p"])))
    nacc, out, idx, flags = oracle.unweight(M, m, cols, UWF_SEED, fid=2, max_value=-1.0)
    assert nacc == int(golden["uwf_c1_nacc"][0])
    assert np.array_equal(idx, golden["uwf_c1_idx"])
    nacc, out, idx, flags = oracle.unweight(M, m, cols, UWF_SEED, fid=1,
                                            fpar=golden["bw3_par"], max_value=40.0)
    assert nacc == int(golden["uwbw_c1_nacc"][0])
    assert np.array_equal(idx, golden["uwbw_c1_idx"])


def test_unweight_is_stable_partition(oracle):
    M, m = DECAYS["c1_3body"]
    cols, _ = oracle.generate(M, m, 3, 3000)
    nacc, out, idx, flags = oracle.unweight(M, m, cols, 11)
    acc = np.nonzero(flags)[0]
    rej = np.nonzero(flags == 0)[0]
    assert np.array_equal(idx, np.concatenate([acc, rej]))
    assert np.array_equal(out, cols[:, idx.astype(np.int64)])


def test_unweight_digest_equals_stored_partition(oracle, golden):
    """The streaming Generate + Unweight digest equals the digest of the
    stored partition's index map (po_unweight, pinned to the reference's
    golden index set) at the golden size and at ragged larger sizes."""
    M, m = DECAYS["c1_3body"]
    d_golden = oracle.permutation_digest(golden["uw_c1_idx"])
    assert oracle.unweight_digest(M, m, SEED, 4096, UW_SEED) == \
        (int(golden["uw_c1_nacc"][0]), d_golden)
    for n, seed in ((1, 5), (1023, 9), (100_003, 11)):
        cols, _ = oracle.generate(M, m, seed, n)
        nacc, _, idx, _ = oracle.unweight(M, m, cols, UW_SEED)
        assert oracle.unweight_digest(M, m, seed, n, UW_SEED) == \
            (nacc, oracle.permutation_digest(idx))


def test_permutation_digest_is_order_sensitive(oracle):
    """Swapping two rows, shifting a run by one position, or dropping the
    rejected half all change the digest (the mistakes a scatter could make)."""
    rng = np.random.default_rng(3)
    idx = rng.permutation(50_000).astype(np.uint64)
    d = oracle.permutation_digest(idx)
    sw = idx.copy()
    sw[[10, 20]] = sw[[20, 10]]
    assert oracle.permutation_digest(sw) != d
    sh = np.concatenate([idx[1:], idx[:1]])
    assert oracle.permutation_digest(sh) != d
    assert oracle.permutation_digest(idx[:25_000]) != d
    assert oracle.permutation_digest(idx) == d


def test_unweight_digest_equals_reference(oracle, reference):
    """Against the reference's own Generate + Unweight() (OMP build) at 2^20."""
    M, m = DECAYS["c1_3body"]
    n = 1 << 20
    nacc, _, idx, _ = reference.unweight(M, m, SEED, n, UW_SEED)
    assert oracle.unweight_digest(M, m, SEED, n, UW_SEED) == \
        (nacc, oracle.permutation_digest(idx))


# ---------------------------------------------------------------- AverageOn
@pytest.mark.parametrize("tag,fid", [("one", 0), ("m2", 2), ("bw3", 1)])
def test_average_on_golden(oracle, golden, tag, fid):
    M, m = DECAYS["c1_3body"]
    P = int(golden["avg_procs"][0])
    st = oracle.average_on(M, m, SEED, 100000, fid=fid, fpar=golden["bw3_par"], intervals=P)
    mean, sig = st[1], math.sqrt(st[2]) / st[0]
    assert np.array_equal(bits(np.array([mean, sig])), bits(golden[f"avg_c1_{tag}"]))


def test_average_on_constant_functor(oracle):
    M, m = DECAYS["c3_10body"]
    st = oracle.average_on(M, m, 5, 5000, fid=0, intervals=3)
    assert st[1] == 1.0 and st[2] == 0.0


def test_stats_merge_closed_form(oracle):
    # two points {f=1,w=1} and {f=3,w=3}: W=4, mean=(1+9)/4, M2 = sum w (f-mean)^2
    r = oracle.stats_merge([1.0, 1.0, 0.0], [3.0, 3.0, 0.0])
    mean = 10.0 / 4.0
    assert r[0] == 4.0 and r[1] == mean
    assert abs(r[2] - (1 * (1 - mean) ** 2 + 3 * (3 - mean) ** 2)) < 1e-15


# ---------------------------------------------------------------- live reference
@pytest.mark.parametrize("N", range(2, 13))
def test_oracle_equals_reference_random_decays(oracle, reference, N):
    rng = np.random.default_rng(N)
    m = list(rng.uniform(0.1, 1.0, N))
    M = sum(m) * rng.uniform(1.05, 3.0)
    seed = int(rng.integers(0, 2**63))
    evts = np.array(list(range(40)) + [2**33 + 1, 2**50], dtype=np.uint64)
    for e in range(6):
        p1, w1 = oracle.decay_events(M, m, seed, evts, engine=e)
        p2, w2 = reference.decay_events(M, m, seed, evts, engine=e)
        assert np.array_equal(bits(p1), bits(p2))
        assert np.array_equal(bits(w1), bits(w2))
    # Unweight over a generated container
    pr, wr, _ = reference.generate(M, m, 31, 700)
    cols = rows_to_cols(pr)
    assert np.array_equal(bits(oracle.event_weight(M, m, cols)), bits(wr))
    nacc, pu, idx, _ = reference.unweight(M, m, 31, 700, seed)
    n2, out, idx2, _ = oracle.unweight(M, m, cols, seed)
    assert nacc == n2 and np.array_equal(idx, idx2)


# ---------------------------------------------------------------- Evaluate
BW3 = [(0.892, 0.05, 0.4967141530, -0.1382643012, 1, 2),
       (1.43, 0.11, 0.6476885381, 1.5230298564, 1, 2),
       (0.775, 0.149, -0.2341533747, -0.2341368981, 0, 2)]


def test_evaluate_golden(oracle, golden):
    M, m = DECAYS["c1_3body"]
    got = oracle.evaluate(M, m, SEED, 1024, [(3, (0, 2)), (3, (1, 2))])
    assert np.array_equal(bits(got), bits(golden["eval_c1_m2"]))
    got = oracle.evaluate(M, m, SEED, 1024, [(4, (1, 2, 0, 3)), (1, BW3)])
    assert np.array_equal(bits(got), bits(golden["eval_c1_cos_bw"]))
    M12, m12 = DECAYS["n12"]
    got = oracle.evaluate(M12, m12, SEED, 256, [(3, (4, 11)), (4, (3, 7, 0, 12))])
    assert np.array_equal(bits(got), bits(golden["eval_n12_m2"]))
    mothers = golden["genm_mothers"]
    got = oracle.evaluate_mothers(m, SEED, mothers, [(3, (1, 3)), (3, (0, 1))])
    assert np.array_equal(bits(got), bits(golden["evalm_c1_m2"]))
    got = oracle.evaluate_mothers(m, SEED, mothers, [(4, (2, 3, 1, 3))])
    assert np.array_equal(bits(got), bits(golden["evalm_c1_cos"]))


def test_evaluate_consistent_with_generate(oracle):
    """Evaluate's weight is Generate's generation weight; its functor columns
    are the functors of Generate's momenta, computed independently here."""
    M, m = DECAYS["c1_3body"]
    n = 2000
    ev = oracle.evaluate(M, m, 77, n, [(3, (0, 2)), (4, (1, 2, 0, 3))], first=5000)
    cols, wt = oracle.generate(M, m, 77, n, first=5000)
    p = cols_to_rows(cols, 3)
    assert np.array_equal(bits(ev[0]), bits(wt))
    s = p[:, 0] + p[:, 2]
    m13 = s[:, 0] ** 2 - (s[:, 1:] ** 2).sum(axis=1)
    assert np.max(np.abs(ev[1] - m13)) <= 1e-12 * M * M
    # helicity cosine from the K direction in the (K pi) rest frame against
    # the (K pi) direction in the B frame: Lorentz-invariant closed form
    P = p.sum(axis=1)
    q = p[:, 1] + p[:, 2]
    d = p[:, 1]

    def dot(a, b):
        return a[:, 0] * b[:, 0] - (a[:, 1:] * b[:, 1:]).sum(axis=1)
    num = dot(P, d) * dot(q, q) - dot(P, q) * dot(q, d)
    den = np.sqrt((dot(P, q) ** 2 - dot(q, q) * dot(P, P)) * (dot(q, d) ** 2 - dot(q, q) * dot(d, d)))
    assert np.max(np.abs(ev[2] - num / den)) <= 1e-9
    assert np.all(np.abs(ev[2]) <= 1.0 + 1e-12)
    # decay in the (K pi) rest frame is isotropic under flat phase space:
    # the weighted mean of cos(theta) vanishes
    assert abs(np.sum(ev[0] * ev[2]) / np.sum(ev[0])) < 0.05


def test_evaluate_mothers_closed_forms(oracle):
    """Mothers at rest with the generator's mass: the weight normalisation and
    fECM are the constructor's (same maximum), momenta sum to the mother, and
    the engine of event k is keyed by the pairing hash of (k, seed)."""
    M, m = DECAYS["c1_3body"]
    n = 500
    mothers = np.zeros((4, n))
    mothers[0] = M
    ev = oracle.evaluate_mothers(m, 99, mothers, [(3, (0, 1)), (1, BW3)])
    assert np.all(ev[0] > 0) and np.all(ev[0] <= 1.0)
    # f = (mother + p1)^2 = M^2 + m1^2 + 2 M E1 with E1 <= M: within bounds
    assert np.all(ev[1] >= (M + m[0]) ** 2 - 1e-9)
    # pairing hash (Szudzik): a >= b -> a^2 + a + b else a + b^2, over 2x values
    for a, b in ((0, 0), (3, 7), (7, 3), (2**31 - 1, 5), (2**31, 5)):
        A = (a if a < 2**31 else a - 2**32) % 2**64 * 2 % 2**64
        B = 2 * b
        want = ((A * A + A + B) if A >= B else (A + B * B)) % 2**64 // 2
        assert oracle.pair_hash(a, b) == want


def test_evaluate_equals_reference(oracle, reference):
    for N in (2, 3, 6, 12):
        rng = np.random.default_rng(40 + N)
        m = list(rng.uniform(0.1, 1.0, N))
        M = sum(m) * rng.uniform(1.05, 3.0)
        F = [(3, (0, N - 1)), (4, (0, N - 1, 0, N))]
        ref, _ = reference.evaluate(M, m, 4321, 300, F)
        assert np.array_equal(bits(oracle.evaluate(M, m, 4321, 300, F)), bits(ref))
        mt = np.zeros((4, 300))
        mt[1:] = rng.normal(0, 2.0, (3, 300))
        mt[0] = np.sqrt((M * rng.uniform(1.0, 1.2, 300)) ** 2 + (mt[1:] ** 2).sum(axis=0))
        Fm = [(3, (1, N)), (4, (1, N, 1, N))]
        ref = reference.evaluate_mothers(M, m, 4321, mt, Fm)
        assert np.array_equal(bits(oracle.evaluate_mothers(m, 4321, mt, Fm)), bits(ref))


# ---------------------------------------------------------------- DenseHistogram
def test_histogram_golden(oracle, golden):
    x = golden["hist_dalitz_x"]
    lo, hi = golden["hist_dalitz_limits"]
    got = oracle.histogram_fill([20, 20], lo, hi, x, golden["gen_c1_w"])
    assert np.array_equal(bits(got), bits(golden["hist_dalitz_w"]))
    got = oracle.histogram_fill([20, 20], lo, hi, x)
    assert np.array_equal(got, golden["hist_dalitz_n"])
    got = oracle.histogram_fill([25], [lo[0]], [hi[0]], golden["hist_1d_x"][None, :])
    assert np.array_equal(got, golden["hist_1d_n"])


def test_histogram_keys_closed_form(oracle):
    """GetGlobalBin by hand: row-major keys, the all-below / all-above slots,
    the upper edge's index grid, truncation of (-1, 0) to 0, undefined cases
    dropped."""
    g, lo, hi = [4, 5], [0.0, 10.0], [2.0, 20.0]
    nb = 20
    assert oracle.hist_key(g, lo, hi, [0.1, 10.1]) == 0
    assert oracle.hist_key(g, lo, hi, [1.9, 19.9]) == 3 * 5 + 4
    assert oracle.hist_key(g, lo, hi, [-1.0, 5.0]) == nb          # all below
    assert oracle.hist_key(g, lo, hi, [3.0, 30.0]) == nb + 1      # all above
    assert oracle.hist_key(g, lo, hi, [-0.2, 15.0]) == 2          # X0 in (-1, 0) -> 0
    assert oracle.hist_key(g, lo, hi, [2.0, 10.0]) == 4 * 5       # X0 == grid -> slot nb
    assert oracle.hist_key(g, lo, hi, [-0.6, 15.0]) == 2**64 - 1  # X0 = -1.2: undefined
    assert oracle.hist_key(g, lo, hi, [float("nan"), 15.0]) == 2**64 - 1
    assert oracle.hist_key(g, lo, hi, [1.0, 40.0]) == 2**64 - 1   # key past nb + 1
    # 1-D: the upper limit itself lands in the underflow slot (index == grid)
    assert oracle.hist_key([10], [0.0], [1.0], [1.0]) == 10
    # sums: unit weights count the events; weights sum per bin
    rng = np.random.default_rng(0)
    x = rng.uniform([0, 10], [2, 20], (5000, 2)).T
    w = rng.uniform(0, 1, 5000)
    h = oracle.histogram_fill(g, lo, hi, x, w)
    assert abs(h.sum() - w.sum()) < 1e-9 and h[nb] == 0 and h[nb + 1] == 0
    assert oracle.histogram_fill(g, lo, hi, x).sum() == 5000


def test_histogram_equals_reference(oracle, reference):
    rng = np.random.default_rng(5)
    for grid in ([30], [17, 9]):
        nd = len(grid)
        lo = rng.uniform(-1, 0, nd)
        hi = lo + rng.uniform(1, 2, nd)
        x = rng.uniform(lo - 0.01 * (hi - lo), hi + 0.2 * (hi - lo), (20000, nd)).T
        w = rng.exponential(1.0, 20000)
        # keep to events whose key the reference defines
        ok = np.array([oracle.hist_key(grid, lo, hi, x[:, k]) != 2**64 - 1 for k in range(20000)])
        x, w = np.ascontiguousarray(x[:, ok]), w[ok]
        for ww in (w, None):
            ref, _ = reference.histogram_fill(grid, lo, hi, x, ww)
            assert np.array_equal(bits(oracle.histogram_fill(grid, lo, hi, x, ww)), bits(ref))


# ---------------------------------------------------------------- hydra::unweight
def test_unweight_sample_equals_reference(oracle, reference):
    """hydra::unweight(device::sys, x, Gaussian, max, seed, jump): the reference
    OMP build's order (sequential swap partition) bit for bit; the stable
    order keeps the same prefix and the same rejected set."""
    rng = np.random.default_rng(1)
    for n, mp, jump in ((100000, -1.0, 0), (50000, 0.3, 200000), (7, -1.0, 5),
                        (12, -1.0, 0), (1, -1.0, 0)):
        x = rng.uniform(0, 20, n)
        kept, xo, w = reference.unweight_sample(x, 10.0, 1.0, mp, 0x8EC74D321E6B5A27, jump)
        k1, o1, _ = oracle.unweight_sample(x[None, :], w, mp, 0x8EC74D321E6B5A27, jump, order=1)
        assert k1 == kept and np.array_equal(bits(o1[0]), bits(xo))
        k0, o0, _ = oracle.unweight_sample(x[None, :], w, mp, 0x8EC74D321E6B5A27, jump, order=0)
        assert np.array_equal(o0[0][:k0], xo[:kept])
        assert np.array_equal(np.sort(o0[0][k0:]), np.sort(xo[kept:]))


def test_unweight_sample_closed_form(oracle):
    """Acceptance u * max <= w by hand; the sequential-partition order by a
    direct simulation of the swap loop."""
    rng = np.random.default_rng(3)
    n = 3000
    w = rng.uniform(0, 2, n)
    words = oracle.rng_words("squares3", 0x8EC74D321E6B5A27, 17, n).astype(np.float64)
    keep = (words / 2**32) * w.max() <= w
    k, out, idx = oracle.unweight_sample(np.arange(n, dtype=float)[None, :], w, -1.0,
                                         0x8EC74D321E6B5A27, 17, order=1)
    assert k == keep.sum()
    a = list(range(n))
    first = 0
    while first < n and keep[first]:
        first += 1
    for nxt in range(first + 1, n):
        if keep[nxt]:
            a[first], a[nxt] = a[nxt], a[first]
            first += 1
    assert np.array_equal(idx, np.array(a, dtype=np.uint64))


def test_unweight_sample_golden(oracle, golden):
    x, w = golden["uws_x"], golden["uws_w"]
    for tag, mp, jump in (("auto", -1.0, 0), ("max", 0.5, 10000)):
        k, out, _ = oracle.unweight_sample(x[None, :], w, mp, 0x8EC74D321E6B5A27, jump, order=1)
        assert k == int(golden[f"uws_{tag}_kept"][0])
        assert np.array_equal(bits(out[0]), bits(golden[f"uws_{tag}_x"]))


# ---------------------------------------------------------------- phase_space_range
from oracle import MINSTD  # noqa: E402

RANGES = ["k3pi", "c1_moving", "n12", "n2"]


def range_case(golden, key):
    v = golden[f"range_{key}_in"]
    seed, first = (int(x) for x in golden[f"range_{key}_seed"])
    m = {"k3pi": [0.13957061] * 3, "c1_moving": DECAYS["c1_3body"][1], "n12": DECAYS["n12"][1],
         "n2": DECAYS["n2"][1]}[key]
    return list(v[:4]), m, seed, first


def test_minstd_known_answer(oracle):
    # minstd_rand's standard check value: the 10000th draw of an engine seeded
    # with 1 is 399268537 (ISO C++ [rand.predef]); a=48271 with a dropped
    # modulus step or the wrong multiplier fails it
    w = oracle.rng_words(MINSTD, 1, 0, 10000)
    assert int(w[-1]) == 399268537
    # brute-force stepping of x -> 48271 x mod (2^31 - 1)
    x = 321
    for k in range(64):
        x = x * 48271 % 2147483647
        assert int(oracle.rng_words(MINSTD, 321, 0, 64)[k]) == x


def test_minstd_seed_and_discard(oracle):
    # seed(s) keeps s mod m (1 for 0); only the low 32 bits of a wider seed
    assert np.array_equal(oracle.rng_words(MINSTD, 0, 0, 8), oracle.rng_words(MINSTD, 1, 0, 8))
    assert np.array_equal(oracle.rng_words(MINSTD, 2147483647, 0, 8),
                          oracle.rng_words(MINSTD, 1, 0, 8))
    assert np.array_equal(oracle.rng_words(MINSTD, 2**32 + 5, 0, 8),
                          oracle.rng_words(MINSTD, 5, 0, 8))
    # discard(n) by modular exponentiation equals n draws
    full = oracle.rng_words(MINSTD, 99, 0, 3000)
    assert np.array_equal(oracle.rng_words(MINSTD, 99, 2500, 500), full[2500:])
    # u = (x - 1) / (m - 1): 0 at x = 1, just below 1 at x = m - 1
    assert oracle.uniform(MINSTD, 1) == 0.0
    assert oracle.uniform(MINSTD, 2147483646) == 2147483645 / 2147483646


@pytest.mark.parametrize("seed", [321, 2**32 + 321, 0])
@pytest.mark.parametrize("disc", [0, 2**40 + 12345])
def test_minstd_words_golden(oracle, golden, seed, disc):
    assert np.array_equal(oracle.rng_words(MINSTD, seed, disc, 64),
                          golden[f"rng_minstd_s{seed}_d{disc}"])


@pytest.mark.parametrize("key", RANGES)
def test_range_golden(oracle, golden, key):
    m4, m, seed, first = range_case(golden, key)
    w, cols = oracle.phase_space_range(m4, m, seed, 512, first=first)
    assert np.array_equal(bits(w), bits(golden[f"range_{key}_w"]))
    assert np.array_equal(bits(cols_to_rows(cols, len(m))), bits(golden[f"range_{key}_p"]))


def test_range_equals_reference(oracle, reference):
    rng = np.random.default_rng(4321)
    for N in range(2, 13):
        m = list(rng.uniform(0.1, 1.0, N))
        M = sum(m) * rng.uniform(1.05, 3.0)
        pv = rng.normal(0, 2.0, 3)
        m4 = [math.sqrt(M * M + (pv ** 2).sum())] + list(pv)
        seed = int(rng.integers(0, 2**64, dtype=np.uint64))
        first = int(rng.integers(0, 2**45))
        w, cols = oracle.phase_space_range(m4, m, seed, 300, first=first)
        wr, pr = reference.phase_space_range(m4, m, seed, 300, first=first)
        assert np.array_equal(bits(w), bits(wr)), N
        assert np.array_equal(bits(cols_to_rows(cols, N)), bits(pr)), N


def test_range_closed_forms(oracle):
    # element k is the DecayMother event k of engine MINSTD with fTeCmTm from
    # the mother's mass; at rest: energy sum M, momentum sum 0, daughter masses
    M = 5.279
    m = DECAYS["c1_3body"][1]
    w, cols = oracle.phase_space_range([M, 0, 0, 0], m, 7, 400)
    p = cols_to_rows(cols, 3)
    assert np.max(np.abs(p[:, :, 0].sum(axis=1) - M)) < 1e-13 * M
    assert np.max(np.abs(p[:, :, 1:].sum(axis=1))) < 1e-13 * M
    mass = np.sqrt(p[:, :, 0] ** 2 - (p[:, :, 1:] ** 2).sum(axis=-1))
    assert np.max(np.abs(mass - np.array(m))) < 1e-10
    # the weights are the generation weights of those events
    rc, ecm, mw = oracle.params(M, m)
    pe, we = oracle.decay_events(M, m, 7, np.arange(400, dtype=np.uint64), engine=MINSTD)
    assert np.array_equal(bits(we), bits(w))
    # two-body: |p| = p*, weight 1 to rounding
    w2, c2 = oracle.phase_space_range([3.0, 0, 0, 0], [1.0, 0.5], 11, 200)
    p2 = cols_to_rows(c2, 2)
    pstar = math.sqrt((9.0 - 1.5 ** 2) * (9.0 - 0.5 ** 2)) / 6.0
    assert np.max(np.abs(np.sqrt((p2[:, 0, 1:] ** 2).sum(axis=-1)) - pstar)) < 1e-14 * 3
    assert np.max(np.abs(w2 - 1.0)) < 1e-14
    # a seed is taken modulo 2^32
    wa, _ = oracle.phase_space_range([M, 0, 0, 0], m, 2**32 + 7, 50)
    assert np.array_equal(bits(wa), bits(w[:50]))


def test_range_mother_mass_and_boost(oracle):
    # fTeCmTm comes from the mother's four-vector; the final boost carries the
    # reference's (E, px, py) / E components
    m4 = [6.0, 0.3, -0.2, 0.5]
    M = math.sqrt(36.0 - 0.09 - 0.04 - 0.25)
    m = DECAYS["c1_3body"][1]
    rc, ecm, mw, beta = oracle.range_params(m4, m)
    rc2, ecm2, mw2 = oracle.params(M, m)
    assert (rc, rc2) == (0, 0) and ecm == ecm2 and mw == mw2
    d3 = math.sqrt(0.09 + 0.04 + 0.25)
    assert np.allclose(beta, np.array([6.0, 0.3, -0.2]) * (d3 / 6.0) / d3, rtol=1e-15)
    assert oracle.range_params([0.3, 0, 0, 0], [0.2, 0.2], )[0] == -2
    assert oracle.range_params([1.0, 0, 0, 0], [0.2, 0.0])[0] == -1


def test_range_histogram_equals_reference(oracle, reference, golden):
    # the range histogram of range_semantics.inl: the oracle's elements, the
    # pair mass (functor 5) and the oracle's fill reproduce the reference's
    # DenseHistogram over the lazy range bit for bit
    mk, mpi = 0.493677, 0.13957061
    w, cols = oracle.phase_space_range([mk, 0, 0, 0], [mpi] * 3, 321, 100000)
    p = cols_to_rows(cols, 3)
    x = np.array([oracle.functor(5, [0, 1], p[k]) for k in range(100000)])
    for wtd in (0, 1):
        got = oracle.histogram_fill([10], [2 * mpi], [mk - mpi], x[None, :],
                                    w if wtd else None)
        assert np.array_equal(bits(got), bits(golden[f"range_hist_k3pi_w{wtd}"]))
    ref, _ = reference.range_histogram3([mk, 0, 0, 0], [mpi] * 3, 321, 100000, 0, 1, 10,
                                        2 * mpi, mk - mpi)
    assert np.array_equal(bits(ref), bits(golden["range_hist_k3pi_w0"]))


def test_mass_functor_closed_form(oracle):
    # (p_a + p_b).mass(): sqrt of a positive square, 0 otherwise
    p = np.array([[5.0, 1.0, 2.0, 2.0], [3.0, -1.0, 0.0, 0.0], [1.0, 0.0, 0.0, 2.0]])
    assert oracle.functor(5, [0, 1], p) == math.sqrt(64.0 - 0.0 - 4.0 - 4.0)
    assert oracle.functor(5, [2, 2], p) == 0.0  # spacelike sum
    assert oracle.functor(5, [0, 1], p) == math.sqrt(oracle.functor(3, [0, 1], p))


# ---------------------------------------------------------------- decay chain (Meld)
B0_CHAIN = (5.27955, [3.0969, 0.493677, 0.13957061])
JPSI_CHAIN = (3.0969, [0.1056583745] * 2)
CHAIN_VARS = [(3, (0, 2)), (3, (1, 2)), (4, (1, 2, 0, 3)), (6, (1, 2, 3))]


def chain_oracle(oracle, golden, n):
    """phsp_chain.inl:234-244 restated: Generate(B0), Generate over the J/psi
    column of each event, the rows melded: (n, 5, 4)."""
    s3, s2 = (int(x) for x in golden["chain_seeds"])
    c3, _ = oracle.generate(B0_CHAIN[0], B0_CHAIN[1], s3, n)
    c2, _ = oracle.generate_mothers(JPSI_CHAIN[0], JPSI_CHAIN[1], s2, c3[0:4])
    return np.concatenate([cols_to_rows(c3, 3), cols_to_rows(c2, 2)], axis=1)


def test_chain_golden(oracle, golden):
    rows = chain_oracle(oracle, golden, 512)
    assert np.array_equal(bits(rows), bits(golden["chain_p"]))
    for j, (fid, t) in enumerate(CHAIN_VARS):
        got = np.array([oracle.functor(fid, list(t), rows[k]) for k in range(512)])
        assert np.array_equal(bits(got), bits(golden["chain_vars"][j])), fid


def test_chain_seeds_equal_reference(reference, golden):
    got = [reference.params(M, m)[3] for M, m in (B0_CHAIN, JPSI_CHAIN)]
    assert got == [int(x) for x in golden["chain_seeds"]]


def test_decay_plane_angle_closed_forms(oracle):
    # K along z: d in the x-z plane, h rotated about z by phi -> chi = phi;
    # h in the plane of (d, K) on d's side -> 0; swapping the sides -> pi;
    # chi is unchanged by boosts of h along K's direction and by scaling h
    K_d = np.array([1.0, 0.3, 0.0, 0.5])
    e = np.array([1.2, -0.3, 0.0, 0.7])
    for phi in (0.3, 1.7, 3.0, 4.4, 6.1):
        h = np.array([0.9, 0.4 * math.cos(phi), 0.4 * math.sin(phi), -0.2])
        p = np.vstack([K_d, e, h])
        chi = oracle.functor(6, [0, 1, 2], p)
        assert abs(chi - phi) < 1e-13, (phi, chi)
        h2 = h.copy()
        h2[3] += 5.0
        h2[1:3] *= 3.0
        assert abs(oracle.functor(6, [0, 1, 2], np.vstack([K_d, e, h2])) - phi) < 1e-13
    assert oracle.functor(6, [0, 1, 2], np.vstack([K_d, e, [1.0, 0.2, 0.0, 0.1]])) == 0.0
    assert abs(oracle.functor(6, [0, 1, 2], np.vstack([K_d, e, [1.0, -0.2, 0.0, 0.1]]))
               - math.pi) < 1e-15
