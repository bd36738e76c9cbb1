"""Parity of the gfx950 kernels (through the C ABI) with the oracle and the
reference's golden vectors.

Bars: engine words, accepted-index sets and every result built only from
+ - * / sqrt (weights, PhaseSpaceWeight, flags, permutations) are compared
bit for bit; momenta, which pass through sin/cos (ocml on the GPU, glibc in
the oracle), within |dp| <= 1e-12 * E per particle; AverageOn within 1e-12
relative on the mean (the reduction order differs from the reference's).
"""
import math

import numpy as np
import pytest
import torch

from oracle import ENGINES, cols_to_rows, rows_to_cols
from gpu_util import (FP64, MIXED, average_on, bw_functor, event_weight, generate,
                      make_decay, make_functor, momentum_error, rng_words, unweight)
from hydra_amd import capi

pytestmark = pytest.mark.gpu

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
P_TOL = 1e-12


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def random_decay(rng, N):
    m = list(rng.uniform(0.1, 1.0, N))
    M = sum(m) * rng.uniform(1.05, 3.0)
    return M, m


# ------------------------------------------------------------------ engines
@pytest.mark.parametrize("engine", list(ENGINES))
def test_rng_words_golden_and_oracle(golden, oracle, engine):
    for disc in (0, 2**40 + 12345):
        got = rng_words(engine, SEED, disc, 64)
        assert np.array_equal(got, golden[f"rng_{engine}_d{disc}"]), (engine, disc)
    for seed, disc, n in ((1, 0, 1000), (2**63 + 5, 2**64 - 7, 333), (7, 123456789, 4097)):
        assert np.array_equal(rng_words(engine, seed, disc, n),
                              oracle.rng_words(engine, seed, disc, n)), (engine, seed, disc)


# ------------------------------------------------------------------ generation
@pytest.mark.parametrize("key", list(DECAYS))
def test_generate_fp64_golden_first_events(golden, key):
    M, m = DECAYS[key]
    N = len(m)
    dec = make_decay(M, m)
    cols, wt, evw = generate(dec, SEED, 256)
    ref_p = golden[f"decay_{key}_squares3_p"][:256]
    err, _ = momentum_error(cols.cpu().numpy(), rows_to_cols(ref_p), N)
    assert err <= P_TOL, err
    assert np.array_equal(bits(wt.cpu().numpy()), bits(golden[f"decay_{key}_squares3_wt"][:256]))


@pytest.mark.parametrize("key", ["c1_3body", "c3_10body"])
@pytest.mark.parametrize("engine", list(ENGINES))
def test_generate_fp64_golden_sparse_high_indices(golden, key, engine):
    M, m = DECAYS[key]
    N = len(m)
    dec = make_decay(M, m, engine=engine)
    evts = golden["evts"]
    ref_p = golden[f"decay_{key}_{engine}_p"]
    ref_w = golden[f"decay_{key}_{engine}_wt"]
    for j in list(range(0, 256, 37)) + list(range(256, len(evts))):
        cols, wt, _ = generate(dec, SEED, 1, first=int(evts[j]))
        err, _ = momentum_error(cols.cpu().numpy(), rows_to_cols(ref_p[j:j + 1]), N)
        assert err <= P_TOL, (j, err)
        assert bits(wt.cpu().numpy())[0] == bits(ref_w[j:j + 1])[0]


@pytest.mark.parametrize("N", range(2, 13))
@pytest.mark.parametrize("engine", list(ENGINES))
def test_generate_fp64_vs_oracle(oracle, N, engine):
    rng = np.random.default_rng(100 * N + ENGINES[engine])
    M, m = random_decay(rng, N)
    seed = int(rng.integers(0, 2**63))
    first = int(rng.integers(0, 2**40))
    n = 3000  # not a multiple of the wave or block size
    dec = make_decay(M, m, engine=engine)
    cols, wt, evw = generate(dec, seed, n, first=first)
    ocols, owt = oracle.generate(M, m, seed, n, first=first, engine=engine)
    g = cols.cpu().numpy()
    err, rms = momentum_error(g, ocols, N)
    assert err <= P_TOL, (err, rms)
    assert np.array_equal(bits(wt.cpu().numpy()), bits(owt))
    # PhaseSpaceWeight of the stored momenta: bit-equal to the oracle's on the
    # same momenta, and to the kernel that recomputes it from the columns
    assert np.array_equal(bits(evw.cpu().numpy()), bits(oracle.event_weight(M, m, g)))
    assert np.array_equal(bits(event_weight(dec, cols).cpu().numpy()), bits(evw.cpu().numpy()))


def test_generate_threshold_weights(oracle):
    # near threshold the weight chain cancels; it is pure + - * / sqrt, so it
    # must still agree bit for bit
    M, m = DECAYS["c3_10body_threshold"]
    dec = make_decay(M, m)
    cols, wt, evw = generate(dec, SEED, 20000)
    ocols, owt = oracle.generate(M, m, SEED, 20000)
    assert np.array_equal(bits(wt.cpu().numpy()), bits(owt))
    err, _ = momentum_error(cols.cpu().numpy(), ocols, 10)
    assert err <= P_TOL, err
    ow = oracle.event_weight(M, m, ocols)
    rel = np.abs(evw.cpu().numpy() - ow) / np.abs(ow)
    # PhaseSpaceWeight recomputed from momenta is ill-conditioned near
    # threshold (masses of running sums cancel in a - b - c).  The scale is
    # set by the reference's own recompute against its generation weight;
    # momenta that differ by ulps (sin/cos) must stay within a few times it.
    ref_scale = np.max(np.abs(ow - owt) / owt)
    assert rel.max() <= 10 * ref_scale, (rel.max(), ref_scale)


def test_generate_moving_mother(golden, oracle):
    M, m = DECAYS["c1_3body"]
    mom = [6.0, 0.3, -0.2, 0.5]
    dec = make_decay(M, m, mother4=mom)
    cols, wt, _ = generate(dec, SEED, 64)
    err, _ = momentum_error(cols.cpu().numpy(), rows_to_cols(golden["decay_moving_p"]), 3)
    assert err <= P_TOL
    # a mother along z alone gives b2 just below 1: the reference applies a
    # (large) boost, the kernel must too
    mom2 = [7.0, 0.0, 0.0, 3.0]
    dec2 = make_decay(M, m, mother4=mom2)
    c2, _, _ = generate(dec2, 5, 500)
    oc2, _ = oracle.generate(M, m, 5, 500, mother4=mom2)
    err, _ = momentum_error(c2.cpu().numpy(), oc2, 3)
    assert err <= 1e-9, err


def test_generate_mothers(golden, oracle):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    mothers = golden["genm_mothers"]
    cols, wt, evw = generate(dec, SEED, mothers.shape[1], mothers=mothers)
    err, _ = momentum_error(cols.cpu().numpy(), rows_to_cols(golden["genm_p"]), 3)
    assert err <= P_TOL, err
    rng = np.random.default_rng(3)
    n = 5000
    mt = np.zeros((4, n))
    mt[1:] = rng.normal(0, 3.0, (3, n))
    mt[0] = np.sqrt(M * M + (mt[1:] ** 2).sum(axis=0))
    cols, wt, _ = generate(dec, 77, n, first=1000, mothers=mt)
    oc, ow = oracle.generate_mothers(M, m, 77, mt, first=1000)
    err, _ = momentum_error(cols.cpu().numpy(), oc, 3)
    assert err <= P_TOL, err
    assert np.array_equal(bits(wt.cpu().numpy()), bits(ow))


def test_generate_chunks_equal_one_call():
    M, m = DECAYS["c3_10body"]
    dec = make_decay(M, m)
    a, _, _ = generate(dec, 9, 10000)
    b, _, _ = generate(dec, 9, 3333, first=0)
    c, _, _ = generate(dec, 9, 6667, first=3333)
    assert torch.equal(a, torch.cat([b, c], dim=1))


def test_generate_ld_and_f32_output(oracle):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    cols, _, _ = generate(dec, 4, 1000, ld=1536)
    ref, _, _ = generate(dec, 4, 1000)
    assert torch.equal(cols[:, :1000], ref)
    c32, _, _ = generate(dec, 4, 1000, out_f32=True)
    assert torch.equal(c32, ref.float())


def test_generate_edge_sizes(oracle):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    for n in (1, 63, 64, 65, 255, 257, 1023):
        cols, wt, _ = generate(dec, 21, n, first=n)
        oc, ow = oracle.generate(M, m, 21, n, first=n)
        err, _ = momentum_error(cols.cpu().numpy()[:, :n], oc, 3)
        assert err <= P_TOL
        assert np.array_equal(bits(wt.cpu().numpy()[:n]), bits(ow))


# ------------------------------------------------------------------ weights / unweighting
def test_event_weight_bit_exact(golden, oracle):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    cols = torch.as_tensor(rows_to_cols(golden["gen_c1_p"]), device="cuda")
    w = event_weight(dec, cols).cpu().numpy()
    assert np.array_equal(bits(w), bits(golden["gen_c1_w"]))
    f = bw_functor(golden["bw3_par"])
    wf = event_weight(dec, cols, f).cpu().numpy()
    owf = oracle.event_weight(M, m, cols.cpu().numpy(), fid=1, fpar=golden["bw3_par"])
    assert np.array_equal(bits(wf), bits(owf))
    M, m = DECAYS["c3_10body_threshold"]
    cols = torch.as_tensor(rows_to_cols(golden["gen_thr_p"]), device="cuda")
    w = event_weight(make_decay(M, m), cols).cpu().numpy()
    assert np.array_equal(bits(w), bits(golden["gen_thr_w"]))


def test_unweight_golden(golden):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    cols = torch.as_tensor(rows_to_cols(golden["gen_c1_p"]), device="cuda")
    nacc, out, idx, _ = unweight(dec, cols, UW_SEED)
    assert nacc == int(golden["uw_c1_nacc"][0])
    assert np.array_equal(idx.cpu().numpy(), golden["uw_c1_idx"].astype(np.int64))
    assert np.array_equal(bits(out.cpu().numpy()), bits(rows_to_cols(golden["uw_c1_p"])))
    f = make_functor(capi.F_MASS2_01)
    nacc, out, idx, _ = unweight(dec, cols, UWF_SEED, f=f, max_value=-1.0)
    assert nacc == int(golden["uwf_c1_nacc"][0])
    assert np.array_equal(idx.cpu().numpy(), golden["uwf_c1_idx"].astype(np.int64))
    nacc, out, idx, _ = unweight(dec, cols, UWF_SEED, f=bw_functor(golden["bw3_par"]),
                                 max_value=40.0)
    assert nacc == int(golden["uwbw_c1_nacc"][0])
    assert np.array_equal(idx.cpu().numpy(), golden["uwbw_c1_idx"].astype(np.int64))


@pytest.mark.parametrize("key", list(DECAYS))
@pytest.mark.parametrize("n", [1, 64, 1000, 1024, 5000])
def test_unweight_vs_oracle(oracle, key, n):
    M, m = DECAYS[key]
    dec = make_decay(M, m)
    cols, wt, evw = generate(dec, 42, n)
    g = cols.cpu().numpy()
    onacc, oout, oidx, oflags = oracle.unweight(M, m, g, UW_SEED)
    for w_in in (None, evw):
        nacc, out, idx, w_out = unweight(dec, cols, UW_SEED, w_in=w_in)
        assert nacc == onacc
        assert np.array_equal(idx.cpu().numpy()[:n], oidx.astype(np.int64))
        assert np.array_equal(bits(out.cpu().numpy()), bits(oout))
        if w_in is not None:
            assert torch.equal(w_out[:n], evw[:n][idx[:n]])


@pytest.mark.parametrize("n", [1, 1000, 5000])
def test_unweight_weights_only_output(oracle, n):
    """w_out requested alone (no columns, no indices) is still written: the
    PhaseSpaceWeight column permuted by the stable partition."""
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    cols, _, evw = generate(dec, 17, n)
    onacc, _, oidx, _ = oracle.unweight(M, m, cols.cpu().numpy(), UW_SEED)
    nacc, out, idx, w_out = unweight(dec, cols, UW_SEED, w_in=evw, want_out=False,
                                     want_idx=False)
    assert out is None and idx is None
    assert nacc == onacc
    expect = evw.cpu().numpy()[:n][oidx.astype(np.int64)]
    assert np.array_equal(bits(w_out.cpu().numpy()[:n]), bits(expect))


def test_unweight_empty_and_extremes(oracle):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    cols = torch.empty((12, 1), dtype=torch.float64, device="cuda")
    nacc, _, _, _ = unweight(dec, cols, 1, n=0, want_out=False, want_idx=False)
    assert nacc == 0
    # N=10 at M=10: prod pd >> 1, every event accepted (saturating predicate)
    M10, m10 = DECAYS["c3_10body"]
    d10 = make_decay(M10, m10)
    cols, _, _ = generate(d10, 1, 3000)
    nacc, out, idx, _ = unweight(d10, cols, 5)
    assert nacc == oracle.unweight(M10, m10, cols.cpu().numpy(), 5)[0]
    # near threshold: prod pd << 1, (almost) nothing accepted
    Mt, mt = DECAYS["c3_10body_threshold"]
    dt = make_decay(Mt, mt)
    cols, _, _ = generate(dt, 1, 3000)
    nacc, out, idx, _ = unweight(dt, cols, 5)
    assert nacc == oracle.unweight(Mt, mt, cols.cpu().numpy(), 5)[0]
    assert torch.equal(idx[:3000].sort().values, torch.arange(3000, device="cuda"))


def test_unweight_functor_max_search(oracle, golden):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    cols, _, evw = generate(dec, 8, 7777)
    f = bw_functor(golden["bw3_par"])
    g = cols.cpu().numpy()
    onacc, oout, oidx, _ = oracle.unweight(M, m, g, 99, fid=1, fpar=golden["bw3_par"],
                                           max_value=-1.0)
    for w_in in (None, evw):
        nacc, out, idx, _ = unweight(dec, cols, 99, f=f, max_value=-1.0, w_in=w_in)
        assert nacc == onacc
        assert np.array_equal(idx.cpu().numpy(), oidx.astype(np.int64))


# ------------------------------------------------------------------ AverageOn
@pytest.mark.parametrize("key", ["c1_3body", "c3_10body", "n2", "n12"])
def test_average_on_vs_oracle(golden, oracle, key):
    M, m = DECAYS[key]
    dec = make_decay(M, m)
    for fid, fpar, f in ((0, None, None), (2, None, make_functor(capi.F_MASS2_01)),
                         (1, golden["bw3_par"], bw_functor(golden["bw3_par"]))):
        if fid == 1 and len(m) < 3:
            continue
        n = 50000
        st = average_on(dec, SEED, n, f=f)
        ost = oracle.average_on(M, m, SEED, n, fid=fid, fpar=fpar, intervals=8)
        assert abs(st[0] - ost[0]) <= 1e-12 * abs(ost[0])
        assert abs(st[1] - ost[1]) <= 1e-12 * abs(ost[1])
        if ost[2] > 0:
            assert abs(st[2] - ost[2]) <= 1e-10 * ost[2]


def test_average_on_golden(golden):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    for tag, f in (("one", None), ("m2", make_functor(capi.F_MASS2_01)),
                   ("bw3", bw_functor(golden["bw3_par"]))):
        st = average_on(dec, SEED, 100000, f=f)
        mean, sig = st[1], math.sqrt(st[2]) / st[0]
        ref = golden[f"avg_c1_{tag}"]
        assert abs(mean - ref[0]) <= 1e-12 * abs(ref[0])
        if ref[1] > 0:
            assert abs(sig - ref[1]) <= 1e-10 * ref[1]


def test_average_on_chunks_and_determinism(oracle):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    f = make_functor(capi.F_MASS2_01)
    a = average_on(dec, 3, 204800, f=f)
    b = average_on(dec, 3, 204800, f=f)
    assert np.array_equal(bits(a), bits(b))
    parts = [average_on(dec, 3, 4096 * 10, f=f, first=4096 * 10 * i) for i in range(5)]
    x = capi.Stats(*parts[0])
    for p in parts[1:]:
        y = capi.Stats(*p)
        r = capi.Stats()
        capi.lib().hpsp_stats_merge(x, y, r)
        x = r
    assert abs(x.mean - a[1]) <= 1e-13 * abs(a[1])
    assert abs(x.w - a[0]) <= 1e-13 * a[0]


def test_average_on_mothers(oracle, golden):
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    mothers = golden["genm_mothers"]
    f = make_functor(capi.F_BW_SUM, [(0.892, 0.05, 1.0, 0.5, 2, 3), (5.0, 1.0, 0.3, 0.0, 0, 1)])
    st = average_on(dec, SEED, mothers.shape[1], f=f, mothers=mothers)
    par = np.array([2, 0.892, 0.05, 1.0, 0.5, 2, 3, 5.0, 1.0, 0.3, 0.0, 0, 1], dtype=float)
    ost = oracle.average_on_mothers(M, m, SEED, mothers, fid=1, fpar=par, intervals=8)
    assert abs(st[1] - ost[1]) <= 1e-12 * abs(ost[1])


# ------------------------------------------------------------------ mixed precision
@pytest.mark.parametrize("key", ["c1_3body", "c3_10body", "c3_10body_threshold", "n12"])
def test_mixed_generate_error_bounds(oracle, key):
    M, m = DECAYS[key]
    N = len(m)
    dec = make_decay(M, m)
    n = 20000
    cols, wt, evw = generate(dec, SEED, n, prec=MIXED, out_f32=True)
    ocols, owt = oracle.generate(M, m, SEED, n)
    # the weight chain stays FP64 (one-ulp sqrt and reciprocal in pdk)
    assert np.max(np.abs(wt.cpu().numpy() - owt) / owt) <= 4e-15
    # the mixed PhaseSpaceWeight column is that generation weight (equal to
    # PhaseSpaceWeight in exact arithmetic, not recomputed from FP32 momenta)
    assert torch.equal(evw, wt)
    err, rms = momentum_error(cols.cpu().numpy(), ocols, N)
    assert err <= 2e-5, (err, rms)
    assert rms <= 1e-6, (err, rms)


@pytest.mark.parametrize("key", ["c1_3body", "c3_10body"])
def test_mixed_average_on_error(golden, oracle, key):
    M, m = DECAYS[key]
    dec = make_decay(M, m)
    f = bw_functor(golden["bw3_par"])
    n = 1 << 22
    a64 = average_on(dec, SEED, n, f=f)
    amx = average_on(dec, SEED, n, f=f, prec=MIXED)
    assert abs(amx[0] - a64[0]) <= 1e-13 * a64[0]  # FP64 weights, few-ulp pdk
    assert abs(amx[1] - a64[1]) <= 1e-6 * abs(a64[1]), (amx, a64)


@pytest.mark.gpu
def test_unweight_accept_test_at_the_boundary(oracle):
    """FlagDaugthers' w / max > u (Decays.inl:75-89) on weights placed within
    three ulp of u * max, where only the correctly rounded quotient decides
    (the kernel settles most events with a product by the reciprocal and
    leaves these to the division): the accepted set equals the one numpy's
    IEEE division gives, row for row in the stable partition."""
    M, m = DECAYS["c1_3body"]
    dec = make_decay(M, m)
    n = 1 << 16
    seed = 0x180EC6D33CFD0ABA
    u = oracle.rng_words(0, seed, 0, n).astype(np.float64) * 2.0 ** -32
    maxw = dec.max_weight
    rng = np.random.default_rng(7)
    w = u * maxw
    steps = rng.integers(-3, 4, n)
    for k in range(1, 4):
        up, down = steps >= k, steps <= -k
        w[up] = np.nextafter(w[up], np.inf)
        w[down] = np.nextafter(w[down], -np.inf)
    far = rng.random(n) < 0.25  # and some far from the boundary
    w[far] = rng.random(int(far.sum())) * maxw
    w[::97] = 0.0
    want = (w / maxw) > u
    boundary = ~far
    assert 0 < int(want[boundary].sum()) < int(boundary.sum())  # both outcomes occur
    cols = torch.zeros((12, n), dtype=torch.float64, device="cuda:0")
    nacc, _, idx, _ = unweight(dec, cols, seed, w_in=torch.as_tensor(w, device="cuda:0"))
    assert nacc == int(want.sum())
    got = idx[:nacc].cpu().numpy()
    assert np.array_equal(got, np.nonzero(want)[0])
