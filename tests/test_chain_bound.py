"""The bound behind hpsp_generate_unweight's flags (generate_unweight.hip),
restated in numpy and checked against the reference's arithmetic as the
oracle carries it out: for each event, the weight chain wt (DecayMother.h:
170-212, from the event's engine words) and PhaseSpaceWeight w recomputed
from the oracle's momenta (Decays.inl:215-231) must satisfy
|w - wt| <= B wt wherever the kernel would trust B (d1, d2 > 0 at every step,
B < 2^-10).  B = sum_i (ka + kc) ((2K + 4) eps S^2 (1/a^2 + 1/c^2) + 4 eps)
+ 40 N eps with K = 2^17 N for the squares engines and, for the Random123
ones, K = N (64 + 4 / sin_min) from the event's own polar angles.

The decays include the headline decay, decays near threshold (where the pdk
terms are ill-conditioned), electrons beside a 90 GeV particle (where no
event is trusted: every one takes the exact test) and 2-, 5-, 7-, 10- and
12-body decays, on every engine.  (A moving mother is not among them: the
reference's mother velocity, DecayMother.h:76-101, has |beta| >= 1 for any
moving mother, so its momenta and weights are NaN and no event is trusted.)
Two further checks keep the test from passing for the wrong reason: the
chain restated here reproduces the oracle's weights (the word mapping is
right), and the conditioning factors ka + kc are needed (without them some
events near threshold fall outside B)."""
import numpy as np
import pytest

from oracle import ENGINES

EPS = 2.0 ** -53
NARROW = {"squares3", "squares4"}

DECAYS = [
    # (name, M, masses, engine, mother4)
    ("headline", 5.279, [3.0969, 0.493677, 0.13957], "squares3", None),
    ("threshold3", 3.0969 + 0.493677 + 0.13957 + 0.001, [3.0969, 0.493677, 0.13957], "squares3",
     None),
    ("threshold10", 10.0, [0.999] * 10, "squares4", None),
    ("light", 100.0, [0.000511, 0.000511, 90.0], "squares4", None),
    ("two_body", 3.0969, [0.1057, 0.1057], "squares3", None),
    ("philox_threshold", 3.0969 + 0.493677 + 0.13957 + 0.01, [3.0969, 0.493677, 0.13957],
     "philox", None),
    ("philox_long5", 12.0, [0.5, 1.2, 0.8, 2.0, 0.3], "philox_long", None),
    ("threefry12", 20.0, [0.5] * 12, "threefry", None),
    ("threefry_long7", 9.0, [0.13957] * 7, "threefry_long", None),
]

# the fraction of events whose flag the chain decides, at least
MIN_TRUSTED = {"headline": 0.999, "threshold3": 0.99, "threshold10": 0.8, "light": 0.0,
               "two_body": 0.999, "philox_threshold": 0.99, "philox_long5": 0.99,
               "threefry12": 0.99, "threefry_long7": 0.99}


def to_unit(words, engine):
    if engine in NARROW:
        return words.astype(np.float64) / 2.0 ** 32
    return words.astype(np.float64) / 2.0 ** 64


def chain_and_bound(oracle, M, masses, seed, n, engine, mother4, K_fixed=None,
                    conditioning=True):
    """(wt, B, trusted) per event, as k_uw_flags_chain forms them (in plain
    FP64 here, without its fused operations and estimates); conditioning=False
    drops the factors ka + kc (a planted mistake)."""
    m = np.asarray(masses, dtype=np.float64)
    N = len(m)
    rc, ecm, mw = oracle.params(M, m)
    assert rc == 0
    W = 3 * N - 4
    # event k's engine is discarded by k + 3N, then gives its W words
    if engine in NARROW:
        # squares: one word per counter, so event k's words are counters
        # k + 3N + j of one stream
        words = oracle.rng_words(engine, seed, 3 * N, n + W - 1)
        ev = np.lib.stride_tricks.sliding_window_view(words, W)[:n]
    else:
        ev = np.stack([oracle.rng_words(engine, seed, k + 3 * N, W) for k in range(n)])
    rno = np.zeros((n, N))
    rno[:, N - 1] = 1.0
    if N > 2:
        rno[:, 1:N - 1] = to_unit(np.sort(ev[:, :N - 2], axis=1), engine)
    csum = np.cumsum(m)
    inv_mas = rno * ecm + csum
    # scale2 bounds E^2 + |p|^2 of every subsystem: M^2 at rest
    Mt = ecm + csum[-1]
    scale2 = Mt * Mt
    if mother4 is not None:
        beta = oracle.mother_beta(mother4)
        b = float(np.sqrt(np.sum(beta * beta)))
        with np.errstate(invalid="ignore"):
            g = 1.0 / np.sqrt(1.0 - b * b)
        scale2 *= (g * (1.0 + b)) ** 2
    if K_fixed is not None:
        ks = np.full(n, (2.0 * K_fixed + 4.0) * EPS * scale2)
    elif engine in NARROW:
        ks = np.full(n, (2.0 * 2.0 ** 17 * N + 4.0) * EPS * scale2)
    else:
        s2 = np.ones(n)
        for i in range(1, N):
            cz = 2 * to_unit(ev[:, N - 2 + 2 * (i - 1)], engine) - 1
            s2 = np.minimum(s2, (1.0 - cz) * (1.0 + cz))
        inv_s = 1.0 / np.sqrt(np.maximum(s2 - 4 * EPS, 2.0 ** -120))
        ks = EPS * scale2 * (2.0 * N * (64.0 + 4.0 * inv_s) + 4.0)
    wt = np.full(n, mw)
    B = np.full(n, 40.0 * N * EPS)
    ic2 = 1.0 / (m[0] * m[0])
    ok = np.ones(n, dtype=bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        for s in range(N - 1):
            a, b, c = inv_mas[:, s + 1], m[s + 1], inv_mas[:, s]
            # the reference's pdk(invMas[s+1], invMas[s], m[s+1])
            wt = wt * (np.sqrt((a - c - b) * (a + c + b) * (a - c + b) * (a + c - b)) / (2 * a))
            spc, smc = b + c, b - c
            d1 = (a - spc) * (a + spc)
            d2 = (a - smc) * (a + smc)
            ok &= np.minimum(d1, d2) > 0
            a2 = a * a
            kk = (d1 * (a2 + c * np.abs(smc)) + d2 * (a2 + c * spc)) / (d1 * d2) + 1.0
            if not conditioning:
                kk = np.ones(n)
            ia2 = 1.0 / a2
            B = B + kk * (ks * (ia2 + ic2) + 4 * EPS)
            ic2 = ia2
    trusted = ok & (B < 2.0 ** -10)
    return wt, B, trusted


@pytest.mark.parametrize("case", DECAYS, ids=[d[0] for d in DECAYS])
def test_phase_space_weight_within_bound(oracle, case):
    name, M, masses, engine, mother4 = case
    n = 1 << 16
    seed = 0x1234567 + 17 * len(masses)
    cols, wt_ref = oracle.generate(M, masses, seed, n, engine=engine, mother4=mother4)
    w = oracle.event_weight(M, masses, cols)
    wt, B, trusted = chain_and_bound(oracle, M, masses, seed, n, engine, mother4)
    # the restated chain is the oracle's (the words and their order are right)
    np.testing.assert_allclose(wt, wt_ref, rtol=8 * EPS, atol=0)
    assert trusted.mean() >= MIN_TRUSTED[name], (name, trusted.mean())
    if not trusted.any():
        return
    rel = np.abs(w - wt_ref)[trusted] / wt_ref[trusted]
    ratio = rel / B[trusted]
    assert ratio.max() <= 1.0, (name, float(ratio.max()))
    # the margin: within 4% of the bound at K = 10 (the kernel's K is at
    # least 19x that)
    _, B10, t10 = chain_and_bound(oracle, M, masses, seed, n, engine, mother4, K_fixed=10.0)
    sel = trusted & t10
    r10 = (np.abs(w - wt_ref)[sel] / wt_ref[sel]) / B10[sel]
    assert r10.max() <= 0.04, (name, float(r10.max()))


def test_conditioning_is_needed(oracle):
    """Without the conditioning factors ka + kc the bound fails near
    threshold, where a - b - c cancels in pdk: the check has the power to
    catch a dropped term."""
    M, masses = 3.0969 + 0.493677 + 0.13957 + 0.001, [3.0969, 0.493677, 0.13957]
    n = 1 << 16
    cols, wt_ref = oracle.generate(M, masses, 99, n, engine="squares3")
    w = oracle.event_weight(M, masses, cols)
    _, Bp, tp = chain_and_bound(oracle, M, masses, 99, n, "squares3", None, conditioning=False)
    rel = np.abs(w - wt_ref)[tp] / wt_ref[tp]
    assert (rel > Bp[tp]).any()


def test_engine_names_cover_the_kernels():
    assert set(ENGINES) == {"squares3", "squares4", "philox", "philox_long", "threefry",
                            "threefry_long"}


# ---- the 3-body bound as a table (generate_unweight.hip, k_chain_table3) ----
BIN3_SIDE = 8 + 28 * 8
BIN3 = 2 * BIN3_SIDE


def bin3(v):
    """Bin of the mass uniform's leading 32 bits v (uint64 array)."""
    v = np.asarray(v, dtype=np.uint64)
    side = v >> np.uint64(31)
    t = np.where(side == 1, np.uint64(0xFFFFFFFF) - v, v)
    e = np.zeros_like(t, dtype=np.int64)
    big = t >= 8
    e[big] = np.floor(np.log2(t[big].astype(np.float64))).astype(np.int64)
    # correct log2 rounding at powers of two
    e[big] -= (np.uint64(1) << e[big].astype(np.uint64)) > t[big]
    e[big] += (np.uint64(1) << (e[big] + 1).astype(np.uint64)) <= t[big]
    idx = t.astype(np.int64)
    sh = (e[big] - 3).astype(np.uint64)
    idx[big] = (e[big] - 2) * 8 + ((t[big] >> sh) & np.uint64(7)).astype(np.int64)
    return side.astype(np.int64) * BIN3_SIDE + idx


def table3(oracle, M, masses):
    """(X, Y) per bin: B <= ks X + Y over the bin."""
    m0, m1, m2 = (float(x) for x in masses)
    rc, ecm, _ = oracle.params(M, masses)
    X = np.zeros(BIN3)
    Y = np.zeros(BIN3)
    s01 = (0.0 + m0) + m1
    A = 1.0 * ecm + (s01 + m2)
    for i in range(BIN3):
        side, idx = divmod(i, BIN3_SIDE)
        if idx < 8:
            tlo = thi = idx
        else:
            e, mb = idx // 8 + 2, idx % 8
            tlo = (8 + mb) << (e - 3)
            thi = ((9 + mb) << (e - 3)) - 1
        if side == 0:
            rlo, rhi = tlo * 2.0 ** -32, (thi + 1) * 2.0 ** -32
        else:
            rlo, rhi = (0xFFFFFFFF - thi) * 2.0 ** -32, (0x100000000 - tlo) * 2.0 ** -32
        alo, ahi = rlo * ecm + s01, rhi * ecm + s01
        spc0, smc0 = m1 + m0, m1 - m0
        den1, den2 = (alo - spc0) * (alo + spc0), (alo - smc0) * (alo + smc0)
        kk0 = ((ahi * ahi + m0 * abs(smc0)) / den2 + (ahi * ahi + m0 * spc0) / den1 + 1.0
               if den1 > 0 and den2 > 0 else np.inf)
        e1 = (A - (m2 + ahi)) * (A + (m2 + ahi))
        dm = max(abs(m2 - alo), abs(m2 - ahi))
        e2 = (A - dm) * (A + dm)
        kk1 = ((A * A + ahi * dm) / e2 + (A * A + ahi * (m2 + ahi)) / e1 + 1.0
               if e1 > 0 and e2 > 0 else np.inf)
        X[i] = kk0 * (1 / (alo * alo) + 1 / (m0 * m0)) + kk1 * (1 / (A * A) + 1 / (alo * alo))
        Y[i] = 4 * EPS * (kk0 + kk1) + 120 * EPS
    return X * (1 + 2.0 ** -40), Y * (1 + 2.0 ** -40)


THREE_BODY = [("headline", 5.279, [3.0969, 0.493677, 0.13957]),
              ("threshold3", 3.0969 + 0.493677 + 0.13957 + 0.001, [3.0969, 0.493677, 0.13957]),
              ("light", 100.0, [0.000511, 0.000511, 90.0]),
              ("heavy_last", 10.0, [0.1, 0.2, 8.0]),
              ("equal", 3.0, [0.5, 0.5, 0.5])]


def test_bin3_edges():
    v = np.array([0, 1, 7, 8, 15, 16, 2 ** 30, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 8, 2 ** 32 - 1],
                 dtype=np.uint64)
    b = bin3(v)
    assert list(b[:6]) == [0, 1, 7, 8, 15, 16]
    assert b[6] == 28 * 8 and b[7] == 28 * 8 + 7
    assert b[8] == BIN3_SIDE + 28 * 8 + 7 and b[-1] == BIN3_SIDE and b[-2] == BIN3_SIDE + 7
    assert b.min() >= 0 and b.max() < BIN3


@pytest.mark.parametrize("case", THREE_BODY, ids=[c[0] for c in THREE_BODY])
def test_table3_bounds_every_event(oracle, case):
    """Over each bin the table's X and Y bound the per-event terms of B (so
    ks X + Y >= B for any ks >= 0): 2^20 random words plus both ends of
    every bin, for 32- and 64-bit words."""
    name, M, masses = case
    m = np.asarray(masses)
    rc, ecm, _ = oracle.params(M, m)
    X, Y = table3(oracle, M, masses)
    rng = np.random.default_rng(11)
    # 64-bit words: random, and the extreme words of every 32-bit bin
    w = rng.integers(0, 2 ** 64, 1 << 20, dtype=np.uint64)
    edges = []
    for i in range(BIN3):
        side, idx = divmod(i, BIN3_SIDE)
        if idx < 8:
            tlo = thi = idx
        else:
            e, mb = idx // 8 + 2, idx % 8
            tlo, thi = (8 + mb) << (e - 3), ((9 + mb) << (e - 3)) - 1
        vlo, vhi = (tlo, thi) if side == 0 else (0xFFFFFFFF - thi, 0xFFFFFFFF - tlo)
        edges += [vlo << 32, (vhi << 32) | 0xFFFFFFFF, vlo << 32 | 0x80000000]
    w = np.concatenate([w, np.array(edges, dtype=np.uint64)])
    for narrow in (True, False):
        if narrow:
            v = w >> np.uint64(32)
            r = v.astype(np.float64) * 2.0 ** -32
        else:
            v = w >> np.uint64(32)
            r = w.astype(np.float64) * 2.0 ** -64
        b = bin3(v)
        s01 = (0.0 + m[0]) + m[1]
        A = 1.0 * ecm + (s01 + m[2])
        a0 = r * ecm + s01
        with np.errstate(divide="ignore", invalid="ignore"):
            xe = np.zeros_like(r)
            ye = np.full_like(r, 120 * EPS)
            ok = np.ones(r.shape, dtype=bool)
            for a, bb, c, ic2 in ((a0, m[1], m[0], 1 / m[0] ** 2), (np.full_like(r, A), m[2], a0,
                                                                      1 / a0 ** 2)):
                spc, smc = bb + c, bb - c
                d1, d2 = (a - spc) * (a + spc), (a - smc) * (a + smc)
                ok &= np.minimum(d1, d2) > 0
                kk = (a * a + c * np.abs(smc)) / d2 + (a * a + c * spc) / d1 + 1
                xe += kk * (1 / a ** 2 + ic2)
                ye += 4 * EPS * kk
        sel = ok
        assert np.all(X[b[sel]] >= xe[sel] * (1 - 1e-12)), name
        assert np.all(Y[b[sel]] >= ye[sel] * (1 - 1e-12)), name
        # events outside the domain (d1 or d2 not positive) land in bins of
        # infinite bound
        assert np.all(np.isinf(X[b[~ok]])), name
    # the table leaves almost every event to the chain on the headline decay
    if name == "headline":
        ks = (2.0 * 2 ** 17 * 3 + 4) * EPS * (ecm + m.sum()) ** 2
        Bt = ks * X + Y
        frac = np.mean(Bt[bin3(w[:1 << 20] >> np.uint64(32))] >= 2.0 ** -10)
        assert frac < 1e-3, frac


@pytest.mark.parametrize("case", THREE_BODY, ids=[c[0] for c in THREE_BODY])
def test_mixed_chain_within_rounding_bound(oracle, case):
    """The mixed form's flags (k_uw_flags_mixed) trust the squared chain where
    |q^2 - u^2| exceeds its margin with B_m = 4 B(K = 0): the mixed weight
    chain (pdk as the reference orders it, (a - b - c)(a + b + c)(a - b + c)
    (a + b - c), with one-ulp square roots and reciprocals) and the squared
    chain (d1 d2 from b + c and b - c) must both lie within B_m of the exact
    chain for the same invariant masses.  Checked in FP64 with correctly
    rounded square roots and divisions (the device's one-ulp ones add an ulp
    or two, far inside B_m's 480 eps floor) against the chain in 64-bit
    extended precision, on 2^18 random mass uniforms and the ends of [0, 1]."""
    name, M, masses = case
    m = np.asarray(masses, dtype=np.float64)
    rc, ecm, mw = oracle.params(M, m)
    rng = np.random.default_rng(5)
    r = np.concatenate([rng.random(1 << 18), [2.0 ** -30, 1e-9, 1 - 2.0 ** -30, 0.5]])
    s01 = (0.0 + m[0]) + m[1]
    A = 1.0 * ecm + (s01 + m[2])
    a0 = r * ecm + s01
    with np.errstate(divide="ignore", invalid="ignore"):
        # the mixed chain: pdk(invMas[n+1], invMas[n], m[n+1]) in the reference's order
        def pdk_ref(a, b, c):
            return np.sqrt((a - b - c) * (a + b + c) * (a - b + c) * (a + b - c)) / (2 * a)
        wt_m = mw * pdk_ref(a0, m[0], m[1]) * pdk_ref(A, a0, m[2])
        # the squared chain (chain3_q2): q^2 max^2 = maxw^2 x0 x1 / (4 a0^2 A^2)
        spc0, smc0 = m[1] + m[0], m[1] - m[0]
        x0 = ((a0 - spc0) * (a0 + spc0)) * ((a0 - smc0) * (a0 + smc0))
        spc1, smc1 = m[2] + a0, m[2] - a0
        x1 = ((A - spc1) * (A + spc1)) * ((A - smc1) * (A + smc1))
        wt_q = np.sqrt(mw * mw / 16.0 / (A * A) * (x0 * x1) / (a0 * a0))
        # the exact chain in extended precision
        L = np.longdouble
        aL, AL = a0.astype(L), L(A)
        def pdk_L(a, b, c):
            b, c = L(b) if np.isscalar(b) else b, L(c)
            return np.sqrt((a * a - (b + c) ** 2) * (a * a - (b - c) ** 2)) / (2 * a)
        wt_x = L(mw) * pdk_L(aL, m[1], m[0]) * pdk_L(AL, m[2], aL)
        # B(K = 0) per event, as chain_q2 forms it without the momenta's term
        B = np.full(r.shape, 120 * EPS)
        ok = np.ones(r.shape, dtype=bool)
        for a, b, c in ((a0, m[1], np.full_like(r, m[0])), (np.full_like(r, A), m[2], a0)):
            spc, smc = b + c, b - c
            d1, d2 = (a - spc) * (a + spc), (a - smc) * (a + smc)
            ok &= np.minimum(d1, d2) > 0
            kk = (a * a + c * np.abs(smc)) / d2 + (a * a + c * spc) / d1 + 1
            B += 4 * EPS * kk
    Bm = 4 * B
    sel = ok & (Bm < 2.0 ** -10) & (wt_x > 0)
    assert sel.mean() > 0.5, name
    em = np.abs((wt_m[sel].astype(L) - wt_x[sel]) / wt_x[sel]).astype(np.float64)
    eq = np.abs((wt_q[sel].astype(L) - wt_x[sel]) / wt_x[sel]).astype(np.float64)
    assert np.all(em + eq <= Bm[sel]), (name, float(np.max((em + eq) / Bm[sel])))
