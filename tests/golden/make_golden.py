"""Write the golden vectors of the phase-space path from the reference itself.

Runs the reference harness (oracle/_ref/libhydra_ref.so, compiled from the
reference headers by oracle/Makefile) and stores inputs and outputs as small
.npz fixtures next to this script.  Only runs where the reference is mounted
and the harness is built:

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from oracle import ENGINES, MINSTD, Reference  # noqa: E402

# evaluate outputs are (1 + nf, n): row 0 the weight, row 1 + j functor j

# Decays of the workloads (3-body B -> J/psi K pi; 10-body flat; near-threshold
# 10-body; 2-body; 12-body with unequal masses).
DECAYS = {
    "c1_3body": (5.279, [3.0969, 0.493677, 0.13957]),
    "c3_10body": (10.0, [0.13957] * 10),
    "c3_10body_threshold": (10.0, [0.999] * 10),
    "n2": (3.0, [1.0, 0.5]),
    "n12": (20.0, [0.1 + 0.1 * i for i in range(12)]),
}
SEED = 159753
UW_SEED = 0x180EC6D33CFD0ABA
UWF_SEED = 0x39ABDC4529B1661C
# Dalitz functor: three relativistic Breit-Wigners (M, Gamma, cr, ci, a, b)
BW3 = [(0.892, 0.05, 0.4967141530, -0.1382643012, 1, 2),
       (1.43, 0.11, 0.6476885381, 1.5230298564, 1, 2),
       (0.775, 0.149, -0.2341533747, -0.2341368981, 0, 2)]


def bw_par(terms):
    flat = [float(len(terms))]
    for t in terms:
        flat.extend(float(x) for x in t)
    return np.array(flat)


def main():
    R = Reference()
    procs = os.cpu_count()
    out = {}

    # engine words: start of the stream and after a large discard
    for name, e in ENGINES.items():
        for disc in (0, 2**40 + 12345):
            out[f"rng_{name}_d{disc}"] = R.rng_words(e, SEED, disc, 64)

    # constructor products and the default seed
    for key, (M, m) in DECAYS.items():
        rc, ecm, mw, seed_full, seed_get = R.params(M, m)
        out[f"params_{key}"] = np.array([ecm, mw])
        out[f"seed_{key}"] = np.array([seed_full, np.uint64(seed_get & 0xFFFFFFFFFFFFFFFF)],
                                      dtype=np.uint64)

    # per-event GENBOD output: first events and sparse high indices, all engines
    evts = np.concatenate([np.arange(256, dtype=np.uint64),
                           np.array([2**20 + 7, 2**26 - 1, 2**27 - 1, 2**30 + 3, 2**40 + 5],
                                    dtype=np.uint64)])
    out["evts"] = evts
    for key, (M, m) in DECAYS.items():
        for name, e in ENGINES.items():
            if e != 0 and key not in ("c1_3body", "c3_10body"):
                continue
            p, wt = R.decay_events(M, m, SEED, evts, engine=e)
            out[f"decay_{key}_{name}_p"] = p
            out[f"decay_{key}_{name}_wt"] = wt
    # a moving mother (the single-mother boost as the reference builds it)
    mom = [6.0, 0.3, -0.2, 0.5]
    p, wt = R.decay_events(5.279, DECAYS["c1_3body"][1], SEED, evts[:64], mother4=mom)
    out["decay_moving_p"] = p
    out["decay_moving_wt"] = wt

    # Generate into Decays + GetEventWeightFunctor, 4096 events
    M, m = DECAYS["c1_3body"]
    p, w, _ = R.generate(M, m, SEED, 4096)
    out["gen_c1_p"] = p
    out["gen_c1_w"] = w
    M10, m10 = DECAYS["c3_10body_threshold"]
    p, w, _ = R.generate(M10, m10, SEED, 1024)
    out["gen_thr_p"] = p
    out["gen_thr_w"] = w
    # default seed (no SetSeed)
    p, w, _ = R.generate(M, m, None, 64)
    out["gen_c1_defaultseed_p"] = p

    # Generate(mothers, daughters)
    rng = np.random.default_rng(7)
    mothers = np.zeros((4, 256))
    pm = rng.normal(0.0, 2.0, size=(3, 256))
    mothers[1:] = pm
    mothers[0] = np.sqrt(M * M + (pm ** 2).sum(axis=0))
    p, w = R.generate_mothers(M, m, SEED, mothers)
    out["genm_mothers"] = mothers
    out["genm_p"] = p
    out["genm_w"] = w

    # Unweight(): plain and with a functor (f = (p0+p1)^2), 4096 events
    nacc, p, idx, _ = R.unweight(M, m, SEED, 4096, UW_SEED)
    out["uw_c1_nacc"] = np.array([nacc], dtype=np.uint64)
    out["uw_c1_idx"] = idx
    out["uw_c1_p"] = p
    nacc, p, idx, _ = R.unweight(M, m, SEED, 4096, UWF_SEED, fid=2, fpar=np.zeros(1),
                                 max_weight=-1.0)
    out["uwf_c1_nacc"] = np.array([nacc], dtype=np.uint64)
    out["uwf_c1_idx"] = idx
    nacc, p, idx, _ = R.unweight(M, m, SEED, 4096, UWF_SEED, fid=1, fpar=bw_par(BW3),
                                 max_weight=40.0)
    out["uwbw_c1_nacc"] = np.array([nacc], dtype=np.uint64)
    out["uwbw_c1_idx"] = idx

    # AverageOn (OMP fold over omp_get_num_procs() intervals)
    for fid, fp, tag in ((0, np.zeros(1), "one"), (2, np.zeros(1), "m2"),
                         (1, bw_par(BW3), "bw3")):
        mean, sig, _ = R.average_on(M, m, SEED, 100000, fid=fid, fpar=fp)
        out[f"avg_c1_{tag}"] = np.array([mean, sig])
    mean, sig, _ = R.average_on(M10, m10, SEED, 20000, fid=2, fpar=np.zeros(1))
    out["avg_thr_m2"] = np.array([mean, sig])
    out["avg_procs"] = np.array([procs])
    out["bw3_par"] = bw_par(BW3)

    # PhaseSpace::Evaluate(mother, result, M13^2, M23^2) and with the helicity
    # cosine of phsp_evaluating_functor.inl (fid 3: (p_a + p_b)^2, fid 4:
    # helicity cosine, par = {d, e, first, count}); 1024 events
    ev_m2 = [(3, (0, 2)), (3, (1, 2))]
    ev_cos = [(4, (1, 2, 0, 3)), (1, BW3)]
    out["eval_c1_m2"], _ = R.evaluate(M, m, SEED, 1024, ev_m2)
    out["eval_c1_cos_bw"], _ = R.evaluate(M, m, SEED, 1024, ev_cos)
    # Evaluate(mothers, result, ...): functors see (mother, daughters...)
    out["evalm_c1_m2"] = R.evaluate_mothers(M, m, SEED, mothers, [(3, (1, 3)), (3, (0, 1))])
    out["evalm_c1_cos"] = R.evaluate_mothers(M, m, SEED, mothers, [(4, (2, 3, 1, 3))])
    p5 = DECAYS["n12"]
    out["eval_n12_m2"], _ = R.evaluate(p5[0], p5[1], SEED, 256, [(3, (4, 11)), (4, (3, 7, 0, 12))])

    # DenseHistogram fills of the Dalitz variables (m^2_AB, m^2_BC) of the 4096
    # generated events, weighted by PhaseSpaceWeight and unweighted, and a
    # one-dimensional fill with the edge cases (upper limit, just below the
    # lower limit, far outside)
    pg = out["gen_c1_p"]
    sab, sbc = pg[:, 0] + pg[:, 1], pg[:, 1] + pg[:, 2]

    def m2(v):
        return v[:, 0] * v[:, 0] - v[:, 1] * v[:, 1] - v[:, 2] * v[:, 2] - v[:, 3] * v[:, 3]
    dal = np.vstack([m2(sab), m2(sbc)])
    lo = [(m[0] + m[1]) ** 2, (m[1] + m[2]) ** 2]
    hi = [(M - m[2]) ** 2, (M - m[0]) ** 2]
    out["hist_dalitz_x"] = dal
    out["hist_dalitz_limits"] = np.array([lo, hi])
    out["hist_dalitz_w"], _ = R.histogram_fill([20, 20], lo, hi, dal, out["gen_c1_w"])
    out["hist_dalitz_n"], _ = R.histogram_fill([20, 20], lo, hi, dal)
    x1 = np.concatenate([dal[0, :1000], [hi[0], lo[0] - 1e-9, lo[0] - 100.0, hi[0] + 5.0]])
    out["hist_1d_x"] = x1
    out["hist_1d_n"], _ = R.histogram_fill([25], [lo[0]], [hi[0]], x1[None, :])

    # hydra::unweight(device::sys, flat, gauss, max, seed, jump) as
    # examples/random/unweight_sample.inl:118-158 runs it (flat sample in
    # [0, 20), Gaussian(10, 1)): the reordered sample and the weights
    xs = np.random.default_rng(11).uniform(0.0, 20.0, 5000)
    out["uws_x"] = xs
    for tag, mp, jump in (("auto", -1.0, 0), ("max", 0.5, 10000)):
        kept, xo, wg = R.unweight_sample(xs, 10.0, 1.0, mp, 0x8EC74D321E6B5A27, jump)
        out[f"uws_{tag}_kept"] = np.array([kept], dtype=np.uint64)
        out[f"uws_{tag}_x"] = xo
        out["uws_w"] = wg

    # phase_space_range: minstd_rand words (a seed past 2^32 keeps its low
    # 32 bits), and range elements read through the range's own iterator
    for seed in (321, 2**32 + 321, 0):
        for disc in (0, 2**40 + 12345):
            out[f"rng_minstd_s{seed}_d{disc}"] = R.rng_words(MINSTD, seed, disc, 64)
    RANGES = {
        # K+ -> pi+ pi+ pi- at rest, seed 321 (examples/misc/range_semantics.inl:154-156)
        "k3pi": ([0.493677, 0.0, 0.0, 0.0], [0.13957061] * 3, 321, 0),
        # a moving B with the J/psi K pi masses, a seed past 2^32, an offset
        "c1_moving": ([6.0, 0.3, -0.2, 0.5], DECAYS["c1_3body"][1], 2**40 + 987654321,
                      2**33 + 5),
        "n12": ([20.0, 0.0, 0.0, 0.0], DECAYS["n12"][1], 159753, 0),
        "n2": ([3.2, 0.0, 0.4, -0.9], DECAYS["n2"][1], 77, 1000),
    }
    for key, (m4, ms, seed, first) in RANGES.items():
        w, p = R.phase_space_range(m4, ms, seed, 512, first=first)
        out[f"range_{key}_in"] = np.array(list(m4) + [float(seed), float(first)])
        out[f"range_{key}_seed"] = np.array([seed, first], dtype=np.uint64)
        out[f"range_{key}_w"] = w
        out[f"range_{key}_p"] = p
    # the range histogram of range_semantics.inl:158-167: m(pi pi) of the first
    # two daughters in 10 bins over [2 m_pi, m_K - m_pi], 100000 elements,
    # unweighted and weighted by the element's weight
    mk, mpi = 0.493677, 0.13957061
    for wtd in (0, 1):
        c, _ = R.range_histogram3([mk, 0, 0, 0], [mpi] * 3, 321, 100000, 0, 1, 10, 2 * mpi,
                                  mk - mpi, weighted=bool(wtd))
        out[f"range_hist_k3pi_w{wtd}"] = c

    # the decay chain of examples/phase_space/phsp_chain.inl:216-244: melded
    # rows (J/psi, K, pi, mu+, mu-), the example's four variables (M^2(J/psi
    # pi), M^2(K pi), cos theta_K* within the B0, the decay-plane angle), the
    # three-body weights, and the default seeds of both PhaseSpace objects
    CHAIN_VARS = [(3, (0, 2)), (3, (1, 2)), (4, (1, 2, 0, 3)), (6, (1, 2, 3))]
    vals = []
    for fv in CHAIN_VARS:
        p, f, w, _ = R.chain(512, fv)
        vals.append(f)
    out["chain_p"] = p
    out["chain_vars"] = np.array(vals)
    out["chain_w"] = w
    seeds = []
    for M, m in ((5.27955, [3.0969, 0.493677, 0.13957061]), (3.0969, [0.1056583745] * 2)):
        seeds.append(R.params(M, m)[3])
    out["chain_seeds"] = np.array(seeds, dtype=np.uint64)

    np.savez_compressed(os.path.join(HERE, "phsp_golden.npz"), **out)
    print("wrote", os.path.join(HERE, "phsp_golden.npz"), len(out), "arrays")


if __name__ == "__main__":
    main()
