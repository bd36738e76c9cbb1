"""The C++ drop-in API (hydra_amd/include/hydra) as a user of the reference
would drive it (tests/cpp/test_api.cpp), checked against the reference's
golden vectors and the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import rows_to_cols

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin", "test_api")
M, MASSES = 5.279, [3.0969, 0.493677, 0.13957]


def run(mode):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (make -C tests/cpp)")
    out = subprocess.run([BIN, mode], check=True, capture_output=True, text=True, timeout=120)
    return json.loads(out.stdout)


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def test_host_engines_constructor_and_seeds(golden):
    r = run("host")
    big = 2**40 + 12345
    pairs = [("rng_squares3_d0", "rng_squares3_d0"), ("rng_squares3_dbig", f"rng_squares3_d{big}"),
             ("rng_squares4_d0", "rng_squares4_d0"), ("rng_philox_d0", "rng_philox_d0"),
             ("rng_philox_long_dbig", f"rng_philox_long_d{big}"),
             ("rng_threefry_d0", "rng_threefry_d0"),
             ("rng_threefry_long_d0", "rng_threefry_long_d0"),
             ("rng_default_d0", "rng_squares3_d0")]
    for ours, ref in pairs:
        assert [int(x) for x in r[ours]] == [int(x) for x in golden[ref]], ours
    assert bits([r["ecm"], r["max_weight"]]).tolist() == bits(golden["params_c1_3body"]).tolist()
    # default seed 1*(1+hash_range(masses)) and GetSeed()'s int truncation
    assert int(r["seed_full"]) == int(golden["seed_c1_3body"][0])
    assert np.int32(r["seed_get"]).astype(np.uint64) == golden["seed_c1_3body"][1]
    assert int(r["copy_seed_full"]) == r["seed_get"] & 0xFFFFFFFFFFFFFFFF
    assert r["err_mass"] == 1 and r["err_energy"] == 1 and r["err_mother"] == 1


@pytest.mark.gpu
def test_gpu_generate_unweight_average(golden, oracle):
    r = run("gpu")
    ref = golden["decay_c1_3body_squares3_p"][:16]
    got = np.array(r["gen_rows16"]).reshape(16, 3, 4)
    E = np.abs(ref[..., 0:1])
    assert np.max(np.abs(got - ref) / E) <= 1e-12
    w = np.array(r["gen_weights"])
    assert w.shape == (4096,)
    # PhaseSpaceWeight recomputed from momenta that differ by ulps (sin/cos):
    # bounded by the conditioning the reference itself shows between its
    # recomputed and its generation weights (test_oracle pins that below 1e-10)
    assert np.max(np.abs(w - golden["gen_c1_w"]) / golden["gen_c1_w"]) <= 1e-10
    pi7 = np.array(r["pion_event7"])
    assert np.max(np.abs(pi7 - golden["gen_c1_p"][7, 2]) / pi7[0]) <= 1e-12
    assert r["unweight_nacc"] == int(golden["uw_c1_nacc"][0])
    first = np.array(r["unweight_first_A"]).reshape(4, 4)
    refA = golden["uw_c1_p"][:4, 0]
    assert np.max(np.abs(first - refA) / refA[:, 0:1]) <= 1e-12
    assert r["unweight_f_nacc"] == int(golden["uwf_c1_nacc"][0])
    # Unweight() regenerates a freshly generated container and moves one whose
    # columns were handed out for writing: the same container either way
    assert r["gen_flags"] == [1, 0] and r["unweight_paths_same"] == 1
    # momenta written through a column pointer: the kept weights are dropped
    # and Unweight() flags with the weights of the new momenta
    assert r["unweight_after_write"] == [0, 1]
    # 1000 events: columns 1024 apart in the container, the bits of the C
    # ABI on columns 1000 apart (Generate, and Generate + Unweight())
    assert r["padded_stride"] == [1, 1, 1]
    # opt-in check of the deferred path: a stale-pointer write is caught
    assert r["deferred_checks"] == [1, 1, 1]
    # Generate + Unweight() cycles on one container reuse its buffers and give
    # a fresh container's rows and weights each time
    assert r["unweight_cycles"] == [1, 1]
    # Generate writes the weights and defers the momenta to their first read
    assert r["lazy_after_generate"] == 1 and r["lazy_after_read"] == 0
    assert r["lazy_same"] == 1
    # once read after a deferred Generate, a container is generated eagerly
    # (momenta and weights in one pass): same bits, Unweight() still regenerates
    assert r["eager_after_read"] == [1, 0, 1, 1, 1]
    for key, gkey in (("avg_bw3", "avg_c1_bw3"), ("integrate_m2", "avg_c1_m2")):
        mean, sig = r[key]
        gm, gs = golden[gkey]
        assert abs(mean - gm) <= 1e-12 * abs(gm), key
        assert abs(sig - gs) <= 1e-10 * abs(gs), key
    assert r["avg_one"] == [1.0, 0.0]
    # Generate(mothers, daughters) against the oracle
    k = np.arange(256)
    px, py, pz = 0.01 * k, -0.02 * k, 0.005 * k
    mothers = np.stack([np.sqrt(M * M + px * px + py * py + pz * pz), px, py, pz])
    cols, _ = oracle.generate_mothers(M, MASSES, 159753, mothers)
    exp = cols.reshape(3, 4, 256)[2][:, [0, 100, 255]].T
    got = np.array(r["mothers_rows_C"]).reshape(3, 4)
    assert np.max(np.abs(got - exp) / exp[:, 0:1]) <= 1e-12
    # Evaluate(mother, dataset, M13^2, M23^2, cos(theta_K*)) against the oracle,
    # which agrees bit for bit with the reference's PhaseSpace::Evaluate
    ev = oracle.evaluate(M, MASSES, 159753, 4096, [(3, (0, 2)), (3, (1, 2)), (4, (1, 2, 0, 3))])
    got = np.array(r["evaluate_rows"]).reshape(4, 4).T
    exp = ev[:, [0, 1, 2, 4095]]
    assert np.array_equal(got[0], exp[0])  # generation weights: bit-exact
    assert np.max(np.abs(got[1:3] - exp[1:3]) / (M * M)) <= 1e-12
    assert np.max(np.abs(got[3] - exp[3])) <= 1e-9
    # Evaluate over the 256 mothers: functors see (mother, A, B, C)
    em = oracle.evaluate_mothers(MASSES, 159753, mothers, [(3, (1, 3))])
    got = np.array(r["evaluate_mothers_rows"]).reshape(3, 2).T
    exp = em[:, [0, 100, 255]]
    assert np.array_equal(got[0], exp[0])
    assert np.max(np.abs(got[1] - exp[1]) / (M * M)) <= 1e-12
    # Dalitz DenseHistogram (phsp_basic.inl:196-214) of the same 4096 events:
    # bins as the reference's (events sit far from bin edges compared with
    # the 1e-12 momentum tolerance), weights as gen_weights above
    assert np.array_equal(np.array(r["hist_dalitz_n"]), golden["hist_dalitz_n"])
    hw = golden["hist_dalitz_w"]
    for key in ("hist_dalitz_w", "hist_dalitz_phsp"):
        got = np.array(r[key])
        assert np.all(np.abs(got - hw) <= 1e-10 * hw), key
    assert r["hist_bin_3_7"] == golden["hist_dalitz_n"][3 * 20 + 7]
    # hydra::unweight(device::sys, events, events.GetEventWeightFunctor(bw)):
    # the oracle (the reference's order bit for bit) on the reference's events
    from oracle import rows_to_cols
    gcols = rows_to_cols(golden["gen_c1_p"])
    wf = oracle.event_weight(M, MASSES, gcols, fid=1, fpar=golden["bw3_par"])
    k, out, _ = oracle.unweight_sample(gcols, wf, -1.0, 0x8EC74D321E6B5A27, 0, order=1)
    assert r["free_unweight_nacc"] == k
    got = np.array(r["free_unweight_rows_C"]).reshape(4, 4)
    exp = out.reshape(3, 4, 4096)[2][:, [0, 1, k, 4095]].T
    assert np.max(np.abs(got - exp) / exp[:, 0:1]) <= 1e-12
    # device -> host -> device copies keep every value
    assert r["host_copy_same"] == 1
    w = np.array(r["gen_weights"])
    assert r["host_copy_weights3"] == [w[0], w[17], w[4095]]


@pytest.mark.gpu
def test_gpu_phase_space_range(golden):
    """phase_space_range as examples/misc/range_semantics.inl uses it: the lazy
    range histogrammed (counts exact, weighted sums to rounding), single
    elements and a copy into a Decays container (weights bit for bit, momenta
    within 1e-12 E), an unbounded range refused by the fill."""
    r = run("gpu")
    assert np.array_equal(np.array(r["range_hist_n"]), golden["range_hist_k3pi_w0"])
    hw = golden["range_hist_k3pi_w1"]
    got = np.array(r["range_hist_w"])
    assert np.all(np.abs(got - hw) <= 1e-13 * np.maximum(hw, 1.0))
    gw, gp = golden["range_k3pi_w"], golden["range_k3pi_p"]
    el = np.array(r["range_elements4"]).reshape(4, 13)
    assert np.array_equal(bits(el[:, 0]), bits(gw[:4]))
    E = np.abs(gp[:4, :, 0:1])
    assert np.max(np.abs(el[:, 1:].reshape(4, 3, 4) - gp[:4]) / E) <= 1e-12
    rows = np.array(r["range_copy_rows"]).reshape(512, 3, 4)
    assert np.max(np.abs(rows - gp) / np.abs(gp[..., 0:1])) <= 1e-12
    assert np.array_equal(bits(r["range_copy_w"]), bits(gw))
    assert r["range_err_unbounded"] == 1


@pytest.mark.gpu
def test_gpu_decay_chain(golden, oracle):
    """phsp_chain.inl end to end through the C++ API: Generate, Generate over
    GetDaugtherRange(_0), Meld, the example's variables copied into a
    multivector, the weights, and a weighted histogram of the two angles."""
    r = run("gpu")
    gp = golden["chain_p"]
    m = np.array(r["chain_melded_0_300"]).reshape(2, 5, 4)
    exp = gp[[0, 300]]
    assert np.max(np.abs(m - exp) / np.abs(exp[..., 0:1])) <= 1e-12
    v = np.array(r["chain_vars"]).reshape(512, 4).T
    ev = golden["chain_vars"]
    M2 = 5.27955 ** 2
    assert np.max(np.abs(v[0] - ev[0])) <= 1e-11 * M2
    assert np.max(np.abs(v[1] - ev[1])) <= 1e-11 * M2
    assert np.max(np.abs(v[2] - ev[2])) <= 1e-9
    d = np.abs(v[3] - ev[3])
    assert np.max(np.minimum(d, 2 * np.pi - d)) <= 1e-9
    assert r["chain_row7"] == list(v[:, 7])
    w = np.array(r["chain_weights"])
    assert np.max(np.abs(w - golden["chain_w"]) / golden["chain_w"]) <= 1e-10
    h = np.array(r["chain_hist_w"])
    grid, lo, hi = [20, 16], [-1.0, 0.0], [1.0, 6.283185307179586]
    exp_h = oracle.histogram_fill(grid, lo, hi, v[2:4], w)
    scale = oracle.histogram_fill(grid, lo, hi, v[2:4], np.abs(w))
    assert np.all(np.abs(h - exp_h) <= 1e-13 * scale + 1e-300)
