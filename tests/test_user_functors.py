"""User functors compiled into the phase-space kernels by hipcc
(tests/cpp/test_user_functors.hip: the reference examples' lambdas through
PhaseSpace::AverageOn/Evaluate, Decays::Unweight(functor),
GetEventWeightFunctor(functor) and `events | lambda`).

Built with contraction off, every functor value equals the described
functor's bit for bit where the described path divides exactly (Evaluate,
Unweight(functor), PhaseSpaceReweight), and the generation is the library's
own; AverageOn, whose described path divides by reciprocals, agrees within
1e-13.  Both agree with the oracle (Evaluate within the momentum tolerance,
the accepted count and index set exactly).  Built with hipcc's default
contraction, the generation weights keep their bits (the library's code turns
contraction off itself) and the functor values agree within 1e-12."""
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import cols_to_rows

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "bin")
UWF_SEED = 0x39ABDC4529B1661C
BW3 = [(0.892, 0.05, 0.4967141530, -0.1382643012, 1, 2),
       (1.43, 0.11, 0.6476885381, 1.5230298564, 1, 2),
       (0.775, 0.149, -0.2341533747, -0.2341368981, 0, 2)]


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


_cache = {}


def run(kind):
    if kind not in _cache:
        path = os.path.join(BIN, f"test_user_functors_{kind}")
        if not os.path.exists(path):
            pytest.fail(f"{path} not built (make -C tests/cpp)")
        print(f"running {path}", flush=True)
        out = subprocess.run([path], check=True, capture_output=True, text=True, timeout=300)
        _cache[kind] = json.loads(out.stdout)
    return _cache[kind]


def bw_par():
    flat = [float(len(BW3))]
    for t in BW3:
        flat.extend(float(x) for x in t)
    return np.array(flat)


@pytest.mark.gpu
def test_average_on_cos_theta_lambda(oracle):
    r = run("exact")
    (mu, sig), (dmu, dsig) = r["avg_cos_user"], r["avg_cos_desc"]
    # the described AverageOn divides through reciprocals: within 1e-13
    assert abs(mu - dmu) <= 1e-13 and abs(sig - dsig) <= 1e-13 * abs(dsig)
    M, m = 5.27955, [3.0969, 0.493677, 0.13957061]
    st = oracle.average_on(M, m, int(r["avg_cos_seed"]), 1000000, fid=4, fpar=[2, 1, 0, 3],
                           intervals=16)
    assert abs(mu - st[1]) <= 1e-12
    assert abs(sig - np.sqrt(st[2]) / st[0]) <= 1e-12 * abs(sig)


@pytest.mark.gpu
def test_evaluate_lambdas(oracle):
    r = run("exact")
    n = int(r["eval_n"])
    u = np.array(r["eval_user"]).reshape(4, n)
    d = np.array(r["eval_desc"]).reshape(4, n)
    assert np.array_equal(bits(u), bits(d))
    M, m = 5.27955, [3.0969, 0.493677, 0.13957061]
    o = oracle.evaluate(M, m, int(r["eval_seed"]), n,
                        [(3, (0, 2)), (3, (1, 2)), (4, (1, 2, 0, 3))])
    assert np.array_equal(bits(u[0]), bits(o[0]))  # generation weights
    scale = np.array([M * M, M * M, 1.0])[:, None]
    assert np.max(np.abs(u[1:] - o[1:]) / scale) <= 1e-12
    # over mothers: the functor sees the mother first
    nm = len(r["evalm_user"]) // 2
    assert np.array_equal(bits(r["evalm_user"]), bits(r["evalm_desc"])) and nm == 3000
    (a, s), (b, t) = r["avgm_user"], r["avgm_desc"]
    assert abs(a - b) <= 1e-13 * abs(b) and abs(s - t) <= 1e-12 * abs(t)


@pytest.mark.gpu
def test_unweight_lambda(oracle):
    r = run("exact")
    na, nd = r["uw_nacc"]
    assert na == nd and r["uw_same"] == 1
    nfix, same = r["uw_fixed_max"]
    assert same == 1
    n = 50000
    cols = np.array(r["uw_events"]).reshape(12, n)
    onacc, _, oidx, _ = oracle.unweight(5.279, [3.0969, 0.493677, 0.13957], cols, UWF_SEED,
                                        fid=1, fpar=bw_par(), max_value=-1.0)
    assert onacc == na
    onfix = oracle.unweight(5.279, [3.0969, 0.493677, 0.13957], cols, 99, fid=1, fpar=bw_par(),
                            max_value=40.0)[0]
    assert onfix == nfix
    # the unweighted sample through the Dalitz lambda: the variables of the
    # accepted events, in order
    sv = np.array(r["sample_vars"]).reshape(2, -1)
    p = cols_to_rows(cols, 3)[oidx[:onacc].astype(np.int64)]
    def m2(x):
        return x[:, 0] * x[:, 0] - x[:, 1] * x[:, 1] - x[:, 2] * x[:, 2] - x[:, 3] * x[:, 3]
    want = np.stack([m2(p[:, 0] + p[:, 1]), m2(p[:, 1] + p[:, 2])])
    assert sv.shape[1] >= onacc
    assert np.array_equal(bits(sv[:, :onacc]), bits(want))


@pytest.mark.gpu
def test_event_pipes_and_histograms():
    r = run("exact")
    # GetEventWeightFunctor(lambda of (B, C)) selects the container's B and C
    assert np.array_equal(bits(r["res_w_user"]), bits(r["res_w_desc"]))
    hu = np.array(r["hist_user_flat"])
    hd = np.array(r["hist_desc_flat"])
    assert np.allclose(hu, hd, rtol=1e-12, atol=0.0)
    assert np.sum(hu) > 0
    hr = np.array(r["hist_user_res"])
    assert np.all(hr >= 0) and np.sum(hr) > 0


@pytest.mark.gpu
def test_default_contraction_keeps_generation_bits():
    ex, df = run("exact"), run("default")
    n = int(ex["eval_n"])
    u = np.array(df["eval_user"]).reshape(4, n)
    e = np.array(ex["eval_user"]).reshape(4, n)
    assert np.array_equal(bits(u[0]), bits(e[0]))  # weights: the library's code
    assert np.max(np.abs(u[1:3] - e[1:3])) <= 1e-12 * 5.3 ** 2
    assert np.max(np.abs(u[3] - e[3])) <= 1e-12
    assert df["uw_nacc"] == ex["uw_nacc"]
    assert np.array_equal(bits(df["uw_events"]), bits(ex["uw_events"]))
