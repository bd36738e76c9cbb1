#!/usr/bin/env python3
"""Phase-space benchmark on one MI355X.

Headline workload (one "step"): the 3-body decay M = 5.279 GeV ->
(3.0969, 0.493677, 0.13957) GeV, FP64, 2^27 events generated into SoA
four-momentum columns (with the PhaseSpaceWeight column the Decays
container keeps), then Decays::Unweight() with the reference's default seed
(accept-reject, stable partition written out of place).  value = events per
second of the whole step; inputs never leave HBM.

Also measured (reported under "extras", not the headline): the 10-body
decay at 2^26 events in FP64 and mixed precision (plus the near-threshold
variant), and AverageOn of the three-Breit-Wigner Dalitz functor over 2^30
events in 2^24-event chunks, FP64 and mixed.

Kernel durations come from HIP events recorded on the stream the kernels are
launched on (torch's current stream, handed to the C ABI).  The CPU baseline
is the reference's own CPU path (oracle/_ref, OMP backend) on a bounded
sample of the same workload, or the oracle's C restatement when the
reference build is absent.
"""
import argparse
import ctypes
import json
import math
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3

M3 = 5.279
MASSES3 = [3.0969, 0.493677, 0.13957]
SEED = 159753
UW_SEED = 0x180EC6D33CFD0ABA
BW3 = [(0.892, 0.05, 0.4967141530, -0.1382643012, 1, 2),
       (1.43, 0.11, 0.6476885381, 1.5230298564, 1, 2),
       (0.775, 0.149, -0.2341533747, -0.2341368981, 0, 2)]


def genbod_flops(N):
    """Algorithmic FP64 operations of one GENBOD event (add/sub/mul/div/sqrt
    each one; sin and cos not counted), from the reference's op sequence."""
    s1 = sum(11 + 12 * (i + 1) for i in range(1, N))
    s2 = sum(5 + 53 * (i + 1) for i in range(1, N - 1))
    return 3 * (3 * N - 4) + 3 * N + 15 * (N - 1) + 4 + s1 + s2 + 5 * N


def weight_flops(N):
    """Algorithmic FP64 operations of PhaseSpaceWeight::Evaluate for one event
    (Decays.inl:215-231): per running sum, R + p_i (4), two masses (4 mul,
    3 sub, sqrt each; Vector4R.inl:196-208), pdk (Decays.inl:245-248: 8
    add/sub, 3 mul, sqrt, 2*a, div), w *= and R += p_i (5): 39."""
    return 39 * (N - 1)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 27, help="events per step (headline)")
    ap.add_argument("--extras", type=int, default=1, help="also time the other workloads")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--only", default="",
                    help="run one workload: c2, c2_cpp (the step through the C++ API), c2_mixed, c3, c4, evaluate, histogram, range, c5 (serving), or copy (a plain 1 GiB device copy "
                         "that calibrates the profiler's byte counters)")
    return ap.parse_args()


class Timer:
    """HIP events on torch's current stream (where the C ABI launches)."""

    def __init__(self, torch):
        self.torch = torch
        self.a = torch.cuda.Event(enable_timing=True)
        self.b = torch.cuda.Event(enable_timing=True)
        self.total = 0.0
        self.count = 0

    def start(self):
        self.a.record()

    def stop(self):
        self.b.record()
        self.b.synchronize()
        self.total += self.a.elapsed_time(self.b)
        self.count += 1

    @property
    def avg_ms(self):
        return self.total / max(self.count, 1)


def run_c2(torch, capi, n, steps, warmup):
    """The headline step as the C++ drop-in runs `phsp.Generate(B0, Events);
    Events.Unweight()`: Generate records (decay, seed) and forms nothing;
    Unweight (hpsp_generate_unweight) takes each event's flag from its weight
    chain (PhaseSpaceWeight only where the chain leaves the test in doubt),
    scans the counts and generates every event once, at its place in the
    partition.  Timed beside it, with the same final container bit for bit:
    the two-pass form (Generate writing the PhaseSpaceWeight column, then
    hpsp_unweight_regenerate), Generate writing the momenta then the same
    Unweight ("eager"), and then the moving Unweight (hpsp_unweight: every
    event read and moved)."""
    lib = capi.lib()
    dec = capi.make_decay(M3, MASSES3)
    N = 3
    dev = torch.device("cuda:0")
    cols = torch.empty((4 * N, n), dtype=torch.float64, device=dev)
    evw = torch.empty(n, dtype=torch.float64, device=dev)
    out = torch.empty((4 * N, n), dtype=torch.float64, device=dev)
    out2 = torch.empty((4 * N, n), dtype=torch.float64, device=dev)
    ws_bytes = max(lib.hpsp_unweight_workspace_size(n), lib.hpsp_generate_unweight_workspace_size(n))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    nacc = torch.zeros(1, dtype=torch.int64, device=dev)
    nacc2 = torch.zeros(1, dtype=torch.int64, device=dev)
    sh = capi.stream_handle()
    T = {k: Timer(torch) for k in ("weights", "gen", "uw", "flags", "move", "cols", "step",
                                   "chain", "store")}

    def step(with_output=True):
        capi.check(lib.hpsp_generate_unweight(ctypes.byref(dec), SEED, 0, n, UW_SEED,
                                              capi.ptr(out2) if with_output else None, n, None,
                                              None, capi.ptr(ws), ws_bytes, capi.ptr(nacc2), sh),
                   "generate_unweight")

    def gen():
        capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, 0, n, capi.FP64, capi.ptr(cols), n,
                                     0, None, capi.ptr(evw), sh), "generate")

    def gen_cols():
        # the momenta alone: what the first read of a deferred container costs
        capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, 0, n, capi.FP64, capi.ptr(cols), n,
                                     0, None, None, sh), "generate (momenta)")

    def gen_weights():
        capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, 0, n, capi.FP64, None, n, 0, None,
                                     capi.ptr(evw), sh), "generate (weights)")

    def unweight(with_output=True):
        capi.check(lib.hpsp_unweight_regenerate(ctypes.byref(dec), SEED, 0, n, capi.ptr(evw),
                                                UW_SEED, capi.ptr(out) if with_output else None,
                                                n, None, None, capi.ptr(ws), ws_bytes,
                                                capi.ptr(nacc), sh), "unweight_regenerate")

    def unweight_move():
        capi.check(lib.hpsp_unweight(ctypes.byref(dec), None, -1.0, UW_SEED, capi.ptr(cols), n,
                                     n, capi.ptr(evw), capi.ptr(out), n, None, None,
                                     capi.ptr(ws), ws_bytes, capi.ptr(nacc), sh), "unweight")

    def timed(key, fn, *args):
        T[key].start()
        fn(*args)
        T[key].stop()

    for _ in range(warmup):
        step()
        gen_weights()
        unweight()
    torch.cuda.synchronize()
    # kernel-level timing (separate pass; events synchronise).  With no
    # outputs requested only the flag and scan kernels run, so the
    # regenerating scatter's time is the full call minus that.
    for _ in range(max(3, min(steps, 10))):
        timed("step", step)
        timed("chain", step, False)
        timed("weights", gen_weights)
        # a plain store of the step's 12.9 GB (torch's fill of 12 columns):
        # what HBM takes for the regenerating scatter's writes alone
        timed("store", cols.fill_, 0.5)
        timed("gen", gen)
        timed("uw", unweight)
        timed("flags", unweight, False)
        timed("move", unweight_move)
        timed("cols", gen_cols)
    torch.cuda.synchronize()
    same = bool(torch.equal(out, out2)) and int(nacc.item()) == int(nacc2.item())
    # groups of 64 events with events the chain left to the exact test: the
    # counter of hpsp_generate_unweight's list, which follows hpsp_unweight's
    # part of the workspace (capi.hip)
    step()
    off = lib.hpsp_unweight_workspace_size(n)
    doubt_groups = int(ws[off:off + 8].cpu().view(torch.int64).item())
    # the timed steps, right after their warm-up steps
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1000.0 / steps
    for _ in range(warmup):
        gen_weights()
        unweight()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(steps):
        gen_weights()
        unweight()
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    ms_two = (t3 - t2) * 1000.0 / steps
    ms_of = {k: t.avg_ms for k, t in T.items()}
    regen_ms = ms_of["step"] - ms_of["chain"]
    regen2_ms = ms_of["uw"] - ms_of["flags"]
    move_ms = ms_of["move"] - ms_of["flags"]
    res = {"ms_per_step": ms, "events_per_s": n / (ms / 1000.0), "n_acc": int(nacc2.item()),
           "same_as_two_pass": same, "doubtful_groups": doubt_groups,
           "chain_flags_scan_ms": ms_of["chain"], "regen_scatter_ms": regen_ms,
           "two_pass_step_ms": ms_two, "generate_weights_ms": ms_of["weights"],
           "flags_scan_ms": ms_of["flags"], "two_pass_regen_scatter_ms": regen2_ms,
           "generate_ms": ms_of["gen"], "eager_step_ms": ms_of["gen"] + ms_of["uw"],
           "move_unweight_ms": ms_of["move"], "move_scatter_ms": move_ms,
           "move_step_ms": ms_of["gen"] + ms_of["move"],
           "generate_momenta_only_ms": ms_of["cols"], "plain_store_ms": ms_of["store"]}
    gen_bytes = n * (32 * N + 8)
    res["generate_weights_GBs"] = n * 8 / (ms_of["weights"] * 1e6)
    regen_bytes = n * 32 * N
    move_bytes = n * 64 * N
    res["generate_GBs"] = gen_bytes / (ms_of["gen"] * 1e6)
    res["regen_scatter_GBs"] = regen_bytes / (regen_ms * 1e6)
    res["plain_store_GBs"] = regen_bytes / (ms_of["store"] * 1e6)
    res["move_scatter_GBs"] = move_bytes / (move_ms * 1e6)
    res["generate_bytes_per_launch"] = gen_bytes
    res["regen_scatter_bytes_per_launch"] = regen_bytes
    res["move_scatter_bytes_per_launch"] = move_bytes
    del cols, evw, out, out2, ws
    torch.cuda.empty_cache()
    return res


def time_generate(torch, capi, M, masses, n, prec, out_f32, reps=5):
    lib = capi.lib()
    dec = capi.make_decay(M, masses)
    N = len(masses)
    cols = torch.empty((4 * N, n), dtype=torch.float32 if out_f32 else torch.float64,
                       device="cuda:0")
    sh = capi.stream_handle()
    t = Timer(torch)
    for i in range(reps + 1):
        if i:
            t.start()
        capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, 0, n, prec, capi.ptr(cols), n,
                                     int(out_f32), None, None, sh), "generate")
        if i:
            t.stop()
    return t.avg_ms, cols


def compare_cols(a, b, N, sample=1 << 20):
    """max / RMS of |p_a - p_b| / E over the first `sample` events."""
    import torch
    x = a[:, :sample].double().reshape(N, 4, -1)
    y = b[:, :sample].double().reshape(N, 4, -1)
    d = (x - y).abs() / y[:, 0:1, :].abs()
    return float(d.max()), float(torch.sqrt((d * d).mean()))


def generate_both(torch, capi, M, masses, n, first=0):
    """FP64 and mixed Generate of the same events: (cols64, wt64, evw64),
    (cols32, wt_mixed, evw_mixed) as tensors."""
    lib = capi.lib()
    dec = capi.make_decay(M, masses)
    N = len(masses)
    sh = capi.stream_handle()
    out = []
    for prec, dt in ((capi.FP64, torch.float64), (capi.MIXED, torch.float32)):
        cols = torch.empty((4 * N, n), dtype=dt, device="cuda:0")
        wt = torch.empty(n, dtype=torch.float64, device="cuda:0")
        evw = torch.empty(n, dtype=torch.float64, device="cuda:0")
        capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, first, n, prec, capi.ptr(cols), n,
                                     int(dt == torch.float32), capi.ptr(wt), capi.ptr(evw), sh),
                   "generate")
        out.append((cols, wt, evw))
    torch.cuda.synchronize()
    return out


def rel_err(a, b):
    """max and RMS of |a - b| / |b| over the entries where b != 0."""
    import torch
    d = (a.double() - b.double()).abs()
    m = b.double().abs() > 0
    r = d[m] / b.double().abs()[m]
    return float(r.max()), float(torch.sqrt((r * r).mean()))


def mixed_errors(torch, capi, M, masses, n):
    """Errors of the mixed form against FP64 on the same n events: momenta
    (|dp|/E per component), generation weight, PhaseSpaceWeight, and the
    mean weight (the phase-space integral)."""
    N = len(masses)
    (c64, w64, e64), (c32, w32, e32) = generate_both(torch, capi, M, masses, n)
    pmax, prms = compare_cols(c32, c64, N, sample=n)
    wmax, wrms = rel_err(w32, w64)
    emax, erms = rel_err(e32, e64)
    i64, i32 = float(w64.double().mean()), float(w32.double().mean())
    return {"sample_events": n, "momentum_err_max": pmax, "momentum_err_rms": prms,
            "gen_weight_rel_err_max": wmax, "gen_weight_rel_err_rms": wrms,
            "phsp_weight_rel_err_max": emax, "phsp_weight_rel_err_rms": erms,
            "integral_rel_err": abs(i32 - i64) / abs(i64)}


def run_c2_cpp(n, steps, warmup):
    """The headline step through the C++ drop-in (tests/cpp/bench_decays:
    `phsp.Generate(B0, Events); Events.Unweight()` on one device Decays, timed
    on the host clock around each step, the accepted count copied back as the
    reference's Range size is).  Reported beside the C ABI's line: the same
    kernels plus the C++ API's host work and synchronisation."""
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "bench_decays")
    if not os.path.exists(exe):
        return {"skipped": "tests/cpp/bin/bench_decays not built"}
    out = subprocess.run([exe, str(n), str(steps), str(warmup)], check=True,
                         capture_output=True, text=True, timeout=600)
    r = json.loads(out.stdout.strip().splitlines()[-1])
    r["events_per_s"] = n / (r["ms_per_step"] * 1e-3)
    return r


def run_c2_mixed(torch, capi, n=1 << 27):
    """The headline 3-body decay, FP64 against mixed (FP32 columns): Generate
    with the PhaseSpaceWeight column, the momenta alone, and the headline step
    (Generate + Unweight() in one call; the mixed form takes its flags from
    the mixed container's weight column, the FP64 generation weight)."""
    N = 3
    res = {"events": n}
    lib = capi.lib()
    dec = capi.make_decay(M3, MASSES3)
    sh = capi.stream_handle()
    # both precisions timed in alternating rounds, so that both see the same clocks
    for tag, with_w in (("generate", True), ("momenta_only", False)):
        gen = {}
        for prec, f32, name in ((capi.FP64, False, "fp64"), (capi.MIXED, True, "mixed")):
            cols = torch.empty((4 * N, n), dtype=torch.float32 if f32 else torch.float64,
                               device="cuda:0")
            evw = torch.empty(n, dtype=torch.float64, device="cuda:0") if with_w else None
            gen[name] = (prec, f32, cols, evw, Timer(torch))

        def gcall(name):
            prec, f32, cols, evw, _ = gen[name]
            capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, 0, n, prec, capi.ptr(cols),
                                         n, int(f32), None, capi.ptr(evw), sh), "generate")

        for name in gen:
            gcall(name)
        for _ in range(3):
            for name, g in gen.items():
                for _ in range(4):
                    g[4].start()
                    gcall(name)
                    g[4].stop()
        for name, (_, f32, _, _, t) in gen.items():
            res[f"{tag}_{name}_ms"] = t.avg_ms
            b = n * ((16 if f32 else 32) * N + (8 if with_w else 0))
            res[f"{tag}_{name}_GBs"] = b / (t.avg_ms * 1e6)
        del gen
        torch.cuda.empty_cache()
        res[f"{tag}_mixed_speedup"] = res[f"{tag}_fp64_ms"] / res[f"{tag}_mixed_ms"]
    # the headline step (Generate + Unweight() in one call) in both forms: FP64
    # columns (hpsp_generate_unweight) and FP32 ones (hpsp_generate_unweight_mixed),
    # timed in alternating rounds so that both see the same clocks
    lib = capi.lib()
    dec = capi.make_decay(M3, MASSES3)
    sh = capi.stream_handle()
    nacc = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    wsb = lib.hpsp_generate_unweight_workspace_size(n)
    ws = torch.empty(wsb, dtype=torch.uint8, device="cuda:0")
    forms = {}
    for name, f32 in (("fp64", False), ("mixed", True)):
        cols = torch.empty((4 * N, n), dtype=torch.float32 if f32 else torch.float64,
                           device="cuda:0")
        fn = lib.hpsp_generate_unweight_mixed if f32 else lib.hpsp_generate_unweight
        forms[name] = (cols, fn, f32, Timer(torch))

    def call(name):
        cols, fn, _, _ = forms[name]
        capi.check(fn(ctypes.byref(dec), SEED, 0, n, UW_SEED, capi.ptr(cols), n, None, None,
                      capi.ptr(ws), wsb, capi.ptr(nacc), sh), "generate_unweight")

    for name in forms:
        call(name)
    for _ in range(3):
        for name, (_, _, _, t) in forms.items():
            for _ in range(4):
                t.start()
                call(name)
                t.stop()
            res[f"step_{name}_n_acc"] = int(nacc.item())
    for name, (_, _, f32, t) in forms.items():
        res[f"step_{name}_ms"] = t.avg_ms
        res[f"step_{name}_events_per_s"] = n / (t.avg_ms * 1e-3)
        res[f"step_{name}_GBs"] = n * (16 if f32 else 32) * N / (t.avg_ms * 1e6)
    del forms, ws
    torch.cuda.empty_cache()
    res["step_mixed_speedup"] = res["step_fp64_ms"] / res["step_mixed_ms"]
    res["errors"] = mixed_errors(torch, capi, M3, MASSES3, 1 << 22)
    torch.cuda.empty_cache()
    return res


def run_c3(torch, capi):
    res = {}
    n = 1 << 26
    N = 10
    fl = genbod_flops(N)
    for tag, masses in (("c3", [0.13957] * 10), ("c3_threshold", [0.999] * 10)):
        ms64, c64 = time_generate(torch, capi, 10.0, masses, n, capi.FP64, False)
        msmx, cmx = time_generate(torch, capi, 10.0, masses, n, capi.MIXED, True)
        err_max, err_rms = compare_cols(cmx, c64, N)
        res[tag] = {
            "events": n, "fp64_ms": ms64, "mixed_ms": msmx,
            "fp64_events_per_s": n / (ms64 / 1e3), "mixed_events_per_s": n / (msmx / 1e3),
            "mixed_speedup": ms64 / msmx,
            "fp64_TFLOPs": n * fl / (ms64 * 1e9), "fp64_vs_fp64_peak": n * fl / (ms64 * 1e9) / FP64_PEAK_TFLOPS,
            "fp64_GBs": n * 32 * N / (ms64 * 1e6), "mixed_GBs": n * 16 * N / (msmx * 1e6),
            "mixed_vs_fp64_momentum_err_max": err_max, "mixed_vs_fp64_momentum_err_rms": err_rms,
            "flops_per_event": fl,
        }
        del c64, cmx
        torch.cuda.empty_cache()
        res[tag]["errors"] = mixed_errors(torch, capi, 10.0, masses, 1 << 22)
        torch.cuda.empty_cache()
    return res


def run_c4(torch, capi, chunk=1 << 24, nchunks=64):
    lib = capi.lib()
    dec = capi.make_decay(M3, MASSES3)
    f = capi.make_functor(capi.F_BW_SUM, BW3)
    ws_bytes = lib.hpsp_average_on_workspace_size(chunk)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda:0")
    outs = torch.empty((nchunks, 3), dtype=torch.float64, device="cuda:0")
    sh = capi.stream_handle()
    res = {"events": chunk * nchunks, "chunk": chunk}
    means = {}
    for prec, tag in ((capi.FP64, "fp64"), (capi.MIXED, "mixed")):
        for rep in range(2):  # first pass warms up
            t = Timer(torch)
            t.start()
            for c in range(nchunks):
                o = ctypes.c_void_p(outs.data_ptr() + 24 * c)
                capi.check(lib.hpsp_average_on(ctypes.byref(dec), ctypes.byref(f), SEED,
                                               c * chunk, chunk, prec, capi.ptr(ws), ws_bytes, o,
                                               sh), "average_on")
            t.stop()
        st = outs.cpu().numpy()
        acc = capi.Stats(*st[0])
        for c in range(1, nchunks):
            r = capi.Stats()
            lib.hpsp_stats_merge(acc, capi.Stats(*st[c]), r)
            acc = r
        means[tag] = (acc.w, acc.mean, acc.m2)
        res[f"{tag}_ms"] = t.avg_ms
        res[f"{tag}_events_per_s"] = res["events"] / (t.avg_ms / 1e3)
        res[f"{tag}_mean"] = acc.mean
        res[f"{tag}_sigma"] = math.sqrt(acc.m2) / acc.w
    res["mixed_speedup"] = res["fp64_ms"] / res["mixed_ms"]
    res["mixed_rel_err_mean"] = abs(means["mixed"][1] - means["fp64"][1]) / abs(means["fp64"][1])
    res["mixed_rel_err_sigma"] = abs(res["mixed_sigma"] - res["fp64_sigma"]) / res["fp64_sigma"]
    fl = genbod_flops(3)
    res["fp64_TFLOPs_genbod_only"] = res["events"] * fl / (res["fp64_ms"] * 1e9)
    return res


def run_evaluate(torch, capi, n=1 << 27, reps=5):
    """PhaseSpace::Evaluate as examples/phase_space/phsp_evaluating_functor.inl
    runs it: B0 -> J/psi K pi with the weight, M13^2, M23^2 and the helicity
    cosine per event (4 FP64 columns, no momenta stored); then over 2^26
    moving mothers (EvalMothers: one engine per event)."""
    import numpy as np
    lib = capi.lib()
    dec = capi.make_decay(M3, MASSES3)
    fs = capi.functor_array([capi.make_functor(capi.F_MASS2, (0, 2)),
                             capi.make_functor(capi.F_MASS2, (1, 2)),
                             capi.make_functor(capi.F_COS_HELICITY, (1, 2))])
    out = torch.empty((4, n), dtype=torch.float64, device="cuda:0")
    sh = capi.stream_handle()

    def ev():
        capi.check(lib.hpsp_evaluate(ctypes.byref(dec), fs, 3, SEED, 0, n, capi.ptr(out), n, sh),
                   "evaluate")
    ev()
    t = Timer(torch)
    for _ in range(reps):
        t.start()
        ev()
        t.stop()
    res = {"events": n, "functors": "M13^2, M23^2, cos(theta_K*)", "ms": t.avg_ms,
           "events_per_s": n / (t.avg_ms / 1e3), "bytes_per_event": 32,
           "GBs": n * 32 / (t.avg_ms * 1e6)}
    nm = n // 2
    rng = np.random.default_rng(11)
    mt = np.zeros((4, nm))
    mt[1:] = rng.normal(0, 2.0, (3, nm))
    mt[0] = np.sqrt(M3 * M3 + (mt[1:] ** 2).sum(axis=0))
    mth = torch.as_tensor(mt, device="cuda:0")
    del mt
    outm = torch.empty((2, nm), dtype=torch.float64, device="cuda:0")
    fm = capi.functor_array([capi.make_functor(capi.F_MASS2, (1, 3))])

    def evm():
        capi.check(lib.hpsp_evaluate_mothers(ctypes.byref(dec), fm, 1, SEED, capi.ptr(mth), nm, 0,
                                             nm, capi.ptr(outm), nm, sh), "evaluate_mothers")
    evm()
    tm = Timer(torch)
    for _ in range(reps):
        tm.start()
        evm()
        tm.stop()
    res.update({"mothers_events": nm, "mothers_ms": tm.avg_ms,
                "mothers_events_per_s": nm / (tm.avg_ms / 1e3),
                "mothers_bytes_per_event": 48, "mothers_GBs": nm * 48 / (tm.avg_ms * 1e6)})
    return res


def run_histogram(torch, capi, n=1 << 27, reps=5):
    """The Dalitz plot of phsp_basic.inl:196-214: a 100 x 100 DenseHistogram
    of (m^2_AB, m^2_BC) weighted by PhaseSpaceWeight, over the headline's
    2^27 generated 3-body events.  Fused: axes and weights computed from the
    event columns in the fill (reads 96 B/event); columns: from two value
    columns and the weight column written beforehand (reads 24 B/event)."""
    lib = capi.lib()
    dec = capi.make_decay(M3, MASSES3)
    sh = capi.stream_handle()
    cols = torch.empty((12, n), dtype=torch.float64, device="cuda:0")
    evw = torch.empty(n, dtype=torch.float64, device="cuda:0")
    capi.check(lib.hpsp_generate(ctypes.byref(dec), SEED, 0, n, capi.FP64, capi.ptr(cols), n, 0,
                                 None, capi.ptr(evw), sh), "generate")
    m = MASSES3
    lo = [(m[0] + m[1]) ** 2, (m[1] + m[2]) ** 2]
    hi = [(M3 - m[2]) ** 2, (M3 - m[0]) ** 2]
    h = capi.make_histogram([100, 100], lo, hi)
    out = torch.empty(10002, dtype=torch.float64, device="cuda:0")
    axes = capi.functor_array([capi.make_functor(capi.F_MASS2, (0, 1)),
                               capi.make_functor(capi.F_MASS2, (1, 2))])
    res = {"events": n, "bins": "100 x 100"}
    for tag, mode, w, nbytes in (("fused_phsp_weight", capi.HIST_WEIGHT_PHSP, None, 96),
                                 ("fused_weight_column", capi.HIST_WEIGHT_COLUMN, evw, 104)):
        def fill():
            capi.check(lib.hpsp_histogram_fill_events(ctypes.byref(h), ctypes.byref(dec), axes,
                                                      capi.ptr(cols), n, n, mode, capi.ptr(w),
                                                      capi.ptr(out), sh), "fill_events")
        fill()
        t = Timer(torch)
        for _ in range(reps):
            t.start()
            fill()
            t.stop()
        res[f"{tag}_ms"] = t.avg_ms
        res[f"{tag}_events_per_s"] = n / (t.avg_ms / 1e3)
        res[f"{tag}_GBs"] = n * nbytes / (t.avg_ms * 1e6)
    res["sum_of_weights"] = float(out.sum().item())
    # the same plot from value columns (as Evaluate writes them)
    x = torch.empty((2, n), dtype=torch.float64, device="cuda:0")
    c = cols.view(3, 4, n)
    for j, (a, b) in enumerate(((0, 1), (1, 2))):
        s = c[a] + c[b]
        x[j] = s[0] * s[0] - s[1] * s[1] - s[2] * s[2] - s[3] * s[3]
        del s
    del cols

    def fillc():
        capi.check(lib.hpsp_histogram_fill(ctypes.byref(h), capi.ptr(x), n, n, capi.ptr(evw),
                                           capi.ptr(out), sh), "fill")
    fillc()
    t = Timer(torch)
    for _ in range(reps):
        t.start()
        fillc()
        t.stop()
    res["columns_ms"] = t.avg_ms
    res["columns_events_per_s"] = n / (t.avg_ms / 1e3)
    res["columns_GBs"] = n * 24 / (t.avg_ms * 1e6)
    return res


def cpu_baseline_histogram(n=1 << 22):
    """The reference's DenseHistogram<double,2>::Fill (OMP, all host threads)
    on the Dalitz variables and weights of 2^22 events of the same decay."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    lib = orc.REF_NATIVE_LIB if os.path.exists(orc.REF_NATIVE_LIB) else orc.REF_LIB
    if not os.path.exists(lib):
        return None
    R = orc.Reference(lib)
    cores = len(os.sched_getaffinity(0))
    R.set_threads(cores)
    p, w, _ = R.generate(M3, MASSES3, SEED, n)
    sab, sbc = p[:, 0] + p[:, 1], p[:, 1] + p[:, 2]
    x = np.vstack([sab[:, 0] ** 2 - (sab[:, 1:] ** 2).sum(1), sbc[:, 0] ** 2 - (sbc[:, 1:] ** 2).sum(1)])
    m = MASSES3
    lo = [(m[0] + m[1]) ** 2, (m[1] + m[2]) ** 2]
    hi = [(M3 - m[2]) ** 2, (M3 - m[0]) ** 2]
    _, sec = R.histogram_fill([100, 100], lo, hi, x, w)
    return {"value": n / sec, "unit": "events/s", "cores": cores, "kind": "reference",
            "sample": f"{n} events: DenseHistogram<double,2>::Fill(dalitz, weights), 100 x 100, "
                      f"reference OMP backend, {cores} threads (sort_by_key + reduce_by_key)"}


MK, MPI = 0.493677, 0.13957061


def run_range(torch, capi, n=1 << 27, reps=5):
    """phase_space_range(K+ at rest, 3 pi, seed 321) of
    examples/misc/range_semantics.inl:154-167 over 2^27 elements: written out
    as columns (weight + 12 momentum doubles, 104 B/element), and histogrammed
    as the example does (m(pi pi) of daughters 0 and 1, 10 bins) with the
    elements generated inside the fill (no memory traffic but the bins)."""
    lib = capi.lib()
    dec = capi.make_range([MK, 0.0, 0.0, 0.0], [MPI] * 3)
    sh = capi.stream_handle()
    res = {"elements": n, "decay": "K+ (at rest) -> 3 pi, seed 321"}
    cols = torch.empty((12, n), dtype=torch.float64, device="cuda:0")
    wt = torch.empty(n, dtype=torch.float64, device="cuda:0")

    def gen():
        capi.check(lib.hpsp_range_generate(ctypes.byref(dec), 321, 0, n, capi.ptr(cols), n,
                                           capi.ptr(wt), sh), "range_generate")
    gen()
    t = Timer(torch)
    for _ in range(reps):
        t.start()
        gen()
        t.stop()
    res["generate_ms"] = t.avg_ms
    res["generate_elements_per_s"] = n / (t.avg_ms / 1e3)
    res["generate_GBs"] = n * 104 / (t.avg_ms * 1e6)
    del cols, wt
    h = capi.make_histogram([10], [2 * MPI], [MK - MPI])
    out = torch.empty(12, dtype=torch.float64, device="cuda:0")
    axes = capi.functor_array([capi.make_functor(capi.F_MASS, (0, 1))])
    for tag, mode in (("hist", capi.HIST_WEIGHT_ONE),
                      ("hist_weighted", capi.HIST_WEIGHT_GENERATION)):
        def fill():
            capi.check(lib.hpsp_histogram_fill_range(ctypes.byref(h), ctypes.byref(dec), 321,
                                                     axes, 0, n, mode, capi.ptr(out), sh),
                       "fill_range")
        fill()
        t = Timer(torch)
        for _ in range(reps):
            t.start()
            fill()
            t.stop()
        res[f"{tag}_ms"] = t.avg_ms
        res[f"{tag}_elements_per_s"] = n / (t.avg_ms / 1e3)
        if mode == capi.HIST_WEIGHT_ONE:
            res["hist_counts"] = float(out[:10].sum().item())
    return res


def cpu_baseline_range(n=1 << 24):
    """The reference's range histogram (DenseHistogram<double,1>::Fill over
    phase_space_range | invariant mass, OMP, all host threads), 2^24 elements."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    lib = orc.REF_NATIVE_LIB if os.path.exists(orc.REF_NATIVE_LIB) else orc.REF_LIB
    if not os.path.exists(lib):
        return None
    R = orc.Reference(lib)
    cores = len(os.sched_getaffinity(0))
    R.set_threads(cores)
    _, sec = R.range_histogram3([MK, 0, 0, 0], [MPI] * 3, 321, n, 0, 1, 10, 2 * MPI, MK - MPI)
    return {"value": n / sec, "unit": "elements/s", "cores": cores, "kind": "reference",
            "sample": f"{n} elements: DenseHistogram<double,1>::Fill(phase_space_range | "
                      f"(A+B).mass()), 10 bins, reference OMP backend, {cores} threads"}


BW2 = [(0.892, 0.05, 1.0, 0.3, 1, 2), (1.43, 0.11, -0.4, 0.8, 0, 1)]


def serving_mix(count, seed=5):
    """Request mix: N uniform in 2..12, daughter masses U(0.1, 1) GeV, mother
    mass U(1.05, 3) x their sum, events log-uniform in 1e3..1e7, 30% AverageOn
    of a two-Breit-Wigner functor, half FP64 and half mixed precision."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        N = int(rng.integers(2, 13))
        m = [float(x) for x in rng.uniform(0.1, 1.0, N)]
        M = sum(m) * float(rng.uniform(1.05, 3.0))
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e7))))
        avg = bool(rng.uniform() < 0.3)
        mixed = bool(rng.uniform() < 0.5)
        out.append(dict(N=N, M=M, m=m, n=n, avg=avg, mixed=mixed,
                        seed=int(rng.integers(1, 2**31 - 1))))
    return out


def served_mixed_errors(torch, capi, nreq=6000, seed=5, count=24, nmax=1 << 20):
    """Errors of the served mix's mixed-precision requests against FP64 (the
    first `count` of each kind, events capped at nmax).  A served request's
    result is bit for bit the direct call's (tests/test_gpu_serve.py), so the
    direct calls are compared here."""
    import numpy as np
    lib = capi.lib()
    sh = capi.stream_handle()
    gen, avg = [], []
    for r in serving_mix(nreq, seed):
        if not r["mixed"]:
            continue
        n = min(r["n"], nmax)
        if r["avg"] and len(avg) < count:
            dec = capi.make_decay(r["M"], r["m"])
            terms = BW2 if r["N"] >= 3 else [(0.892, 0.05, 1.0, 0.3, 0, 1)]
            f = capi.make_functor(capi.F_BW_SUM, terms)
            ws_bytes = lib.hpsp_average_on_workspace_size(n)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda:0")
            out = torch.empty((2, 3), dtype=torch.float64, device="cuda:0")
            for k, prec in enumerate((capi.FP64, capi.MIXED)):
                o = ctypes.c_void_p(out.data_ptr() + 24 * k)
                capi.check(lib.hpsp_average_on(ctypes.byref(dec), ctypes.byref(f), r["seed"], 0, n,
                                               prec, capi.ptr(ws), ws_bytes, o, sh), "average_on")
            st = out.cpu().numpy()
            s64 = math.sqrt(st[0, 2]) / st[0, 0]
            s32 = math.sqrt(st[1, 2]) / st[1, 0]
            # a 2-body request integrates a constant (its one pair mass is the
            # mother's): both sigmas are rounding noise there, so only the mean
            # is compared
            sig = abs(s32 - s64) / s64 if r["N"] >= 3 else None
            avg.append((abs(st[1, 1] - st[0, 1]) / abs(st[0, 1]), sig))
        elif not r["avg"] and len(gen) < count:
            e = mixed_errors(torch, capi, r["M"], r["m"], n)
            gen.append(e)
        if len(gen) >= count and len(avg) >= count:
            break
    torch.cuda.empty_cache()
    g = {k: float(max(e[k] for e in gen)) for k in gen[0] if k != "sample_events"}
    return {"generate_requests": len(gen), "generate_max_over_requests": g,
            "average_requests": len(avg),
            "average_mean_rel_err_max": float(max(a[0] for a in avg)),
            "average_sigma_requests": sum(1 for a in avg if a[1] is not None),
            "average_sigma_rel_err_max": float(max(a[1] for a in avg if a[1] is not None)),
            "events_cap": nmax}


def run_c5(torch, capi, nreq=6000, rate=3000.0, seed=5):
    """Poisson arrivals into the continuous-batching server."""
    import numpy as np
    mix = serving_mix(nreq, seed)
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    reqs = []
    for i, r in enumerate(mix):
        dec = capi.make_decay(r["M"], r["m"])
        terms = BW2 if r["N"] >= 3 else [(0.892, 0.05, 1.0, 0.3, 0, 1)]
        f = capi.make_functor(capi.F_BW_SUM, terms)
        kind = capi.REQ_AVERAGE if r["avg"] else capi.REQ_GENERATE
        prec = capi.MIXED if r["mixed"] else capi.FP64
        reqs.append(capi.make_request(kind, dec, r["seed"], r["n"], prec=prec,
                                      out_f32=r["mixed"], functor=f, tag=i))
    srv = capi.Server(64 << 30, n_streams=4)
    try:
        # warm the code objects of every (N, precision, kind) once
        warm = 0
        for r in reqs[:200]:
            w = capi.Request.from_buffer_copy(r)
            w.n = 4096
            srv.submit(w)
            warm += 1
        got = 0
        while got < warm:
            for c in srv.poll(wait_ms=100):
                got += 1
                if c.kind == capi.REQ_GENERATE:
                    srv.release(c.ticket)
        info0 = srv.info()
        lat = np.zeros(nreq)
        queue = np.zeros(nreq)
        t0 = time.perf_counter()
        i = done = 0
        first_submit = last_done = None
        while done < nreq:
            now = time.perf_counter() - t0
            while i < nreq and arrivals[i] <= now:
                srv.submit(reqs[i])
                i += 1
            for c in srv.poll(wait_ms=0 if i < nreq else 2):
                k = int(c.user_tag)
                lat[k] = c.done_s - c.submit_s
                queue[k] = c.admit_s - c.submit_s
                first_submit = c.submit_s if first_submit is None else min(first_submit, c.submit_s)
                last_done = c.done_s if last_done is None else max(last_done, c.done_s)
                if c.kind == capi.REQ_GENERATE:
                    srv.release(c.ticket)
                done += 1
            if i < nreq:
                gap = arrivals[i] - (time.perf_counter() - t0)
                if gap > 2e-4:
                    time.sleep(min(gap, 1e-3))
        info = srv.info()
    finally:
        srv.close()
    events = sum(r["n"] for r in mix)
    wall = last_done - first_submit
    lat_ms = lat * 1e3
    return {
        "requests": nreq, "offered_rate_per_s": rate, "events": events,
        "events_per_s": events / wall, "wall_s": wall,
        "latency_ms_p50": float(np.percentile(lat_ms, 50)),
        "latency_ms_p99": float(np.percentile(lat_ms, 99)),
        "latency_ms_max": float(lat_ms.max()),
        "admission_wait_ms_p99": float(np.percentile(queue * 1e3, 99)),
        "batches": info["batches"] - info0["batches"],
        "launches": info["launches"] - info0["launches"],
        "arena_peak_GB": info["arena_peak"] / 1e9,
        "mix": "N 2..12, m U(0.1,1) GeV, M U(1.05,3)*sum(m), n log-uniform 1e3..1e7, "
               "30% AverageOn (2 Breit-Wigners), 50% mixed precision (FP32 columns)",
    }


def cpu_baseline_mix(count=40, cap=200000, seed=5):
    """The reference CPU path on the first requests of the serving mix (FP64)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    if not os.path.exists(orc.REF_LIB):
        return None
    R = orc.Reference(orc.REF_LIB)
    cores = len(os.sched_getaffinity(0))
    R.set_threads(cores)
    ev = 0
    sec = 0.0
    for r in serving_mix(count, seed):
        n = min(r["n"], cap)
        if r["avg"]:
            terms = BW2 if r["N"] >= 3 else [(0.892, 0.05, 1.0, 0.3, 0, 1)]
            par = orc.functor_params(1, terms)
            _, _, s = R.average_on(r["M"], r["m"], r["seed"], n, fid=1, fpar=par)
        else:
            _, _, s = R.generate(r["M"], r["m"], r["seed"], n, want_output=False)
        ev += n
        sec += s
    return {"value": ev / sec, "unit": "events/s", "cores": cores, "kind": "reference",
            "sample": f"first {count} requests of the serving mix, events capped at {cap}, "
                      f"FP64, reference OMP backend, {cores} threads"}


def cpu_baseline_evaluate(n=1 << 23):
    """The reference's PhaseSpace::Evaluate (OMP, all host threads) on the
    Evaluate workload's first 2^23 events."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    lib = orc.REF_NATIVE_LIB if os.path.exists(orc.REF_NATIVE_LIB) else orc.REF_LIB
    if not os.path.exists(lib):
        return None
    R = orc.Reference(lib)
    cores = len(os.sched_getaffinity(0))
    R.set_threads(cores)
    F = [(3, (0, 2)), (3, (1, 2))]
    _, sec = R.evaluate(M3, MASSES3, SEED, n, F)
    return {"value": n / sec, "unit": "events/s", "cores": cores, "kind": "reference",
            "sample": f"{n} events, Evaluate(mother, result, M13^2, M23^2) on the reference's OMP "
                      f"backend, {cores} threads (the harness binds two functors; the GPU line "
                      f"evaluates three)"}


def cpu_baseline(n=1 << 25, min_s=10.0, max_reps=16):
    """The reference CPU path (OMP) on a bounded sample of the headline step:
    the harness built with the reference's own flags (-O3, AVX2/FMA), 2^25
    events per repetition, repeated until about ten seconds of wall time."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    cores = len(os.sched_getaffinity(0))
    lib = orc.REF_NATIVE_LIB if os.path.exists(orc.REF_NATIVE_LIB) else orc.REF_LIB
    if os.path.exists(lib):
        R = orc.Reference(lib)
        R.set_threads(cores)
        t0 = time.perf_counter()
        gsec = usec = 0.0
        reps = 0
        while reps < max_reps and (reps == 0 or time.perf_counter() - t0 < min_s):
            _, _, g = R.generate(M3, MASSES3, SEED, n, want_output=False)
            _, _, _, u = R.unweight(M3, MASSES3, SEED, n, UW_SEED, want_output=False)
            gsec += g
            usec += u
            reps += 1
        t1 = time.perf_counter()
        flags = ("--std=c++20 -march=x86-64-v3 -O3" if lib == orc.REF_NATIVE_LIB
                 else "-O2 -ffp-contract=off")
        return {"value": reps * n / (gsec + usec), "unit": "events/s", "cores": cores,
                "kind": "reference",
                "sample": f"{reps} x {n} events of the same 3-body FP64 decay: "
                          f"PhaseSpace<3>::Generate into Decays + Decays::Unweight() on the "
                          f"reference's OMP backend (built from its headers, {flags}), "
                          f"{cores} threads; generate {gsec / reps:.3f} s, unweight "
                          f"{usec / reps:.3f} s per repetition",
                "wall_s": t1 - t0}
    O = orc.Oracle()
    import numpy as np
    t0 = time.perf_counter()
    cols, _ = O.generate(M3, MASSES3, SEED, n // 8, nthreads=cores)
    O.unweight(M3, MASSES3, cols, UW_SEED, nthreads=cores)
    t1 = time.perf_counter()
    return {"value": (n // 8) / (t1 - t0), "unit": "events/s", "cores": cores, "kind": "port",
            "sample": f"{n // 8} events, oracle C restatement (OpenMP generate, serial "
                      f"partition), {cores} threads"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported by this "
              "build (the path shards by event index, see DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    import torch
    from hydra_amd import capi

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    torch.cuda.set_device(0)
    extras = {}
    if args.only == "c3":
        extras["c3"] = run_c3(torch, capi)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "c4":
        extras["c4"] = run_c4(torch, capi)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "evaluate":
        extras["evaluate"] = run_evaluate(torch, capi)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "histogram":
        extras["histogram"] = run_histogram(torch, capi)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "range":
        extras["range"] = run_range(torch, capi)
        if args.cpu_baseline:
            extras["range"]["cpu_baseline"] = cpu_baseline_range()
        print(json.dumps(extras), flush=True)
        return
    if args.only == "c2_cpp":
        extras["c2_cpp"] = run_c2_cpp(args.n, args.steps, args.warmup)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "c2_mixed":
        extras["c2_mixed"] = run_c2_mixed(torch, capi)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "c5":
        extras["c5"] = run_c5(torch, capi)
        extras["c5_saturated"] = run_c5(torch, capi, rate=50000.0)
        extras["c5_mixed_errors"] = served_mixed_errors(torch, capi)
        print(json.dumps(extras), flush=True)
        return
    if args.only == "copy":
        # plain device copy of a known size: calibrates the HBM byte counters
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")
        y = torch.empty_like(x)
        t = Timer(torch)
        for _ in range(3):
            t.start()
            y.copy_(x)
            t.stop()
        print(json.dumps({"copy_bytes_read": x.numel() * 8, "copy_bytes_written": y.numel() * 8,
                          "copy_ms": t.avg_ms,
                          "copy_GBs": 2 * x.numel() * 8 / (t.avg_ms * 1e6)}), flush=True)
        return

    c2 = run_c2(torch, capi, args.n, args.steps, args.warmup)
    print(f"# c2 done: {c2['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if args.extras and not args.only:
        extras["c2_cpp"] = run_c2_cpp(args.n, args.steps, args.warmup)
        print("# c2 through the C++ API done", file=sys.stderr, flush=True)
        extras["c2_mixed"] = run_c2_mixed(torch, capi)
        print("# c2 mixed done", file=sys.stderr, flush=True)
        extras["c3"] = run_c3(torch, capi)
        print("# c3 done", file=sys.stderr, flush=True)
        extras["c4"] = run_c4(torch, capi)
        print("# c4 done", file=sys.stderr, flush=True)
        extras["evaluate"] = run_evaluate(torch, capi)
        print("# evaluate done", file=sys.stderr, flush=True)
        extras["histogram"] = run_histogram(torch, capi)
        print("# histogram done", file=sys.stderr, flush=True)
        extras["range"] = run_range(torch, capi)
        print("# range done", file=sys.stderr, flush=True)
        extras["c5"] = run_c5(torch, capi)
        extras["c5"]["mixed_errors"] = served_mixed_errors(torch, capi)
        extras["c5_saturated"] = run_c5(torch, capi, rate=50000.0)
        print("# c5 done", file=sys.stderr, flush=True)
    extras["c2"] = c2

    # The step is the flags from the weight chain (the engine words and the
    # chain's FP64 arithmetic; no HBM traffic but the flag words) and the
    # regenerating scatter, which writes the step's 12.9 GB and is bound by HBM
    # (the roofline).  The two-pass form's weights pass is reported beside it.
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tf):
        with open(tf) as fh:
            traffic = json.load(fh).get("regen", {}).get("bytes_per_launch")
    w_flops = genbod_flops(3) + weight_flops(3)
    w_tflops = args.n * w_flops / (c2["generate_weights_ms"] * 1e9)
    roof = {"bound": "hbm", "achieved": c2["regen_scatter_GBs"], "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": c2["regen_scatter_GBs"] / HBM_PEAK_GBS, "traffic": traffic,
            "kernel": "k_uw_regen<3, 0>", "ms": c2["regen_scatter_ms"],
            "algorithmic_bytes_per_launch": c2["regen_scatter_bytes_per_launch"],
            "plain_store": {"kernel": "torch fill of the 12 output columns (12.9 GB, no reads)",
                            "ms": c2["plain_store_ms"], "achieved": c2["plain_store_GBs"],
                            "regen_over_plain_store": c2["regen_scatter_GBs"] / c2["plain_store_GBs"]},
            "flags_pass": {"kernel": "k_chain_table3 + k_uw_flags_chain<3, 0> + k_uw_patch_chain"
                                     " + scans",
                           "bound": "valu (engine words, weight chain)",
                           "ms": c2["chain_flags_scan_ms"],
                           "events_per_s": args.n / (c2["chain_flags_scan_ms"] * 1e-3)},
            "two_pass_weights_pass": {"kernel": "k_generate_f64<3, 0, false, false> (no columns)",
                                      "bound": "fp64 issue", "ms": c2["generate_weights_ms"],
                                      "flops_per_event": w_flops, "achieved": w_tflops,
                                      "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                                      "frac": w_tflops / FP64_PEAK_TFLOPS}}
    cpu = None
    if args.cpu_baseline:
        cpu = cpu_baseline()
        if args.extras and "c5" in extras:
            extras["c5"]["cpu_baseline"] = cpu_baseline_mix()
        if args.extras and "evaluate" in extras:
            extras["evaluate"]["cpu_baseline"] = cpu_baseline_evaluate()
        if args.extras and "histogram" in extras:
            extras["histogram"]["cpu_baseline"] = cpu_baseline_histogram()
        if args.extras and "range" in extras:
            extras["range"]["cpu_baseline"] = cpu_baseline_range()
        print("# cpu baseline done", file=sys.stderr, flush=True)

    line = {
        "metric": "phase-space events/s (FP64 generate + unweight, 3-body)",
        "value": c2["events_per_s"],
        "unit": "events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": c2["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (counter-based RNG, seed 159753)",
        "config": {"workload": "3-body decay 5.279 -> (3.0969, 0.493677, 0.13957) GeV, FP64, "
                               "Generate into SoA + Unweight() as Decays runs them (nothing "
                               "formed at Generate; Unweight takes each event's flag from its "
                               "weight chain and writes its momenta once, at its place in the "
                               "unweighted container)", "events_per_step": args.n,
                   "n_daughters": 3, "engine": "squares3", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
