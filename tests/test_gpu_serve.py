"""The continuous-batching server: every request's result equals the direct
call's bit for bit, alone and inside a mix of uneven requests; admission is
by HBM footprint (requests larger than the arena are refused, a stream of
requests larger in total than the arena completes as space is released)."""
import ctypes

import numpy as np
import pytest
import torch

from gpu_util import average_on, generate
from hydra_amd import capi

pytestmark = pytest.mark.gpu

BW2 = [(0.892, 0.05, 1.0, 0.3, 0, 1), (1.43, 0.11, -0.4, 0.8, 1, 2)]


def mix(seed, count, nmax):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        N = int(rng.integers(2, 13))
        m = list(rng.uniform(0.1, 1.0, N))
        M = sum(m) * float(rng.uniform(1.05, 3.0))
        n = int(np.exp(rng.uniform(np.log(1), np.log(nmax))))
        if i == 3:
            n = 0
        kind = capi.REQ_AVERAGE if rng.uniform() < 0.3 else capi.REQ_GENERATE
        prec = capi.MIXED if rng.uniform() < 0.5 else capi.FP64
        dec = capi.make_decay(M, m, engine=int(rng.integers(0, 6)))
        f = capi.make_functor(capi.F_BW_SUM, BW2 if N >= 3 else BW2[:1])
        reqs.append(dict(kind=kind, dec=dec, seed=int(rng.integers(0, 2**63)), n=n,
                         first=int(rng.integers(0, 2**33)), prec=prec,
                         out_f32=(prec == capi.MIXED), want_weights=bool(rng.uniform() < 0.5),
                         f=f, N=N))
    return reqs


def to_request(d, tag):
    return capi.make_request(d["kind"], d["dec"], d["seed"], d["n"], first=d["first"],
                             prec=d["prec"], out_f32=d["out_f32"],
                             want_weights=d["want_weights"], functor=d["f"], tag=tag)


def host_copy(ptr, nbytes):
    out = np.empty(nbytes, dtype=np.uint8)
    if nbytes:
        capi.check(capi.lib().hpsp_memcpy_to_host(out.ctypes.data_as(ctypes.c_void_p),
                                                  ctypes.c_void_p(ptr), nbytes), "copy")
    return out


def direct(d):
    if d["kind"] == capi.REQ_AVERAGE:
        return average_on(d["dec"], d["seed"], d["n"], f=d["f"], first=d["first"],
                          prec=d["prec"])
    if d["n"] == 0:
        return (np.zeros(0, np.uint8), None)
    cols, _, evw = generate(d["dec"], d["seed"], d["n"], first=d["first"], prec=d["prec"],
                            out_f32=d["out_f32"], want_wt=False, want_evw=d["want_weights"])
    c = cols.cpu().numpy().view(np.uint8).reshape(-1)
    w = evw.cpu().numpy() if d["want_weights"] else None
    return (c, w)


def served(c, d):
    if d["kind"] == capi.REQ_AVERAGE:
        return np.array([c.stats.w, c.stats.mean, c.stats.m2])
    esz = 4 if d["out_f32"] else 8
    # the served columns are c.ld apart: keep the first n values of each
    assert c.ld >= d["n"] and c.ld % 64 == 0
    cols = host_copy(c.cols, c.ld * 4 * d["N"] * esz)
    cols = cols.reshape(4 * d["N"], c.ld * esz)[:, :d["n"] * esz].reshape(-1)
    w = None
    if d["want_weights"] and d["n"]:
        w = host_copy(c.weights, d["n"] * 8).view(np.float64)
    return (cols, w)


def same(a, b, d):
    if d["kind"] == capi.REQ_AVERAGE:
        return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))
    if not np.array_equal(a[0], b[0]):
        return False
    if d["want_weights"] and d["n"]:
        return np.array_equal(a[1].view(np.uint64), b[1].view(np.uint64))
    return True


def run_all(server, reqs, release=True):
    tickets = {}
    for i, d in enumerate(reqs):
        tickets[server.submit(to_request(d, i))] = i
    got = {}
    while len(got) < len(reqs):
        for c in server.poll(wait_ms=200):
            assert c.status == 0, c.status
            d = reqs[c.user_tag]
            got[c.user_tag] = served(c, d)
            if d["kind"] == capi.REQ_GENERATE and release:
                server.release(c.ticket)
    return got


def test_mix_equals_direct_and_alone():
    reqs = mix(11, 40, 300000)
    srv = capi.Server(8 << 30, n_streams=4, chunk_events=1 << 16)
    try:
        got = run_all(srv, reqs)
        info = srv.info()
    finally:
        srv.close()
    assert info["completed"] == len(reqs) and info["arena_in_use"] == 0
    # chunks of different requests shared batches
    assert info["launches"] > info["batches"]
    for i, d in enumerate(reqs):
        assert same(got[i], direct(d), d), i
    # each request alone on a fresh server gives the same bits
    for i in (0, 1, 2, 3, 5, 8):
        srv = capi.Server(1 << 30, n_streams=2, chunk_events=1 << 16)
        try:
            alone = run_all(srv, [reqs[i]])
        finally:
            srv.close()
        assert same(alone[0], got[i], reqs[i]), i


def test_admission_by_footprint():
    dec = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957])
    arena = 64 << 20
    srv = capi.Server(arena, n_streams=2, chunk_events=1 << 18)
    try:
        too_big = capi.make_request(capi.REQ_GENERATE, dec, 1, arena // 96 + 1000)
        with pytest.raises(capi.HpspError) as e:
            srv.submit(too_big)
        assert e.value.code == capi.ERR_WORKSPACE
        # 12 requests of ~24 MB each through a 64 MB arena
        reqs = [dict(kind=capi.REQ_GENERATE, dec=dec, seed=7 + i, n=250000, first=i * 250000,
                     prec=capi.FP64, out_f32=False, want_weights=False, f=None, N=3)
                for i in range(12)]
        got = run_all(srv, reqs)
        info = srv.info()
    finally:
        srv.close()
    assert info["arena_peak"] <= arena and info["rejected"] == 1
    for i in (0, 11):
        assert same(got[i], direct(reqs[i]), reqs[i])
