"""Helpers for the GPU tests: call the C ABI with torch-owned device memory."""
import ctypes

import numpy as np
import torch

from hydra_amd import capi
from hydra_amd.capi import FP64, MIXED, check, lib, make_decay, make_functor, ptr, stream_handle

DEV = torch.device("cuda:0")


def generate(dec, seed, n, first=0, prec=FP64, out_f32=False, want_wt=True, want_evw=True,
             mothers=None, ld=None):
    N = dec.n
    ld = ld or max(n, 1)
    cols = torch.empty((4 * N, ld), dtype=torch.float32 if out_f32 else torch.float64,
                       device=DEV)
    wt = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if want_wt else None
    evw = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if want_evw else None
    if mothers is None:
        rc = lib().hpsp_generate(ctypes.byref(dec), seed, first, n, prec, ptr(cols), ld,
                                 int(out_f32), ptr(wt), ptr(evw), stream_handle())
    else:
        mt = torch.as_tensor(np.ascontiguousarray(mothers), dtype=torch.float64, device=DEV)
        rc = lib().hpsp_generate_mothers(ctypes.byref(dec), seed, first, n, ptr(mt),
                                         mt.shape[1], prec, ptr(cols), ld, int(out_f32),
                                         ptr(wt), ptr(evw), stream_handle())
    check(rc, "hpsp_generate")
    torch.cuda.synchronize()
    return cols, wt, evw


def event_weight(dec, cols, f=None):
    n = cols.shape[1]
    w = torch.empty(n, dtype=torch.float64, device=DEV)
    rc = lib().hpsp_event_weight(ctypes.byref(dec), ctypes.byref(f) if f else None, ptr(cols),
                                 cols.shape[1], n, ptr(w), stream_handle())
    check(rc, "hpsp_event_weight")
    torch.cuda.synchronize()
    return w


def unweight(dec, cols, seed, f=None, max_value=-1.0, w_in=None, want_out=True, want_idx=True,
             n=None):
    n = cols.shape[1] if n is None else n
    ws_bytes = lib().hpsp_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty_like(cols) if want_out else None
    idx = torch.empty(max(n, 1), dtype=torch.int64, device=DEV) if want_idx else None
    w_out = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if w_in is not None else None
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    rc = lib().hpsp_unweight(ctypes.byref(dec), ctypes.byref(f) if f else None, max_value,
                             seed, ptr(cols), cols.shape[1], n, ptr(w_in), ptr(out),
                             cols.shape[1], ptr(idx), ptr(w_out), ptr(ws), ws_bytes, ptr(nacc),
                             stream_handle())
    check(rc, "hpsp_unweight")
    torch.cuda.synchronize()
    return int(nacc.item()), out, idx, w_out


def average_on(dec, seed, n, f=None, first=0, prec=FP64, mothers=None):
    ws_bytes = lib().hpsp_average_on_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty(3, dtype=torch.float64, device=DEV)
    if mothers is None:
        rc = lib().hpsp_average_on(ctypes.byref(dec), ctypes.byref(f) if f else None, seed,
                                   first, n, prec, ptr(ws), ws_bytes, ptr(out), stream_handle())
    else:
        mt = torch.as_tensor(np.ascontiguousarray(mothers), dtype=torch.float64, device=DEV)
        rc = lib().hpsp_average_on_mothers(ctypes.byref(dec), ctypes.byref(f) if f else None,
                                           seed, ptr(mt), mt.shape[1], n, prec, ptr(ws),
                                           ws_bytes, ptr(out), stream_handle())
    check(rc, "hpsp_average_on")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def rng_words(engine, seed, discard, n):
    out = torch.empty(n, dtype=torch.int64, device=DEV)
    check(lib().hpsp_rng_words(capi.ENGINES.get(engine, engine), seed, discard, n, ptr(out),
                               stream_handle()), "hpsp_rng_words")
    torch.cuda.synchronize()
    return out.cpu().numpy().view(np.uint64)


def bw_functor(par):
    K = int(par[0])
    terms = [tuple(par[1 + 6 * k: 7 + 6 * k]) for k in range(K)]
    return make_functor(capi.F_BW_SUM, terms)


def momentum_error(a, b, N):
    """Max over events/particles of max_c |a-b| / E (rows from SoA (4N, n))."""
    a = np.asarray(a, dtype=np.float64).reshape(N, 4, -1)
    b = np.asarray(b, dtype=np.float64).reshape(N, 4, -1)
    E = np.abs(b[:, 0:1, :])
    d = np.abs(a - b) / E
    return float(d.max()), float(np.sqrt((d ** 2).mean()))


__all__ = ["DEV", "FP64", "MIXED", "generate", "event_weight", "unweight", "average_on",
           "rng_words", "bw_functor", "momentum_error", "make_decay", "make_functor"]


def evaluate(dec, functors, seed, n, first=0, mothers=None, ld=None):
    """hpsp_evaluate / hpsp_evaluate_mothers -> (1 + nf, ld) float64 tensor."""
    ld = ld or max(n, 1)
    out = torch.empty((1 + len(functors), ld), dtype=torch.float64, device=DEV)
    fs = capi.functor_array(functors)
    if mothers is None:
        rc = lib().hpsp_evaluate(ctypes.byref(dec), fs, len(functors), seed, first, n, ptr(out),
                                 ld, stream_handle())
    else:
        mt = torch.as_tensor(np.ascontiguousarray(mothers), dtype=torch.float64, device=DEV)
        rc = lib().hpsp_evaluate_mothers(ctypes.byref(dec), fs, len(functors), seed, ptr(mt),
                                         mt.shape[1], first, n, ptr(out), ld, stream_handle())
    check(rc, "hpsp_evaluate")
    torch.cuda.synchronize()
    return out


def unweight_sample(cols, w, max_pdf=-1.0, seed=capi.UNWEIGHT_SAMPLE_SEED, jump=0,
                    order=capi.ORDER_STABLE, want_cols=True, want_idx=True, want_w=False):
    """hpsp_unweight_sample on (ncols, n) value columns -> (kept, cols, idx, w)."""
    c = torch.as_tensor(np.ascontiguousarray(np.atleast_2d(cols)), dtype=torch.float64,
                        device=DEV)
    ncols, n = c.shape
    wt = torch.as_tensor(np.ascontiguousarray(w), dtype=torch.float64, device=DEV)
    ws_bytes = lib().hpsp_unweight_sample_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty_like(c) if want_cols else None
    idx = torch.empty(max(n, 1), dtype=torch.int64, device=DEV) if want_idx else None
    w_out = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if want_w else None
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    rc = lib().hpsp_unweight_sample(ncols, ptr(c), n, n, ptr(wt), max_pdf, seed, jump, order,
                                    ptr(out), n, ptr(idx), ptr(w_out), ptr(ws), ws_bytes,
                                    ptr(nacc), stream_handle())
    check(rc, "hpsp_unweight_sample")
    torch.cuda.synchronize()
    return (int(nacc.item()), None if out is None else out.cpu().numpy(),
            None if idx is None else idx.cpu().numpy()[:n],
            None if w_out is None else w_out.cpu().numpy()[:n])


def range_generate(dec, seed, n, first=0, ld=None, want_wt=True):
    """hpsp_range_generate -> (weights (n,), SoA momentum columns (4N, ld))."""
    N = dec.n
    ld = ld or max(n, 1)
    cols = torch.empty((4 * N, ld), dtype=torch.float64, device=DEV)
    wt = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if want_wt else None
    check(lib().hpsp_range_generate(ctypes.byref(dec), seed, first, n, ptr(cols), ld, ptr(wt),
                                    stream_handle()), "hpsp_range_generate")
    torch.cuda.synchronize()
    return (None if wt is None else wt.cpu().numpy()[:n]), cols.cpu().numpy()


def range_histogram(grid, lo, hi, dec, seed, axes, n, first=0, mode=capi.HIST_WEIGHT_ONE):
    """hpsp_histogram_fill_range -> the nbins + 2 contents."""
    h = capi.make_histogram(grid, lo, hi)
    out = torch.empty(int(np.prod(grid)) + 2, dtype=torch.float64, device=DEV)
    check(lib().hpsp_histogram_fill_range(ctypes.byref(h), ctypes.byref(dec), seed,
                                          capi.functor_array(axes), first, n, mode, ptr(out),
                                          stream_handle()), "hpsp_histogram_fill_range")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def unweight_regenerate(dec, gen_seed, first, n, w, seed, ld=None, want_out=True, want_idx=True,
                        want_w=True):
    """hpsp_unweight_regenerate -> (kept, cols, idx, w_out) as tensors."""
    ld = ld or max(n, 1)
    ws_bytes = lib().hpsp_unweight_workspace_size(n)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=DEV)
    out = torch.empty((4 * dec.n, ld), dtype=torch.float64, device=DEV) if want_out else None
    idx = torch.empty(max(n, 1), dtype=torch.int64, device=DEV) if want_idx else None
    w_out = torch.empty(max(n, 1), dtype=torch.float64, device=DEV) if want_w else None
    nacc = torch.zeros(1, dtype=torch.int64, device=DEV)
    check(lib().hpsp_unweight_regenerate(ctypes.byref(dec), gen_seed, first, n, ptr(w), seed,
                                         ptr(out), ld, ptr(idx), ptr(w_out), ptr(ws), ws_bytes,
                                         ptr(nacc), stream_handle()), "hpsp_unweight_regenerate")
    torch.cuda.synchronize()
    return int(nacc.item()), out, idx, w_out
