"""hpsp_average_on_noalloc: AverageOn with its partials in the library's fixed
scratch area instead of a caller's workspace.  Checked bit for bit against
hpsp_average_on with a workspace, on two calls in a row (the second smaller,
so stale partials of the first would show), in both precisions and with a
functor; sizes past the scratch area are refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from hydra_amd import capi
from gpu_util import DEV, FP64, MIXED, average_on, bw_functor, make_decay, make_functor

M3, MASSES3 = 5.279, [3.0969, 0.493677, 0.13957]


def noalloc(dec, seed, n, f=None, first=0, prec=FP64):
    out = torch.full((3,), float("nan"), dtype=torch.float64, device=DEV)
    rc = capi.lib().hpsp_average_on_noalloc(ctypes.byref(dec), ctypes.byref(f) if f else None,
                                            seed, first, n, prec, capi.ptr(out),
                                            capi.stream_handle())
    capi.check(rc, "hpsp_average_on_noalloc")
    torch.cuda.synchronize()
    return out.cpu().numpy()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_noalloc_matches_workspace_call(golden, prec):
    dec = make_decay(M3, MASSES3)
    f = bw_functor(golden["bw3_par"])
    # the first call fills most of the partials it uses; the second, smaller,
    # call must read only its own
    for n, first in ((3 << 20) + 123, 0), (50_001, 12345):
        got = noalloc(dec, 159753, n, f=f, first=first, prec=prec)
        want = average_on(dec, 159753, n, f=f, first=first, prec=prec)
        assert np.array_equal(bits(got), bits(want)), (n, got, want)


@pytest.mark.gpu
def test_noalloc_sizes_and_functors():
    dec = make_decay(10.0, [0.13957] * 10)
    for n, f in ((1 << 22, make_functor(capi.F_MASS2_01)), (4096, None), (1, None), (0, None)):
        got = noalloc(dec, 7, n, f=f)
        want = average_on(dec, 7, n, f=f)
        assert np.array_equal(bits(got), bits(want)), n
    big = capi.NOALLOC_MAX_EVENTS
    out = torch.zeros(3, dtype=torch.float64, device=DEV)
    rc = capi.lib().hpsp_average_on_noalloc(ctypes.byref(dec), None, 7, 0, big + 1, FP64,
                                            capi.ptr(out), capi.stream_handle())
    assert rc == capi.ERR_WORKSPACE
    torch.cuda.synchronize()
    assert torch.equal(out, torch.zeros(3, dtype=torch.float64, device=DEV))
