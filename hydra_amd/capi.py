"""ctypes binding of libhpsp.so (include/hpsp.h) for tests and benchmarks.

Device buffers are passed as raw pointers; the helpers accept torch tensors
(which supply device memory and the current HIP stream) and plain integers.
Loading fails loudly when the library has not been built: there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libhpsp.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "hpsp.h")

MAX_N = 12
MAX_TERMS = 4

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_MASS = 2
ERR_ENERGY = 3
ERR_WORKSPACE = 4
ERR_HIP = 5

ENGINES = {"squares3": 0, "squares4": 1, "philox": 2, "philox_long": 3,
           "threefry": 4, "threefry_long": 5}
MINSTD_RAND = 6  # phase_space_range's engine (hpsp_range_* only)
FP64, MIXED = 0, 1
F_ONE, F_BW_SUM, F_MASS2_01, F_MASS2, F_COS_HELICITY, F_MASS = 0, 1, 2, 3, 4, 5
F_DECAY_PLANE_ANGLE = 6
MAX_PARTICLES = 16
MAX_EVAL_FUNCTORS = 8


class Decay(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("engine", ctypes.c_int32),
                ("mother_mass", ctypes.c_double),
                ("masses", ctypes.c_double * MAX_N),
                ("ecm", ctypes.c_double), ("max_weight", ctypes.c_double),
                ("beta", ctypes.c_double * 3)]


class Functor(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nterms", ctypes.c_int32),
                ("m", ctypes.c_double * MAX_TERMS), ("g", ctypes.c_double * MAX_TERMS),
                ("cr", ctypes.c_double * MAX_TERMS), ("ci", ctypes.c_double * MAX_TERMS),
                ("a", ctypes.c_int32 * MAX_TERMS), ("b", ctypes.c_int32 * MAX_TERMS)]


class Particles(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("cols", ctypes.c_void_p * MAX_PARTICLES),
                ("ld", ctypes.c_uint64 * MAX_PARTICLES)]


class Stats(ctypes.Structure):
    _fields_ = [("w", ctypes.c_double), ("mean", ctypes.c_double), ("m2", ctypes.c_double)]


ORDER_STABLE, ORDER_SEQUENTIAL = 0, 1
UNWEIGHT_SAMPLE_SEED = 0x8EC74D321E6B5A27
HIST_MAX_DIM = 4
HIST_WEIGHT_ONE, HIST_WEIGHT_COLUMN, HIST_WEIGHT_PHSP, HIST_WEIGHT_GENERATION = 0, 1, 2, 3


class Histogram(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int32), ("grid", ctypes.c_uint64 * HIST_MAX_DIM),
                ("lower", ctypes.c_double * HIST_MAX_DIM),
                ("upper", ctypes.c_double * HIST_MAX_DIM)]


def make_histogram(grid, lower, upper):
    h = Histogram()
    h.ndim = len(grid)
    for i, (g, lo, hi) in enumerate(zip(grid, lower, upper)):
        h.grid[i], h.lower[i], h.upper[i] = int(g), float(lo), float(hi)
    return h


class Request(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("prec", ctypes.c_int32),
                ("out_f32", ctypes.c_int32), ("want_weights", ctypes.c_int32),
                ("decay", Decay), ("functor", Functor),
                ("seed", ctypes.c_uint64), ("first", ctypes.c_uint64), ("n", ctypes.c_uint64),
                ("user_tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_uint64), ("user_tag", ctypes.c_uint64),
                ("status", ctypes.c_int32), ("kind", ctypes.c_int32), ("stats", Stats),
                ("cols", ctypes.c_void_p), ("weights", ctypes.c_void_p),
                ("n", ctypes.c_uint64), ("ld", ctypes.c_uint64), ("submit_s", ctypes.c_double),
                ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double)]


class ServerInfo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in
                ("submitted", "admitted", "completed", "rejected", "batches", "launches",
                 "arena_bytes", "arena_in_use", "arena_peak")]


NOALLOC_MAX_EVENTS = 1 << 28  # HPSP_NOALLOC_MAX_EVENTS
MAX_EVENTS = 1 << 48  # HPSP_MAX_EVENTS: events per call of any entry point

REQ_GENERATE, REQ_AVERAGE = 0, 1
ERR_SHUTDOWN = 7


class HpspError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: hpsp status {code}")


_lib = None

_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_I = ctypes.c_int
_D = ctypes.c_double
_SZ = ctypes.c_size_t

_SIGS = {
    "hpsp_decay_init": (_I, [ctypes.POINTER(Decay), _I, _D, _P, _I]),
    "hpsp_decay_set_mother": (_I, [ctypes.POINTER(Decay), _P]),
    "hpsp_stats_merge": (None, [ctypes.POINTER(Stats), ctypes.POINTER(Stats),
                                ctypes.POINTER(Stats)]),
    "hpsp_error_string": (ctypes.c_char_p, [_I]),
    "hpsp_rng_words": (_I, [_I, _U64, _U64, _U64, _P, _P]),
    "hpsp_generate": (_I, [ctypes.POINTER(Decay), _U64, _U64, _U64, _I, _P, _U64, _I, _P, _P,
                           _P]),
    "hpsp_generate_mothers": (_I, [ctypes.POINTER(Decay), _U64, _U64, _U64, _P, _U64, _I, _P,
                                   _U64, _I, _P, _P, _P]),
    "hpsp_event_weight": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _P, _U64, _U64,
                               _P, _P]),
    "hpsp_max_weight": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _P, _U64, _U64, _P,
                             _SZ, _P, _P]),
    "hpsp_unweight_workspace_size": (_SZ, [_U64]),
    "hpsp_unweight": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _D, _U64, _P, _U64,
                           _U64, _P, _P, _U64, _P, _P, _P, _SZ, _P, _P]),
    "hpsp_average_on_workspace_size": (_SZ, [_U64]),
    "hpsp_average_on": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _U64, _U64, _U64,
                             _I, _P, _SZ, _P, _P]),
    "hpsp_average_on_noalloc": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _U64,
                                     _U64, _U64, _I, _P, _P]),
    "hpsp_average_on_mothers": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _U64, _P,
                                     _U64, _U64, _I, _P, _SZ, _P, _P]),
    "hpsp_evaluate": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _I, _U64, _U64, _U64,
                           _P, _U64, _P]),
    "hpsp_evaluate_mothers": (_I, [ctypes.POINTER(Decay), ctypes.POINTER(Functor), _I, _U64, _P,
                                   _U64, _U64, _U64, _P, _U64, _P]),
    "hpsp_unweight_values": (_I, [_I, _P, _D, _U64, _P, _U64, _U64, _P, _U64, _P, _P, _P, _SZ,
                                  _P, _P]),
    "hpsp_stats_reduce": (_I, [_P, _U64, _P, _P]),
    "hpsp_unweight_regenerate": (_I, [ctypes.POINTER(Decay), _U64, _U64, _U64, _P, _U64, _P,
                                      _U64, _P, _P, _P, _SZ, _P, _P]),
    "hpsp_generate_unweight_workspace_size": (_SZ, [_U64]),
    "hpsp_generate_unweight": (_I, [ctypes.POINTER(Decay), _U64, _U64, _U64, _U64, _P, _U64, _P,
                                    _P, _P, _SZ, _P, _P]),
    "hpsp_generate_unweight_mixed": (_I, [ctypes.POINTER(Decay), _U64, _U64, _U64, _U64, _P, _U64,
                                          _P, _P, _P, _SZ, _P, _P]),
    "hpsp_unweight_sample_workspace_size": (_SZ, [_U64]),
    "hpsp_unweight_sample": (_I, [_I, _P, _U64, _U64, _P, _D, _U64, _U64, _I, _P, _U64, _P, _P,
                                  _P, _SZ, _P, _P]),
    "hpsp_histogram_bins": (_U64, [ctypes.POINTER(Histogram)]),
    "hpsp_histogram_fill": (_I, [ctypes.POINTER(Histogram), _P, _U64, _U64, _P, _P, _P]),
    "hpsp_histogram_fill_events": (_I, [ctypes.POINTER(Histogram), ctypes.POINTER(Decay),
                                        ctypes.POINTER(Functor), _P, _U64, _U64, _I, _P, _P, _P]),
    "hpsp_particles_evaluate": (_I, [ctypes.POINTER(Particles), ctypes.POINTER(Functor), _I,
                                     _U64, _P, _U64, _P]),
    "hpsp_histogram_fill_particles": (_I, [ctypes.POINTER(Histogram), ctypes.POINTER(Particles),
                                           ctypes.POINTER(Functor), _U64, _P, _P, _P]),
    "hpsp_range_init": (_I, [ctypes.POINTER(Decay), _P, _I, _P]),
    "hpsp_range_generate": (_I, [ctypes.POINTER(Decay), _U64, _U64, _U64, _P, _U64, _P, _P]),
    "hpsp_histogram_fill_range": (_I, [ctypes.POINTER(Histogram), ctypes.POINTER(Decay), _U64,
                                       ctypes.POINTER(Functor), _U64, _U64, _I, _P, _P]),
    "hpsp_checksum": (_I, [_P, _U64, _P, _P]),
    "hpsp_malloc": (_I, [ctypes.POINTER(_P), _SZ]),
    "hpsp_free": (_I, [_P]),
    "hpsp_memcpy_to_device": (_I, [_P, _P, _SZ]),
    "hpsp_memcpy_to_host": (_I, [_P, _P, _SZ]),
    "hpsp_memcpy_device": (_I, [_P, _P, _SZ, _P]),
    "hpsp_stream_synchronize": (_I, [_P]),
    "hpsp_memcpy2d_to_host": (_I, [_P, _SZ, _P, _SZ, _SZ, _SZ]),
    "hpsp_memcpy2d_to_device": (_I, [_P, _SZ, _P, _SZ, _SZ, _SZ]),
    "hpsp_server_create": (_I, [ctypes.POINTER(_P), _SZ, _I, _U64]),
    "hpsp_server_submit": (_I, [_P, ctypes.POINTER(Request), ctypes.POINTER(_U64)]),
    "hpsp_server_poll": (_I, [_P, ctypes.POINTER(Completion), _I, ctypes.POINTER(_I), _I]),
    "hpsp_server_release": (_I, [_P, _U64]),
    "hpsp_server_info_get": (_I, [_P, ctypes.POINTER(ServerInfo)]),
    "hpsp_server_destroy": (_I, [_P]),
}


def lib():
    """Load libhpsp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C hydra_amd` "
                "(or __graft_entry__.build()); there is no fallback path")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != OK:
        msg = lib().hpsp_error_string(rc).decode()
        raise HpspError(rc, f"{what} ({msg})")


def ptr(t):
    """Device/host address of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def make_decay(mother_mass, masses, engine="squares3", mother4=None):
    d = Decay()
    m = (ctypes.c_double * len(masses))(*[float(x) for x in masses])
    e = ENGINES.get(engine, engine)
    check(lib().hpsp_decay_init(ctypes.byref(d), len(masses), float(mother_mass), m, e),
          "hpsp_decay_init")
    if mother4 is not None:
        m4 = (ctypes.c_double * 4)(*[float(x) for x in mother4])
        check(lib().hpsp_decay_set_mother(ctypes.byref(d), m4), "hpsp_decay_set_mother")
    return d


def make_range(mother4, masses):
    """Decay description of phase_space_range(mother, masses, seed, length)."""
    d = Decay()
    m = (ctypes.c_double * len(masses))(*[float(x) for x in masses])
    m4 = (ctypes.c_double * 4)(*[float(x) for x in mother4])
    check(lib().hpsp_range_init(ctypes.byref(d), m4, len(masses), m), "hpsp_range_init")
    return d


def make_functor(kind=F_ONE, terms=()):
    """F_BW_SUM: terms = [(M, G, cr, ci, a, b), ...]; F_MASS2, F_MASS and
    F_COS_HELICITY: terms = (a, b) (for the helicity cosine, a is the particle
    whose angle is measured and b its partner; (a, b, first, count) names the
    parent's particles); F_DECAY_PLANE_ANGLE: terms = (d, e, h)."""
    f = Functor()
    f.kind = kind
    if kind in (F_MASS2, F_COS_HELICITY, F_MASS):
        f.a[0], f.b[0] = int(terms[0]), int(terms[1])
        if kind == F_COS_HELICITY and len(terms) == 4:
            f.a[1], f.b[1] = int(terms[2]), int(terms[3])
        return f
    if kind == F_DECAY_PLANE_ANGLE:
        f.a[0], f.b[0], f.a[1] = int(terms[0]), int(terms[1]), int(terms[2])
        return f
    f.nterms = len(terms)
    for k, (M, G, cr, ci, a, b) in enumerate(terms):
        f.m[k], f.g[k], f.cr[k], f.ci[k] = M, G, cr, ci
        f.a[k], f.b[k] = int(a), int(b)
    return f


def functor_array(functors):
    """A ctypes array of Functor from make_functor results."""
    arr = (Functor * len(functors))()
    for j, f in enumerate(functors):
        arr[j] = f
    return arr


def exported_symbols():
    """Names of the functions declared in include/hpsp.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|size_t|const char\*)\s+(hpsp_\w+)\s*\(",
                                 text, re.M)))


class Server:
    """Continuous-batching server (hpsp_server_* in include/hpsp.h)."""

    def __init__(self, hbm_bytes, n_streams=4, chunk_events=0):
        self.h = ctypes.c_void_p()
        check(lib().hpsp_server_create(ctypes.byref(self.h), int(hbm_bytes), n_streams,
                                       int(chunk_events)), "hpsp_server_create")

    def submit(self, req):
        t = ctypes.c_uint64()
        rc = lib().hpsp_server_submit(self.h, ctypes.byref(req), ctypes.byref(t))
        check(rc, "hpsp_server_submit")
        return t.value

    def poll(self, max_n=256, wait_ms=10):
        buf = (Completion * max_n)()
        n = ctypes.c_int()
        check(lib().hpsp_server_poll(self.h, buf, max_n, ctypes.byref(n), wait_ms),
              "hpsp_server_poll")
        return [buf[i] for i in range(n.value)]

    def release(self, ticket):
        check(lib().hpsp_server_release(self.h, int(ticket)), "hpsp_server_release")

    def info(self):
        i = ServerInfo()
        check(lib().hpsp_server_info_get(self.h, ctypes.byref(i)), "hpsp_server_info_get")
        return {k: getattr(i, k) for k, _ in ServerInfo._fields_}

    def close(self):
        if self.h:
            check(lib().hpsp_server_destroy(self.h), "hpsp_server_destroy")
            self.h = ctypes.c_void_p()


def make_request(kind, decay, seed, n, first=0, prec=FP64, out_f32=False, want_weights=False,
                 functor=None, tag=0):
    r = Request()
    r.kind, r.prec, r.out_f32, r.want_weights = kind, prec, int(out_f32), int(want_weights)
    r.decay = decay
    if functor is not None:
        r.functor = functor
    r.seed, r.first, r.n, r.user_tag = seed, first, n, tag
    return r
