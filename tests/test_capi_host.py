"""Host-side checks of libhpsp.so that need no GPU: the library loads, exports
every entry point declared in include/hpsp.h, and its host-only set-up
functions reproduce the reference's constructor arithmetic and errors."""
import ctypes
import math

import numpy as np
import pytest

from hydra_amd import capi


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def test_library_exports_declared_symbols():
    L = capi.lib()
    names = capi.exported_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


@pytest.mark.parametrize("M,m", [(5.279, [3.0969, 0.493677, 0.13957]), (10.0, [0.13957] * 10),
                                 (10.0, [0.999] * 10), (3.0, [1.0, 0.5]),
                                 (20.0, [0.1 + 0.1 * i for i in range(12)])])
def test_decay_init_matches_oracle(oracle, M, m):
    d = capi.make_decay(M, m)
    rc, ecm, mw = oracle.params(M, m)
    assert bits(np.array([d.ecm, d.max_weight])).tolist() == bits(np.array([ecm, mw])).tolist()
    assert d.n == len(m) and list(d.beta) == [0.0, 0.0, 0.0]


def test_decay_init_errors():
    L = capi.lib()
    d = capi.Decay()
    m = (ctypes.c_double * 3)(1.0, 0.5, 0.2)
    assert L.hpsp_decay_init(ctypes.byref(d), 3, 0.0, m, 0) == capi.ERR_MASS
    assert L.hpsp_decay_init(ctypes.byref(d), 3, -1.0, m, 0) == capi.ERR_MASS
    bad = (ctypes.c_double * 3)(1.0, 0.0, 0.2)
    assert L.hpsp_decay_init(ctypes.byref(d), 3, 5.0, bad, 0) == capi.ERR_MASS
    assert L.hpsp_decay_init(ctypes.byref(d), 3, 1.7, m, 0) == capi.ERR_ENERGY
    assert L.hpsp_decay_init(ctypes.byref(d), 1, 5.0, m, 0) == capi.ERR_INVALID_ARGUMENT
    assert L.hpsp_decay_init(ctypes.byref(d), 13, 5.0, m, 0) == capi.ERR_INVALID_ARGUMENT
    assert L.hpsp_decay_init(ctypes.byref(d), 3, 5.0, m, 6) == capi.ERR_INVALID_ARGUMENT
    assert L.hpsp_decay_init(None, 3, 5.0, m, 0) == capi.ERR_INVALID_ARGUMENT
    for code in range(7):
        assert L.hpsp_error_string(code)


@pytest.mark.parametrize("mom", [[5.279, 0, 0, 0], [6.0, 0.3, -0.2, 0.5], [7.0, 0.0, 0.0, 3.0]])
def test_set_mother_matches_oracle(oracle, mom):
    d = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957], mother4=mom)
    assert bits(np.array(list(d.beta))).tolist() == bits(oracle.mother_beta(mom)).tolist()


def test_stats_merge_matches_oracle(oracle):
    x = capi.Stats(2.5, 0.7, 0.1)
    y = capi.Stats(1.25, -0.3, 0.05)
    r = capi.Stats()
    capi.lib().hpsp_stats_merge(x, y, r)
    o = oracle.stats_merge([2.5, 0.7, 0.1], [1.25, -0.3, 0.05])
    assert bits(np.array([r.w, r.mean, r.m2])).tolist() == bits(o).tolist()


def test_workspace_sizes_grow_with_n():
    L = capi.lib()
    assert L.hpsp_unweight_workspace_size(0) > 0
    assert L.hpsp_unweight_workspace_size(1 << 27) >= (1 << 27) // 8
    assert L.hpsp_average_on_workspace_size(1 << 30) >= 24 * ((1 << 30) // 4096)


def test_calls_validate_before_touching_the_device():
    L = capi.lib()
    d = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957])
    # n > 0 with null columns, ld < n, bad precision, bad engine: rejected on the host
    assert L.hpsp_generate(ctypes.byref(d), 1, 0, 10, 0, None, 10, 0, None, None, None) == 1
    assert L.hpsp_generate(ctypes.byref(d), 1, 0, 10, 0, ctypes.c_void_p(64), 5, 0, None, None,
                           None) == 1
    assert L.hpsp_generate(ctypes.byref(d), 1, 0, 10, 7, ctypes.c_void_p(64), 10, 0, None, None,
                           None) == 1
    assert L.hpsp_rng_words(9, 1, 0, 10, ctypes.c_void_p(64), None) == 1
    assert L.hpsp_unweight(ctypes.byref(d), None, -1.0, 1, ctypes.c_void_p(64), 10, 10, None,
                           None, 0, None, None, None, 0, ctypes.c_void_p(64), None) == \
        capi.ERR_WORKSPACE
    assert L.hpsp_average_on(ctypes.byref(d), None, 1, 0, 10, 0, None, 0, ctypes.c_void_p(64),
                             None) == capi.ERR_WORKSPACE
    # n = 0 is a no-op
    assert L.hpsp_generate(ctypes.byref(d), 1, 0, 0, 0, None, 0, 0, None, None, None) == 0


def test_evaluate_validates_before_touching_the_device():
    L = capi.lib()
    d = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957])
    out = ctypes.c_void_p(64)
    one = capi.functor_array([capi.make_functor(capi.F_ONE)])
    nine = capi.functor_array([capi.make_functor(capi.F_ONE)] * 9)
    assert L.hpsp_evaluate(ctypes.byref(d), one, 0, 1, 0, 10, out, 10, None) == 1
    assert L.hpsp_evaluate(ctypes.byref(d), nine, 9, 1, 0, 10, out, 10, None) == 1
    assert L.hpsp_evaluate(ctypes.byref(d), None, 1, 1, 0, 10, out, 10, None) == 1
    assert L.hpsp_evaluate(ctypes.byref(d), one, 1, 1, 0, 10, None, 10, None) == 1
    assert L.hpsp_evaluate(ctypes.byref(d), one, 1, 1, 0, 10, out, 5, None) == 1
    # particle indices beyond what the functor sees
    bad = capi.functor_array([capi.make_functor(capi.F_MASS2, (0, 3))])
    assert L.hpsp_evaluate(ctypes.byref(d), bad, 1, 1, 0, 10, out, 10, None) == 1
    bad = capi.functor_array([capi.make_functor(capi.F_COS_HELICITY, (-1, 2))])
    assert L.hpsp_evaluate(ctypes.byref(d), bad, 1, 1, 0, 10, out, 10, None) == 1
    # with mothers the mother is particle 0, so index N is valid ...
    ok = capi.functor_array([capi.make_functor(capi.F_MASS2, (0, 3))])
    assert L.hpsp_evaluate_mothers(ctypes.byref(d), ok, 1, 1, None, 10, 0, 10, out, 10,
                                   None) == 1  # (null mothers)
    assert L.hpsp_evaluate_mothers(ctypes.byref(d), ok, 1, 1, out, 5, 0, 10, out, 10,
                                   None) == 1  # (mld < n)
    # ... and n = 0 is a no-op
    assert L.hpsp_evaluate(ctypes.byref(d), one, 1, 1, 0, 0, None, 0, None) == 0
    assert L.hpsp_evaluate_mothers(ctypes.byref(d), ok, 1, 1, None, 0, 0, 0, None, 0, None) == 0


def test_histogram_validates_before_touching_the_device():
    L = capi.lib()
    out = ctypes.c_void_p(64)
    good = capi.make_histogram([10, 10], [0, 0], [1, 1])
    assert L.hpsp_histogram_bins(good) == 100
    for bad in (capi.make_histogram([], [], []),
                capi.make_histogram([10, 0], [0, 0], [1, 1]),
                capi.make_histogram([10], [1.0], [1.0]),
                capi.make_histogram([10], [2.0], [1.0]),
                capi.make_histogram([2**21, 2**21], [0, 0], [1, 1])):
        assert L.hpsp_histogram_bins(bad) == 0
        assert L.hpsp_histogram_fill(bad, out, 10, 10, None, out, None) == 1
    h5 = capi.make_histogram([2] * 4, [0] * 4, [1] * 4)
    h5.ndim = 5
    assert L.hpsp_histogram_fill(h5, out, 10, 10, None, out, None) == 1
    assert L.hpsp_histogram_fill(good, None, 10, 10, None, out, None) == 1   # null data
    assert L.hpsp_histogram_fill(good, out, 5, 10, None, out, None) == 1     # ldx < n
    assert L.hpsp_histogram_fill(good, out, 10, 10, None, None, None) == 1   # null contents
    d = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957])
    axes = capi.functor_array([capi.make_functor(capi.F_MASS2, (0, 1)),
                               capi.make_functor(capi.F_MASS2, (1, 3))])  # index 3 > N-1
    assert L.hpsp_histogram_fill_events(good, ctypes.byref(d), axes, out, 10, 10, 0, None, out,
                                        None) == 1
    axes = capi.functor_array([capi.make_functor(capi.F_MASS2, (0, 1)),
                               capi.make_functor(capi.F_MASS2, (1, 2))])
    assert L.hpsp_histogram_fill_events(good, ctypes.byref(d), axes, out, 10, 10, 3, None, out,
                                        None) == 1  # weight mode
    assert L.hpsp_histogram_fill_events(good, ctypes.byref(d), axes, out, 10, 10,
                                        capi.HIST_WEIGHT_COLUMN, None, out, None) == 1


def test_unweight_sample_validates_before_touching_the_device():
    L = capi.lib()
    p = ctypes.c_void_p(64)
    n = 100
    wsb = L.hpsp_unweight_sample_workspace_size(n)
    assert wsb > L.hpsp_unweight_workspace_size(n)
    args = dict(ncols=1, cols=p, ld=n, n=n, w=p, mp=-1.0, seed=1, jump=0, order=0, out=p,
                ldo=n, idx=None, wo=None, ws=p, wsb=wsb, nacc=p)

    def call(**kw):
        a = dict(args)
        a.update(kw)
        return L.hpsp_unweight_sample(a["ncols"], a["cols"], a["ld"], a["n"], a["w"], a["mp"],
                                      a["seed"], a["jump"], a["order"], a["out"], a["ldo"],
                                      a["idx"], a["wo"], a["ws"], a["wsb"], a["nacc"], None)
    assert call(order=2) == 1
    assert call(ncols=-1) == 1
    assert call(w=None) == 1
    assert call(nacc=None) == 1
    assert call(ld=50) == 1
    assert call(ldo=50) == 1
    assert call(ncols=0) == 1          # columns out with no columns
    assert call(wsb=wsb - 8) == capi.ERR_WORKSPACE
    assert call(ws=None) == capi.ERR_WORKSPACE


@pytest.mark.parametrize("m4,m", [([0.493677, 0.0, 0.0, 0.0], [0.13957061] * 3),
                                  ([6.0, 0.3, -0.2, 0.5], [3.0969, 0.493677, 0.13957]),
                                  ([20.0, 0.0, 0.0, 0.0], [0.1 + 0.1 * i for i in range(12)])])
def test_range_init_matches_oracle(oracle, m4, m):
    d = capi.make_range(m4, m)
    rc, ecm, mw, beta = oracle.range_params(m4, m)
    assert rc == 0 and d.engine == capi.MINSTD_RAND and d.n == len(m)
    assert bits(np.array([d.ecm, d.max_weight] + list(d.beta))).tolist() == \
        bits(np.array([ecm, mw] + list(beta))).tolist()


def test_range_validates_before_touching_the_device():
    L = capi.lib()
    out = ctypes.c_void_p(64)
    d = capi.Decay()
    m = (ctypes.c_double * 3)(0.2, 0.2, 0.2)
    m4 = (ctypes.c_double * 4)(1.0, 0.0, 0.0, 0.0)
    assert L.hpsp_range_init(ctypes.byref(d), m4, 3, m) == capi.OK
    bad_m = (ctypes.c_double * 3)(0.2, 0.0, 0.2)
    assert L.hpsp_range_init(ctypes.byref(d), m4, 3, bad_m) == capi.ERR_MASS
    light = (ctypes.c_double * 4)(0.5, 0.0, 0.0, 0.0)
    assert L.hpsp_range_init(ctypes.byref(d), light, 3, m) == capi.ERR_ENERGY
    spacelike = (ctypes.c_double * 4)(1.0, 2.0, 0.0, 0.0)  # mass() = 0
    assert L.hpsp_range_init(ctypes.byref(d), spacelike, 3, m) == capi.ERR_MASS
    assert L.hpsp_range_init(ctypes.byref(d), m4, 1, m) == 1
    assert L.hpsp_range_init(ctypes.byref(d), m4, 13, m) == 1
    r = capi.make_range([1.0, 0, 0, 0], [0.2] * 3)
    assert L.hpsp_range_generate(ctypes.byref(r), 1, 0, 0, None, 0, None, None) == capi.OK
    assert L.hpsp_range_generate(ctypes.byref(r), 1, 0, 10, None, 10, None, None) == 1
    assert L.hpsp_range_generate(ctypes.byref(r), 1, 0, 10, out, 5, None, None) == 1
    # first + n + 3N past 2^64 - 1
    assert L.hpsp_range_generate(ctypes.byref(r), 1, 2**64 - 15, 10, out, 10, None, None) == 1
    # a Generate decay is not a range
    g = capi.make_decay(1.0, [0.2] * 3)
    assert L.hpsp_range_generate(ctypes.byref(g), 1, 0, 10, out, 10, None, None) == 1
    h = capi.make_histogram([10], [0.0], [1.0])
    ax = capi.functor_array([capi.make_functor(capi.F_MASS, (0, 1))])
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(g), 1, ax, 0, 10, 0, out, None) == 1
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(r), 1, ax, 0, 10,
                                       capi.HIST_WEIGHT_COLUMN, out, None) == 1
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(r), 1, None, 0, 10, 0, out, None) == 1
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(r), 1, ax, 0, 10, 0, None, None) == 1
    bad_ax = capi.functor_array([capi.make_functor(capi.F_MASS, (0, 3))])
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(r), 1, bad_ax, 0, 10, 0, out, None) == 1
    kind6 = capi.functor_array([capi.make_functor(capi.F_ONE)])
    kind6[0].kind = 6
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(r), 1, kind6, 0, 10, 0, out, None) == 1


def test_particles_validate_before_touching_the_device():
    L = capi.lib()
    out = ctypes.c_void_p(64)
    p = capi.Particles()
    p.n = 3
    for i in range(3):
        p.cols[i] = 64 + 8 * i
        p.ld[i] = 100
    f = capi.functor_array([capi.make_functor(capi.F_DECAY_PLANE_ANGLE, (0, 1, 2))])
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 1, 0, None, 0, None) == capi.OK
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 1, 10, None, 10, None) == 1  # no out
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 1, 10, out, 5, None) == 1    # ld < n
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 1, 200, out, 200, None) == 1  # ld[i] < n
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 0, 10, out, 10, None) == 1
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 9, 10, out, 10, None) == 1
    bad = capi.functor_array([capi.make_functor(capi.F_DECAY_PLANE_ANGLE, (0, 1, 3))])
    assert L.hpsp_particles_evaluate(ctypes.byref(p), bad, 1, 10, out, 10, None) == 1
    cos_bad = capi.functor_array([capi.make_functor(capi.F_COS_HELICITY, (0, 1, 1, 3))])
    assert L.hpsp_particles_evaluate(ctypes.byref(p), cos_bad, 1, 10, out, 10, None) == 1
    p0 = capi.Particles()
    assert L.hpsp_particles_evaluate(ctypes.byref(p0), f, 1, 10, out, 10, None) == 1
    p.n = 17
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 1, 10, out, 10, None) == 1
    p.n = 3
    p.cols[1] = None
    assert L.hpsp_particles_evaluate(ctypes.byref(p), f, 1, 10, out, 10, None) == 1
    p.cols[1] = 72
    h = capi.make_histogram([10], [0.0], [1.0])
    assert L.hpsp_histogram_fill_particles(h, ctypes.byref(p), bad, 10, None, out, None) == 1
    assert L.hpsp_histogram_fill_particles(h, ctypes.byref(p), f, 10, None, None, None) == 1
    assert L.hpsp_histogram_fill_particles(h, ctypes.byref(p0), f, 10, None, out, None) == 1


def test_plane_angle_only_over_stored_particles():
    """The generation, Evaluate and event-column entry points do not compile
    the decay-plane angle and refuse it."""
    L = capi.lib()
    out = ctypes.c_void_p(64)
    d = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957])
    f = capi.functor_array([capi.make_functor(capi.F_DECAY_PLANE_ANGLE, (0, 1, 2))])
    assert L.hpsp_evaluate(ctypes.byref(d), f, 1, 1, 0, 10, out, 10, None) == 1
    h = capi.make_histogram([10], [0.0], [7.0])
    assert L.hpsp_histogram_fill_events(h, ctypes.byref(d), f, out, 10, 10, 0, None, out,
                                        None) == 1
    r = capi.make_range([5.279, 0, 0, 0], [3.0969, 0.493677, 0.13957])
    assert L.hpsp_histogram_fill_range(h, ctypes.byref(r), 1, f, 0, 10, 0, out, None) == 1


def test_noalloc_limit_matches_header():
    import re
    text = open(capi.HEADER_PATH).read()
    m = re.search(r"#define HPSP_NOALLOC_MAX_EVENTS \(1ULL << (\d+)\)", text)
    assert m and capi.NOALLOC_MAX_EVENTS == 1 << int(m.group(1))


def test_event_limit_matches_header():
    import re
    text = open(capi.HEADER_PATH).read()
    m = re.search(r"#define HPSP_MAX_EVENTS \(1ULL << (\d+)\)", text)
    assert m and capi.MAX_EVENTS == 1 << int(m.group(1))


def test_event_counts_beyond_the_limit_are_refused():
    # sizes derived from such a count could wrap: the workspace sizes say so
    # (SIZE_MAX) and every counting entry point refuses it on the host
    L = capi.lib()
    big = capi.MAX_EVENTS + 1
    size_max = (1 << 64) - 1
    for f in (L.hpsp_unweight_workspace_size, L.hpsp_generate_unweight_workspace_size,
              L.hpsp_unweight_sample_workspace_size, L.hpsp_average_on_workspace_size):
        assert f(big) == size_max
        assert f(capi.MAX_EVENTS) < size_max
    assert L.hpsp_unweight_workspace_size((1 << 64) - 1) == size_max
    d = capi.make_decay(5.279, [3.0969, 0.493677, 0.13957])
    p = ctypes.c_void_p(64)
    ws = ctypes.c_void_p(1 << 20)
    one = capi.functor_array([capi.make_functor(capi.F_ONE)])
    bad = capi.ERR_INVALID_ARGUMENT
    for n in (big, (1 << 64) - 1):
        assert L.hpsp_rng_words(0, 1, 0, n, p, None) == bad
        assert L.hpsp_generate(ctypes.byref(d), 1, 0, n, 0, p, n, 0, None, None, None) == bad
        assert L.hpsp_generate_mothers(ctypes.byref(d), 1, 0, n, p, n, 0, p, n, 0, None, None,
                                       None) == bad
        assert L.hpsp_evaluate(ctypes.byref(d), one, 1, 1, 0, n, p, n, None) == bad
        assert L.hpsp_event_weight(ctypes.byref(d), None, p, n, n, p, None) == bad
        assert L.hpsp_max_weight(ctypes.byref(d), None, p, n, n, ws, size_max, p, None) == bad
        assert L.hpsp_unweight(ctypes.byref(d), None, -1.0, 1, p, n, n, p, None, 0, None, None,
                               ws, size_max, p, None) == bad
        assert L.hpsp_unweight_regenerate(ctypes.byref(d), 1, 0, n, p, 1, p, n, None, None, ws,
                                          size_max, p, None) == bad
        assert L.hpsp_generate_unweight(ctypes.byref(d), 1, 0, n, 1, p, n, None, None, ws,
                                        size_max, p, None) == bad
        assert L.hpsp_generate_unweight_mixed(ctypes.byref(d), 1, 0, n, 1, p, n, None, None, ws,
                                              size_max, p, None) == bad
        assert L.hpsp_average_on(ctypes.byref(d), None, 1, 0, n, 0, ws, size_max, p,
                                 None) == bad
    assert L.hpsp_checksum(p, big * 8, p, None) == bad
