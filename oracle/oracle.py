"""ctypes bindings for the oracle libraries.

TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.

* ``Oracle``     -- the C restatement (oracle/lib/libphsp_oracle.so)
* ``Reference``  -- the reference harness built from the reference headers
                    (oracle/_ref/libhydra_ref*.so); present only where it was
                    built.

Layouts: SoA event columns are numpy arrays of shape (4N, n) (row 4*i+c is
component c of particle i); AoS rows are (n, N, 4).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "lib", "libphsp_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libhydra_ref.so")
REF_NATIVE_LIB = os.path.join(HERE, "_ref", "libhydra_ref_native.so")

ENGINES = {"squares3": 0, "squares4": 1, "philox": 2, "philox_long": 3,
           "threefry": 4, "threefry_long": 5}
MINSTD = 6  # thrust's default_random_engine (minstd_rand): phase_space_range

_P = ctypes.c_void_p
_D = ctypes.c_double
_U64 = ctypes.c_uint64
_I = ctypes.c_int


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


FPAR = 32  # stride of one functor's parameters in the evaluate calls


def functor_params(fid, terms=None):
    """Pack functor parameters.  fid 1: terms = [(M, G, cr, ci, a, b), ...];
    fid 3 and 5: terms = (a, b); fid 4: terms = (d, e, first, count);
    fid 6: terms = (d, e, h)."""
    if fid == 1:
        flat = [float(len(terms))]
        for t in terms:
            flat.extend(float(x) for x in t)
        return np.array(flat, dtype=np.float64)
    if fid in (3, 4, 5, 6):
        return np.array([float(x) for x in terms], dtype=np.float64)
    return np.zeros(1)


def pack_functors(functors):
    """[(fid, terms), ...] -> (int32 ids, float64 parameters at stride FPAR)."""
    ids = np.array([f[0] for f in functors], dtype=np.int32)
    par = np.zeros(FPAR * max(1, len(functors)))
    for j, (fid, terms) in enumerate(functors):
        fp = functor_params(fid, terms)
        par[FPAR * j:FPAR * j + len(fp)] = fp
    return ids, par


class Oracle:
    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path}")
        L = ctypes.CDLL(path)
        self.L = L
        L.po_splitmix64.argtypes = [_P]
        L.po_splitmix64.restype = _U64
        L.po_squares3.argtypes = [_U64, _U64]
        L.po_squares3.restype = ctypes.c_uint32
        L.po_squares4.argtypes = [_U64, _U64]
        L.po_squares4.restype = ctypes.c_uint32
        for nm in ("po_philox2x64", "po_philox4x64", "po_threefry2x64", "po_threefry4x64"):
            getattr(L, nm).argtypes = [ctypes.c_uint, _P, _P, _P]
        L.po_rng_words.argtypes = [_I, _U64, _U64, _U64, _P]
        L.po_uniform.argtypes = [_I, _U64]
        L.po_uniform.restype = _D
        L.po_params.argtypes = [_I, _D, _P, _P, _P]
        L.po_mother_beta.argtypes = [_P, _P]
        L.po_range_params.argtypes = [_I, _P, _P, _P, _P, _P]
        L.po_decay_events.argtypes = [_I, _I, _P, _D, _D, _P, _U64, _P, _U64, _P, _P]
        L.po_generate.argtypes = [_I, _I, _P, _D, _D, _P, _U64, _U64, _U64, _P, _P, _I]
        L.po_generate_mothers.argtypes = [_I, _I, _P, _D, _D, _U64, _U64, _U64, _P, _P, _P, _I]
        L.po_functor.argtypes = [_I, _P, _I, _P]
        L.po_functor.restype = _D
        L.po_event_weight.argtypes = [_I, _P, _D, _I, _P, _P, _U64, _P, _I]
        L.po_unweight.argtypes = [_I, _P, _D, _I, _P, _D, _P, _U64, _U64, _P, _P, _P, _I]
        L.po_unweight.restype = _U64
        L.po_average_on.argtypes = [_I, _I, _P, _D, _D, _P, _U64, _U64, _U64, _I, _P, _I, _P]
        L.po_average_on_mothers.argtypes = [_I, _I, _P, _D, _D, _U64, _U64, _P, _I, _P, _I, _P]
        L.po_stats_merge.argtypes = [_P, _P, _P]
        L.po_permutation_digest.argtypes = [_P, _U64]
        L.po_permutation_digest.restype = _U64
        L.po_unweight_digest.argtypes = [_I, _I, _P, _D, _D, _P, _U64, _U64, _U64, _U64, _P, _I]
        L.po_unweight_digest.restype = _U64
        L.po_evaluate.argtypes = [_I, _I, _P, _D, _D, _P, _U64, _U64, _U64, _I, _P, _P, _P, _I]
        L.po_evaluate_mothers.argtypes = [_I, _I, _P, _U64, _U64, _U64, _P, _I, _P, _P, _P, _I]
        L.po_pair_hash.argtypes = [_U64, _U64]
        L.po_pair_hash.restype = _U64
        L.po_unweight_sample.argtypes = [_I, _P, _U64, _P, _D, _U64, _U64, _I, _P, _P, _P]
        L.po_unweight_sample.restype = _U64
        L.po_histogram_fill.argtypes = [_I, _P, _P, _P, _P, _U64, _U64, _P, _P]
        L.po_hist_key.argtypes = [_I, _P, _P, _P, _P]
        L.po_hist_key.restype = _U64

    # -- engines -----------------------------------------------------------
    def splitmix64(self, x):
        v = _U64(x)
        r = self.L.po_splitmix64(ctypes.byref(v))
        return r, v.value

    def rng_words(self, engine, seed, discard, n):
        out = np.zeros(n, dtype=np.uint64)
        rc = self.L.po_rng_words(ENGINES.get(engine, engine), seed, discard, n, _ptr(out))
        if rc:
            raise ValueError("bad engine")
        return out

    def bijection(self, name, rounds, ctr, key):
        ctr = np.asarray(ctr, dtype=np.uint64)
        key = np.asarray(key, dtype=np.uint64)
        out = np.zeros_like(ctr)
        getattr(self.L, "po_" + name)(rounds, _ptr(ctr), _ptr(key), _ptr(out))
        return out

    def uniform(self, engine, word):
        return self.L.po_uniform(ENGINES.get(engine, engine), int(word))

    # -- decay set-up ------------------------------------------------------
    def params(self, M, masses):
        m = _f64(masses)
        ecm, mw = _D(), _D()
        rc = self.L.po_params(len(m), M, _ptr(m), ctypes.byref(ecm), ctypes.byref(mw))
        return rc, ecm.value, mw.value

    def mother_beta(self, mother4):
        m4 = _f64(mother4)
        b = np.zeros(3)
        self.L.po_mother_beta(_ptr(m4), _ptr(b))
        return b

    def decay_events(self, M, masses, seed, evts, engine=0, mother4=None):
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        assert rc == 0
        beta = self.mother_beta(mother4 if mother4 is not None else [M, 0, 0, 0])
        ev = np.ascontiguousarray(evts, dtype=np.uint64)
        p = np.zeros((len(ev), N, 4))
        wt = np.zeros(len(ev))
        self.L.po_decay_events(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw, _ptr(beta),
                               seed, _ptr(ev), len(ev), _ptr(p), _ptr(wt))
        return p, wt

    def generate(self, M, masses, seed, n, first=0, engine=0, mother4=None, nthreads=0):
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        assert rc == 0
        beta = self.mother_beta(mother4 if mother4 is not None else [M, 0, 0, 0])
        cols = np.zeros((4 * N, n))
        wt = np.zeros(n)
        self.L.po_generate(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw, _ptr(beta), seed,
                           first, n, _ptr(cols), _ptr(wt), nthreads)
        return cols, wt

    def range_params(self, mother4, masses):
        """(rc, fTeCmTm, fWtMax, beta) of GenerateDecay(mother, masses, seed)."""
        m = _f64(masses)
        m4 = _f64(mother4)
        ecm, mw = _D(), _D()
        b = np.zeros(3)
        rc = self.L.po_range_params(len(m), _ptr(m4), _ptr(m), ctypes.byref(ecm),
                                    ctypes.byref(mw), _ptr(b))
        return rc, ecm.value, mw.value, b

    def phase_space_range(self, mother4, masses, seed, n, first=0, nthreads=0):
        """Elements first..first+n-1 of phase_space_range(mother, masses, seed):
        (weights (n,), SoA momenta (4N, n)).  The size_t seed reaches the
        engine as (uint32)(int)seed."""
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw, beta = self.range_params(mother4, m)
        assert rc == 0
        cols = np.zeros((4 * N, n))
        wt = np.zeros(n)
        self.L.po_generate(N, MINSTD, _ptr(m), ecm, mw, _ptr(beta), seed & 0xFFFFFFFF, first,
                           n, _ptr(cols), _ptr(wt), nthreads)
        return wt, cols

    def generate_mothers(self, M, masses, seed, mothers, first=0, engine=0, nthreads=0):
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        assert rc == 0
        mc = _f64(mothers)  # (4, n)
        n = mc.shape[1]
        cols = np.zeros((4 * N, n))
        wt = np.zeros(n)
        self.L.po_generate_mothers(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw, seed, first,
                                   n, _ptr(mc), _ptr(cols), _ptr(wt), nthreads)
        return cols, wt

    def functor(self, fid, fpar, p):
        p = _f64(p)
        fp = _f64(fpar)
        return self.L.po_functor(fid, _ptr(fp), p.shape[0], _ptr(p))

    def event_weight(self, M, masses, cols, fid=-1, fpar=None, nthreads=0):
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        c = _f64(cols)
        n = c.shape[1]
        w = np.zeros(n)
        fp = _f64(fpar if fpar is not None else np.zeros(1))
        self.L.po_event_weight(N, _ptr(m), mw, fid, _ptr(fp), _ptr(c), n, _ptr(w), nthreads)
        return w

    def unweight(self, M, masses, cols, seed, fid=-1, fpar=None, max_value=-1.0, nthreads=0):
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        c = _f64(cols)
        n = c.shape[1]
        out = np.zeros_like(c)
        idx = np.zeros(n, dtype=np.uint64)
        flags = np.zeros(n, dtype=np.uint8)
        fp = _f64(fpar if fpar is not None else np.zeros(1))
        nacc = self.L.po_unweight(N, _ptr(m), mw, fid, _ptr(fp), max_value, _ptr(c), n, seed,
                                  _ptr(out), _ptr(idx), _ptr(flags), nthreads)
        return nacc, out, idx, flags

    def permutation_digest(self, idx):
        """Order-sensitive digest of an output-row -> container-index map."""
        a = np.ascontiguousarray(idx, dtype=np.uint64)
        return int(self.L.po_permutation_digest(_ptr(a), len(a)))

    def unweight_digest(self, M, masses, gen_seed, n, uw_seed, first=0, engine=0, nthreads=0):
        """Generate events first..first+n-1 and Unweight(uw_seed) them without
        storing anything: (accepted count, permutation_digest of the stable
        partition's row -> index map)."""
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        assert rc == 0 and n < (1 << 32)
        beta = self.mother_beta([M, 0, 0, 0])
        d = _U64(0)
        nacc = self.L.po_unweight_digest(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw,
                                         _ptr(beta), gen_seed, first, n, uw_seed,
                                         ctypes.byref(d), nthreads)
        return int(nacc), int(d.value)

    def average_on(self, M, masses, seed, n, fid=0, fpar=None, first=0, engine=0,
                   mother4=None, intervals=1):
        """Returns {W, mean, M2} with the OMP fold order for `intervals`."""
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        assert rc == 0
        beta = self.mother_beta(mother4 if mother4 is not None else [M, 0, 0, 0])
        fp = _f64(fpar if fpar is not None else np.zeros(1))
        out = np.zeros(3)
        self.L.po_average_on(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw, _ptr(beta), seed,
                             first, n, fid, _ptr(fp), intervals, _ptr(out))
        return out

    def average_on_mothers(self, M, masses, seed, mothers, fid=0, fpar=None, engine=0,
                           intervals=1):
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        mc = _f64(mothers)
        fp = _f64(fpar if fpar is not None else np.zeros(1))
        out = np.zeros(3)
        self.L.po_average_on_mothers(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw, seed,
                                     mc.shape[1], _ptr(mc), fid, _ptr(fp), intervals, _ptr(out))
        return out

    def evaluate(self, M, masses, seed, n, functors, first=0, engine=0, mother4=None,
                 nthreads=0):
        """PhaseSpace::Evaluate(mother, result, functors): (1+nf, n) columns."""
        m = _f64(masses)
        N = len(m)
        rc, ecm, mw = self.params(M, m)
        assert rc == 0
        beta = self.mother_beta(mother4 if mother4 is not None else [M, 0, 0, 0])
        ids, par = pack_functors(functors)
        out = np.zeros((1 + len(ids), n))
        self.L.po_evaluate(N, ENGINES.get(engine, engine), _ptr(m), ecm, mw, _ptr(beta), seed,
                           first, n, len(ids), _ptr(ids), _ptr(par), _ptr(out), nthreads)
        return out

    def evaluate_mothers(self, masses, seed, mothers, functors, first=0, engine=0, nthreads=0):
        """PhaseSpace::Evaluate(mothers, result, functors); mothers (4, n)."""
        m = _f64(masses)
        N = len(m)
        mc = _f64(mothers)
        n = mc.shape[1]
        ids, par = pack_functors(functors)
        out = np.zeros((1 + len(ids), n))
        self.L.po_evaluate_mothers(N, ENGINES.get(engine, engine), _ptr(m), seed, first, n,
                                   _ptr(mc), len(ids), _ptr(ids), _ptr(par), _ptr(out),
                                   nthreads)
        return out

    def histogram_fill(self, grid, lower, upper, x, w=None):
        """DenseHistogram::Fill: x (ndim, n) values, optional weights; returns
        the nbins + 2 contents (underflow, overflow last)."""
        g = np.ascontiguousarray(grid, dtype=np.uint64)
        lo, hi = _f64(lower), _f64(upper)
        xv = _f64(np.atleast_2d(x))
        ndim, n = xv.shape
        out = np.zeros(int(np.prod(g)) + 2)
        self.L.po_histogram_fill(ndim, _ptr(g), _ptr(lo), _ptr(hi), _ptr(xv), n, n,
                                 _ptr(None if w is None else _f64(w)), _ptr(out))
        return out

    def hist_key(self, grid, lower, upper, x):
        g = np.ascontiguousarray(grid, dtype=np.uint64)
        return self.L.po_hist_key(len(g), _ptr(g), _ptr(_f64(lower)), _ptr(_f64(upper)),
                                  _ptr(_f64(x)))

    def unweight_sample(self, cols, w, max_pdf=-1.0, seed=0x8EC74D321E6B5A27, jump=0, order=0):
        """hydra::unweight on value columns (ncols, n) and weights; returns
        (kept, reordered columns, original index of each output row).
        order 0: both parts stable; 1: the reference OMP build's order."""
        c = _f64(np.atleast_2d(cols))
        ncols, n = c.shape
        out = np.zeros_like(c)
        idx = np.zeros(n, dtype=np.uint64)
        kept = self.L.po_unweight_sample(ncols, _ptr(c), n, _ptr(_f64(w)), max_pdf, seed, jump,
                                         order, _ptr(out), _ptr(idx), None)
        return kept, out, idx

    def pair_hash(self, evt, seed):
        return self.L.po_pair_hash(evt, seed)

    def stats_merge(self, x, y):
        x = _f64(x)
        y = _f64(y)
        o = np.zeros(3)
        self.L.po_stats_merge(_ptr(x), _ptr(y), _ptr(o))
        return o


class Reference:
    """The reference's own templates, compiled from its headers (OMP backend)."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"reference harness not built: {path}")
        L = ctypes.CDLL(path)
        self.L = L
        L.hr_set_threads.argtypes = [_I]
        L.hr_rng_words.argtypes = [_I, _U64, _U64, _U64, _P]
        L.hr_phsp_params.argtypes = [_I, _D, _P, _P, _P, _P, _P]
        L.hr_decay_events.argtypes = [_I, _I, _D, _P, _P, _U64, _P, _U64, _P, _P]
        L.hr_generate.argtypes = [_I, _D, _P, _P, _I, ctypes.c_int32, _U64, _P, _P, _P]
        L.hr_generate_mothers.argtypes = [_I, _D, _P, ctypes.c_int32, _U64, _P, _P, _P]
        L.hr_unweight.argtypes = [_I, _D, _P, _P, ctypes.c_int32, _U64, _I, _I, _P, _D, _U64,
                                  _P, _P, _P, _P]
        L.hr_average_on.argtypes = [_I, _D, _P, _P, _I, ctypes.c_int32, _U64, _I, _P, _P, _P,
                                    _P]
        L.hr_evaluate.argtypes = [_I, _D, _P, _P, _I, ctypes.c_int32, _U64, _I, _P, _P, _P, _P]
        L.hr_evaluate_mothers.argtypes = [_I, _D, _P, ctypes.c_int32, _U64, _P, _I, _P, _P, _P]
        L.hr_histogram_fill.argtypes = [_I, _P, _P, _P, _P, _U64, _P, _P, _P]
        L.hr_unweight_sample.argtypes = [_P, _U64, _D, _D, _D, _U64, _U64, _P, _P, _P]
        L.hr_phase_space_range.argtypes = [_I, _P, _P, _U64, _U64, _U64, _P, _P]
        L.hr_chain.argtypes = [_U64, _I, _P, _P, _P, _P, _P]
        L.hr_range_histogram3.argtypes = [_P, _P, _U64, _U64, _I, _I, _I, _U64, _D, _D, _P, _P]

    def set_threads(self, n):
        return self.L.hr_set_threads(n)

    def phase_space_range(self, mother4, masses, seed, n, first=0):
        """Elements first..first+n-1 of hydra::phase_space_range: (weights (n,),
        AoS momenta (n, N, 4))."""
        m = _f64(masses)
        N = len(m)
        p = np.zeros((n, N, 4))
        w = np.zeros(n)
        rc = self.L.hr_phase_space_range(N, _ptr(_f64(mother4)), _ptr(m), seed, first, n,
                                         _ptr(p), _ptr(w))
        if rc:
            raise ValueError("bad N")
        return w, p

    def chain(self, n, functor=(0, None)):
        """The decay chain of phsp_chain.inl: (melded rows (n, 5, 4), functor
        values (n,), three-body weights (n,), seconds of the two Generates)."""
        fid, terms = functor
        p = np.zeros((n, 5, 4))
        f = np.zeros(n)
        w = np.zeros(n)
        sec = _D()
        self.L.hr_chain(n, fid, _ptr(functor_params(fid, terms)), _ptr(p), _ptr(f), _ptr(w),
                        ctypes.byref(sec))
        return p, f, w, sec.value

    def range_histogram3(self, mother4, masses, seed, n, a, b, bins, lo, hi, weighted=False):
        """DenseHistogram<double,1>.Fill(phase_space_range(...) | (p_a+p_b).mass()
        [, weights]): (contents (bins+2,), seconds of the Fill)."""
        out = np.zeros(bins + 2)
        sec = _D()
        rc = self.L.hr_range_histogram3(_ptr(_f64(mother4)), _ptr(_f64(masses)), seed, n, a, b,
                                        1 if weighted else 0, bins, lo, hi, _ptr(out),
                                        ctypes.byref(sec))
        if rc:
            raise ValueError("bad pair")
        return out, sec.value

    def rng_words(self, engine, seed, discard, n):
        out = np.zeros(n, dtype=np.uint64)
        self.L.hr_rng_words(ENGINES.get(engine, engine), seed, discard, n, _ptr(out))
        return out

    def params(self, M, masses):
        m = _f64(masses)
        ecm, mw, sf, sg = _D(), _D(), _U64(), ctypes.c_int32()
        rc = self.L.hr_phsp_params(len(m), M, _ptr(m), ctypes.byref(ecm), ctypes.byref(mw),
                                   ctypes.byref(sf), ctypes.byref(sg))
        return rc, ecm.value, mw.value, sf.value, sg.value

    def decay_events(self, M, masses, seed, evts, engine=0, mother4=None):
        m = _f64(masses)
        N = len(m)
        m4 = _f64(mother4 if mother4 is not None else [M, 0, 0, 0])
        ev = np.ascontiguousarray(evts, dtype=np.uint64)
        p = np.zeros((len(ev), N, 4))
        wt = np.zeros(len(ev))
        self.L.hr_decay_events(N, ENGINES.get(engine, engine), M, _ptr(m4), _ptr(m), seed,
                               _ptr(ev), len(ev), _ptr(p), _ptr(wt))
        return p, wt

    def generate(self, M, masses, seed, n, mother4=None, want_output=True):
        """Returns (rows (n,N,4), weights, seconds); seed=None uses the default seed."""
        m = _f64(masses)
        N = len(m)
        m4 = _f64(mother4 if mother4 is not None else [M, 0, 0, 0])
        p = np.zeros((n, N, 4)) if want_output else None
        w = np.zeros(n) if want_output else None
        sec = _D()
        use = 0 if seed is None else 1
        self.L.hr_generate(N, M, _ptr(m4), _ptr(m), use, 0 if seed is None else seed, n,
                           _ptr(p), _ptr(w), ctypes.byref(sec))
        return p, w, sec.value

    def generate_mothers(self, M, masses, seed, mothers):
        m = _f64(masses)
        N = len(m)
        mr = np.ascontiguousarray(np.asarray(mothers, dtype=np.float64).T)  # (n,4)
        n = mr.shape[0]
        p = np.zeros((n, N, 4))
        w = np.zeros(n)
        self.L.hr_generate_mothers(N, M, _ptr(m), seed, n, _ptr(mr), _ptr(p), _ptr(w))
        return p, w

    def unweight(self, M, masses, seed, n, uw_seed, fid=None, fpar=None, max_weight=-1.0,
                 mother4=None, want_output=True):
        m = _f64(masses)
        N = len(m)
        m4 = _f64(mother4 if mother4 is not None else [M, 0, 0, 0])
        p = np.zeros((n, N, 4)) if want_output else None
        idx = np.zeros(n, dtype=np.uint64) if want_output else None
        nacc = _U64()
        sec = _D()
        fp = _f64(fpar if fpar is not None else np.zeros(1))
        rc = self.L.hr_unweight(N, M, _ptr(m4), _ptr(m), seed, n, 0 if fid is None else 1,
                                0 if fid is None else fid, _ptr(fp), max_weight, uw_seed,
                                _ptr(p), _ptr(idx), ctypes.byref(nacc), ctypes.byref(sec))
        if rc:
            raise RuntimeError(f"hr_unweight failed: {rc}")
        return nacc.value, p, idx, sec.value

    def average_on(self, M, masses, seed, n, fid=0, fpar=None, mother4=None):
        m = _f64(masses)
        N = len(m)
        m4 = _f64(mother4 if mother4 is not None else [M, 0, 0, 0])
        fp = _f64(fpar if fpar is not None else np.zeros(1))
        mean, sig, sec = _D(), _D(), _D()
        use = 0 if seed is None else 1
        self.L.hr_average_on(N, M, _ptr(m4), _ptr(m), use, 0 if seed is None else seed, n, fid,
                             _ptr(fp), ctypes.byref(mean), ctypes.byref(sig), ctypes.byref(sec))
        return mean.value, sig.value, sec.value


def _ref_evaluate(L, M, masses, seed, n, functors, mother4=None):
    m = _f64(masses)
    N = len(m)
    m4 = _f64(mother4 if mother4 is not None else [M, 0, 0, 0])
    ids, par = pack_functors(functors)
    out = np.zeros((n, 1 + len(ids)))
    sec = _D()
    use = 0 if seed is None else 1
    rc = L.hr_evaluate(N, M, _ptr(m4), _ptr(m), use, 0 if seed is None else seed, n, len(ids),
                       _ptr(ids), _ptr(par), _ptr(out), ctypes.byref(sec))
    if rc:
        raise RuntimeError(f"hr_evaluate failed: {rc}")
    return np.ascontiguousarray(out.T), sec.value


def _ref_evaluate_mothers(L, M, masses, seed, mothers, functors):
    m = _f64(masses)
    N = len(m)
    mr = np.ascontiguousarray(np.asarray(mothers, dtype=np.float64).T)  # (n,4)
    n = mr.shape[0]
    ids, par = pack_functors(functors)
    out = np.zeros((n, 1 + len(ids)))
    rc = L.hr_evaluate_mothers(N, M, _ptr(m), seed, n, _ptr(mr), len(ids), _ptr(ids),
                               _ptr(par), _ptr(out))
    if rc:
        raise RuntimeError(f"hr_evaluate_mothers failed: {rc}")
    return np.ascontiguousarray(out.T)


Reference.evaluate = lambda self, M, masses, seed, n, functors, mother4=None: \
    _ref_evaluate(self.L, M, masses, seed, n, functors, mother4)
Reference.evaluate.__doc__ = "PhaseSpace::Evaluate(mother, result, functors) -> ((1+nf, n), s)"
Reference.evaluate_mothers = lambda self, M, masses, seed, mothers, functors: \
    _ref_evaluate_mothers(self.L, M, masses, seed, mothers, functors)


def _ref_histogram_fill(L, grid, lower, upper, x, w=None):
    g = np.ascontiguousarray(grid, dtype=np.uint64)
    xv = _f64(np.atleast_2d(x))
    ndim, n = xv.shape
    out = np.zeros(int(np.prod(g)) + 2)
    sec = _D()
    rc = L.hr_histogram_fill(ndim, _ptr(g), _ptr(_f64(lower)), _ptr(_f64(upper)), _ptr(xv), n,
                             _ptr(None if w is None else _f64(w)), _ptr(out), ctypes.byref(sec))
    if rc:
        raise RuntimeError(f"hr_histogram_fill failed: {rc}")
    return out, sec.value


Reference.histogram_fill = lambda self, grid, lower, upper, x, w=None: \
    _ref_histogram_fill(self.L, grid, lower, upper, x, w)


def _ref_unweight_sample(L, x, mean, sigma, max_pdf, seed, jump):
    xv = _f64(x)
    n = len(xv)
    xo = np.zeros(n)
    w = np.zeros(n)
    kept = _U64()
    rc = L.hr_unweight_sample(_ptr(xv), n, mean, sigma, max_pdf, seed, jump, _ptr(xo), _ptr(w),
                              ctypes.byref(kept))
    if rc:
        raise RuntimeError(f"hr_unweight_sample failed: {rc}")
    return kept.value, xo, w


Reference.unweight_sample = lambda self, x, mean, sigma, max_pdf=-1.0, \
    seed=0x8EC74D321E6B5A27, jump=0: _ref_unweight_sample(self.L, x, mean, sigma, max_pdf, seed,
                                                          jump)
Reference.unweight_sample.__doc__ = ("hydra::unweight(device::sys, x, Gaussian(mean, sigma), max, "
                                     "seed, jump) -> (kept, reordered x, gauss(x))")


def rows_to_cols(p):
    """(n, N, 4) AoS rows -> (4N, n) SoA columns."""
    n, N, _ = p.shape
    return np.ascontiguousarray(p.reshape(n, 4 * N).T)


def cols_to_rows(c, N):
    return np.ascontiguousarray(np.asarray(c).T.reshape(-1, N, 4))
