// Reference harness: drives the reference Hydra 4.0.1 headers (mounted at
// build time under /root/reference, OMP device backend) through a flat C
// interface so that golden vectors can be written and the reference CPU path
// can be timed.  This file contains no reference code; it only instantiates
// the reference's own templates:
//
//   hydra::PhaseSpace<N,GRND>            hydra/PhaseSpace.h:86, detail/PhaseSpace.inl:42-455
//   hydra::Decays<tuple<P...>,device>    hydra/Decays.h:67-651, detail/Decays.inl:415-511
//   hydra::detail::DecayMother::process  hydra/detail/functors/DecayMother.h:163-276
//   engines squares3/4, philox(_long), threefry(_long)
//                                        hydra/detail/random/{squares3,squares4,EngineR123}.h
//   hydra::phase_space_range             hydra/detail/PhaseSpaceRange.inl:44-72,
//                                        detail/functors/GenerateDecay.h
//
// It is test infrastructure: only tests/, bench.py's CPU-baseline leg and the
// golden-vector script load the library built from it (oracle/_ref/).
#include <hydra/Types.h>
#include <hydra/Vector4R.h>
#include <hydra/PhaseSpace.h>
#include <hydra/Lambda.h>
#include <hydra/Tuple.h>
#include <hydra/host/System.h>
#include <hydra/device/System.h>
#include <hydra/Decays.h>
#include <hydra/multivector.h>
#include <hydra/DenseHistogram.h>
#include <hydra/Random.h>
#include <hydra/Parameter.h>
#include <hydra/functions/Gaussian.h>
#include <hydra/detail/Hash.h>
#include <hydra/Range.h>

#include <omp.h>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

declarg(P0, hydra::Vector4R)
declarg(P1, hydra::Vector4R)
declarg(P2, hydra::Vector4R)
declarg(P3, hydra::Vector4R)
declarg(P4, hydra::Vector4R)
declarg(P5, hydra::Vector4R)
declarg(P6, hydra::Vector4R)
declarg(P7, hydra::Vector4R)
declarg(P8, hydra::Vector4R)
declarg(P9, hydra::Vector4R)
declarg(P10, hydra::Vector4R)
declarg(P11, hydra::Vector4R)

using namespace hydra::arguments;

namespace {

// ---------------------------------------------------------------------------
// Particle type lists P0..P(N-1)
template <size_t I> struct ptype;
template <> struct ptype<0> { using type = P0; };
template <> struct ptype<1> { using type = P1; };
template <> struct ptype<2> { using type = P2; };
template <> struct ptype<3> { using type = P3; };
template <> struct ptype<4> { using type = P4; };
template <> struct ptype<5> { using type = P5; };
template <> struct ptype<6> { using type = P6; };
template <> struct ptype<7> { using type = P7; };
template <> struct ptype<8> { using type = P8; };
template <> struct ptype<9> { using type = P9; };
template <> struct ptype<10> { using type = P10; };
template <> struct ptype<11> { using type = P11; };

template <typename Seq> struct tuple_of;
template <size_t... I> struct tuple_of<std::index_sequence<I...>> {
  using type = hydra::tuple<typename ptype<I>::type...>;
};
template <size_t N>
using decays_t = hydra::Decays<typename tuple_of<std::make_index_sequence<N>>::type,
                               hydra::device::sys_t>;

template <size_t N, typename Row, size_t... I>
inline void row_to_array(Row const& row, double* out, std::index_sequence<I...>) {
  hydra::Vector4R v[N] = {hydra::Vector4R(
      typename ptype<I>::type(hydra::get<I>(row)).Value())...};
  for (size_t p = 0; p < N; ++p)
    for (int c = 0; c < 4; ++c) out[4 * p + c] = v[p].get(c);
}

// Elements 1..N of a (weight, p_1, ..., p_N) tuple into an array.
template <size_t N, typename Tuple, size_t... I>
inline void assign_tail(Tuple const& t, hydra::Vector4R* p, std::index_sequence<I...>) {
  ((p[I] = hydra::get<I + 1>(t)), ...);
}

// ---------------------------------------------------------------------------
// Functors used by AverageOn / Unweight(functor).  Their arithmetic is spelled
// out so the GPU kernels and the C restatement can follow the same sequence.
//   id 0: f = 1
//   id 1: f = |sum_k c_k / (M_k^2 - s_k - i M_k G_k)|^2,  s_k = (p_a + p_b)^2
//         par = {K, (M, G, cr, ci, a, b) x K}
//   id 2: f = (p0 + p1)^2
//   id 3: f = (p_a + p_b)^2                             par = {a, b}
//   id 4: helicity cosine of particle d in the pair (d, e) within
//         P = p_first + ... + p_(first+count-1), written as
//         phsp_evaluating_functor.inl:160-176 writes it  par = {d, e, first, count}
//   id 5: f = (p_a + p_b).mass()                        par = {a, b}
//   id 6: angle between the planes of (d, e) and of h about d + e,
//         written as phsp_chain.inl:191-207 writes it   par = {d, e, h}
struct Functor {
  int id;
  int K;
  double M[4], G[4], cr[4], ci[4];
  int a[4], b[4];

  template <typename V>
  double eval(V const* p) const {
    if (id == 0) return 1.0;
    if (id == 2) {
      hydra::Vector4R s = p[0] + p[1];
      return s.mass2();
    }
    if (id == 3) {
      hydra::Vector4R s = p[a[0]] + p[b[0]];
      return s.mass2();
    }
    if (id == 5) {
      hydra::Vector4R s = p[a[0]] + p[b[0]];
      return s.mass();
    }
    if (id == 6) {
      hydra::Vector4R d = p[a[0]], h = p[a[1]];
      hydra::Vector4R K = d + p[b[0]];
      hydra::Vector4R d_perp = d - (K.dot(d) / K.dot(K)) * K;
      hydra::Vector4R h_perp = h - (K.dot(h) / K.dot(K)) * K;
      hydra::Vector4R d_prime = K.cross(d_perp);
      d_perp = d_perp / d_perp.d3mag();
      d_prime = d_prime / d_prime.d3mag();
      double chi = atan2(d_prime.dot(h_perp), d_perp.dot(h_perp));
      if (chi < 0.0) chi += 2.0 * PI;
      return chi;
    }
    if (id == 4) {
      hydra::Vector4R P = p[a[1]];
      for (int i = a[1] + 1; i < a[1] + b[1]; ++i) P = P + p[i];
      hydra::Vector4R q = p[a[0]] + p[b[0]];
      hydra::Vector4R d = p[a[0]];
      double pd = P * d;
      double pq = P * q;
      double qd = q * d;
      double mp2 = P.mass2();
      double mq2 = q.mass2();
      double md2 = d.mass2();
      return (pd * mq2 - pq * qd) / sqrt((pq * pq - mq2 * mp2) * (qd * qd - mq2 * md2));
    }
    double Ar = 0.0, Ai = 0.0;
    for (int k = 0; k < K; ++k) {
      hydra::Vector4R q = p[a[k]] + p[b[k]];
      double s = q.mass2();
      double dr = M[k] * M[k] - s;
      double di = -(M[k] * G[k]);
      double den = dr * dr + di * di;
      double br = dr / den;
      double bi = -di / den;
      double tr = cr[k] * br - ci[k] * bi;
      double ti = cr[k] * bi + ci[k] * br;
      Ar += tr;
      Ai += ti;
    }
    return Ar * Ar + Ai * Ai;
  }
};

Functor make_functor(int id, const double* par) {
  Functor f{};
  f.id = id;
  if (id == 3 || id == 5) {
    f.a[0] = (int)par[0];
    f.b[0] = (int)par[1];
  }
  if (id == 6) {
    f.a[0] = (int)par[0];
    f.b[0] = (int)par[1];
    f.a[1] = (int)par[2];
  }
  if (id == 4) {
    f.a[0] = (int)par[0];
    f.b[0] = (int)par[1];
    f.a[1] = (int)par[2];
    f.b[1] = (int)par[3];
  }
  if (id == 1) {
    f.K = (int)par[0];
    for (int k = 0; k < f.K; ++k) {
      const double* q = par + 1 + 6 * k;
      f.M[k] = q[0];
      f.G[k] = q[1];
      f.cr[k] = q[2];
      f.ci[k] = q[3];
      f.a[k] = (int)q[4];
      f.b[k] = (int)q[5];
    }
  }
  return f;
}

// Callable handed to PhaseSpace::AverageOn: receives a tuple of N Vector4R.
template <size_t N>
struct TupleFunctor {
  typedef double return_type;  // PhaseSpace::Evaluate reads FUNCTOR::return_type
  Functor f;
  template <typename Tuple, size_t... I>
  double call(Tuple const& t, std::index_sequence<I...>) const {
    hydra::Vector4R v[N] = {hydra::Vector4R(hydra::get<I>(t))...};
    return f.eval(v);
  }
  template <typename Tuple>
  __hydra_host__ __hydra_device__ double operator()(Tuple const& t) const {
    return call(t, std::make_index_sequence<N>{});
  }
};

// Hydra functor (typed arguments) for Decays::Unweight(functor, ...).
template <size_t N, typename Sig> struct HydraF;
template <size_t N, typename... Ps>
struct HydraF<N, void(Ps...)>
    : public hydra::BaseFunctor<HydraF<N, void(Ps...)>, double(Ps...), 0> {
  typedef hydra::BaseFunctor<HydraF<N, void(Ps...)>, double(Ps...), 0> super_type;
  Functor f;
  HydraF() = default;
  explicit HydraF(Functor const& ff) : super_type(), f(ff) {}
  __hydra_host__ __hydra_device__ HydraF(HydraF const& o) : super_type(o), f(o.f) {}
  __hydra_host__ __hydra_device__ HydraF& operator=(HydraF const& o) {
    if (this == &o) return *this;
    super_type::operator=(o);
    f = o.f;
    return *this;
  }
  __hydra_host__ __hydra_device__ double Evaluate(Ps... p) const {
    hydra::Vector4R v[N] = {hydra::Vector4R(p.Value())...};
    return f.eval(v);
  }
};
template <typename Seq> struct sig_of;
template <size_t... I> struct sig_of<std::index_sequence<I...>> {
  using type = void(typename ptype<I>::type...);
};
template <size_t N>
using hydra_functor_t = HydraF<N, typename sig_of<std::make_index_sequence<N>>::type>;

// ---------------------------------------------------------------------------
template <typename GRND>
void engine_words(uint64_t seed, uint64_t discard, uint64_t n, uint64_t* out) {
  GRND eng(seed);
  eng.discard(discard);
  for (uint64_t i = 0; i < n; ++i) out[i] = (uint64_t)eng();
}

template <size_t N, typename GRND>
void decay_events(double M, const double* mother4, const double* masses, uint64_t seed,
                  const uint64_t* evts, uint64_t nevt, double* out_p, double* out_wt) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N, GRND> phsp(M, m);
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  hydra::detail::DecayMother<N, GRND> dm(mother, m, phsp.GetMaxWeight(), phsp.GetECM(),
                                         (size_t)seed);
  for (uint64_t k = 0; k < nevt; ++k) {
    hydra::Vector4R p[N];
    double wt = dm.process((size_t)evts[k], p);
    for (size_t i = 0; i < N; ++i)
      for (int c = 0; c < 4; ++c) out_p[(k * N + i) * 4 + c] = p[i].get(c);
    out_wt[k] = wt;
  }
}

template <size_t N>
void copy_out(decays_t<N>& ev, double* out_p, double* out_w) {
  const size_t n = ev.size();
  if (out_p)
    for (size_t i = 0; i < n; ++i)
      row_to_array<N>(ev[i], out_p + i * N * 4, std::make_index_sequence<N>{});
  if (out_w) {
    auto w = ev | ev.GetEventWeightFunctor();
    for (size_t i = 0; i < n; ++i) out_w[i] = w[i];
  }
}

double now_s() {
  return std::chrono::duration<double>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <size_t N>
void generate(double M, const double* mother4, const double* masses, int use_seed,
              int32_t seed, uint64_t n, double* out_p, double* out_w, double* seconds) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N> phsp(M, m);
  if (use_seed) phsp.SetSeed(seed);
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  decays_t<N> ev(M, m, n);
  double t0 = now_s();
  phsp.Generate(mother, ev);
  double t1 = now_s();
  if (seconds) *seconds = t1 - t0;
  copy_out<N>(ev, out_p, out_w);
}

template <size_t N>
void generate_mothers(double M, const double* masses, int32_t seed, uint64_t n,
                      const double* mothers /* n x 4 */, double* out_p, double* out_w) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N> phsp(M, m);
  phsp.SetSeed(seed);
  hydra::device::vector<hydra::Vector4R> mv(n);
  for (uint64_t i = 0; i < n; ++i)
    mv[i] = hydra::Vector4R(mothers[4 * i], mothers[4 * i + 1], mothers[4 * i + 2],
                            mothers[4 * i + 3]);
  decays_t<N> ev(M, m, n);
  phsp.Generate(mv, ev);
  copy_out<N>(ev, out_p, out_w);
}

template <size_t N>
uint64_t permuted_indices(decays_t<N>& ev, std::vector<double> const& orig,
                          double* out_p, uint64_t* out_idx) {
  const size_t n = ev.size();
  std::map<std::string, uint64_t> where;
  for (size_t i = 0; i < n; ++i)
    where.emplace(std::string((const char*)&orig[i * N * 4], sizeof(double) * N * 4), i);
  std::vector<double> row(N * 4);
  for (size_t i = 0; i < n; ++i) {
    row_to_array<N>(ev[i], row.data(), std::make_index_sequence<N>{});
    if (out_p) std::memcpy(out_p + i * N * 4, row.data(), sizeof(double) * N * 4);
    auto it = where.find(std::string((const char*)row.data(), sizeof(double) * N * 4));
    if (it == where.end()) return 1;
    out_idx[i] = it->second;
  }
  return 0;
}

template <size_t N>
int unweight(double M, const double* mother4, const double* masses, int32_t seed, uint64_t n,
             int with_functor, int fid, const double* fpar, double max_weight,
             uint64_t uw_seed, double* out_p, uint64_t* out_idx, uint64_t* n_acc,
             double* seconds) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N> phsp(M, m);
  phsp.SetSeed(seed);
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  decays_t<N> ev(M, m, n);
  phsp.Generate(mother, ev);
  std::vector<double> orig;
  if (out_idx) {
    orig.resize(n * N * 4);
    copy_out<N>(ev, orig.data(), nullptr);
  }
  double t0 = now_s();
  size_t acc;
  if (with_functor) {
    hydra_functor_t<N> hf(make_functor(fid, fpar));
    acc = ev.Unweight(hf, max_weight, (size_t)uw_seed).size();
  } else {
    acc = ev.Unweight((size_t)uw_seed).size();
  }
  double t1 = now_s();
  if (seconds) *seconds = t1 - t0;
  *n_acc = acc;
  if (out_idx) return (int)permuted_indices<N>(ev, orig, out_p, out_idx);
  return 0;
}

template <size_t N>
void average_on(double M, const double* mother4, const double* masses, int use_seed,
                int32_t seed, uint64_t n, int fid, const double* fpar, double* mean,
                double* sigma, double* seconds) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N> phsp(M, m);
  if (use_seed) phsp.SetSeed(seed);
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  TupleFunctor<N> tf{make_functor(fid, fpar)};
  double t0 = now_s();
  auto r = phsp.AverageOn(hydra::device::sys, mother, tf, n);
  double t1 = now_s();
  if (seconds) *seconds = t1 - t0;
  *mean = r.first;
  *sigma = r.second;
}

constexpr int kFpar = 32;  // stride of a functor's parameters

template <size_t N>
void evaluate(double M, const double* mother4, const double* masses, int use_seed, int32_t seed,
              uint64_t n, int nf, const int* fids, const double* fpar, double* out,
              double* seconds) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N> phsp(M, m);
  if (use_seed) phsp.SetSeed(seed);
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  TupleFunctor<N> f0{make_functor(fids[0], fpar)};
  TupleFunctor<N> f1{make_functor(nf > 1 ? fids[1] : 0, fpar + kFpar)};
  double t0 = now_s();
  if (nf == 1) {
    hydra::multivector<hydra::tuple<double, double>, hydra::device::sys_t> r(n);
    phsp.Evaluate(mother, r, f0);
    if (seconds) *seconds = now_s() - t0;
    for (uint64_t i = 0; i < n; ++i) {
      auto row = r[i];
      out[i * 2] = hydra::get<0>(row);
      out[i * 2 + 1] = hydra::get<1>(row);
    }
  } else {
    hydra::multivector<hydra::tuple<double, double, double>, hydra::device::sys_t> r(n);
    phsp.Evaluate(mother, r, f0, f1);
    if (seconds) *seconds = now_s() - t0;
    for (uint64_t i = 0; i < n; ++i) {
      auto row = r[i];
      out[i * 3] = hydra::get<0>(row);
      out[i * 3 + 1] = hydra::get<1>(row);
      out[i * 3 + 2] = hydra::get<2>(row);
    }
  }
}

template <size_t N>
void evaluate_mothers(double M, const double* masses, int32_t seed, uint64_t n,
                      const double* mothers /* n x 4 */, int nf, const int* fids,
                      const double* fpar, double* out) {
  double m[N];
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::PhaseSpace<N> phsp(M, m);
  phsp.SetSeed(seed);
  hydra::device::vector<hydra::Vector4R> mv(n);
  for (uint64_t i = 0; i < n; ++i)
    mv[i] = hydra::Vector4R(mothers[4 * i], mothers[4 * i + 1], mothers[4 * i + 2],
                            mothers[4 * i + 3]);
  TupleFunctor<N + 1> f0{make_functor(fids[0], fpar)};
  TupleFunctor<N + 1> f1{make_functor(nf > 1 ? fids[1] : 0, fpar + kFpar)};
  if (nf == 1) {
    hydra::multivector<hydra::tuple<double, double>, hydra::device::sys_t> r(n);
    phsp.Evaluate(mv, r, f0);
    for (uint64_t i = 0; i < n; ++i) {
      auto row = r[i];
      out[i * 2] = hydra::get<0>(row);
      out[i * 2 + 1] = hydra::get<1>(row);
    }
  } else {
    hydra::multivector<hydra::tuple<double, double, double>, hydra::device::sys_t> r(n);
    phsp.Evaluate(mv, r, f0, f1);
    for (uint64_t i = 0; i < n; ++i) {
      auto row = r[i];
      out[i * 3] = hydra::get<0>(row);
      out[i * 3 + 1] = hydra::get<1>(row);
      out[i * 3 + 2] = hydra::get<2>(row);
    }
  }
}

// Elements first..first+n-1 of hydra::phase_space_range(mother, masses, seed,
// first + n), read through the range's own iterator: weights and AoS momenta.
template <size_t N>
void phase_space_range(const double* mother4, const double* masses, uint64_t seed,
                       uint64_t first, uint64_t n, double* out_p, double* out_w) {
  std::array<double, N> m;
  for (size_t i = 0; i < N; ++i) m[i] = masses[i];
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  auto events = hydra::phase_space_range(mother, m, (size_t)seed, (size_t)(first + n));
  auto it = events.begin() + first;
  for (uint64_t k = 0; k < n; ++k, ++it) {
    auto t = *it;
    out_w[k] = hydra::get<0>(t);
    hydra::Vector4R p[N];
    assign_tail<N>(t, p, std::make_index_sequence<N>{});
    for (size_t i = 0; i < N; ++i)
      for (int c = 0; c < 4; ++c) out_p[(k * N + i) * 4 + c] = p[i].get(c);
  }
}

// Runtime N dispatch (2..12).
template <template <size_t> class Op, typename... Args>
int dispatch(int N, Args&&... args) {
  switch (N) {
    case 2: Op<2>::run(args...); return 0;
    case 3: Op<3>::run(args...); return 0;
    case 4: Op<4>::run(args...); return 0;
    case 5: Op<5>::run(args...); return 0;
    case 6: Op<6>::run(args...); return 0;
    case 7: Op<7>::run(args...); return 0;
    case 8: Op<8>::run(args...); return 0;
    case 9: Op<9>::run(args...); return 0;
    case 10: Op<10>::run(args...); return 0;
    case 11: Op<11>::run(args...); return 0;
    case 12: Op<12>::run(args...); return 0;
    default: return -1;
  }
}

template <size_t N> struct GenOp {
  template <typename... A> static void run(A... a) { generate<N>(a...); }
};
template <size_t N> struct GenMothersOp {
  template <typename... A> static void run(A... a) { generate_mothers<N>(a...); }
};
template <size_t N> struct AvgOp {
  template <typename... A> static void run(A... a) { average_on<N>(a...); }
};
template <size_t N> struct UwOp {
  template <typename... A> static void run(int* rc, A... a) { *rc = unweight<N>(a...); }
};
template <size_t N> struct EvalOp {
  template <typename... A> static void run(A... a) { evaluate<N>(a...); }
};
template <size_t N> struct RangeOp {
  template <typename... A> static void run(A... a) { phase_space_range<N>(a...); }
};
template <size_t N> struct EvalMothersOp {
  template <typename... A> static void run(A... a) { evaluate_mothers<N>(a...); }
};
template <size_t N> struct ParamsOp {
  static void run(double M, const double* masses, double* ecm, double* maxw,
                  uint64_t* seed_full, int32_t* seed_get) {
    double m[N];
    for (size_t i = 0; i < N; ++i) m[i] = masses[i];
    hydra::PhaseSpace<N> phsp(M, m);
    *ecm = phsp.GetECM();
    *maxw = phsp.GetMaxWeight();
    *seed_full = (uint64_t)(1 * (1 + hydra::detail::hash_range(&m[0], &m[0] + N)));
    *seed_get = phsp.GetSeed();
  }
};
template <size_t N> struct DecayEvtOp {
  static void run(int engine, double M, const double* mother4, const double* masses,
                  uint64_t seed, const uint64_t* evts, uint64_t nevt, double* out_p,
                  double* out_wt) {
    switch (engine) {
      case 0: decay_events<N, hydra::random::squares3>(M, mother4, masses, seed, evts, nevt, out_p, out_wt); break;
      case 1: decay_events<N, hydra::random::squares4>(M, mother4, masses, seed, evts, nevt, out_p, out_wt); break;
      case 2: decay_events<N, hydra::random::philox>(M, mother4, masses, seed, evts, nevt, out_p, out_wt); break;
      case 3: decay_events<N, hydra::random::philox_long>(M, mother4, masses, seed, evts, nevt, out_p, out_wt); break;
      case 4: decay_events<N, hydra::random::threefry>(M, mother4, masses, seed, evts, nevt, out_p, out_wt); break;
      case 5: decay_events<N, hydra::random::threefry_long>(M, mother4, masses, seed, evts, nevt, out_p, out_wt); break;
      default: break;
    }
  }
};

}  // namespace

extern "C" {

int hr_set_threads(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  return omp_get_max_threads();
}

// Engine words: GRND eng(seed); eng.discard(discard); out[i] = eng().
// engine: 0 squares3, 1 squares4, 2 philox, 3 philox_long, 4 threefry, 5 threefry_long,
// 6 thrust's default_random_engine (minstd_rand, seeded with the low 32 bits)
int hr_rng_words(int engine, uint64_t seed, uint64_t discard, uint64_t n, uint64_t* out) {
  switch (engine) {
    case 0: engine_words<hydra::random::squares3>(seed, discard, n, out); return 0;
    case 1: engine_words<hydra::random::squares4>(seed, discard, n, out); return 0;
    case 2: engine_words<hydra::random::philox>(seed, discard, n, out); return 0;
    case 3: engine_words<hydra::random::philox_long>(seed, discard, n, out); return 0;
    case 4: engine_words<hydra::random::threefry>(seed, discard, n, out); return 0;
    case 5: engine_words<hydra::random::threefry_long>(seed, discard, n, out); return 0;
    case 6: engine_words<hydra::thrust::random::default_random_engine>(seed, discard, n, out);
      return 0;
    default: return -1;
  }
}

// PhaseSpace<N> constructor products: fECM, fMaxWeight, the full 64-bit default
// seed (1 * (1 + hash_range(masses))) and what GetSeed() returns for it.
int hr_phsp_params(int N, double M, const double* masses, double* ecm, double* maxw,
                   uint64_t* seed_full, int32_t* seed_get) {
  try {
    return dispatch<ParamsOp>(N, M, masses, ecm, maxw, seed_full, seed_get);
  } catch (std::exception const&) {
    return -2;
  }
}

// DecayMother<N,GRND>::process on explicit event indices: momenta (AoS,
// nevt x N x 4) and the generation weight wt.
int hr_decay_events(int N, int engine, double M, const double* mother4, const double* masses,
                    uint64_t seed, const uint64_t* evts, uint64_t nevt, double* out_p,
                    double* out_wt) {
  return dispatch<DecayEvtOp>(N, engine, M, mother4, masses, seed, evts, nevt, out_p, out_wt);
}

// PhaseSpace<N>::Generate(mother, Decays) then GetEventWeightFunctor weights.
int hr_generate(int N, double M, const double* mother4, const double* masses, int use_seed,
                int32_t seed, uint64_t n, double* out_p, double* out_w, double* seconds) {
  return dispatch<GenOp>(N, M, mother4, masses, use_seed, seed, n, out_p, out_w, seconds);
}

int hr_generate_mothers(int N, double M, const double* masses, int32_t seed, uint64_t n,
                        const double* mothers, double* out_p, double* out_w) {
  return dispatch<GenMothersOp>(N, M, masses, seed, n, mothers, out_p, out_w);
}

// Generate then Decays::Unweight(seed) (with_functor=0) or
// Decays::Unweight(functor, max_weight, seed) (with_functor=1).  Returns the
// permuted container and, for each output row, its original event index.
int hr_unweight(int N, double M, const double* mother4, const double* masses, int32_t seed,
                uint64_t n, int with_functor, int fid, const double* fpar, double max_weight,
                uint64_t uw_seed, double* out_p, uint64_t* out_idx, uint64_t* n_acc,
                double* seconds) {
  int rc = 0;
  int d = dispatch<UwOp>(N, &rc, M, mother4, masses, seed, n, with_functor, fid, fpar,
                         max_weight, uw_seed, out_p, out_idx, n_acc, seconds);
  return d ? d : rc;
}

// PhaseSpace<N>::AverageOn(device::sys, mother, functor, n) -> (mean, sigma).
int hr_average_on(int N, double M, const double* mother4, const double* masses, int use_seed,
                  int32_t seed, uint64_t n, int fid, const double* fpar, double* mean,
                  double* sigma, double* seconds) {
  return dispatch<AvgOp>(N, M, mother4, masses, use_seed, seed, n, fid, fpar, mean, sigma,
                         seconds);
}

// PhaseSpace<N>::Evaluate(mother, result, f0[, f1]) with result a
// multivector<tuple<double, double[, double]>>: rows of (weight, f0[, f1]).
// nf is 1 or 2; functor j's parameters at fpar + 32*j.
int hr_evaluate(int N, double M, const double* mother4, const double* masses, int use_seed,
                int32_t seed, uint64_t n, int nf, const int* fids, const double* fpar,
                double* out, double* seconds) {
  if (nf < 1 || nf > 2) return -3;
  return dispatch<EvalOp>(N, M, mother4, masses, use_seed, seed, n, nf, fids, fpar, out,
                          seconds);
}

// PhaseSpace<N>::Evaluate(mothers, result, f0[, f1]); the functors receive the
// mother followed by the N daughters.
int hr_evaluate_mothers(int N, double M, const double* masses, int32_t seed, uint64_t n,
                        const double* mothers, int nf, const int* fids, const double* fpar,
                        double* out) {
  if (nf < 1 || nf > 2) return -3;
  return dispatch<EvalMothersOp>(N, M, masses, seed, n, mothers, nf, fids, fpar, out);
}

// DenseHistogram<double, ndim>::Fill(data[, weights]) with ndim 1 or 2: data
// x holds ndim columns of n values; contents receives GetBinContent(b) for
// b = 0 .. nbins + 1 (the last two are the underflow and overflow slots).
int hr_histogram_fill(int ndim, const uint64_t* grid, const double* lower, const double* upper,
                      const double* x, uint64_t n, const double* w, double* contents,
                      double* seconds) {
  if (ndim == 1) {
    hydra::DenseHistogram<double, 1, hydra::device::sys_t> H((size_t)grid[0], lower[0], upper[0]);
    hydra::device::vector<double> data(x, x + n);
    double t0 = now_s();
    if (w) {
      hydra::device::vector<double> wv(w, w + n);
      H.Fill(data.begin(), data.end(), wv.begin());
    } else {
      H.Fill(data.begin(), data.end());
    }
    if (seconds) *seconds = now_s() - t0;
    for (uint64_t b = 0; b < grid[0] + 2; ++b) contents[b] = H.GetBinContent((size_t)b);
    return 0;
  }
  if (ndim == 2) {
    std::array<size_t, 2> g{(size_t)grid[0], (size_t)grid[1]};
    std::array<double, 2> lo{lower[0], lower[1]}, hi{upper[0], upper[1]};
    hydra::DenseHistogram<double, 2, hydra::device::sys_t> H(g, lo, hi);
    hydra::multivector<hydra::tuple<double, double>, hydra::device::sys_t> data(n);
    for (uint64_t i = 0; i < n; ++i) data[i] = hydra::make_tuple(x[i], x[n + i]);
    double t0 = now_s();
    if (w) {
      hydra::device::vector<double> wv(w, w + n);
      H.Fill(data.begin(), data.end(), wv.begin());
    } else {
      H.Fill(data.begin(), data.end());
    }
    if (seconds) *seconds = now_s() - t0;
    const uint64_t nb = grid[0] * grid[1];
    for (uint64_t b = 0; b < nb + 2; ++b) contents[b] = H.GetBinContent((size_t)b);
    return 0;
  }
  return -1;
}

// hydra::unweight(device::sys, data, gauss, max_pdf, rng_seed, rng_jump) with
// gauss = hydra::Gaussian<double>(mean, sigma), as examples/random/
// unweight_sample.inl:118-158 runs it (the data-plus-weights overload does not
// compile in the reference: Random.inl:68).  w_out receives gauss(x_k) as the
// same functor computes it; x_out the reordered data.
int hr_unweight_sample(const double* x, uint64_t n, double mean, double sigma, double max_pdf,
                       uint64_t seed, uint64_t jump, double* x_out, double* w_out,
                       uint64_t* n_kept) {
  auto Mean = hydra::Parameter::Create("mean").Value(mean);
  auto Sigma = hydra::Parameter::Create("sigma").Value(sigma);
  auto gauss = hydra::Gaussian<double>(Mean, Sigma);
  for (uint64_t k = 0; k < n; ++k) w_out[k] = gauss(x[k]);
  hydra::device::vector<double> data(x, x + n);
  auto r = hydra::unweight(hydra::device::sys, data, gauss, max_pdf, (size_t)seed, (size_t)jump);
  *n_kept = (uint64_t)r.size();
  for (uint64_t i = 0; i < n; ++i) x_out[i] = data[i];
  return 0;
}

// The decay chain of examples/phase_space/phsp_chain.inl:216-244:
// PhaseSpace<3>(B0 -> J/psi K pi) generates n events of a B0 at rest,
// PhaseSpace<2>(J/psi -> mu+ mu-) decays each event's J/psi
// (Generate(ThreeBodyDecay.GetDaugtherRange(_0), TwoBodyDecay)), both with
// their default seeds, and ThreeBodyDecay.Meld(TwoBodyDecay) joins them.
// out_p receives the melded rows (n x 5 x 4: J/psi, K, pi, mu+, mu-), out_f
// functor fid of each melded row (n values; par as make_functor), out_w the
// three-body weights (ThreeBodyDecay | GetEventWeightFunctor()).
int hr_chain(uint64_t n, int fid, const double* fpar, double* out_p, double* out_f,
             double* out_w, double* seconds) {
  const double B0_mass = 5.27955, Jpsi_mass = 3.0969, K_mass = 0.493677,
               pi_mass = 0.13957061, mu_mass = 0.1056583745;
  hydra::Vector4R B0(B0_mass, 0.0, 0.0, 0.0);
  hydra::PhaseSpace<3> three(B0_mass, {Jpsi_mass, K_mass, pi_mass});
  hydra::PhaseSpace<2> two(Jpsi_mass, {mu_mass, mu_mass});
  hydra::Decays<hydra::tuple<P0, P1, P2>, hydra::device::sys_t> tb(
      B0_mass, {Jpsi_mass, K_mass, pi_mass}, n);
  hydra::Decays<hydra::tuple<P3, P4>, hydra::device::sys_t> db(Jpsi_mass, {mu_mass, mu_mass},
                                                               n);
  double t0 = now_s();
  three.Generate(B0, tb);
  two.Generate(tb.GetDaugtherRange(hydra::placeholders::_0), db);
  if (seconds) *seconds = now_s() - t0;
  auto chain = tb.Meld(db);
  Functor f = make_functor(fid, fpar);
  auto it = chain.begin();
  for (uint64_t k = 0; k < n; ++k, ++it) {
    auto row = *it;
    hydra::Vector4R v[5] = {hydra::Vector4R(P0(hydra::get<0>(row)).Value()),
                            hydra::Vector4R(P1(hydra::get<1>(row)).Value()),
                            hydra::Vector4R(P2(hydra::get<2>(row)).Value()),
                            hydra::Vector4R(P3(hydra::get<3>(row)).Value()),
                            hydra::Vector4R(P4(hydra::get<4>(row)).Value())};
    for (int i = 0; i < 5; ++i)
      for (int c = 0; c < 4; ++c) out_p[(k * 5 + i) * 4 + c] = v[i].get(c);
    out_f[k] = f.eval(v);
  }
  auto w = tb | tb.GetEventWeightFunctor();
  for (uint64_t k = 0; k < n; ++k) out_w[k] = w[k];
  return 0;
}

// phase_space_range elements (weights, AoS momenta n x N x 4).
int hr_phase_space_range(int N, const double* mother4, const double* masses, uint64_t seed,
                         uint64_t first, uint64_t n, double* out_p, double* out_w) {
  return dispatch<RangeOp>(N, mother4, masses, seed, first, n, out_p, out_w);
}

// The range histogram of examples/misc/range_semantics.inl:154-167 for three
// daughters: DenseHistogram<double, 1, device>(bins, lo, hi).Fill(events | m_ab)
// with m_ab = (p_a + p_b).mass() over the first n elements of
// phase_space_range(mother, masses, seed, n); weighted by each element's weight
// when weighted != 0.  contents receives bins + 2 values (underflow, overflow
// last); seconds the time of the Fill.
int hr_range_histogram3(const double* mother4, const double* masses, uint64_t seed, uint64_t n,
                        int a, int b, int weighted, uint64_t bins, double lo, double hi,
                        double* contents, double* seconds) {
  if (a < 0 || a > 2 || b < 0 || b > 2) return -1;
  std::array<double, 3> m{masses[0], masses[1], masses[2]};
  hydra::Vector4R mother(mother4[0], mother4[1], mother4[2], mother4[3]);
  auto events = hydra::phase_space_range(mother, m, (size_t)seed, (size_t)n);
  auto mass_ab = hydra::wrap_lambda(
      [a, b] __hydra_dual__(double w, hydra::Vector4R A, hydra::Vector4R B, hydra::Vector4R C) {
        hydra::Vector4R v[3] = {A, B, C};
        return (v[a] + v[b]).mass();
      });
  auto weight = hydra::wrap_lambda(
      [] __hydra_dual__(double w, hydra::Vector4R A, hydra::Vector4R B, hydra::Vector4R C) {
        return w;
      });
  hydra::DenseHistogram<double, 1, hydra::device::sys_t> H((size_t)bins, lo, hi);
  double t0 = now_s();
  if (weighted)
    H.Fill(events | mass_ab, events | weight);
  else
    H.Fill(events | mass_ab);
  if (seconds) *seconds = now_s() - t0;
  for (uint64_t k = 0; k < bins + 2; ++k) contents[k] = H.GetBinContent((size_t)k);
  return 0;
}

}  // extern "C"
