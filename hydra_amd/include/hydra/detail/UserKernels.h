// User functors compiled into the phase-space kernels (reference: the functor
// plug-in of hydra/Lambda.h:133-200, called per event at
// hydra/detail/functors/AverageMother.h:286, EvalMother.h:283-306 and
// Decays.inl:362-379).
//
// A translation unit built by a HIP compiler that passes PhaseSpace or Decays
// a callable other than a described functor (hydra/PhaseSpaceFunctors.h)
// instantiates here the library's own kernel templates (genbod_f64.h) with
// that callable: the same GENBOD, engine-word sharing and reductions as
// libhpsp.so, so the events and weights are bit for bit those of the library,
// and only the functor's value comes from the user's code.  Partial results
// are merged and events unweighted through the C ABI.
//
// Calling convention.  A functor receives the event's four-vectors:
//   * positionally, each converted to the functor's parameter type (Vector4R,
//     or a declarg of Vector4R: a lambda (A a, B b, C c) over three
//     daughters), or
//   * over a Decays container whose particles are declarg types, by type: a
//     lambda (B b, C c) receives the container's B and C particles
//     (examples/phase_space/phsp_reweighting.inl:172-174), or
//   * as one hydra::tuple of them.
// The functor's result converts to double; over stored events (events |
// functor) it may also be a tuple of arithmetic values, one column each.
//
// Arithmetic.  The library's code turns floating-point contraction off in
// every function, so weights and momenta keep their bits under hipcc's
// default contraction (fast-honor-pragmas), under -ffp-contract=on and off;
// -ffp-contract=fast ignores pragmas and is not supported.  The user's own
// functor body is compiled under the user's flags.
#pragma once

#if !defined(__HIP__)
#error "hydra/detail/UserKernels.h compiles user functors into device kernels: build with hipcc"
#endif

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <tuple>
#include <type_traits>
#include <utility>

#include <hydra/Lambda.h>
#include <hydra/Tuple.h>
#include <hydra/Vector4R.h>
#include <hydra/detail/Runtime.h>

#include "../../../csrc/genbod_f64.h"
#include "../../../csrc/hpsp_host.h"
#include "hpsp.h"

namespace hydra {
namespace detail {

// ---- calling a functor with an event's four-vectors ---------------------

template <typename T, typename... Ts>
struct index_of;
template <typename T, typename... Ts>
struct index_of<T, T, Ts...> : std::integral_constant<int, 0> {};
template <typename T, typename U, typename... Ts>
struct index_of<T, U, Ts...> : std::integral_constant<int, 1 + index_of<T, Ts...>::value> {};
template <typename T>
struct index_of<T> : std::integral_constant<int, -10000> {};

template <typename T, typename... Ts>
struct count_of : std::integral_constant<int, (0 + ... + (std::is_same<T, Ts>::value ? 1 : 0))> {};

template <typename F, typename Args, typename = void>
struct invocable_with : std::false_type {};
template <typename F, typename... A>
struct invocable_with<F, hydra::tuple<A...>,
                      std::void_t<decltype(std::declval<F const&>()(std::declval<A const&>()...))>>
    : std::true_type {};

// every parameter type of F names exactly one of the particle types
template <typename F, typename Params, typename Types>
struct selectable_by_type : std::false_type {};
template <typename F, typename... A, typename... T>
struct selectable_by_type<F, hydra::tuple<A...>, hydra::tuple<T...>>
    : std::integral_constant<bool, sizeof...(A) >= 1 &&
                                       ((count_of<A, T...>::value == 1) && ...)> {};

template <typename F, typename... T>
struct functor_call_kind {
  using params = typename lambda_args<F>::type;
  static constexpr bool positional = invocable_with<F, hydra::tuple<T...>>::value;
  static constexpr bool by_type =
      !positional && lambda_args<F>::valid &&
      selectable_by_type<F, params, hydra::tuple<T...>>::value;
  static constexpr bool as_tuple =
      !positional && !by_type && invocable_with<F, hydra::tuple<hydra::tuple<T...>>>::value;
  static constexpr bool ok = positional || by_type || as_tuple;
};

template <int P>
__device__ __forceinline__ Vector4R vec_of(const double (&p)[P][4], int i) {
  return Vector4R(p[i][0], p[i][1], p[i][2], p[i][3]);
}

template <typename... T, typename F, int P, size_t... I>
__device__ __forceinline__ auto call_positional(const F& f, const double (&p)[P][4],
                                                std::index_sequence<I...>) {
  return f(T(vec_of(p, (int)I))...);
}

template <typename... T, typename F, int P, typename... A>
__device__ __forceinline__ auto call_by_type(const F& f, const double (&p)[P][4],
                                             hydra::tuple<A...>*) {
  return f(A(vec_of(p, index_of<A, T...>::value))...);
}

template <typename... T, typename F, int P, size_t... I>
__device__ __forceinline__ auto call_tuple(const F& f, const double (&p)[P][4],
                                           std::index_sequence<I...>) {
  return f(hydra::tuple<T...>(T(vec_of(p, (int)I))...));
}

// The functor's value for the P four-vectors p, carrying the types T...
template <typename... T, typename F, int P>
__device__ __forceinline__ auto invoke_event(const F& f, const double (&p)[P][4]) {
  static_assert(sizeof...(T) == P, "one particle type per four-vector");
  using K = functor_call_kind<F, T...>;
  static_assert(K::ok,
                "the functor cannot be called with this event's four-vectors: it takes them "
                "(or declarg types of them) in order, a subset of the container's declarg "
                "types, or one tuple of them");
  if constexpr (K::positional)
    return call_positional<T...>(f, p, std::make_index_sequence<P>{});
  else if constexpr (K::by_type)
    return call_by_type<T...>(f, p, (typename K::params*)nullptr);
  else
    return call_tuple<T...>(f, p, std::make_index_sequence<P>{});
}

// The result type of invoke_event<T...>(f, p), named without instantiating it.
template <typename F, typename Params, typename Types, int Kind>
struct call_result_impl;
template <typename F, typename... A, typename... T>
struct call_result_impl<F, hydra::tuple<A...>, hydra::tuple<T...>, 0> {
  using type = decltype(std::declval<F const&>()(std::declval<T const&>()...));
};
template <typename F, typename... A, typename... T>
struct call_result_impl<F, hydra::tuple<A...>, hydra::tuple<T...>, 1> {
  using type = decltype(std::declval<F const&>()(std::declval<A const&>()...));
};
template <typename F, typename... A, typename... T>
struct call_result_impl<F, hydra::tuple<A...>, hydra::tuple<T...>, 2> {
  using type = decltype(std::declval<F const&>()(std::declval<hydra::tuple<T...> const&>()));
};
template <typename F, typename... A, typename... T>
struct call_result_impl<F, hydra::tuple<A...>, hydra::tuple<T...>, 3> {
  using type = double;  // not callable: invoke_event reports it
};
template <typename F, typename... T>
struct call_result {
  using K = functor_call_kind<F, T...>;
  static constexpr int kind = K::positional ? 0 : (K::by_type ? 1 : (K::as_tuple ? 2 : 3));
  using type = std::decay_t<
      typename call_result_impl<F, typename K::params, hydra::tuple<T...>, kind>::type>;
};

template <typename R>
struct is_tuple : std::false_type {};
template <typename... R>
struct is_tuple<std::tuple<R...>> : std::true_type {};

// Columns a result fills: one for a value, one per element for a tuple.
template <typename R>
struct result_width : std::integral_constant<int, 1> {};
template <typename... R>
struct result_width<std::tuple<R...>> : std::integral_constant<int, (int)sizeof...(R)> {};

template <typename R, typename Out, size_t... I>
__device__ __forceinline__ void emit_tuple(const R& r, int j0, double scale, Out& out,
                                           std::index_sequence<I...>) {
#pragma clang fp contract(off)
  (out(j0 + (int)I, scale * (double)std::get<I>(r)), ...);
}

// Stores functor result r as column j0 (a value) or columns j0.. (a tuple),
// each multiplied by scale.
template <typename R, typename Out>
__device__ __forceinline__ void emit(const R& r, int j0, double scale, Out& out) {
#pragma clang fp contract(off)
  if constexpr (is_tuple<std::decay_t<R>>::value) {
    emit_tuple(r, j0, scale, out,
               std::make_index_sequence<(size_t)result_width<std::decay_t<R>>::value>{});
  } else {
    out(j0, scale * (double)r);
  }
}

// A pack of functors (trivially copyable, so it can be a kernel argument).
template <typename... F>
struct FnPack;
template <>
struct FnPack<> {};
template <typename F0, typename... F>
struct FnPack<F0, F...> {
  F0 head;
  FnPack<F...> tail;
};

template <typename... F>
FnPack<F...> make_pack(F const&... f);
template <>
inline FnPack<> make_pack() {
  return FnPack<>{};
}
template <typename F0, typename... F>
FnPack<F0, F...> make_pack(F0 const& f0, F const&... f) {
  return FnPack<F0, F...>{f0, make_pack(f...)};
}

// Columns the functors of a pack write for the particle types T...
template <typename Types, typename Pack>
struct pack_width;
template <typename... T>
struct pack_width<hydra::tuple<T...>, FnPack<>> : std::integral_constant<int, 0> {};
template <typename... T, typename F0, typename... F>
struct pack_width<hydra::tuple<T...>, FnPack<F0, F...>>
    : std::integral_constant<int, result_width<typename call_result<F0, T...>::type>::value +
                                      pack_width<hydra::tuple<T...>, FnPack<F...>>::value> {};

}  // namespace detail
}  // namespace hydra

namespace hpsp {

// A user's callable as the functor of the AverageOn kernel; T... are the
// types of the four-vectors it receives.
template <typename F, typename... T>
struct UserFn {
  F f;
};

template <typename F, typename... T, int P>
__device__ __forceinline__ double avg_value(const UserFn<F, T...>& u, const double (&p)[P][4]) {
  return (double)hydra::detail::invoke_event<T...>(u.f, p);
}

// A pack of user callables as the functor set of the Evaluate kernel (one
// column per functor) and of the stored-event kernel (one column per value).
template <typename Types, typename... F>
struct UserSet {
  hydra::detail::FnPack<F...> fs;
};

template <typename... T, int P, typename Out, typename F0, typename... F>
__device__ __forceinline__ void emit_pack(const hydra::detail::FnPack<F0, F...>& pk,
                                          const double (&p)[P][4], int j0, double scale,
                                          Out& out) {
  const auto r = hydra::detail::invoke_event<T...>(pk.head, p);
  hydra::detail::emit(r, j0, scale, out);
  if constexpr (sizeof...(F) > 0)
    emit_pack<T...>(pk.tail, p, j0 + hydra::detail::result_width<std::decay_t<decltype(r)>>::value,
                    scale, out);
}
template <typename... T, int P, typename Out>
__device__ __forceinline__ void emit_pack(const hydra::detail::FnPack<>&, const double (&)[P][4],
                                          int, double, Out&) {}

template <typename... T, typename... F, int P, typename Out>
__device__ __forceinline__ void eval_set(const UserSet<hydra::tuple<T...>, F...>& s,
                                         const double (&p)[P][4], Out&& out) {
  emit_pack<T...>(s.fs, p, 0, 1.0, out);
}

// The same with every value multiplied by scale.
template <typename... T, typename... F, int P, typename Out>
__device__ __forceinline__ void eval_set_scaled(const UserSet<hydra::tuple<T...>, F...>& s,
                                                const double (&p)[P][4], double scale,
                                                Out&& out) {
  emit_pack<T...>(s.fs, p, 0, scale, out);
}

// Functor values of stored events (4N FP64 columns of ld events), optionally
// times the event's PhaseSpaceWeight (PhaseSpaceReweight, Decays.inl:362-379).
template <typename FS>
struct UserValuesArgs {
  DecayK d;  // masses and fMaxWeight
  FS fs;
  const double* cols;
  uint64_t ld, n;
  double* out;  // value columns of ldo doubles
  uint64_t ldo;
  int with_weight;
};

template <int N, typename FS>
__global__ __launch_bounds__(kBlock) void k_user_values(const UserValuesArgs<FS> a) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < a.n; k += stride) {
    double p[N][4];
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) p[i][q] = a.cols[(uint64_t)(4 * i + q) * a.ld + k];
    const double w = a.with_weight ? event_weight_f64<N>(p, a.d.m, a.d.maxw) : 1.0;
    eval_set_scaled(a.fs, p, w, [&](int j, double v) { a.out[(uint64_t)j * a.ldo + k] = v; });
  }
}

}  // namespace hpsp

namespace hydra {
namespace detail {

inline void hip_check(hipError_t e, const char* where) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("[hydra::") + where + "]: " + hipGetErrorString(e));
}

template <typename T, size_t N>
struct repeat_type {
  template <size_t... I>
  static auto make(std::index_sequence<I...>) -> hydra::tuple<decltype((void)I, T())...>;
  using type = decltype(make(std::make_index_sequence<N>{}));
};

// AverageOn (PhaseSpace.inl:255-287) with a user functor over events
// first..first+n-1 of decay d (or over n mothers, the functor then seeing
// the mother first): the library's AverageOn kernel instantiated with the
// functor, its partials merged in the library's fixed order.  The result
// depends only on (d, functor, seed, first, n), as the library's does.
template <size_t N, int ENGINE, bool MOTHERS, typename... T, typename F>
hpsp_stats user_average_on_t(hydra::tuple<T...>*, hpsp_decay const& d, uint64_t seed,
                             F const& f, uint64_t first, uint64_t n, const double* mothers,
                             uint64_t mld) {
  using FK = hpsp::UserFn<F, T...>;
  const uint64_t m = (n + hpsp::kAvgChunk - 1) / hpsp::kAvgChunk;
  device_buffer<hpsp::Stats> partials(m ? m : 1);
  device_buffer<hpsp_stats> out(1);
  if (n > 0) {
    hpsp::AvgArgsT<FK> a{hpsp::make_decayk(&d, seed), FK{f}, first, n, mothers, mld,
                         partials.data()};
    const unsigned g = (unsigned)(m < 8192 ? m : 8192);
    hipLaunchKernelGGL((hpsp::k_average_f64<(int)N, ENGINE, MOTHERS, FK>), dim3(g),
                       dim3(hpsp::kBlock), 0, 0, a);
    hip_check(hipGetLastError(), "PhaseSpace::AverageOn");
  }
  hpsp_check(hpsp_stats_reduce(reinterpret_cast<const hpsp_stats*>(partials.data()), m,
                               out.data(), nullptr),
             "PhaseSpace::AverageOn");
  hpsp_stats st;
  hpsp_check(hpsp_memcpy_to_host(&st, out.data(), sizeof(st)), "PhaseSpace::AverageOn");
  return st;
}

template <size_t N, int ENGINE, bool MOTHERS, typename F>
hpsp_stats user_average_on(hpsp_decay const& d, uint64_t seed, F const& f, uint64_t first,
                           uint64_t n, const double* mothers, uint64_t mld, int precision) {
  if (precision != HPSP_FP64)
    throw std::invalid_argument(
        "[hydra::PhaseSpace::AverageOn]: a user functor runs in FP64 only");
  using Types = typename repeat_type<Vector4R, N + (MOTHERS ? 1 : 0)>::type;
  return user_average_on_t<N, ENGINE, MOTHERS>((Types*)nullptr, d, seed, f, first, n, mothers,
                                               mld);
}

// Evaluate (PhaseSpace.inl:289-348) with user functors: column 0 the
// generation weight, column 1 + j functor j's value.
template <size_t N, int ENGINE, bool MOTHERS, typename... T, typename... F>
void user_evaluate_t(hydra::tuple<T...>*, hpsp_decay const& d, uint64_t seed,
                     const double* mothers, uint64_t mld, uint64_t first, uint64_t n,
                     double* out, uint64_t ld, F const&... f) {
  using FS = hpsp::UserSet<hydra::tuple<T...>, F...>;
  static_assert(pack_width<hydra::tuple<T...>, FnPack<F...>>::value == (int)sizeof...(F),
                "Evaluate: each functor returns one value");
  if (n == 0) return;
  const uint64_t ngroups = (n + hpsp::kWave - 1) / hpsp::kWave;
  const uint32_t run = MOTHERS ? 1u : hpsp::run_groups(ngroups);
  hpsp::EvalArgsT<FS> a{hpsp::make_decayk(&d, seed), FS{make_pack(f...)}, seed, first, n,
                        mothers, mld, out, ld, run};
  const unsigned g = hpsp::grid_for_runs((ngroups + run - 1) / run);
  hipLaunchKernelGGL((hpsp::k_evaluate_f64<(int)N, ENGINE, MOTHERS, FS>), dim3(g),
                     dim3(hpsp::kBlock), 0, 0, a);
  hip_check(hipGetLastError(), "PhaseSpace::Evaluate");
  hpsp_sync("PhaseSpace::Evaluate");
}

template <size_t N, int ENGINE, bool MOTHERS, typename... F>
void user_evaluate(hpsp_decay const& d, uint64_t seed, const double* mothers, uint64_t mld,
                   uint64_t first, uint64_t n, double* out, uint64_t ld, F const&... f) {
  using Types = typename repeat_type<Vector4R, N + (MOTHERS ? 1 : 0)>::type;
  user_evaluate_t<N, ENGINE, MOTHERS>((Types*)nullptr, d, seed, mothers, mld, first, n, out, ld,
                                      f...);
}

// Values of user functors over stored events (4N columns of ld doubles,
// particles of types P...), times PhaseSpaceWeight when with_weight: into
// `width` columns of ldo doubles.
template <typename... P, typename... F>
void user_event_values(hydra::tuple<P...>*, hpsp_decay const& d, const double* cols,
                       uint64_t ld, uint64_t n, bool with_weight, double* out, uint64_t ldo,
                       F const&... f) {
  constexpr int N = (int)sizeof...(P);
  using FS = hpsp::UserSet<hydra::tuple<P...>, F...>;
  if (n == 0) return;
  hpsp::UserValuesArgs<FS> a{hpsp::make_decayk(&d, 0), FS{make_pack(f...)}, cols, ld, n, out,
                             ldo, with_weight ? 1 : 0};
  hipLaunchKernelGGL((hpsp::k_user_values<N, FS>), dim3(hpsp::grid_for(n)),
                     dim3(hpsp::kBlock), 0, 0, a);
  hip_check(hipGetLastError(), "events | functor");
  hpsp_sync("events | functor");
}

template <typename Types, typename... F>
constexpr int user_width() {
  return pack_width<Types, FnPack<F...>>::value;
}

}  // namespace detail
}  // namespace hydra
