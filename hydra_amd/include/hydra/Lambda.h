// hydra::Lambda and hydra::wrap_lambda (reference: hydra/Lambda.h:52-531):
// a user's callable wrapped so the phase-space calls can compile it into
// their kernels.  Built by a HIP compiler, the wrapped lambda is marked
// __hydra_dual__ ([] __hydra_dual__ (A a, B b, C c) { ... }) and PhaseSpace's
// AverageOn and Evaluate, Decays' Unweight(functor) and GetEventWeightFunctor
// (functor), and `events | functor` run it on the device
// (hydra/detail/UserKernels.h).  Called with
//   * arguments convertible to the lambda's parameters (a Vector4R to a
//     declarg of Vector4R), or
//   * one tuple of them, whose elements are passed in order.
// The reference's parametrised form (wrap_lambda(lambda, Parameter...)) is not
// part of the phase-space path and is not provided.
#pragma once

#include <cstddef>
#include <tuple>
#include <type_traits>
#include <utility>

#include <hydra/Tuple.h>
#include <hydra/Types.h>

namespace hydra {

template <typename LambdaType, size_t NPARAM = 0>
class Lambda;

template <typename LambdaType>
class Lambda<LambdaType, 0> {
 public:
  typedef void hydra_lambda_type;
  typedef LambdaType lambda_type;

  Lambda() = delete;
  explicit Lambda(LambdaType const& lambda) : fLambda(lambda) {}

  __hydra_dual__ const LambdaType& GetLambda() const { return fLambda; }

  // arguments convertible to the lambda's parameters
  template <typename... T>
  __hydra_dual__ auto operator()(T const&... x) const -> decltype(std::declval<LambdaType const&>()(x...)) {
    return fLambda(x...);
  }

  // one tuple holding them, passed element by element
  template <typename... T, typename = decltype(std::declval<LambdaType const&>()(std::declval<T const&>()...))>
  __hydra_dual__ auto operator()(hydra::tuple<T...> const& t) const {
    return call_tuple(t, std::index_sequence_for<T...>{});
  }

 private:
  template <typename Tup, size_t... I>
  __hydra_dual__ auto call_tuple(Tup const& t, std::index_sequence<I...>) const {
    return fLambda(std::get<I>(t)...);
  }

  LambdaType fLambda;
};

template <typename LambdaType>
hydra::Lambda<LambdaType, 0> wrap_lambda(LambdaType const& lambda) {
  return hydra::Lambda<LambdaType, 0>(lambda);
}

namespace detail {

// Parameter types of a callable with one, non-template call operator (a
// lambda, a wrapped lambda): lambda_args<F>::type is a tuple of the decayed
// parameter types; valid is false for callables whose parameters cannot be
// named (generic lambdas, overload sets).
template <typename F, typename = void>
struct lambda_args {
  static constexpr bool valid = false;
  using type = hydra::tuple<>;
};

template <typename R, typename C, typename... A>
struct lambda_args_of_member {
  static constexpr bool valid = true;
  using type = hydra::tuple<std::decay_t<A>...>;
};

template <typename M>
struct member_args;
template <typename R, typename C, typename... A>
struct member_args<R (C::*)(A...) const> : lambda_args_of_member<R, C, A...> {};
template <typename R, typename C, typename... A>
struct member_args<R (C::*)(A...)> : lambda_args_of_member<R, C, A...> {};

template <typename F>
struct lambda_args<F, std::void_t<decltype(&F::operator())>> : member_args<decltype(&F::operator())> {};

template <typename L>
struct lambda_args<hydra::Lambda<L, 0>, void> : lambda_args<L> {};

}  // namespace detail
}  // namespace hydra
