// Typed arguments (reference: declarg, hydra/detail/FunctionArgument.h:241).
// declarg(A, hydra::Vector4R) declares hydra::arguments::A, a four-vector
// type used to name the daughters of a Decays container; declarg(W, double)
// declares a named double, used to name the columns of a multivector.
#pragma once

#include <type_traits>

#include <hydra/Vector4R.h>

namespace hydra {
namespace detail {

template <typename T, bool = std::is_class<T>::value>
struct named_arg;

// class types: the argument is-a T
template <typename T>
struct named_arg<T, true> : public T {
  named_arg() = default;
  __hydra_dual__ named_arg(T const& x) : T(x) {}
  __hydra_dual__ T Value() const { return *this; }
  __hydra_dual__ T const& operator()() const { return *this; }
};

// arithmetic types: the argument holds a T and converts to it
template <typename T>
struct named_arg<T, false> {
  named_arg() = default;
  __hydra_dual__ named_arg(T x) : value(x) {}
  __hydra_dual__ operator T() const { return value; }
  __hydra_dual__ T Value() const { return value; }
  __hydra_dual__ T operator()() const { return value; }
  T value{};
};

}  // namespace detail
}  // namespace hydra

#define declarg(NAME, TYPE)                                   \
  namespace hydra {                                           \
  namespace arguments {                                       \
  struct NAME : public hydra::detail::named_arg<TYPE> {       \
    using value_type = TYPE;                                  \
    NAME() = default;                                         \
    __hydra_dual__ NAME(TYPE const& x) : hydra::detail::named_arg<TYPE>(x) {} \
  };                                                          \
  }                                                           \
  }
