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
  named_arg(T const& x) : T(x) {}
  T Value() const { return *this; }
};

// arithmetic types: the argument holds a T and converts to it
template <typename T>
struct named_arg<T, false> {
  named_arg() = default;
  named_arg(T x) : value(x) {}
  operator T() const { return value; }
  T Value() const { return value; }
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
    NAME(TYPE const& x) : hydra::detail::named_arg<TYPE>(x) {} \
  };                                                          \
  }                                                           \
  }
