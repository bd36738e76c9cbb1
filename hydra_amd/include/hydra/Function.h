// Typed particle arguments (reference: declarg, hydra/detail/FunctionArgument.h:241).
// declarg(A, hydra::Vector4R) declares hydra::arguments::A, a four-vector
// type used to name the daughters of a Decays container.
#pragma once

#include <hydra/Vector4R.h>

#define declarg(NAME, TYPE)                       \
  namespace hydra {                               \
  namespace arguments {                           \
  struct NAME : public TYPE {                     \
    NAME() = default;                             \
    NAME(TYPE const& x) : TYPE(x) {}              \
    TYPE Value() const { return *this; }          \
  };                                              \
  }                                               \
  }
