// Tuples of the hydra API (reference: hydra/Tuple.h); std::tuple underneath.
#pragma once

#include <tuple>

namespace hydra {
using std::get;
using std::make_tuple;
using std::tuple;
using std::tuple_size;
}  // namespace hydra
