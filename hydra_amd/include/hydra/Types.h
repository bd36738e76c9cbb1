// Basic types of the hydra API (reference: hydra/Types.h:59-75).
#pragma once

#include <cstddef>
#include <cstdint>

namespace hydra {

typedef double GReal_t;
typedef int GInt_t;
typedef unsigned int GUInt_t;
typedef long GLong_t;
typedef bool GBool_t;

// hydra/Types.h:75
constexpr GReal_t PI = 3.1415926535897932384626422832795028841971;

}  // namespace hydra
