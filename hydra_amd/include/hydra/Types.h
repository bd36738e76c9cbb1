// Basic types of the hydra API (reference: hydra/Types.h:59-75).
#pragma once

#include <cstddef>
#include <cstdint>

// Functions that user functors call in device code carry __hydra_dual__
// (hydra/detail/Config.h:100-104): host and device under a HIP compiler,
// plain host functions under any other.
#if defined(__HIP__) || defined(__HIPCC__)
#define __hydra_host__ __host__
#define __hydra_device__ __device__
#define __hydra_dual__ __host__ __device__
#else
#define __hydra_host__
#define __hydra_device__
#define __hydra_dual__
#endif

// The reference's floating-point operations one by one inside the library's
// own arithmetic (four-vector algebra, helicity angles): no contraction into
// fused multiply-adds under any flag that honours pragmas.
#if defined(__clang__)
#define HYDRA_EXACT_FP _Pragma("clang fp contract(off)")
#else
#define HYDRA_EXACT_FP
#endif

namespace hydra {

typedef double GReal_t;
typedef int GInt_t;
typedef unsigned int GUInt_t;
typedef long GLong_t;
typedef bool GBool_t;

// hydra/Types.h:75
constexpr GReal_t PI = 3.1415926535897932384626422832795028841971;

}  // namespace hydra
