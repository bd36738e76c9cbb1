// The FP64 weight chain of the mixed-precision kernels (one-ulp square root
// and reciprocal in the reference's order), shared by the mixed Generate,
// AverageOn and regenerating scatter (genbod_mixed.hip) and the mixed
// Unweight's flags (generate_unweight.hip).  Every function turns
// contraction off itself, so the bits do not depend on the translation
// unit's flags.
#pragma once

#include <type_traits>

#include "hpsp_launch.h"

namespace hpsp {

// pdk in FP64 with the reference's subtraction order (so the a - b - c
// cancellation near threshold is the FP64 one) and a one-ulp square root and
// reciprocal instead of correctly rounded ones.
__device__ __forceinline__ double pdk_fast(double a, double b, double c) {
#pragma clang fp contract(off)
  const double x = (a - b - c) * (a + b + c) * (a - b + c) * (a + b - c);
  return sqrt_fast(x) * rcp_fast(2 * a);
}

// FP64 weight chain of DecayMother::process (DecayMother.h:170-212):
// sorted uniforms, invMas, pdk and the weight product in the reference's
// order; the momenta pd are handed on in FP32.
template <int N, int ENGINE, typename WordFn>
__device__ __forceinline__ double weight_chain(const DecayK& d, const WordFn& word,
                                               double (&invMas)[N], float (&pd)[N - 1]) {
#pragma clang fp contract(off)
  double rno[N];
  rno[0] = 0.0;
  rno[N - 1] = 1.0;
  if constexpr (N > 2) {
    using WT = std::conditional_t<is_squares<ENGINE>(), uint32_t, uint64_t>;
    WT v[N - 2];
#pragma unroll
    for (int k = 0; k < N - 2; ++k) v[k] = (WT)word(k);
    sort_net(v);
#pragma unroll
    for (int k = 0; k < N - 2; ++k) rno[k + 1] = to_unit<ENGINE>(v[k]);
  }
  double sum = 0.0;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    sum += d.m[n];
    invMas[n] = rno[n] * d.ecm + sum;
  }
  double wt = d.maxw;
#pragma unroll
  for (int n = 0; n < N - 1; ++n) {
    const double q = pdk_fast(invMas[n + 1], invMas[n], d.m[n + 1]);
    wt *= q;
    pd[n] = (float)q;
  }
  return wt;
}

}  // namespace hpsp
