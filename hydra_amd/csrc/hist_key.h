// GetGlobalBin on the device (GetGlobalBin.h:41-209), shared by the
// histogram fills over columns and events (histogram.hip) and over
// phase_space_range elements (range.hip).  Built with -ffp-contract=off.
#pragma once

#include "hpsp_launch.h"

namespace hpsp {

constexpr uint64_t kDrop = ~0ULL;

// X = (x - lower) * grid / delta per axis; all axes below -> underflow slot
// nbins, all above grid -> overflow slot nbins + 1, else the row-major index
// of the truncated X.  Where the reference's conversion or index is undefined
// (an axis at or below -1, NaN, an index past nbins + 1) the event is dropped.
__device__ __forceinline__ uint64_t hist_key(const HistK& h, const double (&x)[HPSP_HIST_MAX_DIM]) {
#pragma clang fp contract(off)
  double X[HPSP_HIST_MAX_DIM];
  bool under = true, over = true;
#pragma unroll
  for (int d = 0; d < HPSP_HIST_MAX_DIM; ++d) {
    if (d < h.ndim) {
      X[d] = (x[d] - h.lower[d]) * h.gridf[d] / h.delta[d];
      under = under && (X[d] < 0.0);
      over = over && (X[d] > h.gridf[d]);
    }
  }
  if (under) return h.nbins;
  if (over) return h.nbins + 1;
  uint64_t key = 0;
#pragma unroll
  for (int d = 0; d < HPSP_HIST_MAX_DIM; ++d) {
    if (d < h.ndim) {
      if (!(X[d] > -1.0) || X[d] >= 0x1p64) return kDrop;
      const uint64_t idx = X[d] < 0.0 ? 0 : (uint64_t)X[d];
      key += h.stride[d] * idx;
    }
  }
  return key <= h.nbins + 1 ? key : kDrop;
}

}  // namespace hpsp
