// Destinations of the stable partition (Decays::Unweight, Decays.inl:
// 434-453) from the flag words and tile offsets, shared by the moving scatter
// (unweight.hip) and the regenerating one (genbod_f64.hip).
#pragma once

#include "hpsp_launch.h"

namespace hpsp {

constexpr int kItems = kUwTile / kBlock;  // 4 groups of 64 events per wave

// Destinations of one group of 64 samples in the stable partition: kept lanes
// to goff + (kept lanes below), the others to nacc + (k0 - goff) + (rejected
// lanes below).
__device__ __forceinline__ uint64_t stable_dst(uint64_t fw, uint64_t below, int lane,
                                               uint64_t goff, uint64_t nacc, uint64_t k0,
                                               bool* kept) {
  const uint32_t acc_below = (uint32_t)__popcll(fw & below);
  *kept = (fw >> lane) & 1ULL;
  return *kept ? goff + acc_below : nacc + k0 - goff + ((uint32_t)lane - acc_below);
}

// Group offsets of a tile in LDS (lanes 0..15 of wave 0 scan the popcounts).
__device__ __forceinline__ void tile_group_offsets(const UnweightArgs& a, uint64_t tile,
                                                   uint64_t* goff) {
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  const uint64_t ngroups = kUwTile / kWave;
  if (wave == 0) {
    const uint64_t gi = tile * ngroups + (uint64_t)lane;
    const uint32_t own =
        (lane < (int)ngroups && gi * kWave < a.n) ? (uint32_t)__popcll(a.flags[gi]) : 0u;
    uint32_t c = own;
#pragma unroll
    for (int off = 1; off < (int)ngroups; off <<= 1) {
      const uint32_t o = __shfl_up(c, off, kWave);
      if (lane >= off) c += o;
    }
    if (lane < (int)ngroups) goff[lane] = a.tile_off[tile] + (c - own);
  }
  __syncthreads();
}

// Place of a block of kBlock events (one group of 64 per wave, its first
// group g0) in the stable partition.  Lanes 0..15 load the 16 flag words of
// the block's tile in one load and scan their popcounts, so that no wave waits
// on a chain of dependent loads (a serial walk over the words before the block
// cost 12% of a partitioned 12-column store).
struct BlockPart {
  uint64_t goff;      // accepted events before the block
  uint64_t fw;        // the flag word of this wave's group
  uint32_t acc_wave;  // accepted events of the block before this wave's group
  uint32_t a_blk;     // accepted events of the block
};

__device__ __forceinline__ BlockPart block_part(const UnweightArgs& a, uint64_t g0, int wave,
                                                uint64_t ngroups) {
  constexpr int kGroupsPerTile = kUwTile / kWave;
  constexpr int kWavesPerBlock = kBlock / kWave;
  static_assert(kGroupsPerTile <= kWave && kGroupsPerTile % kWavesPerBlock == 0,
                "a block lies inside one tile");
  const int lane = __lane_id();
  const uint64_t tile = g0 / kGroupsPerTile;
  const int gb = (int)(g0 - tile * kGroupsPerTile);  // wave-uniform
  const uint64_t gi = tile * kGroupsPerTile + (uint64_t)lane;
  const uint64_t fl = (lane < kGroupsPerTile && gi < ngroups) ? a.flags[gi] : 0ULL;
  uint32_t incl = (uint32_t)__popcll(fl);  // inclusive over lanes 0..15
#pragma unroll
  for (int off = 1; off < kGroupsPerTile; off <<= 1) {
    const uint32_t o = __shfl_up(incl, off, kWave);
    if (lane >= off) incl += o;
  }
  const int lw = gb + wave;
  const uint32_t before = gb ? (uint32_t)__builtin_amdgcn_readlane((int)incl, gb - 1) : 0u;
  BlockPart b;
  b.goff = a.tile_off[tile] + before;
  b.acc_wave = (lw ? (uint32_t)__builtin_amdgcn_readlane((int)incl, lw - 1) : 0u) - before;
  b.a_blk = (uint32_t)__builtin_amdgcn_readlane((int)incl, gb + kWavesPerBlock - 1) - before;
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)fl, lw);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(fl >> 32), lw);
  b.fw = ((uint64_t)hi << 32) | lo;
  return b;
}

// RN(w / maxv) > u, the reference's test, decided without a division where the
// product with the reciprocal settles it.  rmax = RN(1 / maxv) and
// q = RN(w * rmax) put q within 2^-51.3 |q| of w / maxv's rounded value
// (relative errors 2^-53 each, rmax and q normal), so when |q - u| exceeds
// 2^-50 |q| the rounded quotient lies strictly on q's side of u.  Otherwise
// (and for zero, subnormal, infinite or NaN operands) the division decides:
// the same flag as the division in every case.
__device__ __forceinline__ bool ratio_above(double w, double maxv, double rmax, bool rmax_ok,
                                            double u) {
#pragma clang fp contract(off)
  const double q = w * rmax;
  const double gap = q - u;
  const double aq = __builtin_fabs(q);
  const bool settled = rmax_ok && aq >= 0x1p-1000 && aq <= 0x1p+1000 &&
                       __builtin_fabs(gap) > aq * 0x1p-50;
  bool r = gap > 0.0;
  if (!settled) r = (w / maxv) > u;
  return r;
}

// Tile of a block: each XCD takes a contiguous range of tiles, so that
// neighbouring tiles, which write neighbouring pieces of the accepted and
// rejected runs, share an L2.
__device__ __forceinline__ uint64_t xcd_tile(uint64_t ntiles) {
  const uint64_t per_xcd = (ntiles + kXcds - 1) / kXcds;
  return (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
}

}  // namespace hpsp
