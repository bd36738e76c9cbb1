// Accept-reject unweighting as a wave-ballot stream compaction (gfx950).
//
// Decays::Unweight (Decays.inl:415-455, 457-511) is a stable_partition of
// (index, event) by FlagDaugthers (Decays.inl:75-89).  Here:
//   1. k_uw_flags   -- one 1024-event tile per block: weight (read from a
//                      stored column or recomputed from the momenta), the
//                      default engine's word for the container index, the
//                      test  w / max > u; a 64-bit __ballot per wave gives
//                      the flag word, popcounts give the tile count.
//   2. k_uw_scan    -- one block: exclusive prefix of the tile counts and
//                      the accepted total.
//   3. k_uw_scatter -- each event's destination from its tile offset and the
//                      popcount of the lower lanes of its ballot word:
//                      accepted -> off_acc, rejected -> n_acc + (k - off_acc);
//                      all 4N columns (and optional index/weight) move once.
// Event columns are read twice at most (flags need them only when no weight
// column is supplied) and written once.  The scatter keeps several groups'
// loads in flight before storing and maps tiles to XCDs in contiguous ranges.
#include <array>
#include <utility>

#include "uw_dst.h"

namespace hpsp {


template <int N>
__device__ __forceinline__ void load_event_uw(const double* cols, uint64_t ld, uint64_t k,
                                              double (&p)[N][4]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) p[i][q] = cols[(uint64_t)(4 * i + q) * ld + k];
}

// N = 0: the weights come from a.w_in alone (no functor), and the kernel
// holds no momenta: few registers, many waves in flight for the one stream it
// reads.  Tiles are dealt to XCDs in contiguous ranges.
template <int N>
__global__ __launch_bounds__(kBlock) void k_uw_flags(const UnweightArgs a) {
#pragma clang fp contract(off)
  __shared__ uint32_t cnt[kBlock / kWave];
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;  // whole block
  const double maxv = a.maxv_dev ? *a.maxv_dev : a.maxv;
  const double rmax = 1.0 / maxv;
  const bool rmax_ok = __builtin_fabs(rmax) >= 0x1p-1000 && __builtin_fabs(rmax) <= 0x1p+1000;
  constexpr bool w_only = (N == 0);
  // with a weight column, the wave's four loads are issued before any flag
  // word is stored
  double wcol[kItems];
  if constexpr (w_only) {
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
      const uint64_t g = tile * (kUwTile / kWave) + (uint64_t)j * (kBlock / kWave) + wave;
      const uint64_t k = g * kWave + (uint64_t)lane;
      wcol[j] = (k < a.n) ? *(glob(a.w_in) + k) : 0.0;
    }
  }
  uint32_t c = 0;
  // key * (jump + k) of the wave's first group; group j + 1 is kBlock events
  // on, so the product steps by key * kBlock (an add, not a 64-bit multiply)
  uint64_t ykey = a.uw_key * (a.jump + (tile * (kUwTile / kWave) + wave) * kWave + lane);
  const uint64_t ystep = a.uw_key * (uint64_t)kBlock;
#pragma unroll
  for (int j = 0; j < kItems; ++j, ykey += ystep) {
    const uint64_t g = tile * (kUwTile / kWave) + (uint64_t)j * (kBlock / kWave) + wave;
    const uint64_t k = g * kWave + (uint64_t)lane;
    bool acc = false;
    if (k < a.n) {
      double w;
      if constexpr (w_only) {
        w = wcol[j];
      } else {
        double p[N][4];
        load_event_uw<N>(a.cols_in, a.ld_in, k, p);
        w = a.w_in ? a.w_in[k] : event_weight_f64<N>(p, a.d.m, a.d.maxw);
        if (a.use_f) w = w * functor_eval<double>(a.f, p);
      }
      const double u = (double)squares3_y(a.uw_key, ykey) * 0x1p-32;  // ctr = jump + k
      // Decays::Unweight (FlagDaugthers, Decays.inl:75-89): w / max > u;
      // hydra::unweight (RndFlag, RandomUtils.h:254-262): u * max <= w
      acc = a.rule == 0 ? ratio_above(w, maxv, rmax, rmax_ok, u) : (u * maxv) <= w;
    }
    const uint64_t ball = __ballot(acc);
    if (lane == 0) *(glob(a.flags) + g) = ball;
    c += (uint32_t)__popcll(ball);
  }
  if (lane == 0) cnt[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) a.tile_count[tile] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
}

// Exclusive scan of the tile counts in three fully parallel steps:
// k_uw_scan_blocks sums each run of kScanRun counts, k_uw_scan_top scans those
// sums in one block (and writes the accepted total), k_uw_scan_tiles scans
// each run again starting from its offset.
constexpr int kScanRun = 1024;  // tile counts per scan block (4 per thread)

__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t* sh, uint64_t* total) {
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  uint64_t x = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint64_t o = __shfl_up(x, off, kWave);
    if (lane >= off) x += o;
  }
  if (lane == kWave - 1) sh[wave] = x;
  __syncthreads();
  uint64_t base = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    base += (w < wave) ? sh[w] : 0;
    tot += sh[w];
  }
  __syncthreads();
  *total = tot;
  return base + x - v;
}

__global__ __launch_bounds__(kBlock) void k_uw_scan_blocks(const uint32_t* tile_count,
                                                           uint64_t ntiles, uint64_t* run_sum) {
  __shared__ uint64_t sh[kBlock / kWave];
  const uint64_t b = (uint64_t)blockIdx.x * kScanRun + (uint64_t)threadIdx.x * 4;
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) s += (b + i < ntiles) ? tile_count[b + i] : 0u;
  uint64_t tot;
  (void)block_exclusive_scan(s, sh, &tot);
  if (threadIdx.x == 0) run_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kBlock) void k_uw_scan_top(uint64_t* run_sum, uint64_t nruns,
                                                        uint64_t* n_acc) {
  __shared__ uint64_t sh[kBlock / kWave];
  uint64_t carry = 0;
  for (uint64_t b = 0; b < nruns; b += kBlock) {
    const uint64_t i = b + threadIdx.x;
    const uint64_t v = (i < nruns) ? run_sum[i] : 0;
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan(v, sh, &tot);
    if (i < nruns) run_sum[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *n_acc = carry;
}

__global__ __launch_bounds__(kBlock) void k_uw_scan_tiles(const uint32_t* tile_count,
                                                          uint64_t ntiles,
                                                          const uint64_t* run_off,
                                                          uint64_t* tile_off) {
  __shared__ uint64_t sh[kBlock / kWave];
  const uint64_t b = (uint64_t)blockIdx.x * kScanRun + (uint64_t)threadIdx.x * 4;
  uint32_t c[4];
  uint64_t s = 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    c[i] = (b + i < ntiles) ? tile_count[b + i] : 0u;
    s += c[i];
  }
  uint64_t tot;
  uint64_t run = run_off[blockIdx.x] + block_exclusive_scan(s, sh, &tot);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (b + i < ntiles) tile_off[b + i] = run;
    run += c[i];
  }
}

// Tiles are mapped so that each XCD takes a contiguous range of them:
// neighbouring tiles write neighbouring pieces of the accepted and rejected
// runs, and the cache lines they share then meet in one L2.
// The scatter for N <= 6: one group of 64 events per wave, 256 per block.
// Every event is read where it lies (lane-contiguous), staged in LDS at its
// rank in the block's partition order, and each column is written by all 256
// threads as the block's two runs (accepted, rejected), so that only the ends
// of a block's runs are partial cache lines.  Blocks are dealt to XCDs in
// contiguous ranges.
template <int N>
__global__ __launch_bounds__(kBlock) void k_uw_scatter_staged(const UnweightArgs a) {
  constexpr int C = 4 * N;
  // all columns staged at once: staging half of them (more waves in flight)
  // made this read-and-write kernel 9% slower at N = 3
  constexpr int CS = C;
  __shared__ double sh[CS][kBlock];
  constexpr int kWavesPerBlock = kBlock / kWave;
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  const uint64_t nblocks = (ngroups + kWavesPerBlock - 1) / kWavesPerBlock;
  const uint64_t per_xcd = (nblocks + kXcds - 1) / kXcds;
  const uint64_t blk = (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (blk >= nblocks) return;  // whole block
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t g0 = blk * kWavesPerBlock;
  const BlockPart bp = block_part(a, g0, wave, ngroups);
  const uint64_t goff = bp.goff;
  const uint32_t acc_wave = bp.acc_wave, a_blk = bp.a_blk;
  const uint64_t nacc = *a.n_acc;
  const uint64_t kb = g0 * kWave;
  const uint64_t g = g0 + (uint64_t)wave;
  const uint64_t k0 = g * kWave;
  const uint64_t k = k0 + (uint64_t)lane;
  const bool live_ev = g < ngroups && k < a.n;
  double v[C];
  uint32_t r = 0;  // the event's rank in the block's partition order
  if (live_ev) {
    const uint64_t fw = bp.fw;
    const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const uint32_t acc_below = acc_wave + (uint32_t)__popcll(fw & below);
    const bool kept = (fw >> lane) & 1ULL;
    const uint32_t t = (uint32_t)(wave * kWave + lane);
    r = kept ? acc_below : a_blk + (t - acc_below);
    const uint64_t dst = (r < a_blk) ? goff + r : nacc + (kb - goff) + (r - a_blk);
    if (a.cols_out) {
#pragma unroll
      for (int c = 0; c < C; ++c) v[c] = *(glob_u(a.cols_in + (uint64_t)c * a.ld_in + k0) + lane);
    }
    if (a.w_out) *(glob(a.w_out) + dst) = *(glob(a.w_in) + k);
    if (a.idx_out) *(glob(a.idx_out) + dst) = k;
  }
  if (!a.cols_out) return;  // kernel-uniform
  const uint64_t live = (a.n - kb) < (uint64_t)kBlock ? (a.n - kb) : (uint64_t)kBlock;
  const uint32_t t = threadIdx.x;
  const uint64_t dst = (t < a_blk) ? goff + t : nacc + (kb - goff) + (t - a_blk);
#pragma unroll
  for (int h = 0; h < C; h += CS) {
    if (h) __syncthreads();  // the previous columns have been read
    if (live_ev) {
#pragma unroll
      for (int c = h; c < h + CS; ++c) sh[c - h][r] = v[c];
    }
    __syncthreads();
    if (t < live) {
#pragma unroll
      for (int c = h; c < h + CS; ++c)
        *(glob_u(a.cols_out + (uint64_t)c * a.ld_out) + dst) = sh[c - h][t];
    }
  }
}

template <int N>
__global__ __launch_bounds__(kBlock) void k_uw_scatter(const UnweightArgs a) {
  __shared__ uint64_t goff[kUwTile / kWave];
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t per_xcd = (ntiles + kXcds - 1) / kXcds;
  const uint64_t tile = (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (tile >= ntiles) return;  // whole block
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ngroups = kUwTile / kWave;
  const uint64_t tile_base = a.tile_off[tile];
  const uint64_t nacc = *a.n_acc;
  // group offsets inside the tile: lanes 0..15 of wave 0 each load one flag
  // word and an inclusive shuffle scan turns the popcounts into offsets
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
    if (lane < (int)ngroups) goff[lane] = tile_base + (c - own);
  }
  __syncthreads();
  const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
  // G groups of 64 events are loaded (every column) before any is stored, so
  // a wave keeps G*4N loads in flight; G is bounded by a 64-value budget.
  constexpr int C = 4 * N;
  constexpr int G = (4 * C <= 64) ? 4 : ((2 * C <= 64) ? 2 : 1);
  const bool move_cols = a.cols_out != nullptr;
  const bool move_w = a.w_out != nullptr;
#pragma unroll
  for (int j0 = 0; j0 < kItems; j0 += G) {
    double v[G][C];
    double wv[G];
    uint64_t dst[G];
    uint64_t src[G];
    bool live[G];
#pragma unroll
    for (int h = 0; h < G; ++h) {
      const uint64_t gl = (uint64_t)(j0 + h) * (kBlock / kWave) + wave;
      const uint64_t g = tile * ngroups + gl;
      const uint64_t k0 = g * kWave;
      src[h] = k0 + (uint64_t)lane;
      live[h] = src[h] < a.n;
      dst[h] = 0;
      if (k0 < a.n) {  // wave-uniform
        // accepted lanes go to goff[gl] + (accepted lanes below), rejected
        // lanes to nacc + (k0 - goff[gl]) + (rejected lanes below)
        const uint64_t fw = a.flags[g];
        const uint32_t acc_below = (uint32_t)__popcll(fw & below);
        const bool acc = (fw >> lane) & 1ULL;
        dst[h] = acc ? goff[gl] + acc_below : nacc + k0 - goff[gl] + ((uint32_t)lane - acc_below);
        if (live[h]) {
          if (move_cols) {
#pragma unroll
            for (int c = 0; c < C; ++c)
              v[h][c] = *(glob_u(a.cols_in + (uint64_t)c * a.ld_in + k0) + lane);
          }
          if (move_w) wv[h] = *(glob(a.w_in) + src[h]);
        }
      }
    }
#pragma unroll
    for (int h = 0; h < G; ++h) {
      if (live[h]) {
        if (move_cols) {
#pragma unroll
          for (int c = 0; c < C; ++c)
            *(glob_u(a.cols_out + (uint64_t)c * a.ld_out) + dst[h]) = v[h][c];
        }
        if (a.idx_out) *(glob(a.idx_out) + dst[h]) = src[h];
        if (move_w) *(glob(a.w_out) + dst[h]) = wv[h];
      }
    }
  }
}

// Order 1: the original index of every kept sample by its rank.
__global__ __launch_bounds__(kBlock) void k_uw_kept_index(const UnweightArgs a) {
  __shared__ uint64_t goff[kUwTile / kWave];
  const uint64_t tile = blockIdx.x;
  tile_group_offsets(a, tile, goff);
  const int lane = __lane_id();
  const int wave = threadIdx.x / kWave;
  const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
  const uint64_t nacc = *a.n_acc;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint64_t gl = (uint64_t)j * (kBlock / kWave) + wave;
    const uint64_t g = tile * (kUwTile / kWave) + gl;
    const uint64_t k = g * kWave + (uint64_t)lane;
    if (k >= a.n) continue;
    bool kept;
    const uint64_t dst = stable_dst(a.flags[g], below, lane, goff[gl], nacc, g * kWave, &kept);
    if (kept) a.tidx[dst] = k;
  }
}

// Order 1 destination of a rejected sample e: in place from nacc on; below
// nacc it follows the swaps of the sequential partition: to the position of
// the kept sample of rank e, then of rank that position, while below nacc
// (each step strictly increases the position).
__device__ __forceinline__ uint64_t swap_chain_dst(const UnweightArgs& a, uint64_t e,
                                                   uint64_t nacc) {
  if (e >= nacc) return e;
  uint64_t p = a.tidx[e];
  // at most nacc steps; every index read is a kept rank below nacc
  for (uint64_t it = 0; p < nacc && it < nacc; ++it) p = a.tidx[p];
  return p;
}

// hydra::unweight's scatter of ncols value columns (any count): the same
// destinations as k_uw_scatter (or the sequential-partition ones), columns
// moved in batches of eight with all four groups' loads of a batch in flight.
__global__ __launch_bounds__(kBlock) void k_uw_scatter_cols(const UnweightArgs a, int ncols) {
  __shared__ uint64_t goff[kUwTile / kWave];
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t per_xcd = (ntiles + kXcds - 1) / kXcds;
  const uint64_t tile = (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (tile >= ntiles) return;
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ngroups = kUwTile / kWave;
  const uint64_t tile_base = a.tile_off[tile];
  const uint64_t nacc = *a.n_acc;
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
    if (lane < (int)ngroups) goff[lane] = tile_base + (c - own);
  }
  __syncthreads();
  const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
  uint64_t dst[kItems], src[kItems], k0s[kItems];
  bool live[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint64_t gl = (uint64_t)j * (kBlock / kWave) + wave;
    const uint64_t g = tile * ngroups + gl;
    const uint64_t k0 = g * kWave;
    k0s[j] = k0;
    src[j] = k0 + (uint64_t)lane;
    live[j] = src[j] < a.n;
    dst[j] = 0;
    if (k0 < a.n) {
      bool kept;
      dst[j] = stable_dst(a.flags[g], below, lane, goff[gl], nacc, k0, &kept);
      if (a.order == 1 && !kept && live[j]) dst[j] = swap_chain_dst(a, src[j], nacc);
    }
  }
  constexpr int B = 8;
  for (int c0 = 0; c0 < ncols; c0 += B) {
    double v[kItems][B];
#pragma unroll
    for (int j = 0; j < kItems; ++j)
#pragma unroll
      for (int c = 0; c < B; ++c)
        if (live[j] && c0 + c < ncols)
          v[j][c] = *(glob_u(a.cols_in + (uint64_t)(c0 + c) * a.ld_in + k0s[j]) + lane);
#pragma unroll
    for (int j = 0; j < kItems; ++j)
#pragma unroll
      for (int c = 0; c < B; ++c)
        if (live[j] && c0 + c < ncols)
          *(glob_u(a.cols_out + (uint64_t)(c0 + c) * a.ld_out) + dst[j]) = v[j][c];
  }
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    if (live[j]) {
      if (a.idx_out) *(glob(a.idx_out) + dst[j]) = src[j];
      if (a.w_out) *(glob(a.w_out) + dst[j]) = a.w_in[src[j]];
    }
  }
}

// Largest value of a column (max_element's value), per block.
__global__ __launch_bounds__(kBlock) void k_max_column(const double* w, uint64_t n,
                                                       double* block_max) {
  __shared__ double sh[kBlock];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  double mx = -__builtin_inf();
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += stride)
    mx = (mx < w[k]) ? w[k] : mx;
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      sh[threadIdx.x] = (sh[threadIdx.x] < sh[threadIdx.x + off]) ? sh[threadIdx.x + off]
                                                                  : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_max[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kBlock) void k_max_of_blocks(const double* block_max, uint64_t m,
                                                          double* out) {
  __shared__ double sh[kBlock];
  double mx = -__builtin_inf();
  for (uint64_t i = threadIdx.x; i < m; i += kBlock) mx = (mx < block_max[i]) ? block_max[i] : mx;
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off)
      sh[threadIdx.x] = (sh[threadIdx.x] < sh[threadIdx.x + off]) ? sh[threadIdx.x + off]
                                                                  : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }

// Order-independent checksum of n 64-bit words: the sum of a mix of each word
// with its index (a write that changes any word changes it with probability
// 1 - 2^-64).  One partial per block, added atomically.
__global__ __launch_bounds__(kBlock) void k_checksum(const uint64_t* w, uint64_t n, uint64_t* sum) {
  __shared__ uint64_t sh[kBlock / kWave];
  uint64_t acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) {
    uint64_t z = w[i] ^ (i * 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    acc += z ^ (z >> 31);
  }
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) acc += __shfl_xor(acc, off, kWave);
  if (__lane_id() == 0) sh[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd((unsigned long long*)sum, (unsigned long long)(sh[0] + sh[1] + sh[2] + sh[3]));
}

int launch_checksum(const uint64_t* w, uint64_t n, uint64_t* sum, hipStream_t s) {
  hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, sum, (uint64_t)0);
  if (n) hipLaunchKernelGGL(k_checksum, dim3(grid_for(n, 8192)), dim3(kBlock), 0, s, w, n, sum);
  return (int)hipGetLastError();
}

constexpr int kNumN = HPSP_MAX_N - 1;

static void launch_flags(int N, const UnweightArgs& a, uint64_t ntiles, hipStream_t s);

int launch_uw_scans(const UnweightArgs& a, hipStream_t s) {
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t nruns = (ntiles + kScanRun - 1) / kScanRun;
  hipLaunchKernelGGL(k_uw_scan_blocks, dim3((unsigned)nruns), dim3(kBlock), 0, s, a.tile_count,
                     ntiles, a.run_sum);
  hipLaunchKernelGGL(k_uw_scan_top, dim3(1), dim3(kBlock), 0, s, a.run_sum, nruns, a.n_acc);
  hipLaunchKernelGGL(k_uw_scan_tiles, dim3((unsigned)nruns), dim3(kBlock), 0, s, a.tile_count,
                     ntiles, a.run_sum, a.tile_off);
  return (int)hipGetLastError();
}

template <int N>
static void uw_one(const UnweightArgs& a, hipStream_t s) {
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  if (a.w_in && !a.use_f)
    launch_flags(0, a, ntiles, s);
  else
    launch_flags(N, a, ntiles, s);
  launch_uw_scans(a, s);
  if ((a.cols_out || a.idx_out || a.w_out) && !a.regen) {
    if constexpr (N <= 6) {
      const uint64_t nblocks = (a.n + kBlock - 1) / kBlock;
      const uint64_t grid = (nblocks + kXcds - 1) / kXcds * kXcds;
      hipLaunchKernelGGL(k_uw_scatter_staged<N>, dim3((unsigned)grid), dim3(kBlock), 0, s, a);
    } else {
      const uint64_t grid = (ntiles + kXcds - 1) / kXcds * kXcds;
      hipLaunchKernelGGL(k_uw_scatter<N>, dim3((unsigned)grid), dim3(kBlock), 0, s, a);
    }
  }
}

template <int... I>
static constexpr auto uw_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, hipStream_t), sizeof...(I)>{&uw_one<2 + I>...};
}
static const auto kUw = uw_table(std::make_integer_sequence<int, kNumN>{});

int launch_unweight_sample(int ncols, const UnweightArgs& in, double* block_max,
                           uint64_t nblocks, double* max_dev, hipStream_t s) {
  if (in.n == 0) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, in.n_acc, (uint64_t)0);
    return (int)hipGetLastError();
  }
  UnweightArgs a = in;
  if (max_dev) {  // max_element of the weights
    hipLaunchKernelGGL(k_max_column, dim3((unsigned)nblocks), dim3(kBlock), 0, s, a.w_in, a.n,
                       block_max);
    hipLaunchKernelGGL(k_max_of_blocks, dim3(1), dim3(kBlock), 0, s, block_max, nblocks, max_dev);
    a.maxv_dev = max_dev;
  }
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  launch_flags(0, a, ntiles, s);
  const uint64_t nruns = (ntiles + kScanRun - 1) / kScanRun;
  hipLaunchKernelGGL(k_uw_scan_blocks, dim3((unsigned)nruns), dim3(kBlock), 0, s, a.tile_count,
                     ntiles, a.run_sum);
  hipLaunchKernelGGL(k_uw_scan_top, dim3(1), dim3(kBlock), 0, s, a.run_sum, nruns, a.n_acc);
  hipLaunchKernelGGL(k_uw_scan_tiles, dim3((unsigned)nruns), dim3(kBlock), 0, s, a.tile_count,
                     ntiles, a.run_sum, a.tile_off);
  if (a.cols_out || a.idx_out || a.w_out) {
    if (a.order == 1)
      hipLaunchKernelGGL(k_uw_kept_index, dim3((unsigned)ntiles), dim3(kBlock), 0, s, a);
    const uint64_t grid = (ntiles + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL(k_uw_scatter_cols, dim3((unsigned)grid), dim3(kBlock), 0, s, a,
                       a.cols_out ? ncols : 0);
  }
  return (int)hipGetLastError();
}

template <int N>
static void flags_one(const UnweightArgs& a, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL(k_uw_flags<N>, dim3(grid), dim3(kBlock), 0, s, a);
}
template <int... I>
static constexpr auto flags_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, unsigned, hipStream_t), sizeof...(I)>{
      &flags_one<(I == 0 ? 0 : I + 1)>...};
}
// index 0: weight column only; N - 1: momenta of N daughters
static const auto kFlags = flags_table(std::make_integer_sequence<int, HPSP_MAX_N>{});

static void launch_flags(int N, const UnweightArgs& a, uint64_t ntiles, hipStream_t s) {
  const unsigned grid = (unsigned)((ntiles + kXcds - 1) / kXcds * kXcds);
  kFlags[N == 0 ? 0 : N - 1](a, grid, s);
}

int launch_max_column(const double* w, uint64_t n, double* block_max, uint64_t nblocks,
                      double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_max_column, dim3((unsigned)nblocks), dim3(kBlock), 0, s, w, n, block_max);
  hipLaunchKernelGGL(k_max_of_blocks, dim3(1), dim3(kBlock), 0, s, block_max, nblocks, out);
  return (int)hipGetLastError();
}

int launch_unweight(int N, const UnweightArgs& a, hipStream_t s) {
  if (a.n == 0) {
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, s, a.n_acc, (uint64_t)0);
    return (int)hipGetLastError();
  }
  kUw[N - 2](a, s);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Engine words (inspection / parity): thread t produces word t (squares) or
// block t (Random123, A words) of the stream after discard(d).

template <int ENGINE>
__global__ __launch_bounds__(kBlock) void k_rng_words(const uint64_t* key_dev, uint64_t k0,
                                                      uint64_t k1, uint64_t k2, uint64_t k3,
                                                      uint64_t d, uint64_t n, uint64_t* out) {
  (void)key_dev;
  const uint64_t key[4] = {k0, k1, k2, k3};
  constexpr int A = engine_arity<ENGINE>();
  const uint64_t nitems = (n + A - 1) / A;
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t < nitems; t += stride) {
    if constexpr (is_squares<ENGINE>()) {
      out[t] = squares_word<ENGINE>(key[0], d + t);
    } else {
      uint64_t ctr[4] = {0, 0, 0, 0};
      ctr_add<A>(ctr, d);
      ctr_add<A>(ctr, t + 1);
      uint64_t blk[4];
      r123_block<ENGINE>(key, ctr, blk);
#pragma unroll
      for (int i = 0; i < A; ++i)
        if (t * A + i < n) out[t * A + i] = blk[i];
    }
  }
}

int launch_rng_words(int engine, const uint64_t* key, uint64_t discard, uint64_t n,
                     uint64_t* out, hipStream_t s) {
  const unsigned g = grid_for(n);
  switch (engine) {
#define HPSP_RNG_CASE(E)                                                                   \
  case E:                                                                                  \
    hipLaunchKernelGGL(k_rng_words<E>, dim3(g), dim3(kBlock), 0, s, nullptr, key[0], key[1], \
                       key[2], key[3], discard, n, out);                                   \
    break;
    HPSP_RNG_CASE(0)
    HPSP_RNG_CASE(1)
    HPSP_RNG_CASE(2)
    HPSP_RNG_CASE(3)
    HPSP_RNG_CASE(4)
    HPSP_RNG_CASE(5)
#undef HPSP_RNG_CASE
    default:
      return (int)hipErrorInvalidValue;
  }
  return (int)hipGetLastError();
}

}  // namespace hpsp
