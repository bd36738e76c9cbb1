// FP64 phase-space kernels for gfx950: GENBOD generation into SoA columns,
// AverageOn (fused generation + functor + StatsPHSP reduction) and the
// per-event weight kernels.  Built with -ffp-contract=off; the per-event
// GENBOD and its numerics are in genbod_f64.h.
#include <array>
#include <type_traits>
#include <utility>

#include "genbod_f64.h"
#include "uw_dst.h"

namespace hpsp {

// Each wave takes runs of a.run consecutive groups of 64 events (grid-stride
// over runs): with the squares engines a group's second 64 words are the next
// group's first, so within a run each lane evaluates one word per group
// instead of two.  COLS = false is the weights pass of a deferred Generate
// (no momentum columns): without the column stores' addresses and store
// paths in its loop, the scalar registers it spills are reloaded outside the
// loop only.
template <int N, int ENGINE, bool MOTHERS, bool COLS = true>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_generate_f64(const GenArgs a) {
  const int lane = __lane_id();
  // wave index as a scalar so that column bases stay in SGPRs and every
  // store is (scalar base) + lane * 8
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  const uint64_t run_len = a.run ? a.run : 1;  // the launcher sets it; 0 reads as 1
  const uint64_t nruns = (ngroups + run_len - 1) / run_len;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t run = (uint64_t)blockIdx.x * (kBlock / kWave) + (uint64_t)wave; run < nruns;
       run += nwaves) {
    const uint64_t g_first = run * run_len;
    const uint64_t g_end = (g_first + run_len < ngroups) ? g_first + run_len : ngroups;
    EventWords<N, ENGINE> words;
    for (uint64_t g = g_first; g < g_end; ++g) {
      const uint64_t k0 = g * kWave;
      const uint64_t k = k0 + (uint64_t)lane;
      const bool live = k < a.n;

      if (g == g_first)
        words.init(a.d, a.first + k0, lane);
      else
        words.advance(a.d, a.first + k0, lane);
      double p[N][4];
      const double wt = genbod_f64<N, ENGINE>(a.d, words, p);

      if constexpr (MOTHERS) {
        // applyBoostTo(mother) with beta = p/E (DecayMothers.h:228-233).
        const uint64_t km = live ? k : 0;
        const double e = a.mothers[km];
        const double bx = a.mothers[a.mld + km] / e;
        const double by = a.mothers[2 * a.mld + km] / e;
        const double bz = a.mothers[3 * a.mld + km] / e;
#pragma unroll
        for (int i = 0; i < N; ++i) boost_general(p[i], bx, by, bz);
      } else {
        if (a.d.do_boost) {
#pragma unroll
          for (int i = 0; i < N; ++i) boost_folded(p[i], a.d);
        }
      }

      if (live) {
        if (!COLS || !a.cols) {
          // weights only: the momenta are formed for them and not stored
        } else if (a.out_f32) {
          float* c = static_cast<float*>(a.cols) + k0;
#pragma unroll
          for (int i = 0; i < N; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              st(uniform_ptr(c + (uint64_t)(4 * i + q) * a.ld) + lane, (float)p[i][q]);
        } else {
          double* c = static_cast<double*>(a.cols) + k0;
#pragma unroll
          for (int i = 0; i < N; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              st(uniform_ptr(c + (uint64_t)(4 * i + q) * a.ld) + lane, p[i][q]);
        }
        if (a.wt) st(uniform_ptr(a.wt + k0) + lane, wt);
        if (a.evw) st(uniform_ptr(a.evw + k0) + lane, event_weight_f64<N>(p, a.d.m, a.d.maxw));
      }
    }
  }
}

// Decays::Unweight on a container that holds events gen_first + k of one
// Generate (the flags and tile offsets already computed from its
// PhaseSpaceWeight column): instead of reading every event and moving it to
// its place in the stable partition, each event is generated again, with the
// same operations and so the same bits, and written straight to its place.
// The events are a function of (decay, seed, index) alone, so the momenta are
// never read back: 32N bytes written per event (+16 with the weight column)
// instead of 64N moved.
//
// One group of 64 events per wave (with the squares engines the block's
// 256 + W - 1 engine words are evaluated once each and shared through LDS, so
// an event costs one word, as in k_generate_f64's runs).  The block's 256
// events fall into two contiguous runs of the output, its accepted events and
// its rejected ones; for N <= 6 they are staged in LDS in partition order and
// every column is written by all 256 threads as those two runs, so that only
// the ends of a block's runs, not of every wave's, are partial cache lines
// (8% faster at N = 3).  Larger N stores from registers (their LDS would cost
// occupancy).  Blocks are dealt to XCDs in contiguous ranges, so that
// neighbouring blocks, which write neighbouring pieces of both runs, meet in
// one L2.  A block's offset comes from the tile offset and a 16-lane scan of
// the popcounts of its tile's flag words (block_part).
// Squares words of one event from the block's words in LDS: thread t's own
// word (counter c0 + t) and words j >= 1 at w[j], w = sw + t.
struct BlockWords {
  const uint32_t* w;
  uint32_t own;
  __device__ __forceinline__ uint64_t operator()(int j) const { return j == 0 ? own : w[j]; }
};

template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_uw_regen(const UnweightArgs a) {
  constexpr int C = 4 * N;
  constexpr bool kStage = N <= 6;
  // columns staged at once: all of them up to 8, else half, so that a block's
  // LDS (12 KB at N = 3) does not cap the waves in flight below the registers
  constexpr int CS = !kStage ? 1 : (C <= 8 ? C : C / 2);
  __shared__ double sh[CS][kStage ? kBlock : 1];
  // squares engines: the block's engine words (see BlockWords)
  constexpr int W = 3 * N - 4;
  __shared__ uint32_t sw[is_squares<ENGINE>() ? kBlock + W - 1 : 1];
  constexpr int kWavesPerBlock = kBlock / kWave;
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  const uint64_t nblocks = (ngroups + kWavesPerBlock - 1) / kWavesPerBlock;
  const uint64_t per_xcd = (nblocks + kXcds - 1) / kXcds;
  const uint64_t blk = (uint64_t)(blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (blk >= nblocks) return;  // whole block
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t g0 = blk * kWavesPerBlock;  // the block's first group
  const BlockPart bp = block_part(a, g0, wave, ngroups);
  const uint64_t goff = bp.goff;  // accepted events before the block
  const uint32_t acc_wave = bp.acc_wave, a_blk = bp.a_blk;
  const uint64_t nacc = *a.n_acc;
  const uint64_t kb = g0 * kWave;  // the block's first event
  const uint64_t g = g0 + (uint64_t)wave;
  const uint64_t k0 = g * kWave;
  const uint64_t k = k0 + (uint64_t)lane;
  const bool live_wave = g < ngroups;  // wave-uniform
  uint32_t own = 0;
  if constexpr (is_squares<ENGINE>()) {
    // event kb + t reads the words at counters c0 + t + j, j < W: thread t
    // evaluates the one at c0 + t, threads t < W - 1 also the one at c0 + 256 + t
    const uint64_t c0 = a.gen_first + kb + 3 * (uint64_t)N;
    const uint32_t t = threadIdx.x;
    own = squares_word_y<ENGINE>(a.d.key[0], a.d.key[0] * (c0 + t));
    sw[t] = own;
    if (t < (uint32_t)(W - 1))
      sw[kBlock + t] = squares_word_y<ENGINE>(a.d.key[0], a.d.key[0] * (c0 + kBlock + t));
    __syncthreads();
  }
  double p[N][4];
  uint32_t r = 0;  // the event's rank in the block's partition order
  if (live_wave) {
    const uint64_t fw = bp.fw;
    if constexpr (is_squares<ENGINE>()) {
      genbod_f64<N, ENGINE>(a.d, BlockWords{sw + threadIdx.x, own}, p);
    } else {
      EventWords<N, ENGINE> words;
      words.init(a.d, a.gen_first + k0, lane);
      genbod_f64<N, ENGINE>(a.d, words, p);
    }
    if (a.d.do_boost) {
#pragma unroll
      for (int i = 0; i < N; ++i) boost_folded(p[i], a.d);
    }
    const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const uint32_t acc_below = acc_wave + (uint32_t)__popcll(fw & below);
    const bool kept = (fw >> lane) & 1ULL;
    const uint32_t t = (uint32_t)(wave * kWave + lane);
    r = kept ? acc_below : a_blk + (t - acc_below);
    if (k < a.n) {
      const uint64_t dst = (r < a_blk) ? goff + r : nacc + (kb - goff) + (r - a_blk);
      if constexpr (!kStage) {
        if (a.cols_out) {
#pragma unroll
          for (int i = 0; i < N; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              *(glob_u(a.cols_out + (uint64_t)(4 * i + q) * a.ld_out) + dst) = p[i][q];
        }
      }
      if (a.w_out)
        *(glob(a.w_out) + dst) =
            a.w_in ? *(glob(a.w_in) + k) : event_weight_f64<N>(p, a.d.m, a.d.maxw);
      if (a.idx_out) *(glob(a.idx_out) + dst) = k;
    }
  }
  if constexpr (kStage) {
    const uint64_t live = (a.n - kb) < (uint64_t)kBlock ? (a.n - kb) : (uint64_t)kBlock;
    const uint32_t t = threadIdx.x;
    const uint64_t dst = (t < a_blk) ? goff + t : nacc + (kb - goff) + (t - a_blk);
#pragma unroll
    for (int h = 0; h < C; h += CS) {
      if (h) __syncthreads();  // the previous columns have been read
      if (live_wave && k < a.n) {
#pragma unroll
        for (int c = h; c < h + CS; ++c) sh[c - h][r] = p[c / 4][c % 4];
      }
      __syncthreads();
      if (a.cols_out && t < live) {
#pragma unroll
        for (int c = h; c < h + CS; ++c)
          *(glob_u(a.cols_out + (uint64_t)c * a.ld_out) + dst) = sh[c - h][t];
      }
    }
  }
}

// Merge m partials in a fixed order: thread t folds a contiguous slice, then
// a pairwise tree over the 256 slice results.
__global__ __launch_bounds__(kBlock) void k_stats_finalize(const Stats* part, uint64_t m,
                                                           Stats* out) {
  __shared__ Stats sh[kBlock];
  const uint64_t per = (m + kBlock - 1) / kBlock;
  const uint64_t b = (uint64_t)threadIdx.x * per;
  const uint64_t e = (b + per < m) ? b + per : m;
  Stats acc = {0.0, 0.0, 0.0};
  for (uint64_t i = b; i < e; ++i) acc = stats_merge(acc, part[i]);
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    if ((threadIdx.x & (2 * off - 1)) == 0) sh[threadIdx.x] = stats_merge(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

template <int N>
__device__ __forceinline__ void load_event(const double* cols, uint64_t ld, uint64_t k,
                                           double (&p)[N][4]) {
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) p[i][q] = __builtin_nontemporal_load(cols + (uint64_t)(4 * i + q) * ld + k);
}

// PhaseSpaceWeight (and PhaseSpaceReweight) of stored events.
template <int N>
__global__ __launch_bounds__(kBlock) void k_event_weight(const WeightArgs a) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < a.n; k += stride) {
    double p[N][4];
    load_event<N>(a.cols, a.ld, k, p);
    double w = event_weight_f64<N>(p, a.d.m, a.d.maxw);
    if (a.use_f) w = w * functor_eval<double>(a.f, p);
    a.w_out[k] = w;
  }
}

// Block maxima of w*f (Decays.inl:488-490), then one block folds them.
template <int N>
__global__ __launch_bounds__(kBlock) void k_max_weight(const WeightArgs a, double* block_max) {
  __shared__ double sh[kBlock];
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  double mx = -__builtin_inf();
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < a.n; k += stride) {
    double p[N][4];
    load_event<N>(a.cols, a.ld, k, p);
    double w = event_weight_f64<N>(p, a.d.m, a.d.maxw);
    if (a.use_f) w = w * functor_eval<double>(a.f, p);
    mx = (mx < w) ? w : mx;
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = (sh[threadIdx.x] < sh[threadIdx.x + off]) ? sh[threadIdx.x + off] : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) block_max[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(kBlock) void k_max_finalize(const double* block_max, uint64_t m,
                                                         double* out) {
  __shared__ double sh[kBlock];
  double mx = -__builtin_inf();
  for (uint64_t i = threadIdx.x; i < m; i += kBlock) mx = (mx < block_max[i]) ? block_max[i] : mx;
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if (threadIdx.x < off) sh[threadIdx.x] = (sh[threadIdx.x] < sh[threadIdx.x + off]) ? sh[threadIdx.x + off] : sh[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// ---------------------------------------------------------------------------
// Launch tables over N = 2..HPSP_MAX_N and the six engines.

constexpr int kNumN = HPSP_MAX_N - 1;
constexpr int kNumE = 6;

template <int N, int E>
static void gen_one(const GenArgs& in, hipStream_t s) {
  GenArgs a = in;
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  // momentum columns of at most 3 daughters (with or without the weight
  // column): bound by the stores (more daughters bring arithmetic enough to
  // gain from 16)
  uint64_t cap = (a.cols && N <= 3) ? kRunMaxStores : kRunMax;
  if (a.run_cap && a.run_cap < cap) cap = a.run_cap;
  a.run = run_groups(ngroups, cap);
  const unsigned g = grid_for_runs((ngroups + a.run - 1) / a.run);
  if (a.mothers)
    hipLaunchKernelGGL((k_generate_f64<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
  else if (!a.cols)
    hipLaunchKernelGGL((k_generate_f64<N, E, false, false>), dim3(g), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_generate_f64<N, E, false>), dim3(g), dim3(kBlock), 0, s, a);
}

template <int N, int E>
static void avg_one(const AvgArgs& a, hipStream_t s) {
  const uint64_t nchunks = (a.n + kAvgChunk - 1) / kAvgChunk;
  const unsigned g = (unsigned)(nchunks < 8192 ? (nchunks ? nchunks : 1) : 8192);
  if (a.mothers)
    hipLaunchKernelGGL((k_average_f64<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_average_f64<N, E, false>), dim3(g), dim3(kBlock), 0, s, a);
}

template <int N, int E>
static void regen_one(const UnweightArgs& a, hipStream_t s) {
  const uint64_t nblocks = (a.n + kBlock - 1) / kBlock;
  const uint64_t grid = (nblocks + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL((k_uw_regen<N, E>), dim3((unsigned)grid), dim3(kBlock), 0, s, a);
}

template <int I>
static constexpr auto gen_at = &gen_one<2 + I / kNumE, I % kNumE>;
template <int I>
static constexpr auto regen_at = &regen_one<2 + I / kNumE, I % kNumE>;
template <int I>
static constexpr auto avg_at = &avg_one<2 + I / kNumE, I % kNumE>;

template <int... I>
static constexpr auto gen_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const GenArgs&, hipStream_t), sizeof...(I)>{gen_at<I>...};
}
template <int... I>
static constexpr auto regen_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, hipStream_t), sizeof...(I)>{regen_at<I>...};
}
template <int... I>
static constexpr auto avg_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const AvgArgs&, hipStream_t), sizeof...(I)>{avg_at<I>...};
}

static const auto kGen = gen_table(std::make_integer_sequence<int, kNumN * kNumE>{});
static const auto kAvg = avg_table(std::make_integer_sequence<int, kNumN * kNumE>{});
static const auto kRegen = regen_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_unweight_regen(int N, int engine, const UnweightArgs& a, hipStream_t s) {
  if (a.n == 0) return (int)hipSuccess;
  kRegen[(N - 2) * kNumE + engine](a, s);
  return (int)hipGetLastError();
}

int launch_generate_f64(int N, int engine, const GenArgs& a, hipStream_t s) {
  kGen[(N - 2) * kNumE + engine](a, s);
  return (int)hipGetLastError();
}

int launch_average_f64(int N, int engine, const AvgArgs& a, hipStream_t s) {
  kAvg[(N - 2) * kNumE + engine](a, s);
  return (int)hipGetLastError();
}

int launch_stats_finalize(const Stats* partials, uint64_t m, Stats* out, hipStream_t s) {
  hipLaunchKernelGGL(k_stats_finalize, dim3(1), dim3(kBlock), 0, s, partials, m, out);
  return (int)hipGetLastError();
}

template <int N>
static void ew_one(const WeightArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_event_weight<N>, dim3(grid_for(a.n)), dim3(kBlock), 0, s, a);
}
template <int N>
static void mx_one(const WeightArgs& a, double* bm, uint64_t nb, hipStream_t s) {
  hipLaunchKernelGGL(k_max_weight<N>, dim3((unsigned)nb), dim3(kBlock), 0, s, a, bm);
}

template <int... I>
static constexpr auto ew_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const WeightArgs&, hipStream_t), sizeof...(I)>{&ew_one<2 + I>...};
}
template <int... I>
static constexpr auto mx_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const WeightArgs&, double*, uint64_t, hipStream_t), sizeof...(I)>{
      &mx_one<2 + I>...};
}
static const auto kEw = ew_table(std::make_integer_sequence<int, kNumN>{});
static const auto kMx = mx_table(std::make_integer_sequence<int, kNumN>{});

int launch_event_weight(int N, const WeightArgs& a, hipStream_t s) {
  kEw[N - 2](a, s);
  return (int)hipGetLastError();
}

int launch_max_weight(int N, const WeightArgs& a, double* block_max, uint64_t nblocks,
                      double* out, hipStream_t s) {
  kMx[N - 2](a, block_max, nblocks, s);
  hipLaunchKernelGGL(k_max_finalize, dim3(1), dim3(kBlock), 0, s, block_max, nblocks, out);
  return (int)hipGetLastError();
}

}  // namespace hpsp
