// Mixed-precision phase-space kernels for gfx950.  The weight chain stays
// FP64 in the reference's order (the uniforms of the invariant masses,
// invMas, pdk -- where a - b - c cancels near threshold -- and the weight
// product), with one-ulp square root and reciprocal, so generation weights
// agree with the FP64 ones to a few ulp.  The kinematics after pdk
// (energies, rotations, boosts, functor) run in FP32 with FMA contraction,
// hardware sin/cos in revolutions, square roots and reciprocals.
// Accumulation of AverageOn stays FP64: shifted sums per lane, then the
// StatsPHSP merges of the FP64 kernel.
#include <array>
#include <type_traits>
#include <utility>

#include "genbod_mixed.h"
#include "hpsp_launch.h"
#include "uw_dst.h"

namespace hpsp {

// The 32 bits of a word that feed the FP32 angles (Random123 words: the top
// 32 bits, so u differs from the FP64 path by less than 2^-32).
template <int ENGINE>
__device__ __forceinline__ uint32_t word32(uint64_t w) {
  if constexpr (is_squares<ENGINE>())
    return (uint32_t)w;
  else
    return (uint32_t)(w >> 32);
}

// cos and sin of the polar angle from u = w / 2^32 without cancellation:
// cZ = 2u - 1 = (2w - 2^32) / 2^32 and sZ = sqrt(1 - cZ^2) = 2 sqrt(w (2^32 - w)) / 2^32,
// both formed from exact integers and rounded once to FP32.  Only 32-bit
// conversions: w - 2^31 is w ^ 2^31 read as signed, and 2^32 - w is 0 - w in
// 32 bits except at w = 0, where a = 0 makes the product 0 either way (the
// same bits as the 64-bit forms for every w).
__device__ __forceinline__ void polar_f32(uint32_t w, float& cZ, float& sZ) {
  cZ = (float)(int32_t)(w ^ 0x80000000u) * 0x1p-31f;
  const float a = (float)w;
  const float b = (float)(0u - w);
  sZ = __builtin_amdgcn_sqrtf(a * b) * 0x1p-31f;
}

template <int N, int ENGINE, typename WordFn>
__device__ __forceinline__ double genbod_mixed(const DecayK& d, const WordFn& word,
                                               float (&p)[N][4]) {
#pragma clang fp contract(off)
  float pd[N - 1];
  double invMas[N];
  const double wt = weight_chain<N, ENGINE>(d, word, invMas, pd);

  const float m0 = (float)d.m[0];
  p[0][0] = __builtin_amdgcn_sqrtf(__builtin_fmaf(pd[0], pd[0], m0 * m0));
  p[0][1] = 0.0f;
  p[0][2] = pd[0];
  p[0][3] = 0.0f;

#pragma unroll
  for (int i = 1; i < N; ++i) {
    const float mi = (float)d.m[i];
    p[i][0] = __builtin_amdgcn_sqrtf(__builtin_fmaf(pd[i - 1], pd[i - 1], mi * mi));
    p[i][1] = 0.0f;
    p[i][2] = -pd[i - 1];
    p[i][3] = 0.0f;

    float cZ, sZ;
    polar_f32(word32<ENGINE>(word(N - 2 + 2 * (i - 1))), cZ, sZ);
    const float rev = (float)word32<ENGINE>(word(N - 2 + 2 * (i - 1) + 1)) * 0x1p-32f;
    // v_sin_f32 / v_cos_f32 take the angle in revolutions: sin(2*pi*rev)
    const float sY = __builtin_amdgcn_sinf(rev);
    const float cY = __builtin_amdgcn_cosf(rev);

#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const bool fresh = (j == i) || (i == 1);
      if (fresh) {
        const float y = p[j][2];
        const float t = -(sZ * y);
        p[j][2] = cZ * y;
        p[j][1] = cY * t;
        p[j][3] = sY * t;
      } else {
        float x = p[j][1];
        const float y = p[j][2];
        p[j][1] = __builtin_fmaf(cZ, x, -(sZ * y));
        p[j][2] = __builtin_fmaf(sZ, x, cZ * y);
        x = p[j][1];
        const float z = p[j][3];
        p[j][1] = __builtin_fmaf(cY, x, -(sY * z));
        p[j][3] = __builtin_fmaf(sY, x, cY * z);
      }
    }
    if (i == N - 1) break;

    // beta = pd/sqrt(pd^2 + M^2); gamma = 1/sqrt(1-beta^2);
    // (gamma-1)/beta^2 written as gamma^2/(gamma+1) (no cancellation).
    const float M = (float)invMas[i];
    const float pi2 = pd[i] * pd[i];
    const float beta = pd[i] * __builtin_amdgcn_rsqf(__builtin_fmaf(M, M, pi2));
    const float b2 = beta * beta;
    const float gamma = __builtin_amdgcn_rsqf(1.0f - b2);
    const float gb2yy = gamma * gamma * __builtin_amdgcn_rcpf(gamma + 1.0f) * b2;
    const float gby = gamma * beta;
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const float e2 = p[j][0], py2 = p[j][2];
      p[j][0] = __builtin_fmaf(gamma, e2, gby * py2);
      p[j][2] = __builtin_fmaf(gby, e2, __builtin_fmaf(gb2yy, py2, py2));
    }
  }
  return wt;
}

__device__ __forceinline__ void boost_general_f32(float (&v)[4], float bx, float by, float bz) {
#pragma clang fp contract(off)
  const float b2 = __builtin_fmaf(bz, bz, __builtin_fmaf(by, by, bx * bx));
  if (b2 > 0.0f && b2 < 1.0f) {
    const float gamma = __builtin_amdgcn_rsqf(1.0f - b2);
    const float gb2 = gamma * gamma * __builtin_amdgcn_rcpf(gamma + 1.0f);
    const float e2 = v[0], px2 = v[1], py2 = v[2], pz2 = v[3];
    const float bp = __builtin_fmaf(bz, pz2, __builtin_fmaf(by, py2, bx * px2));
    v[0] = gamma * (e2 + bp);
    const float k = __builtin_fmaf(gb2, bp, gamma * e2);
    v[1] = __builtin_fmaf(k, bx, px2);
    v[2] = __builtin_fmaf(k, by, py2);
    v[3] = __builtin_fmaf(k, bz, pz2);
  }
}

template <int N, int ENGINE>
struct EventWordsM {
  static constexpr int W = 3 * N - 4;
  std::conditional_t<is_squares<ENGINE>(), WaveWords, char> ww;
  uint64_t y_hi;  // squares: key * counter of this lane's second word
  uint64_t w[is_squares<ENGINE>() ? 1 : W];
  __device__ __forceinline__ void init(const DecayK& d, uint64_t wave_first_evt, int lane) {
    if constexpr (is_squares<ENGINE>()) {
      // key * counter of this lane's two words, 64 counters apart: the
      // second product is the first plus key * 64
      const uint64_t c0 = wave_first_evt + 3 * (uint64_t)N;
      ww.lane = lane;
      const uint64_t y_lo = d.key[0] * (c0 + (uint64_t)lane);
      y_hi = y_lo + (d.key[0] << 6);
      ww.lo = squares_word_y<ENGINE>(d.key[0], y_lo);
      ww.hi = squares_word_y<ENGINE>(d.key[0], y_hi);
    } else {
      r123_words<ENGINE, W>(d.key, wave_first_evt + (uint64_t)lane + 3 * (uint64_t)N, w);
    }
  }
  // The next group (first event wave_first_evt): with the squares engines its
  // first 64 words are this group's second ones.
  __device__ __forceinline__ void advance(const DecayK& d, uint64_t wave_first_evt, int lane) {
    if constexpr (is_squares<ENGINE>()) {
      ww.lo = ww.hi;
      y_hi += d.key[0] << 6;  // counter + 64
      ww.hi = squares_word_y<ENGINE>(d.key[0], y_hi);
    } else {
      init(d, wave_first_evt, lane);
    }
  }
  __device__ __forceinline__ uint64_t operator()(int j) const {
    if constexpr (is_squares<ENGINE>())
      return ww.get(j);
    else
      return w[j];
  }
};

template <int N>
__device__ __forceinline__ void final_boost_f32(float (&p)[N][4], const DecayK& d) {
  if (d.do_boost) {
#pragma unroll
    for (int i = 0; i < N; ++i) boost_general_f32(p[i], (float)d.bx, (float)d.by, (float)d.bz);
  }
}

// Runs of a.run consecutive groups of 64 events per wave, as in
// k_generate_f64 (one squares word per lane and group within a run).
template <int N, int ENGINE, bool MOTHERS>
__global__ __launch_bounds__(kBlock) void k_generate_mixed(const GenArgs a) {
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  const uint64_t run_len = a.run ? a.run : 1;  // the launcher sets it; 0 reads as 1
  const uint64_t nruns = (ngroups + run_len - 1) / run_len;
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  for (uint64_t run = (uint64_t)blockIdx.x * (kBlock / kWave) + (uint64_t)wave; run < nruns;
       run += nwaves) {
    const uint64_t g_first = run * run_len;
    const uint64_t g_end = (g_first + run_len < ngroups) ? g_first + run_len : ngroups;
    EventWordsM<N, ENGINE> words;
    for (uint64_t g = g_first; g < g_end; ++g) {
      const uint64_t k0 = g * kWave;
      const uint64_t k = k0 + (uint64_t)lane;
      const bool live = k < a.n;
      if (g == g_first)
        words.init(a.d, a.first + k0, lane);
      else
        words.advance(a.d, a.first + k0, lane);
      float p[N][4];
      const double wt = genbod_mixed<N, ENGINE>(a.d, words, p);
      if constexpr (MOTHERS) {
        const uint64_t km = live ? k : 0;
        const float ie = __builtin_amdgcn_rcpf((float)a.mothers[km]);
        const float bx = (float)a.mothers[a.mld + km] * ie;
        const float by = (float)a.mothers[2 * a.mld + km] * ie;
        const float bz = (float)a.mothers[3 * a.mld + km] * ie;
#pragma unroll
        for (int i = 0; i < N; ++i) boost_general_f32(p[i], bx, by, bz);
      } else {
        final_boost_f32<N>(p, a.d);
      }
      if (live) {
        if (!a.cols) {
          // weights only
        } else if (a.out_f32) {
          float* c = static_cast<float*>(a.cols) + k0;
#pragma unroll
          for (int i = 0; i < N; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              *(glob_u(c + (uint64_t)(4 * i + q) * a.ld) + lane) = p[i][q];
        } else {
          double* c = static_cast<double*>(a.cols) + k0;
#pragma unroll
          for (int i = 0; i < N; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q)
              *(glob_u(c + (uint64_t)(4 * i + q) * a.ld) + lane) = (double)p[i][q];
        }
        if (a.wt) *(glob_u(a.wt + k0) + lane) = wt;
        // PhaseSpaceWeight of the mixed form: the FP64 generation weight,
        // which it equals in exact arithmetic.  Recomputed from the FP32
        // momenta (masses of running sums, then pdk) it would lose up to all
        // of its digits where a - b - c cancels near threshold.
        if (a.evw) *(glob_u(a.evw + k0) + lane) = wt;
      }
    }
  }
}

// One lane's events of the mixed AverageOn in FP64: the sums of w, w d and
// w d^2 with d = f - K, K the lane's first value (the shifted-data form of
// the weighted variance, exact for a constant functor), with no division per
// event; stats() turns them into the (w, mean, m2) that the StatsPHSP merges
// of the block and chunk stages take.
struct ShiftedSums {
  double k = 0.0, w = 0.0, s1 = 0.0, s2 = 0.0;
  __device__ __forceinline__ void push(double f, double wt) {
#pragma clang fp contract(off)
    if (w == 0.0) k = f;
    const double d = f - k;
    const double wd = wt * d;
    w += wt;
    s1 += wd;
    s2 = __builtin_fma(wd, d, s2);
  }
  __device__ __forceinline__ Stats stats() const {
#pragma clang fp contract(off)
    if (w == 0.0) return Stats{0.0, 0.0, 0.0};
    const double r = s1 / w;
    const double m2 = s2 - s1 * r;
    // rounding can leave a tiny negative sum of squares; NaN passes on
    return Stats{w, k + r, m2 < 0.0 ? 0.0 : m2};
  }
};

template <int N, int ENGINE, bool MOTHERS>
__global__ __launch_bounds__(kBlock) void k_average_mixed(const AvgArgs a) {
  __shared__ Stats sh[kBlock / kWave];
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t nchunks = (a.n + kAvgChunk - 1) / kAvgChunk;
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    ShiftedSums acc;
    // wave w folds groups 16w .. 16w+15 of the chunk in order (as the FP64
    // kernel), carrying the squares words from group to group
    constexpr int kGroupsPerWave = kAvgChunk / kBlock;
    EventWordsM<N, ENGINE> words;
    for (int it = 0; it < kGroupsPerWave; ++it) {
      const uint64_t k0 =
          c * kAvgChunk + ((uint64_t)wave * kGroupsPerWave + (uint64_t)it) * kWave;
      const uint64_t k = k0 + (uint64_t)lane;
      const bool live = k < a.n;
      if (it == 0)
        words.init(a.d, a.first + k0, lane);
      else
        words.advance(a.d, a.first + k0, lane);
      float p[N + (MOTHERS ? 1 : 0)][4];
      double wt;
      if constexpr (MOTHERS) {
        float q[N][4];
        wt = genbod_mixed<N, ENGINE>(a.d, words, q);
        const uint64_t km = live ? k : 0;
        float m4[4];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) m4[cc] = (float)a.mothers[(uint64_t)cc * a.mld + km];
        const float ie = __builtin_amdgcn_rcpf(m4[0]);
        const float bx = m4[1] * ie, by = m4[2] * ie, bz = m4[3] * ie;
#pragma unroll
        for (int i = 0; i < N; ++i) {
          boost_general_f32(q[i], bx, by, bz);
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) p[i + 1][cc] = q[i][cc];
        }
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) p[0][cc] = m4[cc];
      } else {
        wt = genbod_mixed<N, ENGINE>(a.d, words, p);
        final_boost_f32<N>(p, a.d);
      }
      const double f = functor_eval<float, false>(a.f, p);
      if (live) acc.push(f, wt);
    }
    const Stats st = wave_reduce_stats(acc.stats());
    if (lane == 0) sh[wave] = st;
    __syncthreads();
    if (threadIdx.x == 0)
      a.partials[c] = stats_merge(stats_merge(sh[0], sh[1]), stats_merge(sh[2], sh[3]));
    __syncthreads();
  }
}

// ---- Generate followed by Unweight in one call, mixed precision ----------
// The container of a mixed Generate holds FP32 momenta and, as its weight
// column, the FP64 generation weight (k_generate_mixed).  Decays::Unweight's
// test on that column, RN(w / max) > u (FlagDaugthers, Decays.inl:75-89),
// therefore needs only each event's FP64 weight chain (k_uw_flags_mixed,
// generate_unweight.hip); k_uw_regen_mixed then generates every event again,
// with the same code and so the same bits, at its place in the stable
// partition.

// Squares words of one event from the block's words in LDS (as k_uw_regen):
// thread t's own word (counter c0 + t) and words j >= 1 at w[j], w = sw + t.
struct BlockWordsM {
  const uint32_t* w;
  uint32_t own;
  __device__ __forceinline__ uint64_t operator()(int j) const { return j == 0 ? own : w[j]; }
};

// The regenerating scatter of k_uw_regen with the mixed GENBOD: one group of
// 64 events per wave, the block's 256 events staged in LDS in partition
// order (all 4N FP32 columns at once: 12 KB at N = 3) and each column written
// as the block's two runs; cols_out holds FP32 columns (UnweightArgs carries
// it as a double pointer).  w_out receives each output row's weight column
// value, the FP64 generation weight.
template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_uw_regen_mixed(const UnweightArgs a) {
  constexpr int C = 4 * N;
  constexpr int CS = C <= 24 ? C : C / 2;  // columns staged at once
  __shared__ float sh[CS][kBlock];
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
  const uint64_t g0 = blk * kWavesPerBlock;
  const BlockPart bp = block_part(a, g0, wave, ngroups);
  const uint64_t goff = bp.goff;
  const uint32_t acc_wave = bp.acc_wave, a_blk = bp.a_blk;
  const uint64_t nacc = *a.n_acc;
  const uint64_t kb = g0 * kWave;
  const uint64_t g = g0 + (uint64_t)wave;
  const uint64_t k0 = g * kWave;
  const uint64_t k = k0 + (uint64_t)lane;
  const bool live_wave = g < ngroups;  // wave-uniform
  uint32_t own = 0;
  if constexpr (is_squares<ENGINE>()) {
    const uint64_t c0 = a.gen_first + kb + 3 * (uint64_t)N;
    const uint32_t t = threadIdx.x;
    own = squares_word_y<ENGINE>(a.d.key[0], a.d.key[0] * (c0 + t));
    sw[t] = own;
    if (t < (uint32_t)(W - 1))
      sw[kBlock + t] = squares_word_y<ENGINE>(a.d.key[0], a.d.key[0] * (c0 + kBlock + t));
    __syncthreads();
  }
  float p[N][4];
  double wt = 0.0;
  uint32_t r = 0;  // the event's rank in the block's partition order
  if (live_wave) {
    const uint64_t fw = bp.fw;
    if constexpr (is_squares<ENGINE>()) {
      wt = genbod_mixed<N, ENGINE>(a.d, BlockWordsM{sw + threadIdx.x, own}, p);
    } else {
      EventWordsM<N, ENGINE> words;
      words.init(a.d, a.gen_first + k0, lane);
      wt = genbod_mixed<N, ENGINE>(a.d, words, p);
    }
    final_boost_f32<N>(p, a.d);
    const uint64_t below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
    const uint32_t acc_below = acc_wave + (uint32_t)__popcll(fw & below);
    const bool kept = (fw >> lane) & 1ULL;
    const uint32_t t = (uint32_t)(wave * kWave + lane);
    r = kept ? acc_below : a_blk + (t - acc_below);
    if (k < a.n) {
      const uint64_t dst = (r < a_blk) ? goff + r : nacc + (kb - goff) + (r - a_blk);
      if (a.w_out) *(glob(a.w_out) + dst) = wt;
      if (a.idx_out) *(glob(a.idx_out) + dst) = k;
    }
  }
  const uint64_t live = (a.n - kb) < (uint64_t)kBlock ? (a.n - kb) : (uint64_t)kBlock;
  const uint32_t t = threadIdx.x;
  const uint64_t dst = (t < a_blk) ? goff + t : nacc + (kb - goff) + (t - a_blk);
  float* cols = reinterpret_cast<float*>(a.cols_out);
#pragma unroll
  for (int h = 0; h < C; h += CS) {
    if (h) __syncthreads();  // the previous columns have been read
    if (live_wave && k < a.n) {
#pragma unroll
      for (int c = h; c < h + CS; ++c) sh[c - h][r] = p[c / 4][c % 4];
    }
    __syncthreads();
    if (cols && t < live) {
      // one 64-bit address per thread, stepped by the column stride
      gptr<float> q = glob(cols + dst) + (uint64_t)h * a.ld_out;
#pragma unroll
      for (int c = h; c < h + CS; ++c, q += a.ld_out) *q = sh[c - h][t];
    }
  }
}

constexpr int kNumN = HPSP_MAX_N - 1;
constexpr int kNumE = 6;

template <int N, int E>
static void regen_mixed_one(const UnweightArgs& a, hipStream_t s) {
  const uint64_t nblocks = (a.n + kBlock - 1) / kBlock;
  const uint64_t grid = (nblocks + kXcds - 1) / kXcds * kXcds;
  hipLaunchKernelGGL((k_uw_regen_mixed<N, E>), dim3((unsigned)grid), dim3(kBlock), 0, s, a);
}

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
    hipLaunchKernelGGL((k_generate_mixed<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_generate_mixed<N, E, false>), dim3(g), dim3(kBlock), 0, s, a);
}

template <int N, int E>
static void avg_one(const AvgArgs& a, hipStream_t s) {
  const uint64_t nchunks = (a.n + kAvgChunk - 1) / kAvgChunk;
  const unsigned g = (unsigned)(nchunks < 8192 ? (nchunks ? nchunks : 1) : 8192);
  if (a.mothers)
    hipLaunchKernelGGL((k_average_mixed<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_average_mixed<N, E, false>), dim3(g), dim3(kBlock), 0, s, a);
}

template <int I>
static constexpr auto gen_at = &gen_one<2 + I / kNumE, I % kNumE>;
template <int I>
static constexpr auto avg_at = &avg_one<2 + I / kNumE, I % kNumE>;

template <int... I>
static constexpr auto gen_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const GenArgs&, hipStream_t), sizeof...(I)>{gen_at<I>...};
}
template <int... I>
static constexpr auto avg_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const AvgArgs&, hipStream_t), sizeof...(I)>{avg_at<I>...};
}

template <int I>
static constexpr auto rgm_at = &regen_mixed_one<2 + I / kNumE, I % kNumE>;
template <int... I>
static constexpr auto rgm_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, hipStream_t), sizeof...(I)>{rgm_at<I>...};
}

static const auto kGen = gen_table(std::make_integer_sequence<int, kNumN * kNumE>{});
static const auto kAvg = avg_table(std::make_integer_sequence<int, kNumN * kNumE>{});
static const auto kRegenMixed = rgm_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_unweight_regen_mixed(int N, int engine, const UnweightArgs& a, hipStream_t s) {
  if (a.n == 0) return (int)hipSuccess;
  kRegenMixed[(N - 2) * kNumE + engine](a, s);
  return (int)hipGetLastError();
}

int launch_generate_mixed(int N, int engine, const GenArgs& a, hipStream_t s) {
  kGen[(N - 2) * kNumE + engine](a, s);
  return (int)hipGetLastError();
}

int launch_average_mixed(int N, int engine, const AvgArgs& a, hipStream_t s) {
  kAvg[(N - 2) * kNumE + engine](a, s);
  return (int)hipGetLastError();
}

}  // namespace hpsp
