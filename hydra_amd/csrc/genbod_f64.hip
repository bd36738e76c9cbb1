// FP64 phase-space kernels for gfx950: GENBOD generation into SoA columns,
// AverageOn (fused generation + functor + StatsPHSP reduction) and the
// per-event weight kernels.  Built with -ffp-contract=off.  The weight chain
// (uniforms, invMas, pdk, the weight product) and PhaseSpaceWeight repeat the
// reference's double operations one for one with correctly rounded sqrt and
// division, so weights agree bit for bit (DecayMother.h:170-212,
// Decays.inl:215-231).  The momenta (DecayMother.h:218-269) are formed with
// the same operations except: exactly-zero terms are dropped, sin/cos and
// square roots outside the weight chain are evaluated to about one ulp, and
// the boost factors use gamma = E*/M, gamma*beta = p/M.
#include <array>
#include <type_traits>
#include <utility>

#include "hpsp_launch.h"

namespace hpsp {

// GENBOD for one event (DecayMother::process).  word(j) yields the j-th
// engine word of the event's stream (after discard(evt + 3N)).
template <int N, int ENGINE, typename WordFn>
__device__ __forceinline__ double genbod_f64(const DecayK& d, const WordFn& word,
                                             double (&p)[N][4]) {
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

  double invMas[N];
  double sum = 0.0;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    sum += d.m[n];
    invMas[n] = rno[n] * d.ecm + sum;
  }

  double wt = d.maxw;
  double pd[N - 1];
#pragma unroll
  for (int n = 0; n < N - 1; ++n) {
    pd[n] = pdk(invMas[n + 1], invMas[n], d.m[n + 1]);
    wt *= pd[n];
  }

  // Daughter 0 along +y; it is rotated for the first time at i = 1.
  p[0][0] = sqrt_fast(pd[0] * pd[0] + d.m[0] * d.m[0]);
  p[0][1] = 0.0;
  p[0][2] = pd[0];
  p[0][3] = 0.0;

#pragma unroll
  for (int i = 1; i < N; ++i) {
    // keep the scheduler from hoisting later steps' angles and boosts
    // above this one: their results would all be live at once
    if constexpr (N > 4) __builtin_amdgcn_sched_barrier(0);
    p[i][0] = sqrt_fast(pd[i - 1] * pd[i - 1] + d.m[i] * d.m[i]);
    p[i][1] = 0.0;
    p[i][2] = -pd[i - 1];
    p[i][3] = 0.0;

    const double cZ = 2 * to_unit<ENGINE>(word(N - 2 + 2 * (i - 1))) - 1;
    const double sZ = sqrt_fast(1 - cZ * cZ);
    // sin/cos of angY = 2*pi*u
    double sY, cY;
    sincos_2pi(to_unit<ENGINE>(word(N - 2 + 2 * (i - 1) + 1)), sY, cY);

#pragma unroll
    for (int j = 0; j <= i; ++j) {
      const bool fresh = (j == i) || (i == 1);  // px = pz = 0 before this rotation
      if (fresh) {
        const double y = p[j][2];
        const double t = -(sZ * y);  // cZ*0 - sZ*y
        p[j][2] = cZ * y;            // sZ*0 + cZ*y
        p[j][1] = cY * t;            // cY*t - sY*0
        p[j][3] = sY * t;            // sY*t + cY*0
      } else {
        double x = p[j][1];
        const double y = p[j][2];
        p[j][1] = cZ * x - sZ * y;
        p[j][2] = sZ * x + cZ * y;
        x = p[j][1];
        const double z = p[j][3];
        p[j][1] = cY * x - sY * z;
        p[j][3] = sY * x + cY * z;
      }
    }
    if (i == N - 1) break;

    // Boost along +y by beta = pd/sqrt(pd^2 + M^2) (Vector3R(0, beta, 0)):
    // only E and py change.  With E* = sqrt(pd^2 + M^2) the reference's
    // gamma = 1/sqrt(1 - beta^2), gamma*beta and (gamma-1)/beta^2 * beta^2 are
    // E*/M, pd/M and gamma - 1; the momenta are formed as
    // E' = gamma E + gamma beta py and py' = gamma beta E + gamma py.
    if (pd[i] > 0.0) {
      const double inv_m = rcp_fast(invMas[i]);
      const double estar = sqrt_fast(pd[i] * pd[i] + invMas[i] * invMas[i]);
      const double gamma = estar * inv_m;
      const double gby = pd[i] * inv_m;
#pragma unroll
      for (int j = 0; j <= i; ++j) {
        const double e2 = p[j][0], py2 = p[j][2];
        p[j][0] = gamma * e2 + gby * py2;
        p[j][2] = gby * e2 + gamma * py2;
      }
    }
  }
  return wt;
}

template <typename T>
__device__ __forceinline__ void st(T* ptr, T v) {
  __builtin_nontemporal_store(v, ptr);
}

// Stream words of this lane's event: squares engines share the wave's 64+W
// consecutive counters across lanes; Random123 engines evaluate per lane.
template <int N, int ENGINE>
struct EventWords {
  static constexpr int W = 3 * N - 4;
  std::conditional_t<is_squares<ENGINE>(), WaveWords, char> ww;
  uint64_t w[is_squares<ENGINE>() ? 1 : W];

  __device__ __forceinline__ void init(const DecayK& d, uint64_t wave_first_evt, int lane) {
    if constexpr (is_squares<ENGINE>()) {
      const uint64_t c0 = wave_first_evt + 3 * (uint64_t)N;
      ww.lane = lane;
      ww.lo = squares_word<ENGINE>(d.key[0], c0 + (uint64_t)lane);
      ww.hi = squares_word<ENGINE>(d.key[0], c0 + kWave + (uint64_t)lane);
    } else {
      r123_words<ENGINE, W>(d.key, wave_first_evt + (uint64_t)lane + 3 * (uint64_t)N, w);
    }
  }
  __device__ __forceinline__ uint64_t operator()(int j) const {
    if constexpr (is_squares<ENGINE>())
      return ww.get(j);
    else
      return w[j];
  }
};

template <int N, int ENGINE, bool MOTHERS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_generate_f64(const GenArgs a) {
  const int lane = __lane_id();
  // wave index as a scalar so that column bases stay in SGPRs and every
  // store is (scalar base) + lane * 8
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t base = (uint64_t)blockIdx.x * kBlock; base < a.n; base += stride) {
    const uint64_t k0 = base + (uint64_t)wave * kWave;
    const uint64_t k = k0 + (uint64_t)lane;
    const bool live = k < a.n;

    EventWords<N, ENGINE> words;
    words.init(a.d, a.first + k0, lane);
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
      if (a.out_f32) {
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

// AverageOn: chunks of kAvgChunk consecutive events, one partial per chunk.
// Within a chunk thread t folds events t, t+256, ... in order; the waves and
// the four wave results merge in a fixed tree, so a partial depends only on
// its event range.
template <int N, int ENGINE, bool MOTHERS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_average_f64(const AvgArgs a) {
  __shared__ Stats sh[kBlock / kWave];
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t nchunks = (a.n + kAvgChunk - 1) / kAvgChunk;
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    Stats acc = {0.0, 0.0, 0.0};
    for (int it = 0; it < kAvgChunk / kBlock; ++it) {
      const uint64_t k0 = c * kAvgChunk + (uint64_t)it * kBlock + (uint64_t)wave * kWave;
      const uint64_t k = k0 + (uint64_t)lane;
      const bool live = k < a.n;
      EventWords<N, ENGINE> words;
      words.init(a.d, a.first + k0, lane);
      double p[N + (MOTHERS ? 1 : 0)][4];
      double f, wt;
      if constexpr (MOTHERS) {
        double q[N][4];
        wt = genbod_f64<N, ENGINE>(a.d, words, q);
        const uint64_t km = live ? k : 0;
        double m4[4];
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) m4[cc] = a.mothers[(uint64_t)cc * a.mld + km];
        const double bx = m4[1] / m4[0], by = m4[2] / m4[0], bz = m4[3] / m4[0];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          boost_general(q[i], bx, by, bz);
#pragma unroll
          for (int cc = 0; cc < 4; ++cc) p[i + 1][cc] = q[i][cc];
        }
#pragma unroll
        for (int cc = 0; cc < 4; ++cc) p[0][cc] = m4[cc];
      } else {
        wt = genbod_f64<N, ENGINE>(a.d, words, p);
        if (a.d.do_boost) {
#pragma unroll
          for (int i = 0; i < N; ++i) boost_folded(p[i], a.d);
        }
      }
      f = functor_eval<double, false>(a.f, p);
      if (live) stats_push(acc, f, wt);
    }
    acc = wave_reduce_stats(acc);
    if (lane == 0) sh[wave] = acc;
    __syncthreads();
    if (threadIdx.x == 0)
      a.partials[c] = stats_merge(stats_merge(sh[0], sh[1]), stats_merge(sh[2], sh[3]));
    __syncthreads();
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
static void gen_one(const GenArgs& a, hipStream_t s) {
  const unsigned g = grid_for(a.n, 1u << 20);
  if (a.mothers)
    hipLaunchKernelGGL((k_generate_f64<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
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

static const auto kGen = gen_table(std::make_integer_sequence<int, kNumN * kNumE>{});
static const auto kAvg = avg_table(std::make_integer_sequence<int, kNumN * kNumE>{});

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
