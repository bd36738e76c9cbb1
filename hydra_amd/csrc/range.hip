// hydra::phase_space_range on gfx950 (PhaseSpaceRange.inl:44-72): element k
// of the lazy range is GenerateDecay<N, minstd_rand>(mother, masses,
// seed)(k) = (weight, p_1, ..., p_N) (GenerateDecay.h:203-343), the GENBOD
// event of DecayMother with the engine discarded by k + 3N.  Built with
// -ffp-contract=off.
//
// Engine words.  minstd_rand's word j of element first + k is
// x_start * a^(k + j) mod m (a = 48271, m = 2^31 - 1).  A wave's 64 elements
// need the words at offsets k0 + 0 .. k0 + 63 + 3N - 5: each lane holds the
// two words at k0 + lane and k0 + 64 + lane and the wave shares them with
// ds_bpermute (WaveWords), as the squares engines' counters are shared.  A
// lane's first pair costs one power of a (uniform across the wave) and a
// table lookup; each further grid pass multiplies both by a^(grid size).
//
// Kernels:
//   k_range_generate<N>  elements into SoA momentum columns and a weight column
//   k_range_hist<N>      DenseHistogram::Fill(range | axes[, weights]) with
//                        the elements generated inside the fill: nothing but
//                        the bins touches memory.
#include <algorithm>
#include <array>
#include <utility>

#include "genbod_f64.h"
#include "hist_key.h"

namespace hpsp {

// a^l mod m for l = 0 .. 2 * kWave - 1
struct MinstdPowers {
  uint32_t v[2 * kWave];
};
constexpr MinstdPowers minstd_powers() {
  MinstdPowers t{};
  uint64_t x = 1;
  for (int l = 0; l < 2 * kWave; ++l) {
    t.v[l] = (uint32_t)x;
    x = x * kMinstdA % kMinstdM;
  }
  return t;
}
__constant__ MinstdPowers kPow = minstd_powers();

// This lane's two stream words and the wave view genbod_f64 reads.
struct RangeWords {
  uint32_t lo, hi;
  int lane;
  // words at offsets k0 + lane and k0 + 64 + lane
  __device__ __forceinline__ void init(uint32_t x_start, uint64_t k0, int l) {
    lane = l;
    const uint32_t xw = minstd_mul(x_start, minstd_pow(k0));
    lo = minstd_mul(xw, kPow.v[l]);
    hi = minstd_mul(xw, kPow.v[kWave + l]);
  }
  __device__ __forceinline__ void step(uint32_t a) {
    lo = minstd_mul(lo, a);
    hi = minstd_mul(hi, a);
  }
  __device__ __forceinline__ uint64_t operator()(int j) const {
    return WaveWords{lo, hi, lane}.get(j);
  }
};

template <int N>
__device__ __forceinline__ double range_event(const DecayK& d, const RangeWords& w,
                                              double (&p)[N][4]) {
  const double wt = genbod_f64<N, kMinstd>(d, w, p);
  if (d.do_boost) {
#pragma unroll
    for (int i = 0; i < N; ++i) boost_folded(p[i], d);
  }
  return wt;
}

template <int N>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_range_generate(const RangeArgs a) {
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t base = (uint64_t)blockIdx.x * kBlock;
  if (base >= a.n) return;
  RangeWords words;
  words.init(a.x_start, base + (uint64_t)wave * kWave, lane);
  for (; base < a.n; base += stride) {
    const uint64_t k0 = base + (uint64_t)wave * kWave;
    double p[N][4];
    const double wt = range_event<N>(a.d, words, p);
    if (k0 + (uint64_t)lane < a.n) {
      double* c = a.cols + k0;
#pragma unroll
      for (int i = 0; i < N; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          st(uniform_ptr(c + (uint64_t)(4 * i + q) * a.ld) + lane, p[i][q]);
      if (a.wt) st(uniform_ptr(a.wt + k0) + lane, wt);
    }
    words.step(a.a_stride);
  }
}

template <int N>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_range_hist(const RangeArgs a) {
  extern __shared__ double bins[];
  const uint64_t nslots = a.h.nbins + 2;
  if (a.use_lds) {
    for (uint64_t b = threadIdx.x; b < nslots; b += kBlock) bins[b] = 0.0;
    __syncthreads();
  }
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  uint64_t base = (uint64_t)blockIdx.x * kBlock;
  RangeWords words;
  words.init(a.x_start, base + (uint64_t)wave * kWave, lane);
  for (; base < a.n; base += stride) {
    const uint64_t k = base + (uint64_t)wave * kWave + (uint64_t)lane;
    double p[N][4];
    const double wt = range_event<N>(a.d, words, p);
    words.step(a.a_stride);
    if (k >= a.n) continue;
    double x[HPSP_HIST_MAX_DIM];
#pragma unroll
    for (int d = 0; d < HPSP_HIST_MAX_DIM; ++d)
      x[d] = (d < a.h.ndim) ? functor_eval<double>(a.axes[d], p) : 0.0;
    const double w = a.weight_mode == HPSP_HIST_WEIGHT_GENERATION ? wt
                     : a.weight_mode == HPSP_HIST_WEIGHT_PHSP
                         ? event_weight_f64<N>(p, a.d.m, a.d.maxw)
                         : 1.0;
    const uint64_t key = hist_key(a.h, x);
    if (key == kDrop) continue;
    if (a.use_lds)
      atomicAdd(&bins[key], w);
    else
      unsafeAtomicAdd(&a.contents[key], w);
  }
  if (a.use_lds) {
    __syncthreads();
    for (uint64_t b = threadIdx.x; b < nslots; b += kBlock) {
      const double v = bins[b];
      if (v != 0.0) unsafeAtomicAdd(&a.contents[b], v);
    }
  }
}

// ---------------------------------------------------------------------------

// A lane's first word pair costs a power of a by square-and-multiply over the
// bits of its offset (some 30 steps); later grid passes step by one multiply.
// The grid is capped so that each lane runs many passes.
template <int N>
static void gen_one(RangeArgs a, hipStream_t s) {
  const unsigned g = grid_for(a.n, 8192u);
  a.a_stride = minstd_pow((uint64_t)g * kBlock);
  hipLaunchKernelGGL(k_range_generate<N>, dim3(g), dim3(kBlock), 0, s, a);
}

template <int N>
static void hist_one(RangeArgs a, hipStream_t s) {
  const uint64_t nslots = a.h.nbins + 2;
  a.use_lds = nslots <= a.lds_slots ? 1 : 0;
  const size_t lds = a.use_lds ? nslots * sizeof(double) : 0;
  // several blocks per CU while the private bins are small; the per-block
  // merge costs nslots atomics
  const unsigned cap = (a.use_lds && lds > 32768) ? 512u : 2048u;
  const unsigned g = grid_for(a.n, cap);
  a.a_stride = minstd_pow((uint64_t)g * kBlock);
  hipLaunchKernelGGL(k_range_hist<N>, dim3(g), dim3(kBlock), lds, s, a);
}

template <int... I>
static constexpr auto gen_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(RangeArgs, hipStream_t), sizeof...(I)>{&gen_one<I + 2>...};
}
template <int... I>
static constexpr auto hist_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(RangeArgs, hipStream_t), sizeof...(I)>{&hist_one<I + 2>...};
}
static const auto kGen = gen_table(std::make_integer_sequence<int, HPSP_MAX_N - 1>{});
static const auto kHist = hist_table(std::make_integer_sequence<int, HPSP_MAX_N - 1>{});

int launch_range_generate(int N, const RangeArgs& a, hipStream_t s) {
  if (a.n == 0) return (int)hipSuccess;
  kGen[N - 2](a, s);
  return (int)hipGetLastError();
}

int launch_range_histogram(int N, const RangeArgs& a, hipStream_t s) {
  launch_zero(a.contents, a.h.nbins + 2, s);
  if (a.n == 0) return (int)hipGetLastError();
  kHist[N - 2](a, s);
  return (int)hipGetLastError();
}

}  // namespace hpsp
