// Dense histograms on gfx950: DenseHistogram<double, D>::Fill
// (DenseHistogram.inl:49-105, 174-222) for D = 1..HPSP_HIST_MAX_DIM.
//
// The reference computes every event's bin key (GetGlobalBin.h:41-209), sorts
// keys and weights, reduces by key and scatters the sums into nbins + 2
// contents.  Here each block keeps private bins in LDS and adds its events
// with FP64 LDS atomics, then adds its non-empty bins into the contents with
// device atomics; nothing is sorted.  Bin keys are computed with the
// reference's operations, so every event lands in the same bin; the order of
// the additions inside a bin differs (sums agree to rounding, unit-weight
// counts exactly).  Built with -ffp-contract=off.
//
// Sources of the axis values:
//   columns  D value columns (the reference's data range) and an optional
//            weight column;
//   events   the events of a Decays container in SoA columns, each axis a
//            functor of the event (the lazy `events | dalitz_calculator` of
//            phsp_basic.inl:196-214, fused into the fill), weights from a
//            column, the PhaseSpaceWeight of the event, or 1;
//   particles  stored four-vectors given particle by particle (a melded
//            decay chain, phsp_chain.inl:244), read only where a functor
//            uses them; weights from a column or 1.
#include <algorithm>
#include <array>
#include <utility>

#include "hist_key.h"

namespace hpsp {

template <int N>
__device__ __forceinline__ void event_axes(const HistArgs& a, uint64_t k,
                                           double (&x)[HPSP_HIST_MAX_DIM], double& w) {
  double p[N][4];
#pragma unroll
  for (int i = 0; i < N; ++i)
#pragma unroll
    for (int c = 0; c < 4; ++c) p[i][c] = a.cols[(uint64_t)(4 * i + c) * a.ld + k];
#pragma unroll
  for (int d = 0; d < HPSP_HIST_MAX_DIM; ++d)
    x[d] = (d < a.h.ndim) ? functor_eval<double>(a.axes[d], p) : 0.0;
  if (a.weight_mode == HPSP_HIST_WEIGHT_COLUMN)
    w = a.w[k];
  else if (a.weight_mode == HPSP_HIST_WEIGHT_PHSP)
    w = event_weight_f64<N>(p, a.d.m, a.d.maxw);
  else
    w = 1.0;
}

// N = 0: axis values from columns; N >= 2: from the events' momenta.
template <int N>
__global__ __launch_bounds__(kHistBlock) void k_hist_fill(const HistArgs a) {
  extern __shared__ double bins[];
  const uint64_t nslots = a.h.nbins + 2;
  if (a.use_lds) {
    for (uint64_t b = threadIdx.x; b < nslots; b += kHistBlock) bins[b] = 0.0;
    __syncthreads();
  }
  const uint64_t stride = (uint64_t)gridDim.x * kHistBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kHistBlock + threadIdx.x; k < a.n; k += stride) {
    double x[HPSP_HIST_MAX_DIM];
    double w;
    if constexpr (N == 0) {
#pragma unroll
      for (int d = 0; d < HPSP_HIST_MAX_DIM; ++d)
        x[d] = (d < a.h.ndim) ? a.x[(uint64_t)d * a.ldx + k] : 0.0;
      w = a.w ? a.w[k] : 1.0;
    } else if constexpr (N < 0) {
      const MemAcc acc{a.parts, k};
#pragma unroll
      for (int d = 0; d < HPSP_HIST_MAX_DIM; ++d)
        x[d] = (d < a.h.ndim) ? functor_eval_acc<double, true, true>(a.axes[d], acc) : 0.0;
      w = a.w ? a.w[k] : 1.0;
    } else {
      event_axes<N>(a, k, x, w);
    }
    const uint64_t key = hist_key(a.h, x);
    if (key == kDrop) continue;
    if (a.use_lds)
      atomicAdd(&bins[key], w);
    else
      unsafeAtomicAdd(&a.contents[key], w);
  }
  if (a.use_lds) {
    __syncthreads();
    for (uint64_t b = threadIdx.x; b < nslots; b += kHistBlock) {
      const double v = bins[b];
      if (v != 0.0) unsafeAtomicAdd(&a.contents[b], v);
    }
  }
}

__global__ void k_zero(double* p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

void launch_zero(double* p, uint64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<uint64_t>((n + 255) / 256, 1024)),
                     dim3(256), 0, s, p, n);
}

template <int N>
static void hist_one(const HistArgs& a, unsigned grid, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL(k_hist_fill<N>, dim3(grid), dim3(kHistBlock), lds, s, a);
}

template <int... I>
static constexpr auto hist_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const HistArgs&, unsigned, size_t, hipStream_t), sizeof...(I)>{
      &hist_one<(I == 0 ? 0 : I + 1)>...};
}
// index 0: columns; index N - 1 (N = 2..HPSP_MAX_N): events of N daughters
static const auto kHist = hist_table(std::make_integer_sequence<int, HPSP_MAX_N>{});

int launch_histogram_fill(int N, const HistArgs& in, hipStream_t s) {
  HistArgs a = in;
  const uint64_t nslots = a.h.nbins + 2;
  launch_zero(a.contents, nslots, s);
  if (a.n == 0) return (int)hipGetLastError();
  a.use_lds = nslots <= a.lds_slots ? 1 : 0;
  const size_t lds = a.use_lds ? nslots * sizeof(double) : 0;
  // one block per CU keeps the per-block merge small; more while bins are few
  const uint64_t want = (a.n + kHistBlock - 1) / kHistBlock;
  const uint64_t cap = a.use_lds && nslots > 2048 ? 256 : 1024;
  const unsigned grid = (unsigned)std::min<uint64_t>(want, cap);
  if (N < 0)
    hist_one<-1>(a, grid, lds, s);
  else
    kHist[N == 0 ? 0 : N - 1](a, grid, lds, s);
  return (int)hipGetLastError();
}

// events | functors over stored particles: one event per lane, the functor
// columns written lane-contiguous.
__global__ __launch_bounds__(kBlock) void k_particles_eval(const PartEvalArgs a) {
  const uint64_t stride = (uint64_t)gridDim.x * kBlock;
  for (uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x; k < a.n; k += stride) {
    const MemAcc acc{a.p, k};
    for (int j = 0; j < a.nf; ++j)
      a.out[(uint64_t)j * a.ld + k] = functor_eval_acc<double, true, true>(a.f[j], acc);
  }
}

int launch_particles_evaluate(const PartEvalArgs& a, hipStream_t s) {
  if (a.n == 0) return (int)hipSuccess;
  hipLaunchKernelGGL(k_particles_eval, dim3(grid_for(a.n, 1u << 16)), dim3(kBlock), 0, s, a);
  return (int)hipGetLastError();
}

}  // namespace hpsp
