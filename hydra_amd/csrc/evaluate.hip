// PhaseSpace::Evaluate on gfx950: per event the GENBOD weight and the values
// of up to HPSP_MAX_EVAL_FUNCTORS functors, written as SoA columns; the
// momenta stay in registers.  Built with -ffp-contract=off.
//
//   Evaluate(mother, result, functors...)   EvalMother.h:283-306
//     the event's engine is the PhaseSpace seed discarded by evt + 3N, as in
//     Generate: the wave shares its squares words (EventWords).
//   Evaluate(mothers, result, functors...)  EvalMothers.h:145-273
//     each event seeds its own engine with pair_hash(evt, seed) and does not
//     discard, so every lane evaluates its own words (LaneWords); fECM and the
//     weight normalisation come from that event's mother.
#include <array>
#include <utility>

#include "genbod_f64.h"

namespace hpsp {

// Words of an engine seeded per event and not discarded.
template <int N, int ENGINE>
struct LaneWords {
  static constexpr int W = 3 * N - 4;
  uint64_t key[4];
  uint64_t w[is_squares<ENGINE>() ? 1 : W];

  __device__ __forceinline__ void init(uint64_t seed) {
    engine_keys_dev<ENGINE>(seed, key);
    if constexpr (!is_squares<ENGINE>()) r123_words<ENGINE, W>(key, 0, w);
  }
  __device__ __forceinline__ uint64_t operator()(int j) const {
    if constexpr (is_squares<ENGINE>())
      return squares_word<ENGINE>(key[0], (uint64_t)j);
    else
      return w[j];
  }
};

// Runs of a.run consecutive groups of 64 events per wave, as in
// k_generate_f64: over one mother the squares words of a group's second half
// are the next group's first.
template <int N, int ENGINE, bool MOTHERS>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_evaluate_f64(const EvalArgs a) {
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
    EventWords<N, ENGINE> ewords;
    for (uint64_t g = g_first; g < g_end; ++g) {
      const uint64_t k0 = g * kWave;
      const uint64_t k = k0 + (uint64_t)lane;
      const bool live = k < a.n;
      constexpr int P = N + (MOTHERS ? 1 : 0);
      double p[P][4];
      double wt;
      if constexpr (MOTHERS) {
        const uint64_t km = live ? k : 0;
        double m4[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) m4[c] = a.mothers[(uint64_t)c * a.mld + km];
        // fECM = mother.mass() - sum(m); maximum weight of this mother
        // (EvalMothers.h:151-176)
        DecayK d = a.d;
        {
#pragma clang fp contract(off)
          double te = mass_of(m4[0], m4[1], m4[2], m4[3]);
#pragma unroll
          for (int i = 0; i < N; ++i) te -= d.m[i];
          double emmax = te + d.m[0], emmin = 0.0, wtmax = 1.0;
#pragma unroll
          for (int i = 1; i < N; ++i) {
            emmin += d.m[i - 1];
            emmax += d.m[i];
            wtmax *= pdk(emmax, emmin, d.m[i]);
          }
          d.ecm = te;
          d.maxw = 1.0 / wtmax;
        }
        LaneWords<N, ENGINE> words;
        words.init(pair_hash(a.first + k, a.seed));
        double q[N][4];
        wt = genbod_f64<N, ENGINE>(d, words, q);
        const double bx = m4[1] / m4[0], by = m4[2] / m4[0], bz = m4[3] / m4[0];
#pragma unroll
        for (int i = 0; i < N; ++i) {
          boost_general(q[i], bx, by, bz);
#pragma unroll
          for (int c = 0; c < 4; ++c) p[i + 1][c] = q[i][c];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) p[0][c] = m4[c];
      } else {
        if (g == g_first)
          ewords.init(a.d, a.first + k0, lane);
        else
          ewords.advance(a.d, a.first + k0, lane);
        wt = genbod_f64<N, ENGINE>(a.d, ewords, p);
        if (a.d.do_boost) {
#pragma unroll
          for (int i = 0; i < N; ++i) boost_folded(p[i], a.d);
        }
      }
      if (live) {
        double* out = a.out + k0;
        st(uniform_ptr(out) + lane, wt);
        for (int j = 0; j < a.nf; ++j)
          st(uniform_ptr(out + (uint64_t)(1 + j) * a.ld) + lane, functor_eval<double>(a.f[j], p));
      }
    }
  }
}

constexpr int kNumN = HPSP_MAX_N - 1;
constexpr int kNumE = 6;

template <int N, int E>
static void eval_one(bool mothers, const EvalArgs& in, hipStream_t s) {
  EvalArgs a = in;
  // runs over one mother (as Generate); over mothers each lane seeds its own
  // engine, one group per wave
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  a.run = mothers ? 1u : run_groups(ngroups);
  const unsigned g = grid_for_runs((ngroups + a.run - 1) / a.run);
  if (mothers)
    hipLaunchKernelGGL((k_evaluate_f64<N, E, true>), dim3(g), dim3(kBlock), 0, s, a);
  else
    hipLaunchKernelGGL((k_evaluate_f64<N, E, false>), dim3(g), dim3(kBlock), 0, s, a);
}

template <int I>
static constexpr auto eval_at = &eval_one<2 + I / kNumE, I % kNumE>;

template <int... I>
static constexpr auto eval_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(bool, const EvalArgs&, hipStream_t), sizeof...(I)>{eval_at<I>...};
}

static const auto kEval = eval_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_evaluate_f64(int N, int engine, bool mothers, const EvalArgs& a, hipStream_t s) {
  kEval[(N - 2) * kNumE + engine](mothers, a, s);
  return (int)hipGetLastError();
}

}  // namespace hpsp
