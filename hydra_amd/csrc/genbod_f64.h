// FP64 GENBOD for one event, the engine-word sources it reads, and the
// AverageOn and Evaluate kernels, templated over the functor they evaluate:
// instantiated with described functors by the library (genbod_f64.hip,
// evaluate.hip) and with users' callables by translation units that include
// hydra_amd/include/hydra/detail/UserKernels.h.  Every function turns
// floating-point contraction off itself, so the operation sequence holds
// under any contraction flag that honours pragmas.  The weight chain (uniforms,
// invMas, pdk, the weight product) repeats the reference's double operations
// one for one with correctly rounded sqrt and division, so weights agree bit
// for bit (DecayMother.h:170-212).  The momenta (DecayMother.h:218-269) are
// formed with the same operations except: exactly-zero terms are dropped,
// sin/cos and square roots outside the weight chain are evaluated to about
// one ulp, and the boost factors use gamma = E*/M, gamma*beta = p/M.
#pragma once

#include <type_traits>

#include "hpsp_launch.h"

namespace hpsp {

// GENBOD for one event (DecayMother::process).  word(j) yields the j-th
// engine word of the event's stream (after discard(evt + 3N)).
template <int N, int ENGINE, typename WordFn>
__device__ __forceinline__ double genbod_f64(const DecayK& d, const WordFn& word,
                                             double (&p)[N][4]) {
#pragma clang fp contract(off)
  double rno[N];
  rno[0] = 0.0;
  rno[N - 1] = 1.0;
  if constexpr (N > 2) {
    using WT = std::conditional_t<narrow_words<ENGINE>(), uint32_t, uint64_t>;
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
    if constexpr (is_squares<ENGINE>())
      sincos_2pi_w32((uint32_t)word(N - 2 + 2 * (i - 1) + 1), sY, cY);
    else
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
    // For pd = 0 (beta = 0) the reference's boost does nothing; here gamma = 1
    // and gamma beta = 0 are selected instead, which leave E and py as they
    // are (up to the sign of a zero py): a branch costs register moves on
    // every event, selects of the boosted components 4(i+1) instructions.
    {
      const bool boost = pd[i] > 0.0;
      const double inv_m = rcp_fast(invMas[i]);
      const double estar = sqrt_fast(pd[i] * pd[i] + invMas[i] * invMas[i]);
      const double gamma = boost ? estar * inv_m : 1.0;
      const double gby = boost ? pd[i] * inv_m : 0.0;
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

// Column stores of the generation kernels: plain global stores.  Where the
// columns start on cache lines these write whole lines and stream as fast as
// non-temporal ones; where a caller's column stride leaves a group's lines
// partial, plain stores let L2 merge the neighbouring groups' halves, where
// non-temporal ones sent each partial line to HBM on its own (3-body Generate
// of 2^27 events at a stride of n + 1: 3.0-4.1 ms against 4.7 ms).
template <typename T>
__device__ __forceinline__ void st(T* ptr, T v) {
  *glob(ptr) = v;
}

// Stream words of this lane's event: squares engines share the wave's 64+W
// consecutive counters across lanes; Random123 engines evaluate per lane.
template <int N, int ENGINE>
struct EventWords {
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
  // The next group: wave_first_evt is its first event, 64 after this
  // group's.  With the squares engines its first 64 words are this group's
  // second ones, so each lane evaluates one new word instead of two.
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

// AverageOn's value of a described functor: divisions through reciprocals
// (the result is compared within a tolerance, not bit for bit).  A user's
// callable brings its own overload (found by argument-dependent lookup).
template <int P>
__device__ __forceinline__ double avg_value(const FunctorK& f, const double (&p)[P][4]) {
  return functor_eval<double, false>(f, p);
}

// AverageOn: chunks of kAvgChunk consecutive events, one partial per chunk.
// Within a chunk wave w folds its own run of consecutive groups of 64 events
// (groups 16w .. 16w+15), lane l event l of each, in order, so the squares
// words carry from group to group; the four wave results merge in a fixed
// tree, so a partial depends only on its event range.
template <int N, int ENGINE, bool MOTHERS, typename FK = FunctorK>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_average_f64(const AvgArgsT<FK> a) {
  __shared__ Stats sh[kBlock / kWave];
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t nchunks = (a.n + kAvgChunk - 1) / kAvgChunk;
  for (uint64_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
    Stats acc = {0.0, 0.0, 0.0};
    constexpr int kGroupsPerWave = kAvgChunk / kBlock;
    EventWords<N, ENGINE> words;
    for (int it = 0; it < kGroupsPerWave; ++it) {
      const uint64_t k0 =
          c * kAvgChunk + ((uint64_t)wave * kGroupsPerWave + (uint64_t)it) * kWave;
      const uint64_t k = k0 + (uint64_t)lane;
      const bool live = k < a.n;
      if (it == 0)
        words.init(a.d, a.first + k0, lane);
      else
        words.advance(a.d, a.first + k0, lane);
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
      f = avg_value(a.f, p);
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

// Evaluate's values of described functors, exact divisions (the values are
// compared with the reference's bit for bit); out(j, value) stores functor
// j's.  A user's set of callables brings its own overload.
template <int P, typename Out>
__device__ __forceinline__ void eval_set(const DescSet& s, const double (&p)[P][4], Out&& out) {
  for (int j = 0; j < s.nf; ++j) out(j, functor_eval<double>(s.f[j], p));
}

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

// Evaluate (EvalMother.h:283-306, EvalMothers.h:145-273): runs of a.run
// consecutive groups of 64 events per wave, as in k_generate_f64: over one mother the squares words of a group's second half
// are the next group's first.
template <int N, int ENGINE, bool MOTHERS, typename FS = DescSet>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_evaluate_f64(const EvalArgsT<FS> a) {
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
        eval_set(a.fs, p, [&](int j, double v) {
          st(uniform_ptr(out + (uint64_t)(1 + j) * a.ld) + lane, v);
        });
      }
    }
  }
}

}  // namespace hpsp
