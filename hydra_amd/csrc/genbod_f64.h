// FP64 GENBOD for one event and the engine-word sources it reads, shared by
// the generation/AverageOn kernels (genbod_f64.hip) and the Evaluate kernels
// (evaluate.hip).  Built with -ffp-contract=off.  The weight chain (uniforms,
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

}  // namespace hpsp
