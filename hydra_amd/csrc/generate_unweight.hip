// The flags of Decays::Unweight(seed) (FlagDaugthers, Decays.inl:75-89) for
// the events of a Generate that have not been formed: PhaseSpace::Generate
// followed by Unweight() on the same container (PhaseSpace.inl:377-394,
// Decays.inl:415-455) needs, before any event can be placed, every event's
// PhaseSpaceWeight w -- which needs its momenta -- only to compare w / max
// with u.  PhaseSpaceWeight recomputes, from the momenta, the product of
// two-body momenta that the weight chain forms from the invariant masses
// (DecayMother.h:170-212): the two are equal in exact arithmetic, and they
// differ by rounding in the momenta, amplified by the conditioning of each
// pdk term.  So each event's weight chain alone (no rotations, boosts or
// sin/cos) gives wt and a bound B on |w - wt| / wt, and the flag is q > u
// with q = wt / max wherever |q - u| > (B + 2^-48) q.  The chain is formed
// squared, wt^2 = maxw^2 prod_i d1 d2 / (4 a^2) (no square roots, one
// reciprocal), and the test is taken on q^2 and u^2 (both sides are not
// negative): with e = B + 2^-47, |q^2 - u^2| > e (2 + e) q^2 puts u below
// q (1 - e) or above q (1 + e), the margin of the linear test with room for
// the squares' roundings.  Elsewhere (where the
// two could fall on different sides of u, where the bound is large or its
// terms ill-defined) the event's momenta and PhaseSpaceWeight are formed and
// the reference's test decides, in a second kernel over those events.  The
// flags are those of the test on PhaseSpaceWeight, which
// hpsp_unweight_regenerate reads from its column.
//
// The bound.  With a = invMas[i] (the mass of daughters 0..i), b = m[i] and
// c = invMas[i-1], PhaseSpaceWeight's term is pdk(mass(p_0..p_i), b,
// mass(p_0..p_i-1)).  A subsystem mass computed as sqrt(E^2 - |p|^2) from
// components with absolute errors up to K eps S (S^2 = scale2 bounds E^2 +
// |p|^2 of every subsystem: M^2 at rest) has a relative error of at most
// (2K + 4) eps S^2 / m^2 to first order, and pdk's relative change is
// ka da/a + kc dc/c with ka = a^2/d1 + a^2/d2 + 1, kc = c(b+c)/d1 +
// c|b-c|/d2, d1 = a^2 - (b+c)^2, d2 = a^2 - (b-c)^2.  So
//   B = sum_i (ka + kc) ((2K + 4) eps S^2 (1/a^2 + 1/c^2) + 4 eps) + 40 N eps
// (the last terms for the roundings of the two chains themselves).  K is
// set far above the momenta's errors: each rotation and boost adds a few ulp
// of S, sin/cos are within 1.5e-16, and sin of the polar angle,
// sqrt(1 - cos^2), carries an error up to eps / sin, where sin >= 2^-15.5
// for a 32-bit word: K = 2^17 N for the squares engines.  The Random123
// engines' 64-bit words let sin fall to 2^-31.5, so there K is set per event
// from its own polar angles, K = N (64 + 4 / sin_min), with sin_min^2 taken
// 4 eps below its estimate (1 - cos)(1 + cos).  Over 2^20 events of each of
// nine decays (N = 2..12, near threshold, electrons beside a 90 GeV particle)
// the measured |w - wt| / wt stays below 4% of B with K = 10.
#include <array>
#include <type_traits>
#include <utility>

#include "genbod_f64.h"
#include "genbod_mixed.h"
#include "uw_dst.h"

namespace hpsp {

namespace {

// Engine words of one event evaluated by its own lane (the events that take
// the slow path are single lanes of a wave): counters evt + 3N + j.
template <int N, int ENGINE>
struct LaneEventWords {
  static constexpr int W = 3 * N - 4;
  uint64_t key, y0;
  uint64_t w[is_squares<ENGINE>() ? 1 : W];
  __device__ __forceinline__ void init(const DecayK& d, uint64_t evt) {
    if constexpr (is_squares<ENGINE>()) {
      key = d.key[0];
      y0 = key * (evt + 3 * (uint64_t)N);
    } else {
      r123_words<ENGINE, W>(d.key, evt + 3 * (uint64_t)N, w);
    }
  }
  __device__ __forceinline__ uint64_t operator()(int j) const {
    if constexpr (is_squares<ENGINE>())
      return squares_word_y<ENGINE>(key, y0 + key * (uint64_t)j);
    else
      return w[j];
  }
};

// The N - 2 mass words of one event, held in registers (so that a lane can
// run the weight chain on its own, outside the wave's word sharing).
template <int N>
struct MassWords {
  uint64_t w[N > 2 ? N - 2 : 1];
  __device__ __forceinline__ uint64_t operator()(int j) const { return w[j]; }
};

// A wave-uniform 64-bit value from the first active lane (scalar registers).
__device__ __forceinline__ uint64_t first_lane(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// The one uniform of a 3-body event with a squares engine: the lane's own word.
struct OneWord {
  uint32_t w0;
  __device__ __forceinline__ uint64_t operator()(int) const { return w0; }
};

// The weight chain of one event (DecayMother.h:170-212: sorted uniforms,
// invMas, the pdk terms), squared: returns q^2 = C prod_i d1 d2 / a^2 with
// C = (maxw / max)^2 / 4^(N-1), and the bound B of the file comment; sure is
// false where a term of B is not defined (d1 or d2 not positive) or B is not
// below 2^-10 (there the terms the first-order bound leaves out are below
// 2^-10 of it, far inside the margin K leaves).  With c the previous step's
// a (1/m0^2 at the first step), ka + kc is
// (d1 (a^2 + c|b-c|) + d2 (a^2 + c(b+c))) / (d1 d2) + 1, and B needs
// only estimates of 1/(d1 d2) and 1/a^2: it is a bound with a wide margin,
// formed with fused multiply-adds.
template <int N, int ENGINE, bool MOM = true, typename WordFn>
__device__ __forceinline__ double chain_q2(const DecayK& d, const WordFn& word, double C,
                                           double ks_fixed, double ks_unit, double ic2_0,
                                           double& B, bool& sure) {
#pragma clang fp contract(off)
  constexpr double eps = 0x1p-53;
  // ks = (2K + 4) eps S^2: fixed K for the squares engines, per event for
  // the Random123 ones (polar angle words N - 2 + 2 (i - 1), as GENBOD); 0
  // without the momenta's term (MOM false: the rounding terms alone)
  double ks = MOM ? ks_fixed : 0.0;
  if constexpr (MOM && !is_squares<ENGINE>()) {
    double s2 = 1.0;
#pragma unroll
    for (int i = 1; i < N; ++i) {
      const double cz = 2 * to_unit<ENGINE>(word(N - 2 + 2 * (i - 1))) - 1;
      s2 = __builtin_fmin(s2, (1.0 - cz) * (1.0 + cz));
    }
    const double inv_s = __builtin_amdgcn_rsq(__builtin_fmax(s2 - 4.0 * eps, 0x1p-120));
    ks = ks_unit * (2.0 * N * (64.0 + 4.0 * inv_s) + 4.0);
  }
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
  double P = 1.0, A = 1.0;
  double Bs = 40.0 * N * eps;  // the roundings of the two chains
  double ic2 = ic2_0;
  bool ok = true;
#pragma unroll
  for (int n = 0; n < N - 1; ++n) {
    const double a = invMas[n + 1], b = d.m[n + 1], c = invMas[n];
    const double spc = b + c, smc = b - c;
    const double d1 = (a - spc) * (a + spc);
    const double d2 = (a - smc) * (a + smc);
    const double x = d1 * d2;
    const double a2 = a * a;
    P *= x;
    A *= a2;
    ok = ok && __builtin_fmin(d1, d2) > 0.0;
    const double ix = __builtin_amdgcn_rcp(x), ia2 = __builtin_amdgcn_rcp(a2);
    const double kk = __builtin_fma(
        __builtin_fma(d1, __builtin_fma(c, __builtin_fabs(smc), a2), d2 * __builtin_fma(c, spc, a2)),
        ix, 1.0);
    Bs = __builtin_fma(kk, __builtin_fma(ks, ia2 + ic2, 4.0 * eps), Bs);
    ic2 = ia2;
  }
  B = Bs;
  sure = ok && B < 0x1p-10;  // false for NaN
  return C * P * rcp_fast(A);
}

// ---- the 3-body bound as a table ----------------------------------------
// For N = 3 the bound depends on one per-event variable, the mass uniform r:
// a0 = invMas[1] = r ecm + (m0 + m1); step 0 is pdk(a0; m1, m0), step 1
// pdk(A; m2, a0) with A = invMas[2] the same for every event.  r's leading
// 32 bits v pick a bin: the nearer end of [0, 1] (v >> 31) and the distance t
// to it (v or ~v), in octaves of t cut in 8 (each t < 8 a bin of its own), so
// bins are 1/8 of their distance to the end wide where B grows like 1/t, and
// 1/32 of [0, 1] in the middle.  Over a bin r lies in [r_lo, r_hi] (also for
// 64-bit words, whose 53-bit rounding stays in the bin's closed interval) and
// a0 in [a_lo, a_hi], formed by the chain's own rounded operations, which are
// monotonic in r.  Each term of B is bounded over the bin by its numerator at
// a_hi and its denominator at a_lo (step 1: d2's (m2 - c)^2 at whichever end
// is larger, a convex function of c), and 1/a^2, 1/c^2 at a_lo: B <= ks X + Y
// with ks = (2K + 4) eps S^2 (per event for the Random123 engines),
// X = kk0 (1/a_lo^2 + 1/m0^2) + kk1 (1/A^2 + 1/a_lo^2), Y = 4 eps (kk0 + kk1)
// + 120 eps (stored as (X, Y) pairs); a bin where a denominator is not
// positive gets infinity.
__device__ __forceinline__ int bin3(uint32_t v) {
  const uint32_t side = v >> 31;
  const uint32_t t = side ? ~v : v;  // < 2^31
  int idx;
  if (t < 8) {
    idx = (int)t;
  } else {
    const int e = 31 - __builtin_clz(t);  // 3 .. 30
    idx = (e - 2) * 8 + (int)((t >> (e - 3)) & 7u);
  }
  return (int)side * kBin3Side + idx;
}

__global__ __launch_bounds__(kBlock) void k_chain_table3(const DecayK d, double2* XY) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= kBin3) return;
  const int side = i / kBin3Side, idx = i % kBin3Side;
  uint64_t tlo, thi;
  if (idx < 8) {
    tlo = thi = (uint64_t)idx;
  } else {
    const int e = idx / 8 + 2, mb = idx % 8;
    tlo = (uint64_t)(8 + mb) << (e - 3);
    thi = ((uint64_t)(9 + mb) << (e - 3)) - 1;
  }
  double rlo, rhi;
  if (side == 0) {
    rlo = (double)tlo * 0x1p-32;
    rhi = (double)(thi + 1) * 0x1p-32;
  } else {
    rlo = (double)(0xFFFFFFFFull - thi) * 0x1p-32;
    rhi = (double)(0x100000000ull - tlo) * 0x1p-32;
  }
  // the chain's sums and a0 = r ecm + (m0 + m1), A = 1 ecm + (m0 + m1 + m2)
  double sum = 0.0;
  sum += d.m[0];
  sum += d.m[1];
  const double s01 = sum;
  sum += d.m[2];
  const double A = 1.0 * d.ecm + sum;
  const double alo = rlo * d.ecm + s01, ahi = rhi * d.ecm + s01;
  const double m0 = d.m[0], m1 = d.m[1], m2 = d.m[2];
  constexpr double eps = 0x1p-53;
  const double inf = __builtin_huge_val();
  // step 0: a in [alo, ahi], b = m1, c = m0
  const double spc0 = m1 + m0, smc0 = m1 - m0;
  const double den1 = (alo - spc0) * (alo + spc0), den2 = (alo - smc0) * (alo + smc0);
  const double ah2 = ahi * ahi, al2 = alo * alo;
  const double kk0 = (den1 > 0.0 && den2 > 0.0)
                         ? (ah2 + m0 * __builtin_fabs(smc0)) / den2 + (ah2 + m0 * spc0) / den1 + 1.0
                         : inf;
  // step 1: a = A, b = m2, c in [alo, ahi]
  const double A2 = A * A;
  const double e1 = (A - (m2 + ahi)) * (A + (m2 + ahi));
  const double dl = m2 - alo, dh = m2 - ahi;
  const double dm = __builtin_fmax(__builtin_fabs(dl), __builtin_fabs(dh));
  const double e2 = (A - dm) * (A + dm);
  const double kk1 = (e1 > 0.0 && e2 > 0.0)
                         ? (A2 + ahi * dm) / e2 + (A2 + ahi * (m2 + ahi)) / e1 + 1.0
                         : inf;
  const double x = kk0 * (1.0 / al2 + 1.0 / (m0 * m0)) + kk1 * (1.0 / A2 + 1.0 / al2);
  const double y = 4.0 * eps * (kk0 + kk1) + 120.0 * eps;
  // the few roundings of these operations
  XY[i] = make_double2(x * (1.0 + 0x1p-40), y * (1.0 + 0x1p-40));
}

// The 3-body chain: q^2 as chain_q2 forms it, with the step of the common A
// folded into C3 = C / A^2, and B from the table.
template <int ENGINE, bool MOM = true, typename WordFn>
__device__ __forceinline__ double chain3_q2(const DecayK& d, const WordFn& word, const ChainK& ck,
                                            const double2* XY, double& B) {
#pragma clang fp contract(off)
  constexpr double eps = 0x1p-53;
  const uint64_t w0 = word(0);
  const uint32_t v = is_squares<ENGINE>() ? (uint32_t)w0 : (uint32_t)(w0 >> 32);
  const double s01 = (0.0 + d.m[0]) + d.m[1];
  const double A = ck.A3;
  const double a0 = to_unit<ENGINE>(w0) * d.ecm + s01;
  const double m0 = d.m[0], m1 = d.m[1], m2 = d.m[2];
  const double spc0 = m1 + m0, smc0 = m1 - m0;
  const double x0 = ((a0 - spc0) * (a0 + spc0)) * ((a0 - smc0) * (a0 + smc0));
  const double spc1 = m2 + a0, smc1 = m2 - a0;
  const double x1 = ((A - spc1) * (A + spc1)) * ((A - smc1) * (A + smc1));
  double ks = MOM ? ck.ks_fixed : 0.0;
  if constexpr (MOM && !is_squares<ENGINE>()) {
    double s2 = 1.0;
#pragma unroll
    for (int i = 1; i < 3; ++i) {
      const double cz = 2 * to_unit<ENGINE>(word(1 + 2 * (i - 1))) - 1;
      s2 = __builtin_fmin(s2, (1.0 - cz) * (1.0 + cz));
    }
    const double inv_s = __builtin_amdgcn_rsq(__builtin_fmax(s2 - 4.0 * eps, 0x1p-120));
    ks = ck.ks_unit * (2.0 * 3 * (64.0 + 4.0 * inv_s) + 4.0);
  }
  const double2 xy = XY[bin3(v)];
  B = __builtin_fma(ks, xy.x, xy.y);
  return ck.C3 * (x0 * x1) * rcp_fast(a0 * a0);
}

}  // namespace

// One 1024-event tile per block, items as k_uw_flags: thread item j is event
// tile * 1024 + j * 256 + wave * 64 + lane of the container (generation
// index a.gen_first + that); flags as 64-bit ballots, counts per tile.  The
// events the chain leaves in doubt get the chain's flag for now, and each
// group with any is appended to the list dl as (group, mask of those events),
// one atomicAdd on dl_count per block that has any; k_uw_patch_chain decides
// them.
template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_uw_flags_chain(const UnweightArgs a, uint64_t* dl_count, uint64_t* dl, const ChainK ck,
                      const double* tab3) {
#pragma clang fp contract(off)
  __shared__ uint32_t cnt[kBlock / kWave];
  __shared__ uint64_t dm[kUwTile / kWave];  // doubt mask of each group of the tile
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;  // whole block
  // 3 bodies: the table, bins of (X, Y) pairs
  [[maybe_unused]] const double2* XY3 = reinterpret_cast<const double2*>(tab3);
  const uint64_t off = (uint64_t)wave * kWave + (uint64_t)lane;
  // key * (jump + k) of item 0 for the accept word and key * (gen_first + k +
  // 3N) for the generation word; item j + 1 is kBlock events on
  uint64_t ykey = a.uw_key * (a.jump + tile * kUwTile + off);
  const uint64_t ystep = a.uw_key * (uint64_t)kBlock;
  uint64_t ygen = a.d.key[0] * (a.gen_first + tile * kUwTile + off + 3 * (uint64_t)N);
  const uint64_t gstep = a.d.key[0] * (uint64_t)kBlock;
  uint32_t c = 0;
  // unrolled where the body is small enough for it
  constexpr int kUnroll = N <= 6 || is_squares<ENGINE>() ? kItems : 1;
#pragma unroll kUnroll
  for (int j = 0; j < kItems; ++j, ykey += ystep, ygen += gstep) {
    const uint64_t g = tile * (kUwTile / kWave) + (uint64_t)j * (kBlock / kWave) + wave;
    const uint64_t k0 = g * kWave;
    const uint64_t k = k0 + (uint64_t)lane;
    double B, q2;
    bool sure;
    if constexpr (N == 3 && is_squares<ENGINE>()) {
      // the one uniform is the lane's own word: counter gen_first + k + 3N
      const OneWord w{squares_word_y<ENGINE>(a.d.key[0], ygen)};
      q2 = chain3_q2<ENGINE>(a.d, w, ck, XY3, B);
      sure = B < 0x1p-10;  // false for NaN
    } else if constexpr (N == 3) {
      EventWords<N, ENGINE> words;
      words.init(a.d, a.gen_first + k0, lane);
      q2 = chain3_q2<ENGINE>(a.d, words, ck, XY3, B);
      sure = B < 0x1p-10;
    } else {
      EventWords<N, ENGINE> words;
      words.init(a.d, a.gen_first + k0, lane);
      q2 = chain_q2<N, ENGINE>(a.d, words, ck.C, ck.ks_fixed, ck.ks_unit, ck.ic2_0, B, sure);
    }
    const bool live = k < a.n;
    const double u = (double)squares3_y(a.uw_key, ykey) * 0x1p-32;  // ctr = jump + k
    const double u2 = u * u;
    const double e = B + 0x1p-47;
    sure = sure && ck.fast_ok && q2 >= 0x1p-1000 && q2 <= 0x1p+1000 &&
           __builtin_fabs(q2 - u2) > q2 * (e * (2.0 + e));
    const uint64_t ball = __ballot(live && q2 > u2);
    const uint64_t doubt = __ballot(live && !sure);
    if (lane == 0) {
      *(glob(a.flags) + g) = ball;
      dm[j * (kBlock / kWave) + wave] = doubt;
    }
    c += (uint32_t)__popcll(ball);
  }
  if (lane == 0) cnt[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) a.tile_count[tile] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
  if (wave == 0) {
    // lane t < 16: group tile * 16 + t
    constexpr int G = kUwTile / kWave;
    const uint64_t m = lane < G ? dm[lane] : 0;
    const uint64_t has = __ballot(m != 0);
    if (has) {
      uint64_t base = 0;
      if (lane == 0) base = atomicAdd((unsigned long long*)dl_count, (unsigned long long)__popcll(has));
      base = first_lane(base);
      if (m) {
        const uint64_t slot = base + (uint64_t)__popcll(has & ((1ULL << lane) - 1));
        *(glob(dl) + 2 * slot) = tile * G + (uint64_t)lane;
        *(glob(dl) + 2 * slot + 1) = m;
      }
    }
  }
}

// The events left in doubt: each wave takes entries of the list (a group
// and the mask of its doubtful events; a group appears once) in turn, forms
// each doubtful event's momenta and PhaseSpaceWeight in its own lane and
// applies the reference's test (FlagDaugthers, Decays.inl:75-89); the flag
// word is corrected and the tile's count adjusted.  The entry count is read
// once, so every wave ends after at most count / (waves in the grid) + 1
// entries.
template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) void k_uw_patch_chain(const UnweightArgs a,
                                                           const uint64_t* dl_count,
                                                           const uint64_t* dl) {
#pragma clang fp contract(off)
  const int lane = __lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * (kBlock / kWave);
  const uint64_t w0 =
      (uint64_t)blockIdx.x * (kBlock / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t count = first_lane(*(glob(dl_count)));
  const double rmax = 1.0 / a.maxv;
  const bool rmax_ok = __builtin_fabs(rmax) >= 0x1p-1000 && __builtin_fabs(rmax) <= 0x1p+1000;
  for (uint64_t e = w0; e < count; e += nwaves) {
    const uint64_t g = first_lane(*(glob(dl) + 2 * e));
    const uint64_t m = first_lane(*(glob(dl) + 2 * e + 1));
    const bool mine = (m >> lane) & 1ULL;
    bool acc = false;
    if (mine) {
      const uint64_t k = g * kWave + (uint64_t)lane;
      LaneEventWords<N, ENGINE> lw;
      lw.init(a.d, a.gen_first + k);
      double p[N][4];
      genbod_f64<N, ENGINE>(a.d, lw, p);
      if (a.d.do_boost) {
#pragma unroll
        for (int i = 0; i < N; ++i) boost_folded(p[i], a.d);
      }
      const double u = (double)squares3_y(a.uw_key, a.uw_key * (a.jump + k)) * 0x1p-32;
      acc = ratio_above(event_weight_f64<N>(p, a.d.m, a.d.maxw), a.maxv, rmax, rmax_ok, u);
    }
    const uint64_t ball = __ballot(mine && acc);
    if (lane == 0) {
      const uint64_t fw = *(glob(a.flags) + g);
      const uint64_t nw = (fw & ~m) | ball;
      *(glob(a.flags) + g) = nw;
      const int delta = __popcll(nw) - __popcll(fw);
      if (delta) atomicAdd(a.tile_count + g / (kUwTile / kWave), (uint32_t)delta);
    }
  }
}

__global__ void k_zero_count(uint64_t* p) { *p = 0; }

// The flags of the mixed-precision form (hpsp_generate_unweight_mixed): the
// test RN(w / max) > u on the mixed container's weight column, w = the mixed
// Generate's FP64 weight chain (one-ulp square roots and reciprocals,
// genbod_mixed.h).  The squared chain of k_uw_flags_chain, with the bound's
// rounding terms alone (no momenta are involved), gives q^2 within
// B_m = 4 B(K = 0) of both w and the exact chain for the same invariant
// masses: each pdk factor (a - b - c etc.) of either chain carries a rounding
// of a few eps of a, amplified at most 3 (ka + kc) times, and the square
// roots, reciprocals and products a few eps more, which B(K = 0) =
// sum_i 4 eps (ka + kc) + 40 N eps covers four times over.  Where the test on
// q^2 is not settled the lane forms w with the mixed Generate's code and
// applies the reference's test.
template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_uw_flags_mixed(const UnweightArgs a, const ChainK ck, const double* tab3) {
#pragma clang fp contract(off)
  __shared__ uint32_t cnt[kBlock / kWave];
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;  // whole block
  [[maybe_unused]] const double2* XY3 = reinterpret_cast<const double2*>(tab3);
  const uint64_t off = (uint64_t)wave * kWave + (uint64_t)lane;
  uint64_t ykey = a.uw_key * (a.jump + tile * kUwTile + off);
  const uint64_t ystep = a.uw_key * (uint64_t)kBlock;
  uint64_t ygen = a.d.key[0] * (a.gen_first + tile * kUwTile + off + 3 * (uint64_t)N);
  const uint64_t gstep = a.d.key[0] * (uint64_t)kBlock;
  uint32_t c = 0;
  constexpr int kUnroll = N <= 6 || is_squares<ENGINE>() ? kItems : 1;
#pragma unroll kUnroll
  for (int j = 0; j < kItems; ++j, ykey += ystep, ygen += gstep) {
    const uint64_t g = tile * (kUwTile / kWave) + (uint64_t)j * (kBlock / kWave) + wave;
    const uint64_t k0 = g * kWave;
    const uint64_t k = k0 + (uint64_t)lane;
    // the event's mass words, from the wave's shared words where they are
    // shared, before any lane branches off
    MassWords<N> mw;
    if constexpr (N == 3 && is_squares<ENGINE>()) {
      mw.w[0] = squares_word_y<ENGINE>(a.d.key[0], ygen);  // counter gen_first + k + 3N
    } else if constexpr (N > 2) {
      EventWords<N, ENGINE> words;
      words.init(a.d, a.gen_first + k0, lane);
#pragma unroll
      for (int i = 0; i < N - 2; ++i) mw.w[i] = words(i);
    }
    double B, q2;
    bool sure;
    if constexpr (N == 3) {
      q2 = chain3_q2<ENGINE, false>(a.d, mw, ck, XY3, B);
      B *= 4.0;
      sure = B < 0x1p-10;  // false for NaN
    } else {
      q2 = chain_q2<N, ENGINE, false>(a.d, mw, ck.C, 0.0, 0.0, ck.ic2_0, B, sure);
      B *= 4.0;
      sure = sure && B < 0x1p-10;
    }
    const bool live = k < a.n;
    const double u = (double)squares3_y(a.uw_key, ykey) * 0x1p-32;  // ctr = jump + k
    const double u2 = u * u;
    const double e = B + 0x1p-47;
    sure = sure && ck.fast_ok && q2 >= 0x1p-1000 && q2 <= 0x1p+1000 &&
           __builtin_fabs(q2 - u2) > q2 * (e * (2.0 + e));
    bool acc = q2 > u2;
    if (live && !sure) {
      // the container's weight, as the mixed Generate forms it
      double invMas[N];
      float pd[N - 1];
      const double w = weight_chain<N, ENGINE>(a.d, mw, invMas, pd);
      acc = ratio_above(w, a.maxv, ck.rmax, ck.rmax_ok, u);
    }
    const uint64_t ball = __ballot(live && acc);
    if (lane == 0) *(glob(a.flags) + g) = ball;
    c += (uint32_t)__popcll(ball);
  }
  if (lane == 0) cnt[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) a.tile_count[tile] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
}

constexpr int kNumN = HPSP_MAX_N - 1;
constexpr int kNumE = 6;

template <int N, int E>
static void flags_mixed_one(const UnweightArgs& a, double* tab3, const ChainK& ck,
                            hipStream_t s) {
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const unsigned grid = (unsigned)((ntiles + kXcds - 1) / kXcds * kXcds);
  if constexpr (N == 3)
    hipLaunchKernelGGL(k_chain_table3, dim3((kBin3 + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       a.d, reinterpret_cast<double2*>(tab3));
  hipLaunchKernelGGL((k_uw_flags_mixed<N, E>), dim3(grid), dim3(kBlock), 0, s, a, ck,
                     (const double*)tab3);
}

template <int I>
static constexpr auto fm_at = &flags_mixed_one<2 + I / kNumE, I % kNumE>;
template <int... I>
static constexpr auto fm_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, double*, const ChainK&, hipStream_t),
                    sizeof...(I)>{fm_at<I>...};
}
static const auto kFlagsMixed = fm_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_flags_mixed(int N, int engine, const UnweightArgs& a, double* tab3, const ChainK& ck,
                       hipStream_t s) {
  kFlagsMixed[(N - 2) * kNumE + engine](a, tab3, ck, s);
  return (int)hipGetLastError();
}

// Blocks of the list kernel: enough waves to fill the chip at its occupancy
// when every group is in the list, few enough that an almost empty list
// costs little.
constexpr uint64_t kPatchBlocks = 2048;

template <int N, int E>
static void chain_one(const UnweightArgs& a, uint64_t* dl_count, uint64_t* dl, double* tab3,
                      const ChainK& ck, hipStream_t s) {
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const unsigned grid = (unsigned)((ntiles + kXcds - 1) / kXcds * kXcds);
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, s, dl_count);
  if constexpr (N == 3)
    hipLaunchKernelGGL(k_chain_table3, dim3((kBin3 + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       a.d, reinterpret_cast<double2*>(tab3));
  hipLaunchKernelGGL((k_uw_flags_chain<N, E>), dim3(grid), dim3(kBlock), 0, s, a, dl_count, dl,
                     ck, (const double*)tab3);
  const unsigned pgrid = (unsigned)(ntiles < kPatchBlocks ? ntiles : kPatchBlocks);
  hipLaunchKernelGGL((k_uw_patch_chain<N, E>), dim3(pgrid), dim3(kBlock), 0, s, a,
                     (const uint64_t*)dl_count, (const uint64_t*)dl);
}

template <int I>
static constexpr auto chain_at = &chain_one<2 + I / kNumE, I % kNumE>;

template <int... I>
static constexpr auto chain_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, uint64_t*, uint64_t*, double*, const ChainK&,
                              hipStream_t),
                    sizeof...(I)>{chain_at<I>...};
}

static const auto kChain = chain_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_flags_chain(int N, int engine, const UnweightArgs& a, uint64_t* dl_count, uint64_t* dl,
                       double* tab3, const ChainK& ck, hipStream_t s) {
  kChain[(N - 2) * kNumE + engine](a, dl_count, dl, tab3, ck, s);
  return (int)hipGetLastError();
}

}  // namespace hpsp
