// Device-side building blocks shared by the gfx950 kernels: counter-based
// engines, the word -> [0,1) map, the GENBOD pieces and the StatsPHSP merge.
// Written for CDNA4 wave64; nothing here is a CUDA translation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "hpsp.h"
#include "sincos_table.h"

namespace hpsp {

constexpr int kBlock = 256;  // threads per block (4 waves of 64)
constexpr int kWave = 64;

// Minimum waves per SIMD requested from the register allocator for the
// one-event-per-lane kernels: the FP64 chains are long and dependent, so
// occupancy hides their latency; larger N needs more live momenta.
constexpr int waves_for(int N) { return N <= 4 ? 4 : (N <= 6 ? 3 : (N <= 10 ? 2 : 1)); }

// Kernel-side decay description; everything that depends only on the decay
// is folded on the host (engine keys, final boost constants).
struct DecayK {
  double m[HPSP_MAX_N];
  double ecm;
  double maxw;
  uint64_t key[4];  // engine key words, already splitmix-ed
  // Final boost of a single mother (Vector4R.inl:312-369) with its derived
  // constants, valid when do_boost != 0 (0 < b2 < 1).
  double bx, by, bz, bxx, byy, bzz, gamma, gb2, gb2xy, gb2xz, gb2yz, gbx, gby, gbz;
  int do_boost;
};

struct FunctorK {
  int kind;
  int nterms;
  double m[HPSP_MAX_TERMS], g[HPSP_MAX_TERMS], cr[HPSP_MAX_TERMS], ci[HPSP_MAX_TERMS];
  int a[HPSP_MAX_TERMS], b[HPSP_MAX_TERMS];
};

// A wave-uniform pointer pinned to SGPRs, so that per-lane accesses are
// (scalar base) + lane offset and no 64-bit VGPR address is kept per column.
template <typename T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
  const uint64_t v = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return reinterpret_cast<T*>(((uint64_t)hi << 32) | lo);
}

// Global-address-space views of a pointer: loads and stores through them are
// global_* instructions (not flat), so they do not wait on LDS counters.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glob(T* p) {
  return (gptr<T>)p;
}
template <typename T>
__device__ __forceinline__ gptr<T> glob_u(T* p) {
  return (gptr<T>)uniform_ptr(p);
}

// ---------------------------------------------------------------------------
// Engines

__device__ __forceinline__ uint64_t rot32(uint64_t x) { return (x >> 32) | (x << 32); }

// squares3 word for counter ctr (squares3.h:105-119)
// The same from y = key * ctr, for callers that step the counter by a fixed
// amount and y with it (an add instead of a 64-bit multiply).
__device__ __forceinline__ uint32_t squares3_y(uint64_t key, uint64_t y) {
  uint64_t x = y;
  uint64_t z = y + key;
  x = x * x + y;
  x = rot32(x);
  x = x * x + z;
  x = rot32(x);
  return (uint32_t)((x * x + y) >> 32);
}
__device__ __forceinline__ uint32_t squares3(uint64_t key, uint64_t ctr) {
  return squares3_y(key, key * ctr);
}

// squares4 word (squares4.h:106-122)
__device__ __forceinline__ uint32_t squares4_y(uint64_t key, uint64_t y) {
  uint64_t x = y;
  uint64_t z = y + key;
  x = x * x + y;
  x = rot32(x);
  x = x * x + z;
  x = rot32(x);
  x = x * x + y;
  x = rot32(x);
  return (uint32_t)((x * x + z) >> 32);
}
__device__ __forceinline__ uint32_t squares4(uint64_t key, uint64_t ctr) {
  return squares4_y(key, key * ctr);
}

__device__ __forceinline__ uint64_t rotl64(uint64_t x, unsigned n) {
  return (x << n) | (x >> ((64 - n) & 63));
}

// Philox2x64, R rounds (Salmon et al. SC'11): multiply-hi/lo round, Weyl key bump.
template <int R>
__device__ __forceinline__ void philox2x64(uint64_t c0, uint64_t c1, uint64_t k,
                                           uint64_t& o0, uint64_t& o1) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r > 0) k += 0x9E3779B97F4A7C15ULL;
    uint64_t lo = 0xD2B74407B1CE6E93ULL * c0;
    uint64_t hi = __umul64hi(0xD2B74407B1CE6E93ULL, c0);
    c0 = hi ^ k ^ c1;
    c1 = lo;
  }
  o0 = c0;
  o1 = c1;
}

template <int R>
__device__ __forceinline__ void philox4x64(uint64_t (&c)[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B97F4A7C15ULL;
      k1 += 0xBB67AE8584CAA73BULL;
    }
    uint64_t lo0 = 0xD2E7470EE14C6C93ULL * c[0];
    uint64_t hi0 = __umul64hi(0xD2E7470EE14C6C93ULL, c[0]);
    uint64_t lo1 = 0xCA5A826395121157ULL * c[2];
    uint64_t hi1 = __umul64hi(0xCA5A826395121157ULL, c[2]);
    uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

constexpr uint64_t kSkeinParity = 0x1BD11BDAA9FC1A22ULL;

// Threefry2x64, R rounds (Threefish-style mix with key injection every 4 rounds).
template <int R>
__device__ __forceinline__ void threefry2x64(uint64_t c0, uint64_t c1, uint64_t k0, uint64_t k1,
                                             uint64_t& o0, uint64_t& o1) {
  constexpr unsigned rot[8] = {16, 42, 12, 31, 16, 32, 24, 21};
  const uint64_t ks[3] = {k0, k1, kSkeinParity ^ k0 ^ k1};
  uint64_t x0 = c0 + ks[0], x1 = c1 + ks[1];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    x0 += x1;
    x1 = rotl64(x1, rot[r % 8]);
    x1 ^= x0;
    if ((r % 4) == 3) {
      const int s = (r + 1) / 4;
      x0 += ks[s % 3];
      x1 += ks[(s + 1) % 3];
      x1 += (uint64_t)s;
    }
  }
  o0 = x0;
  o1 = x1;
}

template <int R>
__device__ __forceinline__ void threefry4x64(uint64_t (&x)[4], const uint64_t (&k)[4]) {
  constexpr unsigned rot[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                  {25, 33}, {46, 12}, {58, 22}, {32, 32}};
  const uint64_t ks[5] = {k[0], k[1], k[2], k[3], kSkeinParity ^ k[0] ^ k[1] ^ k[2] ^ k[3]};
#pragma unroll
  for (int i = 0; i < 4; ++i) x[i] += ks[i];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if ((r % 2) == 0) {
      x[0] += x[1]; x[1] = rotl64(x[1], rot[r % 8][0]); x[1] ^= x[0];
      x[2] += x[3]; x[3] = rotl64(x[3], rot[r % 8][1]); x[3] ^= x[2];
    } else {
      x[0] += x[3]; x[3] = rotl64(x[3], rot[r % 8][0]); x[3] ^= x[0];
      x[2] += x[1]; x[1] = rotl64(x[1], rot[r % 8][1]); x[1] ^= x[2];
    }
    if ((r % 4) == 3) {
      const int s = (r + 1) / 4;
#pragma unroll
      for (int i = 0; i < 4; ++i) x[i] += ks[(s + i) % 5];
      x[3] += (uint64_t)s;
    }
  }
}

// Multi-word counter increment with carry (array.h:151-183 semantics).
template <int W>
__device__ __forceinline__ void ctr_add(uint64_t* v, uint64_t n) {
  v[0] += n;
  bool carry = v[0] < n;
#pragma unroll
  for (int i = 1; i < W; ++i) {
    v[i] += carry ? 1ULL : 0ULL;
    carry = carry && (v[i] == 0);
  }
}

// One 2- or 4-word block of a Random123 engine for a counter value.
template <int ENGINE>
__device__ __forceinline__ void r123_block(const uint64_t* key, const uint64_t (&ctr)[4],
                                           uint64_t (&out)[4]) {
  if constexpr (ENGINE == HPSP_PHILOX) {
    philox2x64<7>(ctr[0], ctr[1], key[0], out[0], out[1]);
  } else if constexpr (ENGINE == HPSP_PHILOX_LONG) {
    uint64_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    philox4x64<7>(c, key[0], key[1]);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = c[i];
  } else if constexpr (ENGINE == HPSP_THREEFRY) {
    threefry2x64<12>(ctr[0], ctr[1], key[0], key[1], out[0], out[1]);
  } else {
    uint64_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    const uint64_t k[4] = {key[0], key[1], key[2], key[3]};
    threefry4x64<12>(c, k);
#pragma unroll
    for (int i = 0; i < 4; ++i) out[i] = c[i];
  }
}

template <int ENGINE>
constexpr int engine_arity() {
  return (ENGINE == HPSP_PHILOX || ENGINE == HPSP_THREEFRY) ? 2
         : (ENGINE == HPSP_PHILOX_LONG || ENGINE == HPSP_THREEFRY_LONG) ? 4
                                                                          : 1;
}

template <int ENGINE>
constexpr bool is_squares() {
  return ENGINE == HPSP_SQUARES3 || ENGINE == HPSP_SQUARES4;
}

template <int ENGINE>
__device__ __forceinline__ uint32_t squares_word(uint64_t key, uint64_t ctr) {
  if constexpr (ENGINE == HPSP_SQUARES3)
    return squares3(key, ctr);
  else
    return squares4(key, ctr);
}
// The same from y = key * ctr.
template <int ENGINE>
__device__ __forceinline__ uint32_t squares_word_y(uint64_t key, uint64_t y) {
  if constexpr (ENGINE == HPSP_SQUARES3)
    return squares3_y(key, y);
  else
    return squares4_y(key, y);
}

// SplitMix64 step (splitmix.h:54-61).
__device__ __forceinline__ uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

// Key words an engine's constructor derives from a seed: squares3/4 one
// splitmix64 draw (squares3.h:81-84), Random123 one draw per key word
// (EngineR123.h:97-112).
template <int ENGINE>
__device__ __forceinline__ void engine_keys_dev(uint64_t seed, uint64_t (&key)[4]) {
  constexpr int nkey = (ENGINE == HPSP_PHILOX_LONG || ENGINE == HPSP_THREEFRY) ? 2
                       : (ENGINE == HPSP_THREEFRY_LONG)                       ? 4
                                                                              : 1;
  uint64_t s = seed;
#pragma unroll
  for (int i = 0; i < 4; ++i) key[i] = (i < nkey) ? splitmix64(s) : 0;
}

// Szudzik pairing of (event index, seed) that seeds EvalMothers' per-event
// engines (EvalMothers.h:131-147); the index passes through an int there.
__device__ __forceinline__ uint64_t pair_hash(uint64_t evt, uint64_t seed) {
  const uint64_t a = (uint64_t)(int64_t)(int32_t)(uint32_t)evt;
  const uint64_t A = 2 * a, B = 2 * seed;
  return (A >= B ? A * A + A + B : A + B * B) / 2;
}

// Word j of the stream after GRND eng(seed); eng.discard(d):
// squares: counter d + j; Random123: block counter d + 1 + j / arity
// (pre-incremented), word j % arity.  Words of one event are produced
// together so the block evaluations are shared.
template <int ENGINE, int W>
__device__ __forceinline__ void r123_words(const uint64_t* key, uint64_t d, uint64_t (&w)[W]) {
  constexpr int A = engine_arity<ENGINE>();
  constexpr int NB = (W + A - 1) / A;
  uint64_t ctr[4] = {0, 0, 0, 0};
  ctr_add<A>(ctr, d);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    ctr_add<A>(ctr, 1);
    uint64_t out[4];
    r123_block<ENGINE>(key, ctr, out);
#pragma unroll
    for (int i = 0; i < A; ++i)
      if (b * A + i < W) w[b * A + i] = out[i];
  }
}

// minstd_rand: x -> 48271 x mod (2^31 - 1) (linear_congruential_engine.inl:
// 53-61, linear_congruential_engine.h:138-142).  With m a Mersenne prime,
// p mod m folds as (p & m) + (p >> 31): for x, y < m the sum is at most
// 2m - 1, so one conditional subtraction completes it.
constexpr uint32_t kMinstdA = 48271u;
constexpr uint32_t kMinstdM = 0x7fffffffu;

__host__ __device__ __forceinline__ uint32_t minstd_mul(uint32_t x, uint32_t y) {
  const uint64_t p = (uint64_t)x * y;
  const uint32_t r = (uint32_t)(p & kMinstdM) + (uint32_t)(p >> 31);
  return r >= kMinstdM ? r - kMinstdM : r;
}

// 48271^e mod m by square-and-multiply (the reference's discard,
// linear_congruential_engine_discard.h:49-80, forms the same power).
__host__ __device__ __forceinline__ uint32_t minstd_pow(uint64_t e) {
  uint32_t r = 1, b = kMinstdA;
  while (e) {
    if (e & 1) r = minstd_mul(r, b);
    b = minstd_mul(b, b);
    e >>= 1;
  }
  return r;
}

// The engine's state after seed(s): s mod m, or 1 where that is 0
// (linear_congruential_engine.inl:40-50).
__host__ __device__ __forceinline__ uint32_t minstd_seed(uint32_t s) {
  const uint32_t x = s % kMinstdM;
  return x ? x : 1u;
}

constexpr int kMinstd = HPSP_MINSTD_RAND;

// Engines whose words fit 32 bits.
template <int ENGINE>
constexpr bool narrow_words() {
  return is_squares<ENGINE>() || ENGINE == kMinstd;
}

// word -> double in [0,1) (uniform_real_distribution.inl:62-80):
// (word - min) / (1 + (max - min)).  minstd_rand has min 1 and max m - 1, so
// (x - 1) / 2147483646, a rounded division.
template <int ENGINE>
__device__ __forceinline__ double to_unit(uint64_t w) {
#pragma clang fp contract(off)
  if constexpr (ENGINE == kMinstd)
    return (double)((uint32_t)w - 1u) / 2147483646.0;
  else if constexpr (is_squares<ENGINE>())
    return (double)(uint32_t)w * 0x1p-32;
  else
    return (double)w * 0x1p-64;
}

// In-register ascending sort of a small fixed-size array (odd-even
// transposition network; any correct sort yields the same values).
template <typename T, int K>
__device__ __forceinline__ void sort_net(T (&v)[K]) {
#pragma unroll
  for (int r = 0; r < K; ++r) {
#pragma unroll
    for (int i = (r & 1); i + 1 < K; i += 2) {
      T a = v[i], b = v[i + 1];
      v[i] = a < b ? a : b;
      v[i + 1] = a < b ? b : a;
    }
  }
}

// Wave-level sharing of squares words.  Lane L's event e0+L needs the words
// at counters c0 + L + k, k = 0..W-1 (W <= 64).  Each lane evaluates the two
// words at c0 + L ("lo") and c0 + 64 + L ("hi").  For word k, lane L reads
// source lane s = (L+k) mod 64; if L+k < 64 it needs lo[s] and then s >= k,
// otherwise hi[s] and then s < k.  So each source lane offers
// (s >= k ? lo : hi) and one ds_bpermute per word suffices.
struct WaveWords {
  uint32_t lo, hi;
  int lane;
  __device__ __forceinline__ uint32_t get(int k) const {
    if (k == 0) return lo;
    const uint32_t src = (lane >= k) ? lo : hi;
    return (uint32_t)__builtin_amdgcn_ds_bpermute(((lane + k) & (kWave - 1)) << 2, (int)src);
  }
};

// ---------------------------------------------------------------------------
// sin and cos of 2*pi*u for u in [0, 1].  u = (j + d) / 256 with j = floor(256 u)
// and d in [0, 1) (both exact: 256 u is u scaled by a power of two); the
// angle is 2*pi*j/256 + x with x = d * pi/128 in [0, pi/128) (a two-term
// pi/128), so sin and cos are the table's correctly rounded values at
// 2*pi*j/256 turned by x: s = sa*cx + ca*sx, c = ca*cx - sa*sx, with sin x
// and cos x from their Taylor series to x^7 / x^6 (the first omitted terms
// are below 1e-20 on [0, pi/128)).  The result is within a few 1e-16 of
// sin/cos(2*pi*u); the reference evaluates sin/cos of the rounded angle
// fl(2*pi*u), which differs from 2*pi*u by at most half an ulp.
__device__ __forceinline__ void sincos_turn(int j, double x, double& s, double& c) {
#pragma clang fp contract(off)
  const double sa = kSinCos256[j][0];
  const double ca = kSinCos256[j][1];
  const double x2 = x * x;
  double ps = __builtin_fma(x2, -1.984126984126984e-04, 8.333333333333333e-03);
  ps = __builtin_fma(x2, ps, -0.16666666666666666);
  const double sx = __builtin_fma(x * x2, ps, x);
  double pc = __builtin_fma(x2, -1.388888888888889e-03, 4.1666666666666664e-02);
  pc = __builtin_fma(x2, pc, -0.5);
  const double cx = __builtin_fma(x2, pc, 1.0);
  s = __builtin_fma(sa, cx, ca * sx);
  c = __builtin_fma(ca, cx, -(sa * sx));
}

__device__ __forceinline__ void sincos_2pi(double u, double& s, double& c) {
#pragma clang fp contract(off)
  const double t = u * 256.0;
  const double jf = __builtin_floor(t);
  const double d = t - jf;
  const double x = __builtin_fma(d, 0.02454369260617026, d * 9.567553118338697e-19);
  sincos_turn((int)jf, x, s, c);  // j in 0..256
}

// The same for u = w / 2^32 (a squares engine's word): j = w >> 24 and
// d = (w mod 2^24) / 2^24 taken from the word, the 2^-24 folded into the two
// constants of x (powers of two scale exactly: the same bits as sincos_2pi).
__device__ __forceinline__ void sincos_2pi_w32(uint32_t w, double& s, double& c) {
#pragma clang fp contract(off)
  const double dd = (double)(w & 0xFFFFFFu);
  const double x = __builtin_fma(dd, 0.02454369260617026 * 0x1p-24,
                                 dd * (9.567553118338697e-19 * 0x1p-24));
  sincos_turn((int)(w >> 24), x, s, c);
}

// ---------------------------------------------------------------------------
// FP64 kinematics with the reference's operation order.  Each helper turns
// contraction off itself, so the operation sequence holds in any translation
// unit.

// Correctly rounded FP64 square root: reciprocal-square-root estimate, one
// Newton step on (g ~ sqrt(x), h ~ 1/(2 sqrt(x))), then two residual
// corrections g += (x - g*g) * h (Markstein's scheme, as the AMDGPU backend
// expands llvm.sqrt.f64).  The backend's rescaling of tiny inputs and its
// zero/infinity fix-ups are replaced by sending inputs outside
// [2^-767, 2^1000] to the library sqrt.
__device__ __forceinline__ double sqrt_cr(double x) {
  if (!(x >= 0x1p-767 && x <= 0x1p+1000)) return __builtin_sqrt(x);
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y;
  double h = 0.5 * y;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-g, g, x);
  g = __builtin_fma(d, h, g);
  d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Faithful (within about one ulp) square root for x >= 0 where only a
// tolerance applies (energies, sin of the polar angle, boost factors): the
// same scheme with one residual correction; zero stays zero because the
// estimate is taken of max(x, 2^-1000).
__device__ __forceinline__ double sqrt_fast(double x) {
  const double y = __builtin_amdgcn_rsq(__builtin_fmax(x, 0x1p-1000));
  double g = x * y;
  double h = 0.5 * y;
  const double r = __builtin_fma(-h, g, 0.5);
  g = __builtin_fma(g, r, g);
  h = __builtin_fma(h, r, h);
  const double d = __builtin_fma(-g, g, x);
  return __builtin_fma(d, h, g);
}

// Reciprocal within about one ulp: estimate and two Newton steps.
__device__ __forceinline__ double rcp_fast(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double pdk(double a, double b, double c) {
#pragma clang fp contract(off)
  return sqrt_cr((a - b - c) * (a + b + c) * (a - b + c) * (a + b - c)) / (2 * a);
}

__device__ __forceinline__ double mass_of(double e, double x, double y, double z) {
#pragma clang fp contract(off)
  const double m2 = e * e - x * x - y * y - z * z;
  return (m2 > 0.0) ? sqrt_cr(m2) : 0.0;
}

// Vector4R::applyBoostTo(bx,by,bz) forward (Vector4R.inl:312-369) for a
// per-event boost vector.
__device__ __forceinline__ void boost_general(double (&v)[4], double bx, double by, double bz) {
#pragma clang fp contract(off)
  const double bxx = bx * bx;
  const double byy = by * by;
  const double bzz = bz * bz;
  const double b2 = bxx + byy + bzz;
  if (b2 > 0.0 && b2 < 1.0) {
    const double gamma = 1.0 / sqrt(1.0 - b2);
    const double gb2 = (gamma - 1.0) / b2;
    const double gb2xy = gb2 * bx * by;
    const double gb2xz = gb2 * bx * bz;
    const double gb2yz = gb2 * by * bz;
    const double gbx = gamma * bx;
    const double gby = gamma * by;
    const double gbz = gamma * bz;
    const double e2 = v[0], px2 = v[1], py2 = v[2], pz2 = v[3];
    v[0] = gamma * e2 + gbx * px2 + gby * py2 + gbz * pz2;
    v[1] = gbx * e2 + gb2 * bxx * px2 + px2 + gb2xy * py2 + gb2xz * pz2;
    v[2] = gby * e2 + gb2 * byy * py2 + py2 + gb2xy * px2 + gb2yz * pz2;
    v[3] = gbz * e2 + gb2 * bzz * pz2 + pz2 + gb2yz * py2 + gb2xz * px2;
  }
}

// The same boost with host-folded constants (single-mother final boost).
__device__ __forceinline__ void boost_folded(double (&v)[4], const DecayK& d) {
#pragma clang fp contract(off)
  const double e2 = v[0], px2 = v[1], py2 = v[2], pz2 = v[3];
  v[0] = d.gamma * e2 + d.gbx * px2 + d.gby * py2 + d.gbz * pz2;
  v[1] = d.gbx * e2 + d.gb2 * d.bxx * px2 + px2 + d.gb2xy * py2 + d.gb2xz * pz2;
  v[2] = d.gby * e2 + d.gb2 * d.byy * py2 + py2 + d.gb2xy * px2 + d.gb2yz * pz2;
  v[3] = d.gbz * e2 + d.gb2 * d.bzz * pz2 + pz2 + d.gb2yz * py2 + d.gb2xz * px2;
}

// PhaseSpaceWeight::Evaluate on register momenta (Decays.inl:215-231).
// R + p_i of one iteration and R += p_i round identically, so the mass of the
// running sum is computed once and reused as R.mass() of the next iteration.
template <int N>
__device__ __forceinline__ double event_weight_f64(const double (&p)[N][4], const double* m,
                                                   double maxw) {
#pragma clang fp contract(off)
  double R0 = p[0][0], R1 = p[0][1], R2 = p[0][2], R3 = p[0][3];
  double w = maxw;
  double mR = mass_of(R0, R1, R2, R3);
#pragma unroll
  for (int i = 1; i < N; ++i) {
    R0 += p[i][0];
    R1 += p[i][1];
    R2 += p[i][2];
    R3 += p[i][3];
    const double mS = mass_of(R0, R1, R2, R3);
    w *= pdk(mS, m[i], mR);
    mR = mS;
  }
  return w;
}

// Functor evaluation (same arithmetic sequence as the oracle's).  P is the
// number of four-vectors visible to the functor.  Exact = true keeps the
// divisions correctly rounded (Unweight(functor) and PhaseSpaceReweight must
// reproduce the reference's bits); otherwise the Breit-Wigner divisions use
// a reciprocal (AverageOn, compared within a tolerance).
// Four-vector p[i] for a runtime (wave-uniform) index i, selected through
// the unrolled register array with bit masks: a chain of value selects is
// folded by the optimizer into an indexed load, which moves the whole array
// to scratch memory.
template <typename Real, int P>
__device__ __forceinline__ void pick4(const Real (&p)[P][4], int i, Real (&v)[4]) {
  using U = std::conditional_t<sizeof(Real) == 8, uint64_t, uint32_t>;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    U x = __builtin_bit_cast(U, p[0][c]);
#pragma unroll
    for (int j = 1; j < P; ++j) {
      const U m = (U)0 - (U)(i == j);
      x = (x & ~m) | (__builtin_bit_cast(U, p[j][c]) & m);
    }
    v[c] = __builtin_bit_cast(Real, x);
  }
}

// Vector4R product and mass2 (Vector4R.h:171-176, Vector4R.inl:56-60).
template <typename Real>
__device__ __forceinline__ Real dot4(const Real (&a)[4], const Real (&b)[4]) {
#pragma clang fp contract(off)
  return a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
}

// One of K scalars for a run-time (wave-uniform) index, through bit masks.
template <typename Real, int K>
__device__ __forceinline__ Real pick1(const Real (&v)[K], int i) {
  using U = std::conditional_t<sizeof(Real) == 8, uint64_t, uint32_t>;
  U x = __builtin_bit_cast(U, v[0]);
#pragma unroll
  for (int j = 1; j < K; ++j) {
    const U m = (U)0 - (U)(i == j);
    x = (x & ~m) | (__builtin_bit_cast(U, v[j]) & m);
  }
  return __builtin_bit_cast(Real, x);
}

// (p[a] + p[b])^2, the pair's invariant mass squared (Vector4R.inl:56-60),
// with the sum taken component by component (a + b and b + a round alike).
template <typename Real>
__device__ __forceinline__ Real pair_mass2(const Real (&u)[4], const Real (&w)[4]) {
#pragma clang fp contract(off)
  Real v[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] = u[c] + w[c];
  return v[0] * v[0] - v[1] * v[1] - v[2] * v[2] - v[3] * v[3];
}

// Four-vectors of one event in registers (the generation kernels).
template <typename Real, int P>
struct RegAcc {
  const Real (&p)[P][4];
  // With at most four four-vectors, the functors of particle pairs take the
  // pair masses from a table of all P(P-1)/2 pairs formed once per event
  // (the pair named by wave-uniform indices is then one of 3 or 6 scalars)
  // instead of selecting two four-vectors per term.
  static constexpr bool kPairTable = P >= 2 && P <= 4;
  static constexpr int kPairs = P * (P - 1) / 2;
  __device__ __forceinline__ void pair_table(Real (&m2)[kPairs > 0 ? kPairs : 1]) const {
    int k = 0;
#pragma unroll
    for (int a = 0; a < P; ++a)
#pragma unroll
      for (int b = a + 1; b < P; ++b) m2[k++] = pair_mass2(p[a], p[b]);
  }
  // index in pair_table of the pair (a, b), a != b
  __device__ __forceinline__ static int pair_index(int a, int b) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo * (2 * P - lo - 1) / 2 + (hi - lo - 1);
  }
  __device__ __forceinline__ void get(int i, Real (&v)[4]) const { pick4(p, i, v); }
  // p[first] + ... + p[last - 1], left to right, through the unrolled array
  __device__ __forceinline__ void sum(int first, int last, Real (&Pv)[4]) const {
#pragma clang fp contract(off)
#pragma unroll
    for (int c = 0; c < 4; ++c) Pv[c] = 0;
#pragma unroll
    for (int i = 0; i < P; ++i) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const Real s = Pv[c] + p[i][c];
        Pv[c] = (i == first) ? p[i][c] : ((i > first && i < last) ? s : Pv[c]);
      }
    }
  }
};

// Three-momentum products of Vector4R (Vector4R.inl:372-386, 426-454).
template <typename Real>
__device__ __forceinline__ Real dot3(const Real (&a)[4], const Real (&b)[4]) {
#pragma clang fp contract(off)
  Real t = a[1] * b[1];
  t += a[2] * b[2];
  t += a[3] * b[3];
  return t;
}

// v / v.d3mag(): operator/ multiplies by the reciprocal (Vector4R.inl:76-86)
template <typename Real>
__device__ __forceinline__ void unit3(Real (&v)[4]) {
#pragma clang fp contract(off)
  const Real m2 = v[1] * v[1] + v[2] * v[2] + v[3] * v[3];
  Real m;
  if constexpr (sizeof(Real) == 4)
    m = __builtin_sqrtf(m2);
  else
    m = sqrt_cr(m2);
  const Real inv = (Real)1.0 / m;
#pragma unroll
  for (int c = 0; c < 4; ++c) v[c] *= inv;
}

// Whether an accessor offers the table of pair masses (RegAcc with P <= 4).
template <typename Acc, typename = void>
struct has_pair_table : std::false_type {};
template <typename Acc>
struct has_pair_table<Acc, std::void_t<decltype(Acc::kPairTable)>>
    : std::integral_constant<bool, Acc::kPairTable> {};
template <typename Acc, bool = has_pair_table<Acc>::value>
struct pair_table_size : std::integral_constant<int, 1> {};
template <typename Acc>
struct pair_table_size<Acc, true> : std::integral_constant<int, Acc::kPairs> {};
template <typename Real, typename Acc>
__device__ __forceinline__ Real pair_from_table(const Acc& acc, int a, int b) {
  if constexpr (has_pair_table<Acc>::value) {
    Real tab[Acc::kPairs];
    acc.pair_table(tab);
    return pick1(tab, Acc::pair_index(a, b));
  } else {
    return Real(0);
  }
}

// Functor value of one event; acc.get(i, v) yields four-vector i and
// acc.sum(first, last, P) the left-to-right sum of four-vectors [first, last).
// The decay-plane angle (atan2 and two normalisations) is compiled only where
// Plane is set, the kernels over stored particles: in the generation kernels
// its registers would be reserved for every functor.
template <typename Real, bool Exact, bool Plane = false, typename Acc>
__device__ __forceinline__ double functor_eval_acc(const FunctorK& f, const Acc& acc) {
#pragma clang fp contract(off)
  if (f.kind == HPSP_F_ONE) return 1.0;
  if (f.kind == HPSP_F_MASS2_01 || f.kind == HPSP_F_MASS2 || f.kind == HPSP_F_MASS) {
    const int ia = f.kind == HPSP_F_MASS2_01 ? 0 : f.a[0];
    const int ib = f.kind == HPSP_F_MASS2_01 ? 1 : f.b[0];
    Real m2;
    if (has_pair_table<Acc>::value && ia != ib) {
      m2 = pair_from_table<Real>(acc, ia, ib);
    } else {
      Real u[4], v[4];
      acc.get(ia, u);
      acc.get(ib, v);
      m2 = pair_mass2(u, v);
    }
    if (f.kind != HPSP_F_MASS) return (double)m2;
    // Vector4R::mass: 0 unless the square is positive
    if constexpr (sizeof(Real) == 4)
      return m2 > 0.0f ? (double)__builtin_sqrtf(m2) : 0.0;
    else if constexpr (Exact)
      return m2 > 0.0 ? sqrt_cr(m2) : 0.0;
    else
      return m2 > 0.0 ? sqrt_fast(m2) : 0.0;
  }
  if (f.kind == HPSP_F_COS_HELICITY) {
    // P = p[first] + ... + p[first + count - 1] left to right (first, count
    // in a[1], b[1]); q = d + p[b0]; d = p[a0]
    Real Pv[4];
    acc.sum(f.a[1], f.a[1] + f.b[1], Pv);
    Real d[4], e[4], q[4];
    acc.get(f.a[0], d);
    acc.get(f.b[0], e);
#pragma unroll
    for (int c = 0; c < 4; ++c) q[c] = d[c] + e[c];
    const Real pd = dot4(Pv, d);
    const Real pq = dot4(Pv, q);
    const Real qd = dot4(q, d);
    const Real mp2 = dot4(Pv, Pv);
    const Real mq2 = dot4(q, q);
    const Real md2 = dot4(d, d);
    const Real num = pd * mq2 - pq * qd;
    const Real den = (pq * pq - mq2 * mp2) * (qd * qd - mq2 * md2);
    if constexpr (Exact) {
      return (double)(num / sqrt_cr(den));
    } else {
      if constexpr (sizeof(Real) == 4)
        return (double)(num * __builtin_amdgcn_rsqf(den));
      else
        return num * rcp_fast(sqrt_fast(den));
    }
  }
  if (Plane && f.kind == HPSP_F_DECAY_PLANE_ANGLE) {
    // angle between the plane of (d, e) and the plane of h about K = d + e,
    // in [0, 2 pi) (phsp_chain.inl:191-207): the components of d and h
    // transverse to K's three-momentum, d' = K x d_perp, chi = atan2(y, x)
    Real d[4], e[4], h[4], K[4];
    acc.get(f.a[0], d);
    acc.get(f.b[0], e);
    acc.get(f.a[1], h);
#pragma unroll
    for (int c = 0; c < 4; ++c) K[c] = d[c] + e[c];
    const Real kk = dot3(K, K);
    const Real cd = dot3(K, d) / kk;
    const Real chh = dot3(K, h) / kk;
    Real dp[4], hp[4], dq[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      dp[c] = d[c] - K[c] * cd;
      hp[c] = h[c] - K[c] * chh;
    }
    dq[0] = 0;
    dq[1] = K[2] * dp[3] - K[3] * dp[2];
    dq[2] = K[3] * dp[1] - K[1] * dp[3];
    dq[3] = K[1] * dp[2] - K[2] * dp[1];
    unit3(dp);
    unit3(dq);
    const double x = (double)dot3(dp, hp), y = (double)dot3(dq, hp);
    double chi = atan2(y, x);
    if (chi < 0.0) chi += 2.0 * 3.141592653589793;
    return chi;
  }
  Real Ar = 0, Ai = 0;
  Real tab[pair_table_size<Acc>::value];
  if constexpr (has_pair_table<Acc>::value) acc.pair_table(tab);
  for (int k = 0; k < f.nterms; ++k) {
    Real s = 0;
    bool tabled = false;
    if constexpr (has_pair_table<Acc>::value) {
      if (f.a[k] != f.b[k]) {
        s = pick1(tab, Acc::pair_index(f.a[k], f.b[k]));
        tabled = true;
      }
    }
    if (!tabled) {
      Real u[4], w[4];
      acc.get(f.a[k], u);
      acc.get(f.b[k], w);
      s = pair_mass2(u, w);
    }
    const Real M = (Real)f.m[k], G = (Real)f.g[k];
    const Real dr = M * M - s;
    const Real di = -(M * G);
    const Real den = dr * dr + di * di;
    Real br, bi;
    if constexpr (Exact) {
      br = dr / den;
      bi = -di / den;
    } else {
      Real inv;
      if constexpr (sizeof(Real) == 4)
        inv = __builtin_amdgcn_rcpf(den);
      else
        inv = rcp_fast(den);
      br = dr * inv;
      bi = -di * inv;
    }
    const Real cr = (Real)f.cr[k], ci = (Real)f.ci[k];
    const Real tr = cr * br - ci * bi;
    const Real ti = cr * bi + ci * br;
    Ar += tr;
    Ai += ti;
  }
  return (double)(Ar * Ar + Ai * Ai);
}

template <typename Real, bool Exact = true, int P>
__device__ __forceinline__ double functor_eval(const FunctorK& f, const Real (&p)[P][4]) {
  return functor_eval_acc<Real, Exact>(f, RegAcc<Real, P>{p});
}

// ---------------------------------------------------------------------------
// StatsPHSP

struct Stats {
  double w, mean, m2;
};

// AddStatsPHSP (StatsPHSP.h:62-86); an empty side (w == 0) leaves the other.
__device__ __forceinline__ Stats stats_merge(const Stats& x, const Stats& y) {
#pragma clang fp contract(off)
  if (x.w == 0.0) return y;
  if (y.w == 0.0) return x;
  Stats r;
  const double w = x.w + y.w;
  const double delta = y.mean - x.mean;
  const double delta2 = delta * delta;
  r.w = w;
  r.mean = (x.mean * x.w + y.mean * y.w) / w;
  r.m2 = x.m2 + y.m2;
  r.m2 += delta2 * x.w * y.w / w;
  return r;
}

// Fold one event {f, wt, 0} into a running accumulator (weighted West
// update; equal to AddStatsPHSP with a single-point right side).
__device__ __forceinline__ void stats_push(Stats& s, double f, double wt) {
#pragma clang fp contract(off)
  if (s.w == 0.0) {
    s.w = wt;
    s.mean = f;
    s.m2 = 0.0;
    return;
  }
  const double w = s.w + wt;
  const double delta = f - s.mean;
  const double r = wt * rcp_fast(w);
  const double mean = s.mean + delta * r;
  s.m2 = s.m2 + wt * delta * (f - mean);
  s.mean = mean;
  s.w = w;
}

__device__ __forceinline__ Stats shfl_stats(const Stats& s, int src) {
  Stats r;
  r.w = __shfl(s.w, src, kWave);
  r.mean = __shfl(s.mean, src, kWave);
  r.m2 = __shfl(s.m2, src, kWave);
  return r;
}

// Deterministic wave reduction: lane 0 ends with the merge tree
// ((l0 l1)(l2 l3))... over the 64 lanes.
__device__ __forceinline__ Stats wave_reduce_stats(Stats s) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    Stats o = shfl_stats(s, lane + off);
    if ((lane & (2 * off - 1)) == 0) s = stats_merge(s, o);
  }
  return s;
}

}  // namespace hpsp
