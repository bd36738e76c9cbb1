// Host-side counter-based engines of the hydra API
// (reference: hydra/detail/random/{squares3,squares4,splitmix,EngineR123}.h,
// hydra/detail/PRNGTypedefs.h:59).  They produce the same words as the
// device kernels (the phase-space kernels key each event's stream the same
// way: GRND eng(seed); eng.discard(evt + 3N)), and they tell PhaseSpace which
// engine to run on the device (engine_id).
#pragma once

#include <array>
#include <cstdint>
#include <limits>

#include "hpsp.h"

namespace hydra {
namespace random {

// SplitMix64 (splitmix.h:54-61): advances x and returns the mixed value.
inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

namespace detail {
inline uint64_t rot32(uint64_t x) { return (x >> 32) | (x << 32); }
inline uint64_t rotl64(uint64_t x, unsigned n) { return (x << n) | (x >> ((64 - n) & 63)); }
inline void mulhilo(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
  const unsigned __int128 p = (unsigned __int128)a * b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
}
}  // namespace detail

// Widynski's counter-based middle-square engines (squares3.h:67-159,
// squares4.h:68-163): key = splitmix64(seed) at construction, SetSeed sets the
// key as given, each call uses then increments the counter.
template <int Rounds, int Id>
class squares_engine {
 public:
  typedef uint64_t state_type;
  typedef uint64_t seed_type;
  typedef uint64_t advance_type;
  typedef uint32_t result_type;
  static constexpr int engine_id = Id;

  squares_engine() = delete;
  explicit squares_engine(seed_type s) : fState(0), fSeed(splitmix64(s)) {}

  result_type operator()() {
    uint64_t y = fSeed * fState;
    uint64_t x = y;
    const uint64_t z = y + fSeed;
    x = x * x + y;
    x = detail::rot32(x);
    x = x * x + z;
    x = detail::rot32(x);
    ++fState;
    if (Rounds == 3) return (result_type)((x * x + y) >> 32);
    x = x * x + y;
    x = detail::rot32(x);
    return (result_type)((x * x + z) >> 32);
  }
  void discard(advance_type n) { fState += n; }
  seed_type GetSeed() const { return fSeed; }
  void SetSeed(seed_type s) { fSeed = s; }
  state_type GetState() const { return fState; }
  void SetState(state_type s) { fState = s; }

  static constexpr result_type min = 0;
  static constexpr result_type max = std::numeric_limits<result_type>::max();

 private:
  state_type fState;
  seed_type fSeed;
};

typedef squares_engine<3, HPSP_SQUARES3> squares3;
typedef squares_engine<4, HPSP_SQUARES4> squares4;

// Random123 bijections (Salmon et al., SC'11) at the reference's round counts
// (Philox 7, Threefry 12; philox.h:266-275, threefry.h:183-192).
struct Philox2x64 {
  static constexpr int A = 2, K = 1;
  static void apply(std::array<uint64_t, 4>& c, const std::array<uint64_t, 4>& k0) {
    uint64_t k = k0[0];
    for (int r = 0; r < 7; ++r) {
      if (r > 0) k += 0x9E3779B97F4A7C15ULL;
      uint64_t hi, lo;
      detail::mulhilo(0xD2B74407B1CE6E93ULL, c[0], hi, lo);
      c[0] = hi ^ k ^ c[1];
      c[1] = lo;
    }
  }
};
struct Philox4x64 {
  static constexpr int A = 4, K = 2;
  static void apply(std::array<uint64_t, 4>& c, const std::array<uint64_t, 4>& k) {
    uint64_t k0 = k[0], k1 = k[1];
    for (int r = 0; r < 7; ++r) {
      if (r > 0) {
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
      }
      uint64_t hi0, lo0, hi1, lo1;
      detail::mulhilo(0xD2E7470EE14C6C93ULL, c[0], hi0, lo0);
      detail::mulhilo(0xCA5A826395121157ULL, c[2], hi1, lo1);
      const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
      c = {n0, lo1, n2, lo0};
    }
  }
};
struct Threefry2x64 {
  static constexpr int A = 2, K = 2;
  static void apply(std::array<uint64_t, 4>& c, const std::array<uint64_t, 4>& k) {
    static const unsigned rot[8] = {16, 42, 12, 31, 16, 32, 24, 21};
    const uint64_t ks[3] = {k[0], k[1], 0x1BD11BDAA9FC1A22ULL ^ k[0] ^ k[1]};
    uint64_t x0 = c[0] + ks[0], x1 = c[1] + ks[1];
    for (int r = 0; r < 12; ++r) {
      x0 += x1;
      x1 = detail::rotl64(x1, rot[r % 8]);
      x1 ^= x0;
      if (r % 4 == 3) {
        const int s = (r + 1) / 4;
        x0 += ks[s % 3];
        x1 += ks[(s + 1) % 3] + (uint64_t)s;
      }
    }
    c[0] = x0;
    c[1] = x1;
  }
};
struct Threefry4x64 {
  static constexpr int A = 4, K = 4;
  static void apply(std::array<uint64_t, 4>& x, const std::array<uint64_t, 4>& k) {
    static const unsigned rot[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                       {25, 33}, {46, 12}, {58, 22}, {32, 32}};
    const uint64_t ks[5] = {k[0], k[1], k[2], k[3],
                            0x1BD11BDAA9FC1A22ULL ^ k[0] ^ k[1] ^ k[2] ^ k[3]};
    for (int i = 0; i < 4; ++i) x[i] += ks[i];
    for (int r = 0; r < 12; ++r) {
      const unsigned* q = rot[r % 8];
      if (r % 2 == 0) {
        x[0] += x[1]; x[1] = detail::rotl64(x[1], q[0]); x[1] ^= x[0];
        x[2] += x[3]; x[3] = detail::rotl64(x[3], q[1]); x[3] ^= x[2];
      } else {
        x[0] += x[3]; x[3] = detail::rotl64(x[3], q[0]); x[3] ^= x[0];
        x[2] += x[1]; x[1] = detail::rotl64(x[1], q[1]); x[1] ^= x[2];
      }
      if (r % 4 == 3) {
        const int s = (r + 1) / 4;
        for (int i = 0; i < 4; ++i) x[i] += ks[(s + i) % 5];
        x[3] += (uint64_t)s;
      }
    }
  }
};

// EngineR123 (EngineR123.h:74-233): key words are successive splitmix64
// draws of the seed, the counter block is pre-incremented (with carry) before
// each evaluation, words are served from the cached block, discard(n) skips
// n blocks and drops the cache.
template <typename B, int Id>
class EngineR123 {
 public:
  typedef uint64_t result_type;
  typedef uint64_t advance_type;
  static constexpr int engine_id = Id;

  EngineR123() = delete;
  explicit EngineR123(result_type seed) : fCtr{}, fCache{}, fKey{}, fTrig(B::A) {
    for (int i = 0; i < B::K; ++i) fKey[i] = splitmix64(seed);
  }
  result_type operator()() {
    if (fTrig == (unsigned)B::A) {
      incr(1);
      fCache = fCtr;
      B::apply(fCache, fKey);
      fTrig = 0;
    }
    return fCache[fTrig++];
  }
  void discard(advance_type n) {
    incr(n);
    fTrig = B::A;
  }
  static constexpr result_type min = 0;
  static constexpr result_type max = std::numeric_limits<result_type>::max();

 private:
  void incr(uint64_t n) {
    fCtr[0] += n;
    if (n <= fCtr[0]) return;
    for (int i = 1; i < B::A; ++i)
      if (++fCtr[i] != 0) return;
  }
  std::array<uint64_t, 4> fCtr, fCache, fKey;
  unsigned fTrig;
};

typedef EngineR123<Philox2x64, HPSP_PHILOX> philox;
typedef EngineR123<Philox4x64, HPSP_PHILOX_LONG> philox_long;
typedef EngineR123<Threefry2x64, HPSP_THREEFRY> threefry;
typedef EngineR123<Threefry4x64, HPSP_THREEFRY_LONG> threefry_long;

}  // namespace random

// hydra/detail/PRNGTypedefs.h:59
typedef random::squares3 default_random_engine;

}  // namespace hydra
