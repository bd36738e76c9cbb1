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
// with q = wt / max wherever |q - u| > (B + 2^-48) q.  Elsewhere (where the
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
// for a 32-bit word and >= 2^-31.5 for a 64-bit one: K = 2^17 N for the
// squares engines, 2^33 N for the Random123 ones.  Over 2^20 events of each
// of nine decays (N = 2..12, near threshold, electrons beside a 90 GeV
// particle) the measured |w - wt| / wt stays below 4% of B with K = 10.
#include <array>
#include <type_traits>
#include <utility>

#include "genbod_f64.h"
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

// The one uniform of a 3-body event with a squares engine: the lane's own word.
struct OneWord {
  uint32_t w0;
  __device__ __forceinline__ uint64_t operator()(int) const { return w0; }
};

// The weight chain of one event (DecayMother.h:170-212: sorted uniforms,
// invMas, pdk, the product) with one-ulp square roots and reciprocals, and
// the bound B of the file comment; sure is false where a term of B is not
// defined (d1 or d2 not positive) or B is not small.  The bound's
// reciprocals come from those of the chain: 1/(d1 d2) = rsq(x)^2 with
// x = d1 d2 under pdk's square root, 1/a^2 = (2 rcp(2a))^2, and c is the
// previous step's a (1/m0^2 = inv_m0sq at the first step).
template <int N, int ENGINE, typename WordFn>
__device__ __forceinline__ double chain_weight(const DecayK& d, const WordFn& word,
                                               double scale2, double kc, double inv_m0sq,
                                               double& B, bool& sure) {
#pragma clang fp contract(off)
  constexpr double eps = 0x1p-53;
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
  double Bs = 0.0;
  double inv_c2 = inv_m0sq;
  bool ok = true;
  const double ks = kc * eps * scale2;
#pragma unroll
  for (int n = 0; n < N - 1; ++n) {
    const double a = invMas[n + 1], b = d.m[n + 1], c = invMas[n];
    const double spc = b + c, smc = b - c;
    const double d1 = (a - spc) * (a + spc);
    const double d2 = (a - smc) * (a + smc);
    const double x = d1 * d2;
    // sqrt(x) from y ~ 1/sqrt(x) with one Newton step on (g, h) and a residual
    // correction (sqrt_fast's scheme), keeping y
    const double y = __builtin_amdgcn_rsq(__builtin_fmax(x, 0x1p-1000));
    double g = x * y, h = 0.5 * y;
    const double r = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    g = __builtin_fma(__builtin_fma(-g, g, x), h, g);
    const double ra = rcp_fast(2.0 * a);
    wt *= g * ra;
    ok = ok && d1 > 0.0 && d2 > 0.0;
    const double inv_x = y * y;
    const double inv_a2 = 4.0 * ra * ra;
    const double ka = a * a * (d1 + d2) * inv_x + 1.0;
    const double kcc = c * (spc * d2 + __builtin_fabs(smc) * d1) * inv_x;
    Bs += (ka + kcc) * (ks * (inv_a2 + inv_c2) + 4.0 * eps);
    inv_c2 = inv_a2;
  }
  B = Bs + 40.0 * N * eps;
  sure = ok && B < 0x1p-20;  // false for NaN
  return wt;
}

}  // namespace

// One 1024-event tile per block, items as k_uw_flags: thread item j is event
// tile * 1024 + j * 256 + wave * 64 + lane of the container (generation
// index a.gen_first + that); flags as 64-bit ballots, counts per tile.  The
// events the chain leaves in doubt get the chain's flag for now and a bit in
// unc[group]; k_uw_patch_chain decides them.
template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(waves_for(N), 8)))
void k_uw_flags_chain(const UnweightArgs a, uint64_t* unc, double scale2, double kc) {
#pragma clang fp contract(off)
  __shared__ uint32_t cnt[kBlock / kWave];
  const int lane = __lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const uint64_t tile = xcd_tile(ntiles);
  if (tile >= ntiles) return;  // whole block
  const double rmax = 1.0 / a.maxv;
  const bool rmax_ok = __builtin_fabs(rmax) >= 0x1p-1000 && __builtin_fabs(rmax) <= 0x1p+1000;
  const double inv_m0sq = 1.0 / (a.d.m[0] * a.d.m[0]);
  const uint64_t off = (uint64_t)wave * kWave + (uint64_t)lane;
  // key * (jump + k) of item 0; item j + 1 is kBlock events on
  uint64_t ykey = a.uw_key * (a.jump + tile * kUwTile + off);
  const uint64_t ystep = a.uw_key * (uint64_t)kBlock;
  uint32_t c = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j, ykey += ystep) {
    const uint64_t g = tile * (kUwTile / kWave) + (uint64_t)j * (kBlock / kWave) + wave;
    const uint64_t k0 = g * kWave;
    const uint64_t k = k0 + (uint64_t)lane;
    double B, wt;
    bool sure;
    if constexpr (N == 3 && is_squares<ENGINE>()) {
      // the one uniform is the lane's own word: counter gen_first + k + 3N
      const OneWord w{squares_word_y<ENGINE>(
          a.d.key[0], a.d.key[0] * (a.gen_first + k + 3 * (uint64_t)N))};
      wt = chain_weight<N, ENGINE>(a.d, w, scale2, kc, inv_m0sq, B, sure);
    } else {
      EventWords<N, ENGINE> words;
      words.init(a.d, a.gen_first + k0, lane);
      wt = chain_weight<N, ENGINE>(a.d, words, scale2, kc, inv_m0sq, B, sure);
    }
    const bool live = k < a.n;
    const double u = (double)squares3_y(a.uw_key, ykey) * 0x1p-32;  // ctr = jump + k
    const double q = wt * rmax;
    const double aq = __builtin_fabs(q);
    sure = sure && rmax_ok && aq >= 0x1p-1000 && aq <= 0x1p+1000 &&
           __builtin_fabs(q - u) > aq * (B + 0x1p-48);
    const uint64_t ball = __ballot(live && q > u);
    const uint64_t doubt = __ballot(live && !sure);
    if (lane == 0) {
      *(glob(a.flags) + g) = ball;
      *(glob(unc) + g) = doubt;
    }
    c += (uint32_t)__popcll(ball);
  }
  if (lane == 0) cnt[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) a.tile_count[tile] = cnt[0] + cnt[1] + cnt[2] + cnt[3];
}

// The events left in doubt: one thread per group of 64 (almost all have none),
// each doubtful event's momenta and PhaseSpaceWeight formed and the
// reference's test applied (FlagDaugthers, Decays.inl:75-89); the flag word
// corrected and the tile's count adjusted.
template <int N, int ENGINE>
__global__ __launch_bounds__(kBlock) void k_uw_patch_chain(const UnweightArgs a,
                                                           const uint64_t* unc) {
#pragma clang fp contract(off)
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
  if (g >= ngroups) return;
  uint64_t m = *(glob(unc) + g);
  if (m == 0) return;
  const double rmax = 1.0 / a.maxv;
  const bool rmax_ok = __builtin_fabs(rmax) >= 0x1p-1000 && __builtin_fabs(rmax) <= 0x1p+1000;
  uint64_t fw = *(glob(a.flags) + g);
  int delta = 0;
  while (m) {
    const int l = __builtin_ctzll(m);
    m &= m - 1;
    const uint64_t k = g * kWave + (uint64_t)l;
    LaneEventWords<N, ENGINE> lw;
    lw.init(a.d, a.gen_first + k);
    double p[N][4];
    genbod_f64<N, ENGINE>(a.d, lw, p);
    if (a.d.do_boost) {
#pragma unroll
      for (int i = 0; i < N; ++i) boost_folded(p[i], a.d);
    }
    const double u = (double)squares3_y(a.uw_key, a.uw_key * (a.jump + k)) * 0x1p-32;
    const bool acc =
        ratio_above(event_weight_f64<N>(p, a.d.m, a.d.maxw), a.maxv, rmax, rmax_ok, u);
    const bool was = (fw >> l) & 1ULL;
    if (acc != was) {
      fw ^= 1ULL << l;
      delta += acc ? 1 : -1;
    }
  }
  *(glob(a.flags) + g) = fw;
  if (delta) atomicAdd(a.tile_count + g / (kUwTile / kWave), (uint32_t)delta);
}

constexpr int kNumN = HPSP_MAX_N - 1;
constexpr int kNumE = 6;

template <int N, int E>
static void chain_one(const UnweightArgs& a, uint64_t* unc, double scale2, double kc,
                      hipStream_t s) {
  const uint64_t ntiles = (a.n + kUwTile - 1) / kUwTile;
  const unsigned grid = (unsigned)((ntiles + kXcds - 1) / kXcds * kXcds);
  hipLaunchKernelGGL((k_uw_flags_chain<N, E>), dim3(grid), dim3(kBlock), 0, s, a, unc, scale2,
                     kc);
  const uint64_t ngroups = (a.n + kWave - 1) / kWave;
  hipLaunchKernelGGL((k_uw_patch_chain<N, E>), dim3((unsigned)((ngroups + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, s, a, (const uint64_t*)unc);
}

template <int I>
static constexpr auto chain_at = &chain_one<2 + I / kNumE, I % kNumE>;

template <int... I>
static constexpr auto chain_table(std::integer_sequence<int, I...>) {
  return std::array<void (*)(const UnweightArgs&, uint64_t*, double, double, hipStream_t),
                    sizeof...(I)>{chain_at<I>...};
}

static const auto kChain = chain_table(std::make_integer_sequence<int, kNumN * kNumE>{});

int launch_flags_chain(int N, int engine, const UnweightArgs& a, uint64_t* unc, double scale2,
                       double kc, hipStream_t s) {
  kChain[(N - 2) * kNumE + engine](a, unc, scale2, kc, s);
  return (int)hipGetLastError();
}

}  // namespace hpsp
