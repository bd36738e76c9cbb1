/*
 * CPU restatement of the Hydra 4.0.1 phase-space hot path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP kernels in hydra_amd/.
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * the library built from this file; the product never links or calls it.
 *
 * Every function follows the reference's IEEE operation order (build with
 * -ffp-contract=off) so that, compiled by the same gcc against the same
 * glibc libm as the reference harness, it reproduces the reference bit for bit.
 * Reference files are cited relative to the Hydra source root.
 */
#include "phsp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Types.h:75 -- PI rounds to M_PI in double */
static const double PO_PI = 3.1415926535897932384626422832795028841971;

/* ------------------------------------------------------------------------ */
/* SplitMix64 (Steele, Lea, Flood 2014), as used for key derivation at
 * hydra/detail/random/splitmix.h:54-61: pre-increment by the golden gamma,
 * two xor-shift-multiply mixes, final xor-shift.                            */
uint64_t po_splitmix64(uint64_t* x) {
  uint64_t z = (*x += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline uint64_t rot32(uint64_t x) { return (x >> 32) | (x << 32); }

/* Widynski's counter-based middle-square generator, 3 rounds
 * (hydra/detail/random/squares3.h:105-119; arXiv:2004.06278).               */
uint32_t po_squares3(uint64_t key, uint64_t ctr) {
  uint64_t x, y, z;
  y = x = key * ctr;
  z = y + key;
  x = x * x + y;
  x = rot32(x);
  x = x * x + z;
  x = rot32(x);
  return (uint32_t)((x * x + y) >> 32);
}

/* 4-round variant (hydra/detail/random/squares4.h:106-122).                 */
uint32_t po_squares4(uint64_t key, uint64_t ctr) {
  uint64_t x, y, z;
  y = x = key * ctr;
  z = y + key;
  x = x * x + y;
  x = rot32(x);
  x = x * x + z;
  x = rot32(x);
  x = x * x + y;
  x = rot32(x);
  return (uint32_t)((x * x + z) >> 32);
}

/* ------------------------------------------------------------------------ */
/* Random123 bijections (Salmon, Moraes, Dror, Shaw, SC'11), restated from the
 * published algorithm; constants as in hydra/detail/external/hydra_R123/
 * philox.h:223-258 and threefry.h:66-90,177-178.                           */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
}

void po_philox2x64(unsigned R, const uint64_t ctr[2], const uint64_t key[1], uint64_t out[2]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], k = key[0];
  for (unsigned r = 0; r < R; ++r) {
    if (r > 0) k += 0x9E3779B97F4A7C15ULL;
    uint64_t hi, lo;
    mulhilo64(0xD2B74407B1CE6E93ULL, c0, &hi, &lo);
    c0 = hi ^ k ^ c1;
    c1 = lo;
  }
  out[0] = c0;
  out[1] = c1;
}

void po_philox4x64(unsigned R, const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (unsigned r = 0; r < R; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B97F4A7C15ULL;
      k1 += 0xBB67AE8584CAA73BULL;
    }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(0xD2E7470EE14C6C93ULL, c0, &hi0, &lo0);
    mulhilo64(0xCA5A826395121157ULL, c2, &hi1, &lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint64_t rotl64(uint64_t x, unsigned n) {
  return (x << (n & 63)) | (x >> ((64 - n) & 63));
}

#define PO_SKEIN_PARITY 0x1BD11BDAA9FC1A22ULL

void po_threefry2x64(unsigned R, const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2]) {
  static const unsigned rot[8] = {16, 42, 12, 31, 16, 32, 24, 21};
  uint64_t ks[3];
  ks[0] = key[0];
  ks[1] = key[1];
  ks[2] = PO_SKEIN_PARITY ^ key[0] ^ key[1];
  uint64_t x0 = ctr[0] + ks[0], x1 = ctr[1] + ks[1];
  for (unsigned r = 0; r < R; ++r) {
    x0 += x1;
    x1 = rotl64(x1, rot[r % 8]);
    x1 ^= x0;
    if ((r % 4) == 3) {
      unsigned s = (r + 1) / 4;
      x0 += ks[s % 3];
      x1 += ks[(s + 1) % 3];
      x1 += s;
    }
  }
  out[0] = x0;
  out[1] = x1;
}

void po_threefry4x64(unsigned R, const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]) {
  static const unsigned rot[8][2] = {{14, 16}, {52, 57}, {23, 40}, {5, 37},
                                     {25, 33}, {46, 12}, {58, 22}, {32, 32}};
  uint64_t ks[5];
  ks[4] = PO_SKEIN_PARITY;
  for (int i = 0; i < 4; ++i) {
    ks[i] = key[i];
    ks[4] ^= key[i];
  }
  uint64_t x[4];
  for (int i = 0; i < 4; ++i) x[i] = ctr[i] + ks[i];
  for (unsigned r = 0; r < R; ++r) {
    const unsigned* q = rot[r % 8];
    if ((r % 2) == 0) {
      x[0] += x[1]; x[1] = rotl64(x[1], q[0]); x[1] ^= x[0];
      x[2] += x[3]; x[3] = rotl64(x[3], q[1]); x[3] ^= x[2];
    } else {
      x[0] += x[3]; x[3] = rotl64(x[3], q[0]); x[3] ^= x[0];
      x[2] += x[1]; x[1] = rotl64(x[1], q[1]); x[1] ^= x[2];
    }
    if ((r % 4) == 3) {
      unsigned s = (r + 1) / 4;
      for (int i = 0; i < 4; ++i) x[i] += ks[(s + i) % 5];
      x[3] += s;
    }
  }
  for (int i = 0; i < 4; ++i) out[i] = x[i];
}

/* ------------------------------------------------------------------------ */
/* Engine objects with the reference's stream semantics.
 *   squares3/4: key = splitmix64(seed) (squares3.h:81-84), counter from 0,
 *               operator() uses then post-increments the counter,
 *               discard(n) adds n (squares3.h:105-125).
 *   EngineR123: key words are successive splitmix64(seed) draws
 *               (EngineR123.h:97-112); counter block starts at 0 and is
 *               pre-incremented (with carry, array.h:151-183) before each
 *               block evaluation; words are served from the cached block;
 *               discard(n) adds n blocks and drops the cache (EngineR123.h:148-173).
 */
/*   minstd_rand (thrust's default_random_engine, the engine of
 *               phase_space_range): linear_congruential_engine<uint32, 48271,
 *               0, 2^31-1> (linear_congruential_engine.h:138-142,
 *               linear_congruential_engine.inl:40-61): seed(s) keeps
 *               s mod m, or 1 when that is 0; operator() steps
 *               x = 48271 x mod m and returns it; discard(z) multiplies by
 *               48271^z mod m (linear_congruential_engine_discard.h:49-80).
 *               The state lives in ctr[0].                                  */
#define PO_MINSTD_A 48271ULL
#define PO_MINSTD_M 2147483647ULL

typedef struct {
  int engine;
  unsigned arity;
  uint64_t key[4];
  uint64_t ctr[4];
  uint64_t cache[4];
  unsigned trig;
} po_engine;

static unsigned engine_arity(int e) {
  switch (e) {
    case 2: case 4: return 2;
    case 3: case 5: return 4;
    default: return 1;
  }
}

static void engine_init(po_engine* g, int engine, uint64_t seed) {
  memset(g, 0, sizeof(*g));
  g->engine = engine;
  g->arity = engine_arity(engine);
  uint64_t s = seed;
  if (engine == 6) {
    uint64_t x = (uint64_t)(uint32_t)seed % PO_MINSTD_M;
    g->ctr[0] = x ? x : 1;
  } else if (engine == 0 || engine == 1) {
    g->key[0] = po_splitmix64(&s);
  } else {
    unsigned nkey = (engine == 2) ? 1 : (engine == 3) ? 2 : (engine == 4) ? 2 : 4;
    for (unsigned i = 0; i < nkey; ++i) g->key[i] = po_splitmix64(&s);
    g->trig = g->arity;
  }
}

static void ctr_incr(uint64_t* v, unsigned width, uint64_t n) {
  v[0] += n;
  if (width == 1 || n <= v[0]) return;
  for (unsigned i = 1; i < width; ++i) {
    ++v[i];
    if (v[i] != 0) return;
  }
}

/* a^z mod m by square-and-multiply over the bits of z */
static uint64_t minstd_pow(uint64_t z) {
  uint64_t r = 1, b = PO_MINSTD_A;
  while (z) {
    if (z & 1) r = r * b % PO_MINSTD_M;
    b = b * b % PO_MINSTD_M;
    z >>= 1;
  }
  return r;
}

static void engine_discard(po_engine* g, uint64_t n) {
  if (g->engine == 6) {
    g->ctr[0] = minstd_pow(n) * g->ctr[0] % PO_MINSTD_M;
  } else if (g->engine == 0 || g->engine == 1) {
    g->ctr[0] += n;
  } else {
    ctr_incr(g->ctr, g->arity, n);
    g->trig = g->arity;
  }
}

static void engine_block(po_engine* g) {
  switch (g->engine) {
    case 2: po_philox2x64(7, g->ctr, g->key, g->cache); break;
    case 3: po_philox4x64(7, g->ctr, g->key, g->cache); break;
    case 4: po_threefry2x64(12, g->ctr, g->key, g->cache); break;
    case 5: po_threefry4x64(12, g->ctr, g->key, g->cache); break;
    default: break;
  }
}

static uint64_t engine_next(po_engine* g) {
  if (g->engine == 6) {
    g->ctr[0] = g->ctr[0] * PO_MINSTD_A % PO_MINSTD_M;
    return g->ctr[0];
  }
  if (g->engine == 0) return po_squares3(g->key[0], g->ctr[0]++);
  if (g->engine == 1) return po_squares4(g->key[0], g->ctr[0]++);
  if (g->trig == g->arity) {
    ctr_incr(g->ctr, g->arity, 1);
    engine_block(g);
    g->trig = 0;
  }
  return g->cache[g->trig++];
}

/* uniform_real_distribution<double>(0,1) over the engine's word
 * (hydra_thrust/random/detail/uniform_real_distribution.inl:62-80):
 * double(word - min) / (1 + double(max - min)), then *(b-a)+a with a=0,b=1.
 * For 32-bit words this is word * 2^-32 exactly; for 64-bit words the
 * conversion rounds and the result can equal 1.0.  minstd_rand has min 1 and
 * max m-1: (word - 1) / (m - 1), a rounded division.                        */
double po_uniform(int engine, uint64_t word) {
  if (engine == 6) {
    double r = (double)(uint32_t)((uint32_t)word - 1u);
    r /= (1.0 + (double)(uint32_t)(2147483646u - 1u));
    return r * (1.0 - 0.0) + 0.0;
  }
  double r = (double)word;
  if (engine == 0 || engine == 1)
    r /= (1.0 + (double)UINT32_MAX);
  else
    r /= (1.0 + (double)UINT64_MAX);
  return r * (1.0 - 0.0) + 0.0;
}

static inline double engine_uniform(po_engine* g) {
  return po_uniform(g->engine, engine_next(g));
}

int po_rng_words(int engine, uint64_t seed, uint64_t discard, uint64_t n, uint64_t* out) {
  if (engine < 0 || engine > 6) return -1;
  po_engine g;
  engine_init(&g, engine, seed);
  engine_discard(&g, discard);
  for (uint64_t i = 0; i < n; ++i) out[i] = engine_next(&g);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Kinematics helpers                                                         */

/* DecayMother.h:118-126 -- left-to-right product, then sqrt / (2a) */
static inline double pdk(double a, double b, double c) {
  return sqrt((a - b - c) * (a + b + c) * (a - b + c) * (a + b - c)) / (2 * a);
}

/* Vector4R::applyBoostTo(bx,by,bz), forward direction (Vector4R.inl:312-369) */
static void boost(double* v, double bx, double by, double bz) {
  double bxx = bx * bx;
  double byy = by * by;
  double bzz = bz * bz;
  double b2 = bxx + byy + bzz;
  if (b2 > 0.0 && b2 < 1.0) {
    double gamma = 1.0 / sqrt(1.0 - b2);
    double gb2 = (gamma - 1.0) / b2;
    double gb2xy = gb2 * bx * by;
    double gb2xz = gb2 * bx * bz;
    double gb2yz = gb2 * by * bz;
    double gbx = gamma * bx;
    double gby = gamma * by;
    double gbz = gamma * bz;
    double e2 = v[0], px2 = v[1], py2 = v[2], pz2 = v[3];
    v[0] = gamma * e2 + gbx * px2 + gby * py2 + gbz * pz2;
    v[1] = gbx * e2 + gb2 * bxx * px2 + px2 + gb2xy * py2 + gb2xz * pz2;
    v[2] = gby * e2 + gb2 * byy * py2 + py2 + gb2xy * px2 + gb2yz * pz2;
    v[3] = gbz * e2 + gb2 * bzz * pz2 + pz2 + gb2yz * py2 + gb2xz * px2;
  }
}

/* Vector4R::mass2 / mass (Vector4R.inl:56-60,196-208) */
static inline double mass2(const double* v) {
  return v[0] * v[0] - v[1] * v[1] - v[2] * v[2] - v[3] * v[3];
}
static inline double mass(const double* v) {
  double m2 = v[0] * v[0] - v[1] * v[1] - v[2] * v[2] - v[3] * v[3];
  return (m2 > 0.0) ? sqrt(m2) : 0.0;
}

/* PhaseSpace ctor (PhaseSpace.inl:42-90): validation, fECM, fMaxWeight.
 * Returns 0, -1 for a non-positive mass (std::invalid_argument), -2 for
 * fECM <= 0 (std::runtime_error), -3 for N out of range.                   */
int po_params(int N, double M, const double* masses, double* ecm, double* max_weight) {
  if (N < 2 || N > PO_MAXN) return -3;
  if (M <= 0.0) return -1;
  for (int i = 0; i < N; ++i)
    if (masses[i] <= 0.0) return -1;
  double e = M;
  for (int n = 0; n < N; ++n) e -= masses[n];
  if (e <= 0.0) return -2;
  double emmax = e + masses[0];
  double emmin = 0.0;
  double wtmax = 1.0;
  for (int n = 1; n < N; ++n) {
    emmin += masses[n - 1];
    emmax += masses[n];
    double x = (emmax - emmin - masses[n]) * (emmax + emmin + masses[n]) *
               (emmax - emmin + masses[n]) * (emmax + emmin - masses[n]);
    x = sqrt(x) / (2 * emmax);
    wtmax *= x;
  }
  *ecm = e;
  *max_weight = 1.0 / wtmax;
  return 0;
}

/* DecayMother ctor boost vector (DecayMother.h:88-98), including the
 * reference's (E, px, py)/E component choice.                               */
void po_mother_beta(const double* m4, double* b) {
  double d3 = sqrt(m4[1] * m4[1] + m4[2] * m4[2] + m4[3] * m4[3]);
  double beta = d3 / m4[0];
  if (beta != 0.0) {
    double w = beta / d3;
    b[0] = m4[0] * w;
    b[1] = m4[1] * w;
    b[2] = m4[2] * w;
  } else {
    b[0] = b[1] = b[2] = 0.0;
  }
}

/* GenerateDecay<N, minstd_rand>(mother, masses, seed), the functor behind
 * phase_space_range (PhaseSpaceRange.inl:44-72, GenerateDecay.h:64-101):
 * fTeCmTm = mother.mass() - sum(m), fWtMax from the same pdk product as the
 * PhaseSpace constructor, and the DecayMother boost vector.  The reference
 * checks nothing here; -1 / -2 flag what the PhaseSpace constructor would
 * reject (a non-positive mass, fTeCmTm <= 0).  The element k of the range is
 * then the DecayMother event k of engine 6 seeded with (uint32)(int)seed.   */
int po_range_params(int N, const double* mother4, const double* masses, double* ecm,
                    double* max_weight, double* beta3) {
  int rc = po_params(N, mass(mother4), masses, ecm, max_weight);
  if (rc) return rc;
  po_mother_beta(mother4, beta3);
  return 0;
}

/* GENBOD body: DecayMother::process (DecayMother.h:163-276) and
 * DecayMothers::process (DecayMothers.h:131-242).  final_mode 0 boosts by
 * beta3 (single mother); 1 boosts by the per-event mother four-vector
 * (applyBoostTo(Vector4R), Vector4R.inl:237-247).                           */
static double decay_core(int N, int engine, const double* m, double ecm, double maxw,
                         int final_mode, const double* fin, uint64_t seed, uint64_t discard,
                         double (*p)[4]) {
  po_engine g;
  engine_init(&g, engine, seed);
  engine_discard(&g, discard);

  double rno[PO_MAXN];
  rno[0] = 0.0;
  rno[N - 1] = 1.0;
  if (N > 2) {
    for (int n = 1; n < N - 1; ++n) rno[n] = engine_uniform(&g);
    /* bubble sort of rno[1..N-2] (DecayMother.h:128-151) */
    int cnt = N - 2;
    for (int c = 0; c < cnt; ++c) {
      int nswap = 0;
      for (int d = 0; d < cnt - c - 1; ++d) {
        if (rno[1 + d] > rno[1 + d + 1]) {
          double t = rno[1 + d];
          rno[1 + d] = rno[1 + d + 1];
          rno[1 + d + 1] = t;
          nswap++;
        }
      }
      if (nswap == 0) break;
    }
  }

  double invMas[PO_MAXN], sum = 0.0;
  for (int n = 0; n < N; ++n) {
    sum += m[n];
    invMas[n] = rno[n] * ecm + sum;
  }

  double wt = maxw;
  double pd[PO_MAXN];
  for (int n = 0; n < N - 1; ++n) {
    pd[n] = pdk(invMas[n + 1], invMas[n], m[n + 1]);
    wt *= pd[n];
  }

  p[0][0] = sqrt(pd[0] * pd[0] + m[0] * m[0]);
  p[0][1] = 0.0;
  p[0][2] = pd[0];
  p[0][3] = 0.0;

  for (int i = 1; i < N; ++i) {
    p[i][0] = sqrt(pd[i - 1] * pd[i - 1] + m[i] * m[i]);
    p[i][1] = 0.0;
    p[i][2] = -pd[i - 1];
    p[i][3] = 0.0;

    double cZ = 2 * engine_uniform(&g) - 1;
    double sZ = sqrt(1 - cZ * cZ);
    double angY = 2 * PO_PI * engine_uniform(&g);
    double cY = cos(angY);
    double sY = sin(angY);
    for (int j = 0; j <= i; ++j) {
      double x = p[j][1];
      double y = p[j][2];
      p[j][1] = cZ * x - sZ * y;
      p[j][2] = sZ * x + cZ * y;
      x = p[j][1];
      double z = p[j][3];
      p[j][1] = cY * x - sY * z;
      p[j][3] = sY * x + cY * z;
    }
    if (i == N - 1) break;
    double beta = pd[i] / sqrt(pd[i] * pd[i] + invMas[i] * invMas[i]);
    for (int j = 0; j <= i; ++j) boost(p[j], 0, beta, 0);
  }

  if (final_mode == 0) {
    for (int n = 0; n < N; ++n) boost(p[n], fin[0], fin[1], fin[2]);
  } else {
    double e = fin[0];
    double bx = fin[1] / e, by = fin[2] / e, bz = fin[3] / e;
    for (int n = 0; n < N; ++n) boost(p[n], bx, by, bz);
  }
  return wt;
}

/* DecayMother / DecayMothers / AverageMother(s) seed their engine with the
 * PhaseSpace seed and discard evt + 3N (DecayMother.h:166-167).             */
static double decay_one(int N, int engine, const double* m, double ecm, double maxw,
                        int final_mode, const double* fin, uint64_t seed, uint64_t evt,
                        double (*p)[4]) {
  return decay_core(N, engine, m, ecm, maxw, final_mode, fin, seed, evt + 3 * (uint64_t)N, p);
}

void po_decay_events(int N, int engine, const double* masses, double ecm, double max_weight,
                     const double* beta3, uint64_t seed, const uint64_t* evts, uint64_t nevt,
                     double* out_p, double* out_wt) {
  for (uint64_t k = 0; k < nevt; ++k) {
    double p[PO_MAXN][4];
    double wt = decay_one(N, engine, masses, ecm, max_weight, 0, beta3, seed, evts[k], p);
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 4; ++c) out_p[(k * N + i) * 4 + c] = p[i][c];
    if (out_wt) out_wt[k] = wt;
  }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* PhaseSpace::Generate(mother, Decays) over events first..first+n-1
 * (PhaseSpace.inl:377-394, launch_decayers.inl:133-139); SoA output.       */
void po_generate(int N, int engine, const double* masses, double ecm, double max_weight,
                 const double* beta3, uint64_t seed, uint64_t first, uint64_t n, double* cols,
                 double* wt, int nthreads) {
  set_threads(nthreads);
  long long nn = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nn; ++k) {
    double p[PO_MAXN][4];
    double w = decay_one(N, engine, masses, ecm, max_weight, 0, beta3, seed,
                         first + (uint64_t)k, p);
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 4; ++c) cols[(uint64_t)(4 * i + c) * n + (uint64_t)k] = p[i][c];
    if (wt) wt[k] = w;
  }
}

/* PhaseSpace::Generate(mothers, daughters) (PhaseSpace.inl:396-415,
 * launch_decayers.inl:153-166): event k uses counter index first+k and
 * mother column row k.                                                      */
void po_generate_mothers(int N, int engine, const double* masses, double ecm,
                         double max_weight, uint64_t seed, uint64_t first, uint64_t n,
                         const double* mother_cols, double* cols, double* wt, int nthreads) {
  set_threads(nthreads);
  long long nn = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nn; ++k) {
    double p[PO_MAXN][4];
    double m4[4];
    for (int c = 0; c < 4; ++c) m4[c] = mother_cols[(uint64_t)c * n + (uint64_t)k];
    double w = decay_one(N, engine, masses, ecm, max_weight, 1, m4, seed,
                         first + (uint64_t)k, p);
    for (int i = 0; i < N; ++i)
      for (int c = 0; c < 4; ++c) cols[(uint64_t)(4 * i + c) * n + (uint64_t)k] = p[i][c];
    if (wt) wt[k] = w;
  }
}

/* Functors (arithmetic order fixed; see header).                            */
/* Vector4R product (Vector4R.h:171-176) */
static inline double dot4(const double* a, const double* b) {
  return a[0] * b[0] - a[1] * b[1] - a[2] * b[2] - a[3] * b[3];
}

/* Vector4R::dot: the three-momentum product (Vector4R.inl:441-454) */
static inline double dot3(const double* a, const double* b) {
  double t = a[1] * b[1];
  t += a[2] * b[2];
  t += a[3] * b[3];
  return t;
}

double po_functor(int fid, const double* fpar, int N, const double* p) {
  (void)N;
  if (fid == 0) return 1.0;
  if (fid == 2) {
    double s[4];
    for (int c = 0; c < 4; ++c) s[c] = p[c] + p[4 + c];
    return mass2(s);
  }
  if (fid == 3) {
    int a = (int)fpar[0], b = (int)fpar[1];
    double s[4];
    for (int c = 0; c < 4; ++c) s[c] = p[4 * a + c] + p[4 * b + c];
    return mass2(s);
  }
  if (fid == 6) {
    /* angle between the decay planes, phsp_chain.inl:191-207 with
     * Kstar = d + e and the three-momentum products of Vector4R::dot,
     * cross, d3mag and operator/ (Vector4R.inl:76-86, 372-386, 426-454)   */
    int d = (int)fpar[0], e = (int)fpar[1], h = (int)fpar[2];
    const double* D = p + 4 * d;
    const double* H = p + 4 * h;
    double K[4], dp[4], hp[4], dq[4];
    for (int c = 0; c < 4; ++c) K[c] = D[c] + p[4 * e + c];
    double kk = dot3(K, K);
    double cd = dot3(K, D) / kk;
    double ch = dot3(K, H) / kk;
    for (int c = 0; c < 4; ++c) {
      dp[c] = D[c] - K[c] * cd;
      hp[c] = H[c] - K[c] * ch;
    }
    dq[0] = 0.0;
    dq[1] = K[2] * dp[3] - K[3] * dp[2];
    dq[2] = K[3] * dp[1] - K[1] * dp[3];
    dq[3] = K[1] * dp[2] - K[2] * dp[1];
    double inv = 1.0 / sqrt(dp[1] * dp[1] + dp[2] * dp[2] + dp[3] * dp[3]);
    for (int c = 0; c < 4; ++c) dp[c] *= inv;
    inv = 1.0 / sqrt(dq[1] * dq[1] + dq[2] * dq[2] + dq[3] * dq[3]);
    for (int c = 0; c < 4; ++c) dq[c] *= inv;
    double x = dot3(dp, hp);
    double y = dot3(dq, hp);
    double chi = atan2(y, x);
    if (chi < 0.0) chi += 2.0 * PO_PI;
    return chi;
  }
  if (fid == 5) {
    /* (p_a + p_b).mass(): Vector4R::mass (Vector4R.inl:196-208) */
    int a = (int)fpar[0], b = (int)fpar[1];
    double s[4];
    for (int c = 0; c < 4; ++c) s[c] = p[4 * a + c] + p[4 * b + c];
    return mass(s);
  }
  if (fid == 4) {
    /* helicity cosine of particle d in the pair (d, e) within the sum of
     * particles first..first+count-1, as phsp_evaluating_functor.inl:160-176
     * writes it for B0 -> J/psi K pi (d = K, e = pi, the sum of all three) */
    int d = (int)fpar[0], e = (int)fpar[1], first = (int)fpar[2], count = (int)fpar[3];
    double P[4], Q[4];
    for (int c = 0; c < 4; ++c) P[c] = p[4 * first + c];
    for (int i = first + 1; i < first + count; ++i)
      for (int c = 0; c < 4; ++c) P[c] = P[c] + p[4 * i + c];
    for (int c = 0; c < 4; ++c) Q[c] = p[4 * d + c] + p[4 * e + c];
    const double* D = p + 4 * d;
    double pd = dot4(P, D);
    double pq = dot4(P, Q);
    double qd = dot4(Q, D);
    double mp2 = mass2(P);
    double mq2 = mass2(Q);
    double md2 = mass2(D);
    return (pd * mq2 - pq * qd) / sqrt((pq * pq - mq2 * mp2) * (qd * qd - mq2 * md2));
  }
  int K = (int)fpar[0];
  double Ar = 0.0, Ai = 0.0;
  for (int k = 0; k < K; ++k) {
    const double* q = fpar + 1 + 6 * k;
    int a = (int)q[4], b = (int)q[5];
    double v[4];
    for (int c = 0; c < 4; ++c) v[c] = p[4 * a + c] + p[4 * b + c];
    double s = mass2(v);
    double dr = q[0] * q[0] - s;
    double di = -(q[0] * q[1]);
    double den = dr * dr + di * di;
    double br = dr / den;
    double bi = -di / den;
    double tr = q[2] * br - q[3] * bi;
    double ti = q[2] * bi + q[3] * br;
    Ar += tr;
    Ai += ti;
  }
  return Ar * Ar + Ai * Ai;
}

/* PhaseSpaceWeight::Evaluate (Decays.inl:215-231); with fid >= 0 the
 * PhaseSpaceReweight product w * f (Decays.inl:362-379).                    */
static double event_weight_row(int N, const double* m, double maxw, const double* p) {
  double R[4] = {p[0], p[1], p[2], p[3]};
  double w = maxw;
  for (int i = 1; i < N; ++i) {
    double s[4];
    for (int c = 0; c < 4; ++c) s[c] = R[c] + p[4 * i + c];
    w *= pdk(mass(s), m[i], mass(R));
    for (int c = 0; c < 4; ++c) R[c] += p[4 * i + c];
  }
  return w;
}

static void gather_row(int N, const double* cols, uint64_t n, uint64_t k, double* p) {
  for (int i = 0; i < N; ++i)
    for (int c = 0; c < 4; ++c) p[4 * i + c] = cols[(uint64_t)(4 * i + c) * n + k];
}

void po_event_weight(int N, const double* masses, double max_weight, int fid,
                     const double* fpar, const double* cols, uint64_t n, double* w_out,
                     int nthreads) {
  set_threads(nthreads);
  long long nn = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nn; ++k) {
    double p[PO_MAXN * 4];
    gather_row(N, cols, n, (uint64_t)k, p);
    double w = event_weight_row(N, masses, max_weight, p);
    if (fid >= 0) w = w * po_functor(fid, fpar, N, p);
    w_out[k] = w;
  }
}

/* Decays::Unweight (Decays.inl:415-455) / Unweight(functor, max, seed)
 * (Decays.inl:457-511): flag event k iff weight/max_value > u_k with
 * u_k from squares3(seed).discard(k) (FlagDaugthers, Decays.inl:75-89),
 * then a stable partition (accepted first, both halves in original order).
 * max_value: fMaxWeight for the plain form; for the functor form the given
 * value if > 0, else the maximum of w*f over the container.               */
uint64_t po_unweight(int N, const double* masses, double max_weight, int fid,
                     const double* fpar, double max_value, const double* cols_in, uint64_t n,
                     uint64_t seed, double* cols_out, uint64_t* idx_out, uint8_t* flags_out,
                     int nthreads) {
  double* w = (double*)malloc(sizeof(double) * (n ? n : 1));
  uint8_t* f = flags_out ? flags_out : (uint8_t*)malloc(n ? n : 1);
  po_event_weight(N, masses, max_weight, fid, fpar, cols_in, n, w, nthreads);
  double mv = max_weight;
  if (fid >= 0) {
    if (max_value > 0.0) {
      mv = max_value;
    } else {
      mv = w[0];
      for (uint64_t k = 1; k < n; ++k)
        if (mv < w[k]) mv = w[k];
    }
  }
  uint64_t key = seed;
  key = po_splitmix64(&key);
  uint64_t nacc = 0;
  for (uint64_t k = 0; k < n; ++k) {
    double u = po_uniform(0, po_squares3(key, k));
    f[k] = (w[k] / mv > u) ? 1 : 0;
    nacc += f[k];
  }
  if (cols_out || idx_out) {
    uint64_t a = 0, r = nacc;
    for (uint64_t k = 0; k < n; ++k) {
      uint64_t dst = f[k] ? a++ : r++;
      if (idx_out) idx_out[dst] = k;
      if (cols_out)
        for (int c = 0; c < 4 * N; ++c)
          cols_out[(uint64_t)c * n + dst] = cols_in[(uint64_t)c * n + k];
    }
  }
  free(w);
  if (!flags_out) free(f);
  return nacc;
}

/* Order-sensitive digest of a permutation: the sum over output positions
 * pos of mix(idx[pos] << 32 | pos), mix the SplitMix64 output function (a
 * bijection of 64-bit words).  Positions and indices are below 2^32.        */
static inline uint64_t digest_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

uint64_t po_permutation_digest(const uint64_t* idx, uint64_t n) {
  uint64_t d = 0;
  long long nn = (long long)n;
#pragma omp parallel for schedule(static) reduction(+ : d)
  for (long long pos = 0; pos < nn; ++pos) d += digest_mix((idx[pos] << 32) | (uint64_t)pos);
  return d;
}

/* Generate (events first..first+n-1, single mother) followed by
 * Decays::Unweight(uw_seed) on the generated container (Decays.inl:415-455),
 * without storing the events: each event is generated, its PhaseSpaceWeight
 * recomputed from its momenta (Decays.inl:215-231), flagged against
 * squares3(uw_seed) word k (Decays.inl:75-89), and the stable partition's
 * permutation (accepted first, then rejected, both in order) is reduced to
 * po_permutation_digest of the output-row -> container-index map.  Returns
 * the accepted count; n must be below 2^32.  One byte of flags per event is
 * kept between the two passes.                                             */
uint64_t po_unweight_digest(int N, int engine, const double* masses, double ecm,
                            double max_weight, const double* beta3, uint64_t gen_seed,
                            uint64_t first, uint64_t n, uint64_t uw_seed, uint64_t* digest,
                            int nthreads) {
  *digest = 0;
  if (n == 0 || n >= (1ULL << 32)) return 0;
  set_threads(nthreads);
  uint8_t* f = (uint8_t*)malloc(n);
  uint64_t key = uw_seed;
  key = po_splitmix64(&key);
  long long nn = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nn; ++k) {
    double p[PO_MAXN][4];
    decay_one(N, engine, masses, ecm, max_weight, 0, beta3, gen_seed, first + (uint64_t)k, p);
    const double w = event_weight_row(N, masses, max_weight, &p[0][0]);
    const double u = po_uniform(0, po_squares3(key, (uint64_t)k));
    f[k] = (w / max_weight > u) ? 1 : 0;
  }
  /* accepted count per block, then each block's destinations */
  const long long B = 1024;
  const long long nb = (nn + B - 1) / B;
  uint64_t* acc_before = (uint64_t*)calloc((size_t)nb + 1, sizeof(uint64_t));
#pragma omp parallel for schedule(static)
  for (long long b = 0; b < nb; ++b) {
    uint64_t c = 0;
    for (long long k = b * B; k < nn && k < (b + 1) * B; ++k) c += f[k];
    acc_before[b + 1] = c;
  }
  for (long long b = 0; b < nb; ++b) acc_before[b + 1] += acc_before[b];
  const uint64_t nacc = acc_before[nb];
  uint64_t d = 0;
#pragma omp parallel for schedule(static) reduction(+ : d)
  for (long long b = 0; b < nb; ++b) {
    uint64_t a = acc_before[b];
    uint64_t r = nacc + (uint64_t)(b * B) - acc_before[b];
    for (long long k = b * B; k < nn && k < (b + 1) * B; ++k) {
      const uint64_t dst = f[k] ? a++ : r++;
      d += digest_mix(((uint64_t)k << 32) | dst);
    }
  }
  free(acc_before);
  free(f);
  *digest = d;
  return nacc;
}

/* AddStatsPHSP (StatsPHSP.h:62-86); stats are {W, mean, M2}.                */
void po_stats_merge(const double* x, const double* y, double* out) {
  double w = x[0] + y[0];
  double delta = y[1] - x[1];
  double delta2 = delta * delta;
  double mean = (x[1] * x[0] + y[1] * y[0]) / w;
  double m2 = x[2] + y[2];
  m2 += delta2 * x[0] * y[0] / w;
  out[0] = w;
  out[1] = mean;
  out[2] = m2;
}

/* transform_reduce on the OMP system (system/omp/detail/reduce.inl:40-63,
 * reduce_intervals.inl:66-80, internal/decompose.h:53-90): P intervals of a
 * uniform decomposition, each a left fold started at its first element,
 * then a left fold of {init, partials...}.  With intervals = P this repeats
 * the reference's order exactly when it ran with omp_get_num_procs() = P.   */
static void average_core(int N, int engine, const double* masses, double ecm, double maxw,
                         int final_mode, const double* fin, const double* mother_cols,
                         uint64_t seed, uint64_t first, uint64_t n, int fid,
                         const double* fpar, int intervals, double* out3) {
  long long P = intervals > 0 ? intervals : 1;
  long long nn = (long long)n;
  long long nint = nn, small = 1, large = 0, thr = 0;
  if (nint > P) {
    small = nn / P;
    large = small + 1;
    thr = nn % P;
    nint = P;
  }
  double* part = (double*)calloc((size_t)(nint + 1) * 3, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
  for (long long i = 0; i < nint; ++i) {
    long long b, e;
    if (i < thr) {
      b = large * i;
      e = b + large;
    } else {
      b = large * thr + small * (i - thr);
      e = (b + small < nn) ? b + small : nn;
    }
    double acc[3] = {0, 0, 0};
    for (long long k = b; k < e; ++k) {
      double p[PO_MAXN + 1][4];
      double wt, fv;
      uint64_t evt = first + (uint64_t)k;
      if (final_mode == 0) {
        wt = decay_one(N, engine, masses, ecm, maxw, 0, fin, seed, evt, p);
        fv = po_functor(fid, fpar, N, &p[0][0]);
      } else {
        double m4[4];
        for (int c = 0; c < 4; ++c) m4[c] = mother_cols[(uint64_t)c * n + (uint64_t)k];
        wt = decay_one(N, engine, masses, ecm, maxw, 1, m4, seed, evt, &p[1]);
        for (int c = 0; c < 4; ++c) p[0][c] = m4[c];
        fv = po_functor(fid, fpar, N + 1, &p[0][0]);
      }
      double s[3] = {wt, fv, 0.0};
      if (k == b) {
        acc[0] = s[0]; acc[1] = s[1]; acc[2] = s[2];
      } else {
        po_stats_merge(acc, s, acc);
      }
    }
    part[(i + 1) * 3 + 0] = acc[0];
    part[(i + 1) * 3 + 1] = acc[1];
    part[(i + 1) * 3 + 2] = acc[2];
  }
  double r[3] = {part[0], part[1], part[2]};
  for (long long i = 1; i <= nint; ++i) po_stats_merge(r, &part[i * 3], r);
  out3[0] = r[0];
  out3[1] = r[1];
  out3[2] = r[2];
  free(part);
}

/* PhaseSpace::AverageOn(policy, mother, functor, n) (PhaseSpace.inl:255-273,
 * AverageMother.h:272-292): per event {f, wt, 0}; returns {W, mean, M2};
 * the reference reports (mean, sqrt(M2)/W).                                 */
void po_average_on(int N, int engine, const double* masses, double ecm, double max_weight,
                   const double* beta3, uint64_t seed, uint64_t first, uint64_t n, int fid,
                   const double* fpar, int intervals, double* out3) {
  average_core(N, engine, masses, ecm, max_weight, 0, beta3, NULL, seed, first, n, fid, fpar,
               intervals, out3);
}

/* PhaseSpace::AverageOn(begin, end, functor) over mothers
 * (PhaseSpace.inl:275-287, AverageMothers.h:245-283): the functor sees the
 * mother as particle 0 followed by the N daughters.                         */
void po_average_on_mothers(int N, int engine, const double* masses, double ecm,
                           double max_weight, uint64_t seed, uint64_t n,
                           const double* mother_cols, int fid, const double* fpar,
                           int intervals, double* out3) {
  average_core(N, engine, masses, ecm, max_weight, 1, NULL, mother_cols, seed, 0, n, fid, fpar,
               intervals, out3);
}

/* ------------------------------------------------------------------------ */
/* Evaluate                                                                   */

/* PhaseSpace::Evaluate(mother, result, functors...) (PhaseSpace.inl:316-331;
 * EvalMother.h:283-306, launch_decayers.inl:66-72): event k (index first+k)
 * is DecayMother::process's event; out column 0 holds its weight, column
 * 1+j functor j of the N daughters.  fpar: functor j's parameters at
 * fpar + PO_FPAR * j.                                                        */
void po_evaluate(int N, int engine, const double* masses, double ecm, double max_weight,
                 const double* beta3, uint64_t seed, uint64_t first, uint64_t n, int nf,
                 const int* fids, const double* fpar, double* out, int nthreads) {
  set_threads(nthreads);
  long long nn = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nn; ++k) {
    double p[PO_MAXN][4];
    double w = decay_one(N, engine, masses, ecm, max_weight, 0, beta3, seed,
                         first + (uint64_t)k, p);
    out[(uint64_t)k] = w;
    for (int j = 0; j < nf; ++j)
      out[(uint64_t)(1 + j) * n + (uint64_t)k] = po_functor(fids[j], fpar + PO_FPAR * j, N, &p[0][0]);
  }
}

/* Szudzik pairing of (event index, seed) that keys EvalMothers' engines
 * (EvalMothers.h:131-143, applied at :147); the index arrives as an int
 * (process(const GInt_t evt, ...)), the seed as size_t.                    */
uint64_t po_pair_hash(uint64_t evt, uint64_t seed) {
  uint64_t a = (uint64_t)(int64_t)(int32_t)(uint32_t)evt;
  uint64_t A = 2 * a, B = 2 * seed;
  return (A >= B ? A * A + A + B : A + B * B) / 2;
}

/* PhaseSpace::Evaluate(mothers, result, functors...) (PhaseSpace.inl:333-348;
 * EvalMothers.h:145-273, launch_decayers.inl:75-89): event k decays mother
 * row k with an engine seeded by po_pair_hash(first+k, seed) and not
 * discarded; fECM is the mother's mass minus the daughters' and the weight is
 * normalised by that mother's maximum; the final boost is by the mother.
 * Functors see the mother as particle 0 followed by the N daughters.        */
void po_evaluate_mothers(int N, int engine, const double* masses, uint64_t seed,
                         uint64_t first, uint64_t n, const double* mother_cols, int nf,
                         const int* fids, const double* fpar, double* out, int nthreads) {
  set_threads(nthreads);
  long long nn = (long long)n;
#pragma omp parallel for schedule(static)
  for (long long k = 0; k < nn; ++k) {
    double m4[4];
    for (int c = 0; c < 4; ++c) m4[c] = mother_cols[(uint64_t)c * n + (uint64_t)k];
    double te = mass(m4);
    for (int i = 0; i < N; ++i) te -= masses[i];
    double emmax = te + masses[0], emmin = 0.0, wtmax = 1.0;
    for (int i = 1; i < N; ++i) {
      emmin += masses[i - 1];
      emmax += masses[i];
      wtmax *= pdk(emmax, emmin, masses[i]);
    }
    double p[PO_MAXN + 1][4];
    double w = decay_core(N, engine, masses, te, 1.0 / wtmax, 1, m4,
                          po_pair_hash(first + (uint64_t)k, seed), 0, &p[1]);
    for (int c = 0; c < 4; ++c) p[0][c] = m4[c];
    out[(uint64_t)k] = w;
    for (int j = 0; j < nf; ++j)
      out[(uint64_t)(1 + j) * n + (uint64_t)k] =
          po_functor(fids[j], fpar + PO_FPAR * j, N + 1, &p[0][0]);
  }
}

/* ------------------------------------------------------------------------ */
/* Dense histograms                                                           */

/* GetGlobalBin (GetGlobalBin.h:41-150, one dimension :154-209): per axis
 * X = (x - lower) * grid / (upper - lower); all axes below -> the underflow
 * slot nbins, all above grid -> the overflow slot nbins+1, otherwise the
 * row-major index of size_t(X).  Where the reference's conversion or index is
 * undefined (an axis at or below -1 or NaN; an index past nbins+1) the event
 * is dropped: returns UINT64_MAX.                                           */
uint64_t po_hist_key(int ndim, const uint64_t* grid, const double* lower, const double* upper,
                     const double* x) {
  uint64_t nbins = 1;
  for (int i = 0; i < ndim; ++i) nbins *= grid[i];
  double X[8];
  int under = 1, over = 1;
  for (int i = 0; i < ndim; ++i) {
    double delta = upper[i] - lower[i];
    X[i] = (x[i] - lower[i]) * (double)grid[i] / delta;
    under = under && (X[i] < 0.0);
    over = over && (X[i] > (double)grid[i]);
  }
  if (under) return nbins;
  if (over) return nbins + 1;
  uint64_t key = 0;
  for (int i = 0; i < ndim; ++i) {
    if (!(X[i] > -1.0) || X[i] >= 18446744073709551616.0) return UINT64_MAX;
    uint64_t idx = (uint64_t)X[i];  /* truncation; X in (-1, 0) gives 0 */
    uint64_t stride = 1;
    for (int j = ndim - 1; j > i; --j) stride *= grid[j];
    key += stride * idx;
  }
  return key <= nbins + 1 ? key : UINT64_MAX;
}

typedef struct {
  uint64_t key;
  uint64_t idx;
} po_key_idx;

static int cmp_key_idx(const void* a, const void* b) {
  const po_key_idx* x = (const po_key_idx*)a;
  const po_key_idx* y = (const po_key_idx*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* DenseHistogram::Fill (DenseHistogram.inl:49-105, 174-222): keys, a stable
 * sort by key, reduce_by_key of the weights (1.0 each without weights) in
 * that order, scattered into nbins + 2 zeroed contents.  x: ndim columns of
 * n values (column d at x + d*ldx).                                          */
void po_histogram_fill(int ndim, const uint64_t* grid, const double* lower, const double* upper,
                       const double* x, uint64_t ldx, uint64_t n, const double* w,
                       double* contents) {
  uint64_t nbins = 1;
  for (int i = 0; i < ndim; ++i) nbins *= grid[i];
  for (uint64_t b = 0; b < nbins + 2; ++b) contents[b] = 0.0;
  po_key_idx* ki = (po_key_idx*)malloc(sizeof(po_key_idx) * (n ? n : 1));
  uint64_t m = 0;
  for (uint64_t k = 0; k < n; ++k) {
    double v[8];
    for (int d = 0; d < ndim; ++d) v[d] = x[(uint64_t)d * ldx + k];
    uint64_t key = po_hist_key(ndim, grid, lower, upper, v);
    if (key == UINT64_MAX) continue;
    ki[m].key = key;
    ki[m].idx = k;
    ++m;
  }
  qsort(ki, m, sizeof(po_key_idx), cmp_key_idx);
  for (uint64_t a = 0; a < m;) {
    uint64_t b = a;
    double s = w ? w[ki[a].idx] : 1.0;
    for (b = a + 1; b < m && ki[b].key == ki[a].key; ++b) s += w ? w[ki[b].idx] : 1.0;
    contents[ki[a].key] = s;
    a = b;
  }
  free(ki);
}

/* hydra::unweight(policy, data, functor, max_pdf, rng_seed, rng_jump)
 * (Random.inl:134-177) with w_k = functor(data_k) precomputed: max = max_pdf
 * if > 0 else the largest weight (max_element); sample k is kept iff
 * u * max <= w_k with u the uniform of squares3(rng_seed) after
 * discard(rng_jump + k) (RndFlag, RandomUtils.h:254-262:
 * uniform_real_distribution(0, max) gives u * (max - 0) + 0).  The order:
 *   order 0: kept first, both parts in their original order;
 *   order 1: the reference's OMP build, whose thrust::partition runs the
 *            sequential swap algorithm: the kept part in order, each
 *            rejected sample before position nacc moved to the position of
 *            the kept sample it is swapped with (followed while that is
 *            below nacc), the others in place.                              */
uint64_t po_unweight_sample(int ncols, const double* cols_in, uint64_t n, const double* w,
                            double max_pdf, uint64_t seed, uint64_t jump, int order,
                            double* cols_out, uint64_t* idx_out, uint8_t* flags_out) {
  double mx = max_pdf;
  if (!(max_pdf > 0.0) && n > 0) {
    mx = w[0];
    for (uint64_t k = 1; k < n; ++k)
      if (mx < w[k]) mx = w[k];
  }
  uint64_t key = seed;
  key = po_splitmix64(&key);
  uint8_t* f = flags_out ? flags_out : (uint8_t*)malloc(n ? n : 1);
  uint64_t nacc = 0;
  for (uint64_t k = 0; k < n; ++k) {
    double u = po_uniform(0, po_squares3(key, jump + k));
    double urnd = u * (mx - 0.0) + 0.0;
    f[k] = (urnd <= w[k]) ? 1 : 0;
    nacc += f[k];
  }
  uint64_t* pos = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1)); /* output row of k */
  if (order == 0) {
    uint64_t a = 0, r = nacc;
    for (uint64_t k = 0; k < n; ++k) pos[k] = f[k] ? a++ : r++;
  } else {
    /* forward swap partition: first = trues seen; a true at next swaps with
     * position first.  Simulated on a position -> sample map.              */
    uint64_t* at = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint64_t k = 0; k < n; ++k) at[k] = k;
    uint64_t first = 0;
    while (first < n && f[first]) ++first;
    for (uint64_t next = first + 1; next < n; ++next) {
      if (f[next]) {
        uint64_t t = at[first];
        at[first] = at[next];
        at[next] = t;
        ++first;
      }
    }
    for (uint64_t p = 0; p < n; ++p) pos[at[p]] = p;
    free(at);
  }
  for (uint64_t k = 0; k < n; ++k) {
    if (idx_out) idx_out[pos[k]] = k;
    if (cols_out)
      for (int c = 0; c < ncols; ++c) cols_out[(uint64_t)c * n + pos[k]] = cols_in[(uint64_t)c * n + k];
  }
  free(pos);
  if (!flags_out) free(f);
  return nacc;
}
