// C ABI of libhpsp.so (declared in include/hpsp.h): argument checking,
// host-side folding of per-decay constants, workspace carving and kernel
// dispatch.  Host arithmetic here follows the reference's order
// (built with -ffp-contract=off).
#include <math.h>
#include <algorithm>
#include <mutex>
#include <string.h>

#include "hpsp.h"
#include "hpsp_host.h"
#include "hpsp_launch.h"

using namespace hpsp;

namespace {

constexpr size_t kAlign = 256;
size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// n beyond the per-call limit: no size or grid derived from it may wrap
bool too_many(uint64_t n) { return n > HPSP_MAX_EVENTS; }

struct UwLayout {
  size_t flags, tile_count, tile_off, run_sum, maxv, block_max, total;
};

UwLayout uw_layout(uint64_t n) {
  const uint64_t ngroups = (n + kWave - 1) / kWave + kUwTile / kWave;
  const uint64_t ntiles = (n + kUwTile - 1) / kUwTile + 1;
  const uint64_t nblocks = grid_for(n);
  UwLayout L;
  L.flags = 0;
  L.tile_count = L.flags + align_up(ngroups * sizeof(uint64_t));
  L.tile_off = L.tile_count + align_up(ntiles * sizeof(uint32_t));
  L.run_sum = L.tile_off + align_up(ntiles * sizeof(uint64_t));
  L.maxv = L.run_sum + align_up((ntiles / 1024 + 2) * sizeof(uint64_t));
  L.block_max = L.maxv + align_up(sizeof(double));
  L.total = L.block_max + align_up(nblocks * sizeof(double));
  return L;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Partials of hpsp_average_on_noalloc: one per kAvgChunk events of a call,
// HPSP_NOALLOC_MAX_EVENTS / kAvgChunk of them (1.5 MiB).  Each GPU holds its
// own copy at its own address; host code uses that address (scratch_on_device),
// never the array's name, which in host code is only a stand-in.
constexpr uint64_t kNoAllocPartials = HPSP_NOALLOC_MAX_EVENTS / kAvgChunk;
static __device__ __align__(256) Stats scratch[kNoAllocPartials];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device (with that device current) and kept.
int scratch_on_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return HPSP_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return HPSP_ERR_HIP;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HPSP_OK;
}

}  // namespace

extern "C" {

int hpsp_decay_init(hpsp_decay* d, int n, double mother_mass, const double* masses,
                    int engine) {
  if (!d || !masses || n < 2 || n > HPSP_MAX_N || !valid_engine(engine))
    return HPSP_ERR_INVALID_ARGUMENT;
  memset(d, 0, sizeof(*d));
  if (mother_mass <= 0.0) return HPSP_ERR_MASS;
  for (int i = 0; i < n; ++i)
    if (masses[i] <= 0.0) return HPSP_ERR_MASS;
  d->n = n;
  d->engine = engine;
  d->mother_mass = mother_mass;
  for (int i = 0; i < n; ++i) d->masses[i] = masses[i];
  double ecm = mother_mass;
  for (int i = 0; i < n; ++i) ecm -= masses[i];
  if (ecm <= 0.0) return HPSP_ERR_ENERGY;
  double emmax = ecm + masses[0];
  double emmin = 0.0;
  double wtmax = 1.0;
  for (int i = 1; i < n; ++i) {
    emmin += masses[i - 1];
    emmax += masses[i];
    double x = (emmax - emmin - masses[i]) * (emmax + emmin + masses[i]) *
               (emmax - emmin + masses[i]) * (emmax + emmin - masses[i]);
    x = sqrt(x) / (2 * emmax);
    wtmax *= x;
  }
  d->ecm = ecm;
  d->max_weight = 1.0 / wtmax;
  return HPSP_OK;
}

int hpsp_decay_set_mother(hpsp_decay* d, const double mother4[4]) {
  if (!d || !mother4) return HPSP_ERR_INVALID_ARGUMENT;
  const double e = mother4[0];
  const double d3 = sqrt(mother4[1] * mother4[1] + mother4[2] * mother4[2] + mother4[3] * mother4[3]);
  const double beta = d3 / e;
  if (beta != 0.0) {
    const double w = beta / d3;
    d->beta[0] = mother4[0] * w;
    d->beta[1] = mother4[1] * w;
    d->beta[2] = mother4[2] * w;
  } else {
    d->beta[0] = d->beta[1] = d->beta[2] = 0.0;
  }
  return HPSP_OK;
}

void hpsp_stats_merge(const hpsp_stats* x, const hpsp_stats* y, hpsp_stats* out) {
  const double w = x->w + y->w;
  const double delta = y->mean - x->mean;
  const double delta2 = delta * delta;
  hpsp_stats r;
  r.w = w;
  r.mean = (x->mean * x->w + y->mean * y->w) / w;
  r.m2 = x->m2 + y->m2;
  r.m2 += delta2 * x->w * y->w / w;
  *out = r;
}

const char* hpsp_error_string(int status) {
  switch (status) {
    case HPSP_OK: return "ok";
    case HPSP_ERR_INVALID_ARGUMENT: return "invalid argument";
    case HPSP_ERR_MASS: return "mass of mother or daughter particle is negative or zero";
    case HPSP_ERR_ENERGY: return "not enough energy to generate this decay (fECM <= 0)";
    case HPSP_ERR_WORKSPACE: return "workspace missing or too small";
    case HPSP_ERR_HIP: return "HIP runtime error";
    case HPSP_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

int hpsp_rng_words(int engine, uint64_t seed, uint64_t discard, uint64_t n, uint64_t* out,
                   void* stream) {
  if (!valid_engine(engine) || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n == 0) return HPSP_OK;
  if (!out) return HPSP_ERR_INVALID_ARGUMENT;
  uint64_t key[4];
  engine_keys(engine, seed, key);
  return hip_status(launch_rng_words(engine, key, discard, n, out, as_stream(stream)));
}

static int generate_common(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n,
                           const double* mothers, uint64_t mld, int prec, void* cols,
                           uint64_t ld, int out_f32, double* wt, double* evw, void* stream) {
  int rc = check_decay(d);
  if (rc) return rc;
  if (prec != HPSP_FP64 && prec != HPSP_MIXED) return HPSP_ERR_INVALID_ARGUMENT;
  if (too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n == 0) return HPSP_OK;
  if (!cols && !wt && !evw) return HPSP_ERR_INVALID_ARGUMENT;
  if (cols && ld < n) return HPSP_ERR_INVALID_ARGUMENT;
  if (mothers && mld < n) return HPSP_ERR_INVALID_ARGUMENT;
  GenArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, seed);
  a.first = first;
  a.n = n;
  a.cols = cols;
  a.ld = ld;
  a.out_f32 = out_f32 ? 1 : 0;
  a.wt = wt;
  a.evw = evw;
  a.mothers = mothers;
  a.mld = mld;
  const int e = (prec == HPSP_FP64) ? launch_generate_f64(d->n, d->engine, a, as_stream(stream))
                                    : launch_generate_mixed(d->n, d->engine, a, as_stream(stream));
  return hip_status(e);
}

int hpsp_generate(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n, int prec,
                  void* cols, uint64_t ld, int out_f32, double* wt, double* evw, void* stream) {
  return generate_common(d, seed, first, n, nullptr, 0, prec, cols, ld, out_f32, wt, evw, stream);
}

int hpsp_generate_mothers(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n,
                          const double* mothers, uint64_t mld, int prec, void* cols,
                          uint64_t ld, int out_f32, double* wt, double* evw, void* stream) {
  if (n > 0 && !mothers) return HPSP_ERR_INVALID_ARGUMENT;
  return generate_common(d, seed, first, n, mothers, mld, prec, cols, ld, out_f32, wt, evw,
                         stream);
}

namespace {

int evaluate_common(const hpsp_decay* d, const hpsp_functor* fs, int nf, uint64_t seed,
                    const double* mothers, uint64_t mld, uint64_t first, uint64_t n,
                    double* out, uint64_t ld, void* stream) {
  int rc = check_decay(d);
  if (rc) return rc;
  if (nf < 1 || nf > HPSP_MAX_EVAL_FUNCTORS || !fs || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n == 0) return HPSP_OK;
  if (!out || ld < n) return HPSP_ERR_INVALID_ARGUMENT;
  if (mothers && mld < n) return HPSP_ERR_INVALID_ARGUMENT;
  EvalArgs a;
  memset(&a, 0, sizeof(a));
  // Evaluate(mother): the PhaseSpace seed keys the engine as in Generate;
  // Evaluate(mothers): it is paired with each event index on the device
  a.d = make_decayk(d, seed);
  const int nvec = d->n + (mothers ? 1 : 0);
  for (int j = 0; j < nf; ++j) {
    rc = make_functork(&fs[j], nvec, &a.fs.f[j], mothers ? 1 : 0);
    if (rc) return rc;
  }
  a.fs.nf = nf;
  a.seed = seed;
  a.first = first;
  a.n = n;
  a.mothers = mothers;
  a.mld = mld;
  a.out = out;
  a.ld = ld;
  return hip_status(launch_evaluate_f64(d->n, d->engine, mothers != nullptr, a, as_stream(stream)));
}

}  // namespace

int hpsp_evaluate(const hpsp_decay* d, const hpsp_functor* fs, int nf, uint64_t seed,
                  uint64_t first, uint64_t n, double* out, uint64_t ld, void* stream) {
  return evaluate_common(d, fs, nf, seed, nullptr, 0, first, n, out, ld, stream);
}

int hpsp_evaluate_mothers(const hpsp_decay* d, const hpsp_functor* fs, int nf, uint64_t seed,
                          const double* mothers, uint64_t mld, uint64_t first, uint64_t n,
                          double* out, uint64_t ld, void* stream) {
  if (n > 0 && !mothers) return HPSP_ERR_INVALID_ARGUMENT;
  return evaluate_common(d, fs, nf, seed, mothers, mld, first, n, out, ld, stream);
}

namespace {

int make_histk(const hpsp_histogram* h, HistK* k) {
  if (!h || h->ndim < 1 || h->ndim > HPSP_HIST_MAX_DIM) return HPSP_ERR_INVALID_ARGUMENT;
  memset(k, 0, sizeof(*k));
  k->ndim = h->ndim;
  uint64_t nb = 1;
  for (int i = 0; i < h->ndim; ++i) {
    if (h->grid[i] == 0 || !(h->upper[i] > h->lower[i])) return HPSP_ERR_INVALID_ARGUMENT;
    if (nb > (1ULL << 40) / h->grid[i]) return HPSP_ERR_INVALID_ARGUMENT;
    nb *= h->grid[i];
    k->lower[i] = h->lower[i];
    k->delta[i] = h->upper[i] - h->lower[i];  // GetGlobalBin's fDelta
    k->gridf[i] = (double)h->grid[i];
  }
  for (int i = 0; i < h->ndim; ++i) {
    uint64_t st = 1;
    for (int j = h->ndim - 1; j > i; --j) st *= h->grid[j];
    k->stride[i] = st;
  }
  k->nbins = nb;
  return HPSP_OK;
}

// LDS available to one block, in doubles (the device's per-block limit, at
// most 128 KiB so that the fill kernel keeps room for its own use).
uint64_t hist_lds_slots() {
  static const uint64_t slots = [] {
    int dev = 0, bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess)
      bytes = 65536;
    return (uint64_t)std::min(bytes, 131072) / sizeof(double);
  }();
  return slots;
}

}  // namespace

uint64_t hpsp_histogram_bins(const hpsp_histogram* h) {
  HistK k;
  return make_histk(h, &k) == HPSP_OK ? k.nbins : 0;
}

int hpsp_histogram_fill(const hpsp_histogram* h, const double* x, uint64_t ldx, uint64_t n,
                        const double* w, double* contents, void* stream) {
  HistArgs a;
  memset(&a, 0, sizeof(a));
  int rc = make_histk(h, &a.h);
  if (rc) return rc;
  if (!contents || too_many(n) || (n > 0 && (!x || ldx < n))) return HPSP_ERR_INVALID_ARGUMENT;
  a.x = x;
  a.ldx = ldx;
  a.w = w;
  a.weight_mode = w ? HPSP_HIST_WEIGHT_COLUMN : HPSP_HIST_WEIGHT_ONE;
  a.n = n;
  a.contents = contents;
  a.lds_slots = hist_lds_slots();
  return hip_status(launch_histogram_fill(0, a, as_stream(stream)));
}

int hpsp_histogram_fill_events(const hpsp_histogram* h, const hpsp_decay* d,
                               const hpsp_functor* axes, const double* cols, uint64_t ld,
                               uint64_t n, int weight_mode, const double* w, double* contents,
                               void* stream) {
  HistArgs a;
  memset(&a, 0, sizeof(a));
  int rc = make_histk(h, &a.h);
  if (rc) return rc;
  rc = check_decay_masses(d);
  if (rc) return rc;
  if (!contents || !axes) return HPSP_ERR_INVALID_ARGUMENT;
  if (weight_mode < HPSP_HIST_WEIGHT_ONE || weight_mode > HPSP_HIST_WEIGHT_PHSP)
    return HPSP_ERR_INVALID_ARGUMENT;
  if (too_many(n) || (n > 0 && (!cols || ld < n))) return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0 && weight_mode == HPSP_HIST_WEIGHT_COLUMN && !w) return HPSP_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < h->ndim; ++i) {
    rc = make_functork(&axes[i], d->n, &a.axes[i]);
    if (rc) return rc;
  }
  a.d = make_decayk(d, 0);
  a.cols = cols;
  a.ld = ld;
  a.weight_mode = weight_mode;
  a.w = w;
  a.n = n;
  a.contents = contents;
  a.lds_slots = hist_lds_slots();
  return hip_status(launch_histogram_fill(d->n, a, as_stream(stream)));
}

// ---- stored events as particles ---------------------------------------------

namespace {

int make_particlesk(const hpsp_particles* p, uint64_t n, ParticlesK* k) {
  if (!p || p->n < 1 || p->n > HPSP_MAX_PARTICLES || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  memset(k, 0, sizeof(*k));
  k->n = p->n;
  for (int i = 0; i < p->n; ++i) {
    if (n > 0 && (!p->cols[i] || p->ld[i] < n)) return HPSP_ERR_INVALID_ARGUMENT;
    k->cols[i] = p->cols[i];
    k->ld[i] = p->ld[i];
  }
  return HPSP_OK;
}

}  // namespace

int hpsp_particles_evaluate(const hpsp_particles* p, const hpsp_functor* fs, int nf, uint64_t n,
                            double* out, uint64_t ld, void* stream) {
  PartEvalArgs a;
  memset(&a, 0, sizeof(a));
  int rc = make_particlesk(p, n, &a.p);
  if (rc) return rc;
  if (nf < 1 || nf > HPSP_MAX_EVAL_FUNCTORS || !fs) return HPSP_ERR_INVALID_ARGUMENT;
  for (int j = 0; j < nf; ++j) {
    rc = make_functork(&fs[j], p->n, &a.f[j], 0, true);
    if (rc) return rc;
  }
  if (n > 0 && (!out || ld < n)) return HPSP_ERR_INVALID_ARGUMENT;
  a.nf = nf;
  a.n = n;
  a.out = out;
  a.ld = ld;
  return hip_status(launch_particles_evaluate(a, as_stream(stream)));
}

int hpsp_histogram_fill_particles(const hpsp_histogram* h, const hpsp_particles* p,
                                  const hpsp_functor* axes, uint64_t n, const double* w,
                                  double* contents, void* stream) {
  HistArgs a;
  memset(&a, 0, sizeof(a));
  int rc = make_histk(h, &a.h);
  if (rc) return rc;
  rc = make_particlesk(p, n, &a.parts);
  if (rc) return rc;
  if (!contents || !axes) return HPSP_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < h->ndim; ++i) {
    rc = make_functork(&axes[i], p->n, &a.axes[i], 0, true);
    if (rc) return rc;
  }
  a.weight_mode = w ? HPSP_HIST_WEIGHT_COLUMN : HPSP_HIST_WEIGHT_ONE;
  a.w = w;
  a.n = n;
  a.contents = contents;
  a.lds_slots = hist_lds_slots();
  return hip_status(launch_histogram_fill(-1, a, as_stream(stream)));
}

// ---- phase_space_range ----------------------------------------------------

int hpsp_range_init(hpsp_decay* d, const double mother4[4], int n, const double* masses) {
  if (!d || !mother4 || !masses || n < 2 || n > HPSP_MAX_N) return HPSP_ERR_INVALID_ARGUMENT;
  // mother.mass() (Vector4R.inl:196-208)
  const double m2 = mother4[0] * mother4[0] - mother4[1] * mother4[1] -
                    mother4[2] * mother4[2] - mother4[3] * mother4[3];
  const double M = m2 > 0.0 ? sqrt(m2) : 0.0;
  // fTeCmTm and fWtMax follow the PhaseSpace constructor's arithmetic
  // (GenerateDecay.h:64-86)
  int rc = hpsp_decay_init(d, n, M, masses, HPSP_SQUARES3);
  if (rc) return rc;
  d->engine = HPSP_MINSTD_RAND;
  return hpsp_decay_set_mother(d, mother4);
}

namespace {

int range_args(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n, RangeArgs* a) {
  if (!d || d->engine != HPSP_MINSTD_RAND || d->n < 2 || d->n > HPSP_MAX_N || too_many(n))
    return HPSP_ERR_INVALID_ARGUMENT;
  const uint64_t skip = 3 * (uint64_t)d->n;
  if (first > UINT64_MAX - skip || n > UINT64_MAX - skip - first) return HPSP_ERR_INVALID_ARGUMENT;
  memset(a, 0, sizeof(*a));
  hpsp_decay dd = *d;
  dd.engine = HPSP_SQUARES3;  // no key words: the state below is the whole engine
  a->d = make_decayk(&dd, 0);
  // GRND eng((int)seed); eng.discard(first + 3N); then one step for word 0
  // (GenerateDecay.h:208-209)
  const uint32_t x0 = minstd_seed((uint32_t)seed);
  a->x_start = minstd_mul(minstd_mul(x0, minstd_pow(first + skip)), kMinstdA);
  a->n = n;
  return HPSP_OK;
}

}  // namespace

int hpsp_range_generate(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n,
                        double* cols, uint64_t ld, double* wt, void* stream) {
  RangeArgs a;
  int rc = range_args(d, seed, first, n, &a);
  if (rc) return rc;
  if (n == 0) return HPSP_OK;
  if (!cols || ld < n) return HPSP_ERR_INVALID_ARGUMENT;
  a.cols = cols;
  a.ld = ld;
  a.wt = wt;
  return hip_status(launch_range_generate(d->n, a, as_stream(stream)));
}

int hpsp_histogram_fill_range(const hpsp_histogram* h, const hpsp_decay* d, uint64_t seed,
                              const hpsp_functor* axes, uint64_t first, uint64_t n,
                              int weight_mode, double* contents, void* stream) {
  RangeArgs a;
  int rc = range_args(d, seed, first, n, &a);
  if (rc) return rc;
  rc = make_histk(h, &a.h);
  if (rc) return rc;
  if (!contents || !axes) return HPSP_ERR_INVALID_ARGUMENT;
  if (weight_mode != HPSP_HIST_WEIGHT_ONE && weight_mode != HPSP_HIST_WEIGHT_GENERATION &&
      weight_mode != HPSP_HIST_WEIGHT_PHSP)
    return HPSP_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < h->ndim; ++i) {
    rc = make_functork(&axes[i], d->n, &a.axes[i]);
    if (rc) return rc;
  }
  a.weight_mode = weight_mode;
  a.contents = contents;
  a.lds_slots = hist_lds_slots();
  return hip_status(launch_range_histogram(d->n, a, as_stream(stream)));
}

int hpsp_event_weight(const hpsp_decay* d, const hpsp_functor* f, const double* cols,
                      uint64_t ld, uint64_t n, double* w_out, void* stream) {
  int rc = check_decay_masses(d);
  if (rc) return rc;
  if (too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n == 0) return HPSP_OK;
  if (!cols || !w_out || ld < n) return HPSP_ERR_INVALID_ARGUMENT;
  WeightArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, 0);
  rc = make_functork(f, d->n, &a.f);
  if (rc) return rc;
  a.use_f = f ? 1 : 0;
  a.cols = cols;
  a.ld = ld;
  a.n = n;
  a.w_out = w_out;
  return hip_status(launch_event_weight(d->n, a, as_stream(stream)));
}

size_t hpsp_unweight_workspace_size(uint64_t n) {
  return too_many(n) ? SIZE_MAX : uw_layout(n).total;
}

int hpsp_max_weight(const hpsp_decay* d, const hpsp_functor* f, const double* cols,
                    uint64_t ld, uint64_t n, void* ws, size_t ws_bytes, double* max_dev,
                    void* stream) {
  int rc = check_decay_masses(d);
  if (rc) return rc;
  if (n == 0 || too_many(n) || !cols || !max_dev || ld < n) return HPSP_ERR_INVALID_ARGUMENT;
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < L.total) return HPSP_ERR_WORKSPACE;
  WeightArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, 0);
  rc = make_functork(f, d->n, &a.f);
  if (rc) return rc;
  a.use_f = f ? 1 : 0;
  a.cols = cols;
  a.ld = ld;
  a.n = n;
  char* base = static_cast<char*>(ws);
  return hip_status(launch_max_weight(d->n, a, reinterpret_cast<double*>(base + L.block_max),
                                      grid_for(n), max_dev, as_stream(stream)));
}

int hpsp_unweight(const hpsp_decay* d, const hpsp_functor* f, double max_value, uint64_t seed,
                  const double* cols_in, uint64_t ld_in, uint64_t n, const double* w_in,
                  double* cols_out, uint64_t ld_out, uint64_t* idx_out, double* w_out,
                  void* ws, size_t ws_bytes, uint64_t* n_acc_dev, void* stream) {
  int rc = check_decay_masses(d);
  if (rc) return rc;
  if (!n_acc_dev || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0) {
    if (!cols_in || ld_in < n) return HPSP_ERR_INVALID_ARGUMENT;
    if (cols_out && ld_out < n) return HPSP_ERR_INVALID_ARGUMENT;
    if (w_out && !w_in) return HPSP_ERR_INVALID_ARGUMENT;
  }
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < L.total) return HPSP_ERR_WORKSPACE;
  char* base = static_cast<char*>(ws);
  hipStream_t s = as_stream(stream);

  UnweightArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, 0);
  rc = make_functork(f, d->n, &a.f);
  if (rc) return rc;
  a.use_f = f ? 1 : 0;
  a.maxv = d->max_weight;
  a.maxv_dev = nullptr;
  if (f) {
    if (max_value > 0.0) {
      a.maxv = max_value;
    } else if (n > 0) {
      double* mv = reinterpret_cast<double*>(base + L.maxv);
      rc = hpsp_max_weight(d, f, cols_in, ld_in, n, ws, ws_bytes, mv, stream);
      if (rc) return rc;
      a.maxv_dev = mv;
    }
  }
  uint64_t s0 = seed;
  a.uw_key = splitmix64_host(&s0);
  a.cols_in = cols_in;
  a.ld_in = ld_in;
  a.n = n;
  a.w_in = w_in;
  a.cols_out = cols_out;
  a.ld_out = ld_out;
  a.idx_out = idx_out;
  a.w_out = w_out;
  a.flags = reinterpret_cast<uint64_t*>(base + L.flags);
  a.tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
  a.tile_off = reinterpret_cast<uint64_t*>(base + L.tile_off);
  a.run_sum = reinterpret_cast<uint64_t*>(base + L.run_sum);
  a.n_acc = n_acc_dev;
  return hip_status(launch_unweight(d->n, a, s));
}

int hpsp_unweight_values(int n_particles, const double* values, double max_value, uint64_t seed,
                         const double* cols_in, uint64_t ld_in, uint64_t n, double* cols_out,
                         uint64_t ld_out, uint64_t* idx_out, double* values_out, void* ws,
                         size_t ws_bytes, uint64_t* n_acc_dev, void* stream) {
  if (n_particles < 2 || n_particles > HPSP_MAX_N || !n_acc_dev || too_many(n))
    return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0) {
    if (!values) return HPSP_ERR_INVALID_ARGUMENT;
    if (cols_out && (!cols_in || ld_in < n || ld_out < n)) return HPSP_ERR_INVALID_ARGUMENT;
  }
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < L.total) return HPSP_ERR_WORKSPACE;
  char* base = static_cast<char*>(ws);
  hipStream_t s = as_stream(stream);
  UnweightArgs a;
  memset(&a, 0, sizeof(a));
  a.f.kind = HPSP_F_ONE;
  a.maxv = max_value;
  if (!(max_value > 0.0) && n > 0) {
    double* mv = reinterpret_cast<double*>(base + L.maxv);
    const int e = launch_max_column(values, n, reinterpret_cast<double*>(base + L.block_max),
                                    grid_for(n), mv, s);
    if (e) return HPSP_ERR_HIP;
    a.maxv_dev = mv;
  }
  uint64_t s0 = seed;
  a.uw_key = splitmix64_host(&s0);
  a.cols_in = cols_in;
  a.ld_in = ld_in;
  a.n = n;
  a.w_in = values;
  a.cols_out = cols_out;
  a.ld_out = ld_out;
  a.idx_out = idx_out;
  a.w_out = values_out;
  a.flags = reinterpret_cast<uint64_t*>(base + L.flags);
  a.tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
  a.tile_off = reinterpret_cast<uint64_t*>(base + L.tile_off);
  a.run_sum = reinterpret_cast<uint64_t*>(base + L.run_sum);
  a.n_acc = n_acc_dev;
  return hip_status(launch_unweight(n_particles, a, s));
}

int hpsp_unweight_regenerate(const hpsp_decay* d, uint64_t gen_seed, uint64_t first,
                             uint64_t n, const double* w, uint64_t seed, double* cols_out,
                             uint64_t ld_out, uint64_t* idx_out, double* w_out, void* ws,
                             size_t ws_bytes, uint64_t* n_acc_dev, void* stream) {
  int rc = check_decay(d);
  if (rc) return rc;
  if (!n_acc_dev || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0 && (!w || (cols_out && ld_out < n))) return HPSP_ERR_INVALID_ARGUMENT;
  if (first > UINT64_MAX - n) return HPSP_ERR_INVALID_ARGUMENT;
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < L.total) return HPSP_ERR_WORKSPACE;
  char* base = static_cast<char*>(ws);
  hipStream_t s = as_stream(stream);
  UnweightArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, gen_seed);  // masses, maxw and the generation keys
  a.f.kind = HPSP_F_ONE;
  a.maxv = d->max_weight;
  uint64_t s0 = seed;
  a.uw_key = splitmix64_host(&s0);
  a.n = n;
  a.w_in = w;
  a.ld_in = n;
  a.cols_out = cols_out;
  a.ld_out = ld_out;
  a.idx_out = idx_out;
  a.w_out = w_out;
  a.flags = reinterpret_cast<uint64_t*>(base + L.flags);
  a.tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
  a.tile_off = reinterpret_cast<uint64_t*>(base + L.tile_off);
  a.run_sum = reinterpret_cast<uint64_t*>(base + L.run_sum);
  a.n_acc = n_acc_dev;
  a.regen = 1;
  a.gen_first = first;
  rc = hip_status(launch_unweight(d->n, a, s));  // flags and offsets
  if (rc) return rc;
  if (!cols_out && !idx_out && !w_out) return HPSP_OK;
  return hip_status(launch_unweight_regen(d->n, d->engine, a, s));
}

// hpsp_unweight's workspace, a counter, a list of two words per group of 64
// events (the groups with events the weight chain leaves in doubt) and the
// 3-body bound's table
static size_t gu_list_offset(uint64_t n) { return uw_layout(n).total + align_up(sizeof(uint64_t)); }
static size_t gu_table_offset(uint64_t n) {
  const uint64_t ngroups = (n + kWave - 1) / kWave;
  return gu_list_offset(n) + align_up(2 * ngroups * sizeof(uint64_t));
}
size_t hpsp_generate_unweight_workspace_size(uint64_t n) {
  if (too_many(n)) return SIZE_MAX;
  return gu_table_offset(n) + align_up(2 * kBin3 * sizeof(double));
}

// The chain flags kernels' constants, formed once here instead of by every
// wave (FP64 divisions); scale2 bounds E^2 + p^2 of any subsystem; K of the
// squares engines (the Random123 ones set theirs per event).
static ChainK make_chaink(const hpsp_decay* d, double scale2) {
  const double K = 0x1p17 * d->n;
  ChainK ck;
  ck.rmax = 1.0 / d->max_weight;
  ck.rmax_ok = fabs(ck.rmax) >= 0x1p-1000 && fabs(ck.rmax) <= 0x1p+1000;
  ck.ks_unit = 0x1p-53 * scale2;
  ck.ks_fixed = (2.0 * K + 4.0) * ck.ks_unit;
  ck.ic2_0 = 1.0 / (d->masses[0] * d->masses[0]);
  ck.C = d->max_weight * d->max_weight * (ck.rmax * ck.rmax) * ldexp(1.0, -2 * (d->n - 1));
  // the squared test needs q >= 0: max and maxw positive
  ck.fast_ok = ck.rmax_ok && d->max_weight > 0.0 && ck.C >= 0x1p-900 && ck.C <= 0x1p+900;
  ck.A3 = 0.0;
  ck.C3 = 0.0;
  if (d->n == 3) {
    // invMas[2] as the chain forms it: 1 * ecm + ((0 + m0) + m1 + m2)
    double sum = 0.0;
    for (int i = 0; i < 3; ++i) sum += d->masses[i];
    ck.A3 = 1.0 * d->ecm + sum;
    ck.C3 = ck.C / (ck.A3 * ck.A3);
  }
  return ck;
}

int hpsp_generate_unweight(const hpsp_decay* d, uint64_t gen_seed, uint64_t first, uint64_t n,
                           uint64_t seed, double* cols_out, uint64_t ld_out, uint64_t* idx_out,
                           double* w_out, void* ws, size_t ws_bytes, uint64_t* n_acc_dev,
                           void* stream) {
  int rc = check_decay(d);
  if (rc) return rc;
  if (!n_acc_dev || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0 && cols_out && ld_out < n) return HPSP_ERR_INVALID_ARGUMENT;
  if (first > UINT64_MAX - n) return HPSP_ERR_INVALID_ARGUMENT;
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < hpsp_generate_unweight_workspace_size(n)) return HPSP_ERR_WORKSPACE;
  char* base = static_cast<char*>(ws);
  uint64_t* dl_count = reinterpret_cast<uint64_t*>(base + uw_layout(n).total);
  uint64_t* dl = reinterpret_cast<uint64_t*>(base + gu_list_offset(n));
  double* tab3 = reinterpret_cast<double*>(base + gu_table_offset(n));
  hipStream_t s = as_stream(stream);
  UnweightArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, gen_seed);
  a.f.kind = HPSP_F_ONE;
  a.maxv = d->max_weight;
  uint64_t s0 = seed;
  a.uw_key = splitmix64_host(&s0);
  a.n = n;
  a.ld_in = n;
  a.cols_out = cols_out;
  a.ld_out = ld_out;
  a.idx_out = idx_out;
  a.w_out = w_out;
  a.flags = reinterpret_cast<uint64_t*>(base + L.flags);
  a.tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
  a.tile_off = reinterpret_cast<uint64_t*>(base + L.tile_off);
  a.run_sum = reinterpret_cast<uint64_t*>(base + L.run_sum);
  a.n_acc = n_acc_dev;
  a.regen = 1;
  a.gen_first = first;
  if (n == 0) return hip_status(launch_unweight(d->n, a, s));  // the count: 0
  // E^2 + |p|^2 of any subsystem is at most (M gamma (1 + beta))^2, M = the
  // decay's fECM + sum of masses; the momenta's error constant 2K + 4
  double M = d->ecm;
  for (int i = 0; i < d->n; ++i) M += d->masses[i];
  double scale2 = M * M;
  if (a.d.do_boost) {
    const double b = sqrt(d->beta[0] * d->beta[0] + d->beta[1] * d->beta[1] +
                          d->beta[2] * d->beta[2]);
    const double f = a.d.gamma * (1.0 + b);
    scale2 *= f * f;
  }
  const ChainK ck = make_chaink(d, scale2);
  rc = hip_status(launch_flags_chain(d->n, d->engine, a, dl_count, dl, tab3, ck, s));
  if (rc) return rc;
  rc = hip_status(launch_uw_scans(a, s));
  if (rc) return rc;
  if (!cols_out && !idx_out && !w_out) return HPSP_OK;
  return hip_status(launch_unweight_regen(d->n, d->engine, a, s));
}

int hpsp_generate_unweight_mixed(const hpsp_decay* d, uint64_t gen_seed, uint64_t first,
                                 uint64_t n, uint64_t seed, float* cols_out, uint64_t ld_out,
                                 uint64_t* idx_out, double* w_out, void* ws, size_t ws_bytes,
                                 uint64_t* n_acc_dev, void* stream) {
  int rc = check_decay(d);
  if (rc) return rc;
  if (!n_acc_dev || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0 && cols_out && ld_out < n) return HPSP_ERR_INVALID_ARGUMENT;
  if (first > UINT64_MAX - n) return HPSP_ERR_INVALID_ARGUMENT;
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < hpsp_generate_unweight_workspace_size(n)) return HPSP_ERR_WORKSPACE;
  char* base = static_cast<char*>(ws);
  double* tab3 = reinterpret_cast<double*>(base + gu_table_offset(n));
  hipStream_t s = as_stream(stream);
  UnweightArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, gen_seed);
  a.f.kind = HPSP_F_ONE;
  a.maxv = d->max_weight;
  uint64_t s0 = seed;
  a.uw_key = splitmix64_host(&s0);
  a.n = n;
  a.ld_in = n;
  // FP32 columns, carried in the double pointer the kernels' arguments hold
  a.cols_out = reinterpret_cast<double*>(cols_out);
  a.ld_out = ld_out;
  a.idx_out = idx_out;
  a.w_out = w_out;
  a.flags = reinterpret_cast<uint64_t*>(base + L.flags);
  a.tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
  a.tile_off = reinterpret_cast<uint64_t*>(base + L.tile_off);
  a.run_sum = reinterpret_cast<uint64_t*>(base + L.run_sum);
  a.n_acc = n_acc_dev;
  a.regen = 1;
  a.gen_first = first;
  if (n == 0) return hip_status(launch_unweight(d->n, a, s));  // the count: 0
  // no momenta are involved in the mixed flags' bound: no scale
  rc = hip_status(launch_flags_mixed(d->n, d->engine, a, tab3, make_chaink(d, 0.0), s));
  if (rc) return rc;
  rc = hip_status(launch_uw_scans(a, s));
  if (rc) return rc;
  if (!cols_out && !idx_out && !w_out) return HPSP_OK;
  return hip_status(launch_unweight_regen_mixed(d->n, d->engine, a, s));
}

size_t hpsp_unweight_sample_workspace_size(uint64_t n) {
  if (too_many(n)) return SIZE_MAX;
  return uw_layout(n).total + align_up((n ? n : 1) * sizeof(uint64_t));
}

int hpsp_unweight_sample(int ncols, const double* cols_in, uint64_t ld_in, uint64_t n,
                         const double* w, double max_pdf, uint64_t seed, uint64_t jump,
                         int order, double* cols_out, uint64_t ld_out, uint64_t* idx_out,
                         double* w_out, void* ws, size_t ws_bytes, uint64_t* n_acc_dev,
                         void* stream) {
  if (ncols < 0 || !n_acc_dev || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (order != HPSP_ORDER_STABLE && order != HPSP_ORDER_SEQUENTIAL)
    return HPSP_ERR_INVALID_ARGUMENT;
  if (n > 0) {
    if (!w) return HPSP_ERR_INVALID_ARGUMENT;
    if (cols_out && (ncols == 0 || !cols_in || ld_in < n || ld_out < n))
      return HPSP_ERR_INVALID_ARGUMENT;
  }
  const UwLayout L = uw_layout(n);
  if (!ws || ws_bytes < hpsp_unweight_sample_workspace_size(n)) return HPSP_ERR_WORKSPACE;
  char* base = static_cast<char*>(ws);
  UnweightArgs a;
  memset(&a, 0, sizeof(a));
  a.rule = 1;
  a.order = order;
  a.tidx = reinterpret_cast<uint64_t*>(base + L.total);
  a.jump = jump;
  a.maxv = max_pdf;
  uint64_t s0 = seed;
  a.uw_key = splitmix64_host(&s0);
  a.cols_in = cols_in;
  a.ld_in = ld_in;
  a.n = n;
  a.w_in = w;
  a.cols_out = cols_out;
  a.ld_out = ld_out;
  a.idx_out = idx_out;
  a.w_out = w_out;
  a.flags = reinterpret_cast<uint64_t*>(base + L.flags);
  a.tile_count = reinterpret_cast<uint32_t*>(base + L.tile_count);
  a.tile_off = reinterpret_cast<uint64_t*>(base + L.tile_off);
  a.run_sum = reinterpret_cast<uint64_t*>(base + L.run_sum);
  a.n_acc = n_acc_dev;
  double* mv = max_pdf > 0.0 ? nullptr : reinterpret_cast<double*>(base + L.maxv);
  double* bm = reinterpret_cast<double*>(base + L.block_max);
  return hip_status(launch_unweight_sample(ncols, a, bm, grid_for(n), mv, as_stream(stream)));
}

size_t hpsp_average_on_workspace_size(uint64_t n) {
  if (too_many(n)) return SIZE_MAX;
  const uint64_t m = (n + kAvgChunk - 1) / kAvgChunk;
  return align_up((m ? m : 1) * sizeof(Stats));
}

static int average_common(const hpsp_decay* d, const hpsp_functor* f, uint64_t seed,
                          uint64_t first, uint64_t n, const double* mothers, uint64_t mld,
                          int prec, void* ws, size_t ws_bytes, hpsp_stats* out_dev,
                          void* stream) {
  int rc = check_decay(d);
  if (rc) return rc;
  if (prec != HPSP_FP64 && prec != HPSP_MIXED) return HPSP_ERR_INVALID_ARGUMENT;
  if (!out_dev || too_many(n)) return HPSP_ERR_INVALID_ARGUMENT;
  if (mothers && mld < n) return HPSP_ERR_INVALID_ARGUMENT;
  if (!ws || ws_bytes < hpsp_average_on_workspace_size(n)) return HPSP_ERR_WORKSPACE;
  AvgArgs a;
  memset(&a, 0, sizeof(a));
  a.d = make_decayk(d, seed);
  rc = make_functork(f, d->n + (mothers ? 1 : 0), &a.f, mothers ? 1 : 0);
  if (rc) return rc;
  a.first = first;
  a.n = n;
  a.mothers = mothers;
  a.mld = mld;
  a.partials = static_cast<Stats*>(ws);
  hipStream_t s = as_stream(stream);
  const uint64_t m = (n + kAvgChunk - 1) / kAvgChunk;
  if (n > 0) {
    const int e = (prec == HPSP_FP64) ? launch_average_f64(d->n, d->engine, a, s)
                                      : launch_average_mixed(d->n, d->engine, a, s);
    if (e) return HPSP_ERR_HIP;
  }
  return hip_status(launch_stats_finalize(a.partials, m, reinterpret_cast<Stats*>(out_dev), s));
}

int hpsp_average_on(const hpsp_decay* d, const hpsp_functor* f, uint64_t seed, uint64_t first,
                    uint64_t n, int prec, void* ws, size_t ws_bytes, hpsp_stats* out_dev,
                    void* stream) {
  return average_common(d, f, seed, first, n, nullptr, 0, prec, ws, ws_bytes, out_dev, stream);
}

int hpsp_stats_reduce(const hpsp_stats* partials, uint64_t m, hpsp_stats* out_dev, void* stream) {
  if (!out_dev || (m > 0 && !partials)) return HPSP_ERR_INVALID_ARGUMENT;
  return hip_status(launch_stats_finalize(reinterpret_cast<const Stats*>(partials), m,
                                          reinterpret_cast<Stats*>(out_dev), as_stream(stream)));
}

int hpsp_average_on_noalloc(const hpsp_decay* d, const hpsp_functor* f, uint64_t seed,
                            uint64_t first, uint64_t n, int prec, hpsp_stats* out_dev,
                            void* stream) {
  if (n > HPSP_NOALLOC_MAX_EVENTS) return HPSP_ERR_WORKSPACE;
  int rc = check_decay(d);
  if (rc) return rc;
  void* ws = nullptr;
  rc = scratch_on_device(&ws);
  if (rc) return rc;
  return average_common(d, f, seed, first, n, nullptr, 0, prec, ws, sizeof(scratch), out_dev,
                        stream);
}

int hpsp_average_on_mothers(const hpsp_decay* d, const hpsp_functor* f, uint64_t seed,
                            const double* mothers, uint64_t mld, uint64_t n, int prec,
                            void* ws, size_t ws_bytes, hpsp_stats* out_dev, void* stream) {
  if (n > 0 && !mothers) return HPSP_ERR_INVALID_ARGUMENT;
  return average_common(d, f, seed, 0, n, mothers, mld, prec, ws, ws_bytes, out_dev, stream);
}

int hpsp_checksum(const void* data, uint64_t bytes, uint64_t* sum_dev, void* stream) {
  if (!sum_dev || (bytes % 8) != 0 || too_many(bytes / 8) || (bytes > 0 && !data))
    return HPSP_ERR_INVALID_ARGUMENT;
  return hip_status(launch_checksum(static_cast<const uint64_t*>(data), bytes / 8, sum_dev,
                                    as_stream(stream)));
}

int hpsp_malloc(void** ptr, size_t bytes) {
  if (!ptr) return HPSP_ERR_INVALID_ARGUMENT;
  return hip_status((int)hipMalloc(ptr, bytes ? bytes : 1));
}

int hpsp_free(void* ptr) { return hip_status((int)hipFree(ptr)); }

int hpsp_memcpy_to_device(void* dst, const void* src, size_t bytes) {
  return hip_status((int)hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

int hpsp_memcpy_to_host(void* dst, const void* src, size_t bytes) {
  return hip_status((int)hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

int hpsp_memcpy_device(void* dst, const void* src, size_t bytes, void* stream) {
  return hip_status(
      (int)hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
}

int hpsp_stream_synchronize(void* stream) {
  return hip_status((int)hipStreamSynchronize(as_stream(stream)));
}

int hpsp_memcpy2d_to_host(void* dst, size_t dpitch, const void* src, size_t spitch,
                          size_t width, size_t height) {
  return hip_status(
      (int)hipMemcpy2D(dst, dpitch, src, spitch, width, height, hipMemcpyDeviceToHost));
}

int hpsp_memcpy2d_to_device(void* dst, size_t dpitch, const void* src, size_t spitch,
                            size_t width, size_t height) {
  return hip_status(
      (int)hipMemcpy2D(dst, dpitch, src, spitch, width, height, hipMemcpyHostToDevice));
}

}  // extern "C"
