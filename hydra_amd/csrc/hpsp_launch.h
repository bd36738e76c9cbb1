// Host-side launch interface between the C ABI (capi.hip) and the kernel
// translation units.  Internal to libhpsp.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hpsp_device.h"

namespace hpsp {

constexpr int kAvgChunk = 4096;  // events per AverageOn reduction chunk
constexpr int kUwTile = 1024;    // events per Unweight tile (256 threads x 4)
constexpr int kXcds = 8;         // MI355X: workgroups are dealt to 8 XCDs in turn
constexpr uint64_t kRunMax = 16;        // longest generation run, in groups of 64 events
constexpr uint64_t kRunMaxStores = 4;   // longest run of a launch bound by its column stores
constexpr uint64_t kRunMinRuns = 8192;  // 8 runs per SIMD of 256 CUs x 4 SIMDs

struct GenArgs {
  DecayK d;
  uint64_t first, n;
  void* cols;
  uint64_t ld;
  int out_f32;
  double* wt;
  double* evw;
  const double* mothers;  // FP64 mother columns (E,px,py,pz) or null
  uint64_t mld;
  uint32_t run;      // consecutive groups of 64 events per wave (>= 1; set by the launcher)
  uint32_t run_cap;  // 0: the launcher's policy; else the longest run allowed
};

// AverageOn arguments; FK is the functor: a FunctorK description here, a
// user's callable in a translation unit that compiles its own kernels
// (hydra_amd/include/hydra/detail/UserKernels.h).
template <typename FK>
struct AvgArgsT {
  DecayK d;
  FK f;
  uint64_t first, n;
  const double* mothers;
  uint64_t mld;
  Stats* partials;  // one per kAvgChunk events
};
using AvgArgs = AvgArgsT<FunctorK>;

// Up to HPSP_MAX_EVAL_FUNCTORS described functors, evaluated one by one.
struct DescSet {
  FunctorK f[HPSP_MAX_EVAL_FUNCTORS];
  int nf;
};

// Evaluate arguments; FS is the set of functors whose values are written.
template <typename FS>
struct EvalArgsT {
  DecayK d;
  FS fs;
  uint64_t seed;            // Evaluate(mothers): the seed paired with each event index
  uint64_t first, n;
  const double* mothers;    // Evaluate(mothers): SoA mother columns
  uint64_t mld;
  double* out;              // 1 + nf columns of ld doubles
  uint64_t ld;
  uint32_t run;             // consecutive groups of 64 events per wave (set by the launcher)
};
using EvalArgs = EvalArgsT<DescSet>;

// Four-vectors of stored events: particle i, component c, event k at
// cols[i][c * ld[i] + k].
struct ParticlesK {
  const double* cols[HPSP_MAX_PARTICLES];
  uint64_t ld[HPSP_MAX_PARTICLES];
  int n;
};

// Reads the particles a functor asks for from memory (wave-uniform particle
// index, lane-contiguous events): only the columns in use are loaded.
struct MemAcc {
  const ParticlesK& p;
  uint64_t k;
  __device__ __forceinline__ void get(int i, double (&v)[4]) const {
    const double* c = p.cols[i] + k;
    const uint64_t ld = p.ld[i];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = c[(uint64_t)q * ld];
  }
  __device__ __forceinline__ void sum(int first, int last, double (&Pv)[4]) const {
#pragma clang fp contract(off)
    get(first, Pv);
    for (int i = first + 1; i < last; ++i) {
      double v[4];
      get(i, v);
#pragma unroll
      for (int q = 0; q < 4; ++q) Pv[q] = Pv[q] + v[q];
    }
  }
};

struct PartEvalArgs {
  ParticlesK p;
  FunctorK f[HPSP_MAX_EVAL_FUNCTORS];
  int nf;
  uint64_t n;
  double* out;  // nf columns of ld doubles
  uint64_t ld;
};

// Dense histogram geometry with the reference's derived quantities:
// delta = upper - lower, gridf = (double)grid, row-major strides.
struct HistK {
  int ndim;
  double lower[HPSP_HIST_MAX_DIM], delta[HPSP_HIST_MAX_DIM], gridf[HPSP_HIST_MAX_DIM];
  uint64_t stride[HPSP_HIST_MAX_DIM];
  uint64_t nbins;
};

constexpr int kHistBlock = 1024;

struct HistArgs {
  HistK h;
  const double* x;          // columns source: ndim value columns
  uint64_t ldx;
  DecayK d;                 // events source: masses and fMaxWeight (PhaseSpaceWeight)
  FunctorK axes[HPSP_HIST_MAX_DIM];
  const double* cols;       // events source: 4N momentum columns
  uint64_t ld;
  ParticlesK parts;         // particles source
  int weight_mode;          // hpsp_hist_weight
  const double* w;
  uint64_t n;
  double* contents;         // nbins + 2
  int use_lds;              // block-private bins in LDS
  uint64_t lds_slots;       // LDS capacity in doubles (launcher)
};

// phase_space_range: the DecayMother events of a minstd_rand stream.  Word j
// of element first + k is x_start * 48271^(k + j) mod (2^31 - 1), x_start
// being the engine's state after discard(first + 3N) and one step.
struct RangeArgs {
  DecayK d;            // key words unused
  uint32_t x_start;
  uint32_t a_stride;   // 48271^(grid * kBlock): the state advance of one grid pass
  uint64_t n;
  double* cols;        // generate: 4N momentum columns of ld doubles
  uint64_t ld;
  double* wt;          // generate: weights or null
  HistK h;             // histogram fill
  FunctorK axes[HPSP_HIST_MAX_DIM];
  int weight_mode;
  double* contents;
  int use_lds;
  uint64_t lds_slots;
};

struct WeightArgs {
  DecayK d;
  FunctorK f;
  int use_f;
  const double* cols;
  uint64_t ld, n;
  double* w_out;
};

struct UnweightArgs {
  DecayK d;
  FunctorK f;
  int use_f;
  double maxv;          // divisor of the acceptance test
  const double* maxv_dev;  // if non-null, divisor read from device memory
  uint64_t uw_key;      // splitmix64(seed) of the default engine
  const double* cols_in;
  uint64_t ld_in, n;
  const double* w_in;
  double* cols_out;
  uint64_t ld_out;
  uint64_t* idx_out;
  double* w_out;
  uint64_t* flags;      // one bit per event, 64 per word
  uint32_t* tile_count; // accepted per tile
  uint64_t* tile_off;   // exclusive prefix of tile_count
  uint64_t* run_sum;    // per-1024-tile sums, then their exclusive prefix
  uint64_t* n_acc;      // total accepted (device)
  int rule;             // 0: w / max > u (Decays::Unweight); 1: u * max <= w (RndFlag)
  uint64_t jump;        // counter offset of the uniforms (rule 1)
  int order;            // 0: both parts stable; 1: sequential swap partition
  uint64_t* tidx;       // order 1: original index of the r-th kept sample
  int regen;            // events regenerated at their destinations, not moved
  uint64_t gen_first;   // regen: generation index of the container's row 0
};

// Kernel launchers; each returns a hipError_t value.
int launch_generate_f64(int N, int engine, const GenArgs& a, hipStream_t s);
int launch_generate_mixed(int N, int engine, const GenArgs& a, hipStream_t s);
int launch_average_f64(int N, int engine, const AvgArgs& a, hipStream_t s);
int launch_average_mixed(int N, int engine, const AvgArgs& a, hipStream_t s);
int launch_evaluate_f64(int N, int engine, bool mothers, const EvalArgs& a, hipStream_t s);
// N = 0: value columns; N >= 2: events of N daughters; N = -1: particles
int launch_histogram_fill(int N, const HistArgs& a, hipStream_t s);
int launch_particles_evaluate(const PartEvalArgs& a, hipStream_t s);
void launch_zero(double* p, uint64_t n, hipStream_t s);
int launch_range_generate(int N, const RangeArgs& a, hipStream_t s);
int launch_range_histogram(int N, const RangeArgs& a, hipStream_t s);
int launch_stats_finalize(const Stats* partials, uint64_t m, Stats* out, hipStream_t s);
int launch_event_weight(int N, const WeightArgs& a, hipStream_t s);
// Exclusive scan of the tile counts into a.tile_off, the total into *a.n_acc.
int launch_uw_scans(const UnweightArgs& a, hipStream_t s);
// Flags of Decays::Unweight for events a.gen_first.. of a Generate with decay
// a.d, decided from each event's generation weight where it is far enough
// from the boundary (generate_unweight.hip); dl_count and dl: a counter and a
// list of two words per group of 64 events (the groups with events left in
// doubt, decided by a second kernel); tab3: 2 kBin3 doubles, the 3-body
// bound's table (written and read by the call when N = 3); ck: the call's
// constants, formed on the host.
constexpr int kBin3Side = 8 + 28 * 8;  // bins per end of [0, 1]
constexpr int kBin3 = 2 * kBin3Side;
struct ChainK {
  double C;         // (maxw / max)^2 / 4^(N-1)
  double C3;        // C / A^2, A = invMas[2] (N = 3)
  double A3;        // that A
  double ks_unit;   // eps S^2, S^2 bounding E^2 + p^2 of any subsystem
  double ks_fixed;  // (2K + 4) eps S^2 with the squares engines' K = 2^17 N
  double ic2_0;     // 1 / m0^2
  int fast_ok;      // max and maxw positive, 1/max and C in range
  double rmax;      // RN(1 / max) and whether it is normal (ratio_above)
  int rmax_ok;
};
int launch_flags_chain(int N, int engine, const UnweightArgs& a, uint64_t* dl_count, uint64_t* dl,
                       double* tab3, const ChainK& ck, hipStream_t s);
// The mixed form: flags from each event's FP64 weight chain, the mixed
// Generate's weight column (generate_unweight.hip; tab3 and ck as for
// launch_flags_chain); then every event generated again at its place, FP32
// columns in a.cols_out, carried as a double pointer (genbod_mixed.hip).
int launch_flags_mixed(int N, int engine, const UnweightArgs& a, double* tab3, const ChainK& ck,
                       hipStream_t s);
int launch_unweight_regen_mixed(int N, int engine, const UnweightArgs& a, hipStream_t s);
int launch_max_weight(int N, const WeightArgs& a, double* block_max, uint64_t nblocks,
                      double* out, hipStream_t s);
int launch_unweight(int N, const UnweightArgs& a, hipStream_t s);
// order-independent checksum of n 64-bit words into *sum (device)
int launch_checksum(const uint64_t* w, uint64_t n, uint64_t* sum, hipStream_t s);
// largest of n values (max_element's value) into *out, through per-block maxima
int launch_max_column(const double* w, uint64_t n, double* block_max, uint64_t nblocks,
                      double* out, hipStream_t s);
// after launch_unweight with regen set: the stable partition of events
// gen_first .. gen_first + n - 1 of the generation stream, generated again
int launch_unweight_regen(int N, int engine, const UnweightArgs& a, hipStream_t s);
// hydra::unweight over ncols value columns and a weight column (rule 1).
int launch_unweight_sample(int ncols, const UnweightArgs& a, double* block_max,
                           uint64_t nblocks, double* max_dev, hipStream_t s);
int launch_rng_words(int engine, const uint64_t* key, uint64_t discard, uint64_t n,
                     uint64_t* out, hipStream_t s);

// Blocks for a grid-stride launch over `items` work items of kBlock each.
// Groups of 64 events per wave in a generation run (k_generate_*,
// k_evaluate_f64 over one mother): a longer run shares more engine words
// between consecutive groups, but the runs must still give every SIMD several
// waves; the longest power of two up to max_run that leaves kRunMinRuns runs.
// A launch bound by its column stores (momenta of at most 3 daughters, with or
// without the weight column) takes max_run = kRunMaxStores: longer runs spread
// the chip's concurrent writes over a wider window of the columns (3-body
// momenta alone at 2^27 events: 2.25 ms with runs of 4, 2.42 ms with 16;
// momenta and weights 2.55 ms against 2.63-2.69), where launches with more
// arithmetic per byte gain from 16 (4 to 8 daughters: 2-4%).
inline uint32_t run_groups(uint64_t ngroups, uint64_t max_run = kRunMax) {
  uint64_t r = 1;
  while (r < max_run && ngroups >= 2 * r * kRunMinRuns) r *= 2;
  return (uint32_t)r;
}

// Blocks of four waves for nruns runs, one run per wave (grid-stride beyond).
inline unsigned grid_for_runs(uint64_t nruns) {
  const uint64_t want = (nruns + kBlock / kWave - 1) / (kBlock / kWave);
  return (unsigned)(want == 0 ? 1 : (want < (1u << 20) ? want : (1u << 20)));
}

inline unsigned grid_for(uint64_t items, unsigned cap = 65536u) {
  const uint64_t b = (items + kBlock - 1) / kBlock;
  return (unsigned)(b == 0 ? 1 : (b < cap ? b : cap));
}

}  // namespace hpsp
