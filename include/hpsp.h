/*
 * hpsp.h -- C ABI of the MI355X (gfx950) phase-space library libhpsp.so.
 *
 * These entry points sit where the reference hands its per-event functors to
 * Thrust (hydra/detail/launch_decayers.inl:66-182, hydra/detail/Decays.inl:
 * 415-511).  The C++ headers under hydra_amd/include/hydra/ (PhaseSpace.h,
 * Decays.h) call them; any other host language binds them with plain pointers
 * and integers.  No HIP or C++ types appear here: streams are passed as
 * opaque `void*` (a hipStream_t, NULL for the default stream).
 *
 * Data layout on the device (all buffers are owned by the caller):
 *   event columns  -- structure of arrays; for particle i (0..N-1) and
 *                     component c (0:E 1:px 2:py 3:pz) the value of event k is
 *                     cols[(4*i + c) * ld + k], ld >= n.  FP64 unless a call
 *                     says otherwise.
 *   weights        -- one double per event.
 *
 * Every entry point is stream-ordered and asynchronous unless it returns a
 * host value (documented per call), re-entrant across host threads, and
 * returns an hpsp_status.  Nothing is allocated by the library except where a
 * call says so.
 */
#ifndef HPSP_H
#define HPSP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPSP_MAX_N 12 /* daughters per decay: 2..HPSP_MAX_N */
#define HPSP_MAX_TERMS 4

typedef enum {
  HPSP_OK = 0,
  HPSP_ERR_INVALID_ARGUMENT = 1, /* null pointer, n/N/ld out of range, bad enum  */
  HPSP_ERR_MASS = 2,             /* mass <= 0: std::invalid_argument in the reference */
  HPSP_ERR_ENERGY = 3,           /* M - sum(m) <= 0: std::runtime_error in the reference */
  HPSP_ERR_WORKSPACE = 4,        /* workspace missing or too small */
  HPSP_ERR_HIP = 5,              /* a HIP runtime call failed */
  HPSP_ERR_NO_DEVICE = 6,
  HPSP_ERR_SHUTDOWN = 7          /* server is shutting down */
} hpsp_status;

/* Counter-based engines (hydra/detail/PRNGTypedefs.h:59 and the headers in hydra/detail/random/). */
typedef enum {
  HPSP_SQUARES3 = 0,      /* hydra::random::squares3 = hydra::default_random_engine */
  HPSP_SQUARES4 = 1,      /* hydra::random::squares4 */
  HPSP_PHILOX = 2,        /* Philox2x64, 7 rounds */
  HPSP_PHILOX_LONG = 3,   /* Philox4x64, 7 rounds */
  HPSP_THREEFRY = 4,      /* Threefry2x64, 12 rounds */
  HPSP_THREEFRY_LONG = 5  /* Threefry4x64, 12 rounds */
} hpsp_engine;

typedef enum {
  HPSP_FP64 = 0,  /* the reference's arithmetic, op for op */
  HPSP_MIXED = 1  /* FP32 kinematics; FP64 pdk, weight product and sums */
} hpsp_prec;

/* Decay description: what PhaseSpace<N,GRND>'s constructor derives
 * (hydra/detail/PhaseSpace.inl:42-90) plus the single-mother boost vector
 * that DecayMother's constructor derives (hydra/detail/functors/
 * DecayMother.h:76-101).  Fill it with hpsp_decay_init / hpsp_decay_set_mother. */
typedef struct {
  int32_t n;           /* number of daughters N */
  int32_t engine;      /* hpsp_engine */
  double mother_mass;
  double masses[HPSP_MAX_N];
  double ecm;          /* fECM = M - sum(m) */
  double max_weight;   /* fMaxWeight = 1 / wtmax */
  double beta[3];      /* final boost; zero for a mother at rest */
} hpsp_decay;

/* Per-event functors evaluated on the device (the reference compiles any
 * callable of a tuple of N Vector4R into its kernels, hydra/Lambda.h:133-200;
 * here a functor is a tagged description).
 *   HPSP_F_ONE      f = 1
 *   HPSP_F_BW_SUM   f = |sum_k c_k / (M_k^2 - s_k - i M_k G_k)|^2 with
 *                   s_k = (p[a_k] + p[b_k])^2 and c_k = cr_k + i ci_k
 *   HPSP_F_MASS2_01 f = (p0 + p1)^2
 *   HPSP_F_MASS2    f = (p[a0] + p[b0])^2
 *   HPSP_F_COS_HELICITY
 *                   helicity cosine of d = p[a0] in q = p[a0] + p[b0] within P,
 *                   the sum of the decay's daughters (left to right):
 *                   (P.d q^2 - P.q q.d) / sqrt((P.q^2 - q^2 P^2)(q.d^2 - q^2 d^2)),
 *                   the cos(theta_K*) of examples/phase_space/
 *                   phsp_evaluating_functor.inl:160-176
 * Particle indices count the four-vectors the functor sees: the N daughters,
 * or for the "mothers" entry points the mother (index 0) then the daughters. */
typedef enum {
  HPSP_F_ONE = 0,
  HPSP_F_BW_SUM = 1,
  HPSP_F_MASS2_01 = 2,
  HPSP_F_MASS2 = 3,
  HPSP_F_COS_HELICITY = 4
} hpsp_functor_kind;

#define HPSP_MAX_EVAL_FUNCTORS 8

typedef struct {
  int32_t kind;
  int32_t nterms;
  double m[HPSP_MAX_TERMS], g[HPSP_MAX_TERMS], cr[HPSP_MAX_TERMS], ci[HPSP_MAX_TERMS];
  int32_t a[HPSP_MAX_TERMS], b[HPSP_MAX_TERMS];
} hpsp_functor;

/* Streaming statistics of AverageOn: StatsPHSP (hydra/detail/functors/
 * StatsPHSP.h:38-59).  Partial results over disjoint event ranges merge with
 * hpsp_stats_merge, the reference's AddStatsPHSP (StatsPHSP.h:62-86).       */
typedef struct {
  double w;    /* sum of weights   (fW)    */
  double mean; /* weighted mean    (fMean) */
  double m2;   /* weighted M2      (fM2)   */
} hpsp_stats;

/* ---- set-up (host only, synchronous) ----------------------------------- */

/* PhaseSpace<N>(M, masses) (PhaseSpace.inl:42-90): validates, fills ecm and
 * max_weight, zero beta.  HPSP_ERR_MASS / HPSP_ERR_ENERGY mirror the
 * constructor's exceptions. */
int hpsp_decay_init(hpsp_decay* d, int n, double mother_mass, const double* masses,
                    int engine);

/* Boost vector for Generate(mother, ...) exactly as DecayMother's constructor
 * computes it (DecayMother.h:88-98, including its (E,px,py)/E components). */
int hpsp_decay_set_mother(hpsp_decay* d, const double mother4[4]);

/* AddStatsPHSP (StatsPHSP.h:62-86). */
void hpsp_stats_merge(const hpsp_stats* x, const hpsp_stats* y, hpsp_stats* out);

const char* hpsp_error_string(int status);

/* ---- engines ------------------------------------------------------------ */

/* Words of one engine stream: GRND eng(seed); eng.discard(discard);
 * out[i] = eng() for i < n (squares3.h:81-125, EngineR123.h:97-173).
 * squares3/4 words are written zero-extended to 64 bits.
 * Replaces: direct use of hydra::default_random_engine and friends. */
int hpsp_rng_words(int engine, uint64_t seed, uint64_t discard, uint64_t n, uint64_t* out,
                   void* stream);

/* ---- generation --------------------------------------------------------- */

/* Events first..first+n-1 of PhaseSpace<N,GRND>::Generate(mother, events)
 * (PhaseSpace.inl:377-394 -> launch_decayer, launch_decayers.inl:133-139 ->
 * DecayMother::operator(), DecayMother.h:279-292).  seed is the PhaseSpace's
 * fSeed.  prec selects the arithmetic; out_f32 != 0 stores FP32 columns
 * (float* cols) instead of FP64.  Optional outputs (NULL to skip):
 *   wt   generation weight of each event (DecayMother::process's return value)
 *   evw  PhaseSpaceWeight of the stored event (Decays.inl:215-231), bit-equal to
 *        hpsp_event_weight on the written FP64 columns.                      */
int hpsp_generate(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n, int prec,
                  void* cols, uint64_t ld, int out_f32, double* wt, double* evw,
                  void* stream);

/* Generate(mothers, daughters) (PhaseSpace.inl:396-415 -> DecayMothers,
 * DecayMothers.h:131-260): event k is boosted by mother k, read from the FP64
 * columns mothers[c * mld + k] (c = E, px, py, pz).  The decay's ecm is used
 * for every mother (PhaseSpace.inl:407).                                     */
int hpsp_generate_mothers(const hpsp_decay* d, uint64_t seed, uint64_t first, uint64_t n,
                          const double* mothers, uint64_t mld, int prec, void* cols,
                          uint64_t ld, int out_f32, double* wt, double* evw, void* stream);

/* ---- weights and unweighting ------------------------------------------- */

/* w_out[k] = PhaseSpaceWeight(event k) (Decays.inl:215-231), times f(event k)
 * when f is non-NULL (PhaseSpaceReweight, Decays.inl:362-379). */
int hpsp_event_weight(const hpsp_decay* d, const hpsp_functor* f, const double* cols,
                      uint64_t ld, uint64_t n, double* w_out, void* stream);

/* Maximum of w*f over the events (Decays.inl:488-490); written to *max_dev. */
int hpsp_max_weight(const hpsp_decay* d, const hpsp_functor* f, const double* cols,
                    uint64_t ld, uint64_t n, void* ws, size_t ws_bytes, double* max_dev,
                    void* stream);

/* Workspace bytes needed by hpsp_unweight and hpsp_max_weight for n events. */
size_t hpsp_unweight_workspace_size(uint64_t n);

/* Decays::Unweight(seed) (Decays.inl:415-455) when f is NULL, else
 * Decays::Unweight(f, max_value, seed) (Decays.inl:457-511).
 * Event k is accepted iff weight_k / max_value > u_k, u_k the k-th word of
 * squares3(seed) mapped to [0,1) (FlagDaugthers, Decays.inl:75-89); weight_k is
 * PhaseSpaceWeight (times f).  For f == NULL max_value is ignored and the
 * decay's max_weight is used, as the reference does; for f != NULL a
 * max_value <= 0 means "maximum over the events".
 * The stable partition (accepted first, both halves in original order) of
 * all 4N columns is written out of place to cols_out (ld_out).  w_in, when
 * non-NULL, supplies the PhaseSpaceWeight column (e.g. hpsp_generate's evw) so
 * it need not be recomputed from the momenta.  Optional outputs: idx_out (the
 * original index of each output row), w_out (w_in permuted alongside).
 * *n_acc_dev (device memory) receives the number accepted.                   */
int hpsp_unweight(const hpsp_decay* d, const hpsp_functor* f, double max_value, uint64_t seed,
                  const double* cols_in, uint64_t ld_in, uint64_t n, const double* w_in,
                  double* cols_out, uint64_t ld_out, uint64_t* idx_out, double* w_out,
                  void* ws, size_t ws_bytes, uint64_t* n_acc_dev, void* stream);

/* ---- integration -------------------------------------------------------- */

/* Workspace bytes needed by hpsp_average_on for n events. */
/* hydra::unweight(policy, data, functor[, max_pdf, rng_seed, rng_jump])
 * (hydra/detail/Random.inl:134-230, defaults in hydra/Random.h:133-330) with
 * the functor's values precomputed in the column w: keeps sample k iff
 * u_k * max <= w_k (RndFlag, RandomUtils.h:254-262) with u_k the uniform of
 * word jump + k of squares3(seed) and max = max_pdf if > 0, else the largest
 * weight.  Kept samples come first, in their original order.  The rejected
 * ones follow in their original order (HPSP_ORDER_STABLE) or in the order the
 * reference's thrust::partition leaves them on its OMP backend, the
 * sequential swap partition (HPSP_ORDER_SEQUENTIAL).  ncols value columns
 * move with their samples (cols_out may be NULL); idx_out / w_out optional.
 * The reference's defaults: seed 0x8ec74d321e6b5a27, jump 0.               */
typedef enum { HPSP_ORDER_STABLE = 0, HPSP_ORDER_SEQUENTIAL = 1 } hpsp_partition_order;

size_t hpsp_unweight_sample_workspace_size(uint64_t n);

int hpsp_unweight_sample(int ncols, const double* cols_in, uint64_t ld_in, uint64_t n,
                         const double* w, double max_pdf, uint64_t seed, uint64_t jump,
                         int order, double* cols_out, uint64_t ld_out, uint64_t* idx_out,
                         double* w_out, void* ws, size_t ws_bytes, uint64_t* n_acc_dev,
                         void* stream);

size_t hpsp_average_on_workspace_size(uint64_t n);

/* PhaseSpace<N,GRND>::AverageOn(policy, mother, f, n) over events
 * first..first+n-1 (PhaseSpace.inl:255-273, AverageMother.h:272-292,
 * launch_reducer, launch_decayers.inl:94-106): generation, functor and the
 * StatsPHSP reduction fused, nothing stored per event.  The result
 * (*out_dev, device memory) depends only on (d, f, seed, first, n, prec);
 * the reference reports (mean, sqrt(m2)/w).                                 */
int hpsp_average_on(const hpsp_decay* d, const hpsp_functor* f, uint64_t seed, uint64_t first,
                    uint64_t n, int prec, void* ws, size_t ws_bytes, hpsp_stats* out_dev,
                    void* stream);

/* AverageOn(mothers_begin, mothers_end, f) (PhaseSpace.inl:275-287,
 * AverageMothers.h:245-283): event k boosted by mother k; the functor sees
 * the mother as particle 0 followed by the daughters (so a_k, b_k index
 * 0..N).  Events are numbered 0..n-1 as in the reference.                  */
int hpsp_average_on_mothers(const hpsp_decay* d, const hpsp_functor* f, uint64_t seed,
                            const double* mothers, uint64_t mld, uint64_t n, int prec,
                            void* ws, size_t ws_bytes, hpsp_stats* out_dev, void* stream);

/* PhaseSpace<N>::Evaluate(mother, result, functors...) (PhaseSpace.inl:289-297,
 * 316-331; EvalMother.h:283-306): per event the generation weight and the
 * value of each functor, no momenta stored.  out holds 1 + nf columns of ld
 * doubles: column 0 the weight (DecayMother::process's return value), column
 * 1 + j functor j.  nf is 1..HPSP_MAX_EVAL_FUNCTORS.  Events first..first+n-1
 * (the reference's are 0..n-1).                                             */
int hpsp_evaluate(const hpsp_decay* d, const hpsp_functor* fs, int nf, uint64_t seed,
                  uint64_t first, uint64_t n, double* out, uint64_t ld, void* stream);

/* PhaseSpace<N>::Evaluate(mothers, result, functors...) (PhaseSpace.inl:333-348;
 * EvalMothers.h:145-273): event k decays the mother in row k of the SoA
 * mother columns (E, px, py, pz at mothers + c*mld).  Unlike Generate(mothers),
 * each event's engine is seeded with the pairing hash of (event index, seed)
 * and not discarded (EvalMothers.h:131-147), fECM is that mother's mass minus
 * the daughters' masses and the weight is normalised by that mother's maximum
 * weight.  Functors see the mother as particle 0 and the daughters as 1..N.   */
int hpsp_evaluate_mothers(const hpsp_decay* d, const hpsp_functor* fs, int nf, uint64_t seed,
                          const double* mothers, uint64_t mld, uint64_t first, uint64_t n,
                          double* out, uint64_t ld, void* stream);

/* ---- dense histograms ------------------------------------------------------
 * hydra::DenseHistogram<double, ndim> (hydra/DenseHistogram.h:54-636,
 * detail/DenseHistogram.inl) as the phase-space examples fill it downstream of
 * Generate (examples/phase_space/phsp_basic.inl:196-214).  contents holds
 * nbins + 2 doubles: bin (i_1, ..., i_ndim) at i_1*(g_2*...*g_ndim) + ... + i_ndim,
 * then the underflow (all axes below range) and overflow (all above) slots.
 * A fill overwrites the contents, as the reference's does.  Bin keys follow
 * GetGlobalBin (GetGlobalBin.h:41-209), including its edge cases: a value
 * exactly at an axis' upper limit has index grid (in one dimension that is the
 * underflow slot).  Events whose key the reference leaves undefined (an axis
 * at or below lower - delta/grid, NaN, a key past nbins + 1) are not counted.
 * Sums inside a bin are accumulated in a different order than the
 * reference's sorted reduce_by_key: equal to rounding; unit weights exactly. */
#define HPSP_HIST_MAX_DIM 4

typedef struct {
  int32_t ndim;                         /* 1..HPSP_HIST_MAX_DIM */
  uint64_t grid[HPSP_HIST_MAX_DIM];     /* bins per axis */
  double lower[HPSP_HIST_MAX_DIM];
  double upper[HPSP_HIST_MAX_DIM];
} hpsp_histogram;

typedef enum {
  HPSP_HIST_WEIGHT_ONE = 0,    /* every event counts 1 (Fill(data)) */
  HPSP_HIST_WEIGHT_COLUMN = 1, /* weights from a column (Fill(data, weights)) */
  HPSP_HIST_WEIGHT_PHSP = 2    /* the event's PhaseSpaceWeight (Decays.inl:215-231) */
} hpsp_hist_weight;

/* Number of bins (product of the grid); contents need two more doubles. */
uint64_t hpsp_histogram_bins(const hpsp_histogram* h);

/* DenseHistogram::Fill(data[, weights]): x holds ndim columns of n values
 * (column d at x + d*ldx); w is a weight column or NULL for unit weights.   */
int hpsp_histogram_fill(const hpsp_histogram* h, const double* x, uint64_t ldx, uint64_t n,
                        const double* w, double* contents, void* stream);

/* The same fill with the axis values computed from the events themselves:
 * axis d is functor axes[d] of the N daughters of the SoA event columns
 * (4N columns of ld doubles, as hpsp_generate writes them), weights per
 * weight_mode (w is the column for HPSP_HIST_WEIGHT_COLUMN).  This is the
 * reference's `events | dalitz_calculator` range histogrammed with
 * `events | GetEventWeightFunctor()`, without materialising either.        */
int hpsp_histogram_fill_events(const hpsp_histogram* h, const hpsp_decay* d,
                               const hpsp_functor* axes, const double* cols, uint64_t ld,
                               uint64_t n, int weight_mode, const double* w, double* contents,
                               void* stream);

/* ---- serving: continuous batching of generation and integration requests --
 * No counterpart in the reference (its calls are synchronous, one at a time).
 * A server owns an HBM arena and a worker thread.  Requests are admitted when
 * their footprint fits the arena, split into event chunks, and packed with
 * the chunks of all other admitted requests into batches launched over
 * several streams.  A request's result is bit-identical to the direct call
 * (hpsp_generate / hpsp_average_on with the same arguments), whatever else
 * runs beside it.  Generation outputs stay in the arena until released.     */
typedef struct hpsp_server hpsp_server;

typedef enum { HPSP_REQ_GENERATE = 0, HPSP_REQ_AVERAGE = 1 } hpsp_request_kind;

typedef struct {
  int32_t kind;         /* hpsp_request_kind */
  int32_t prec;         /* HPSP_FP64 or HPSP_MIXED */
  int32_t out_f32;      /* GENERATE: FP32 columns */
  int32_t want_weights; /* GENERATE: also the PhaseSpaceWeight column */
  hpsp_decay decay;
  hpsp_functor functor; /* AVERAGE */
  uint64_t seed, first, n;
  uint64_t user_tag;
} hpsp_request;

typedef struct {
  uint64_t ticket, user_tag;
  int32_t status; /* hpsp_status of the request */
  int32_t kind;
  hpsp_stats stats; /* AVERAGE */
  void* cols;       /* GENERATE: device columns (ld = n), valid until released */
  double* weights;  /* GENERATE: device PhaseSpaceWeight column or NULL */
  uint64_t n;
  double submit_s, admit_s, done_s; /* host monotonic clock, seconds */
} hpsp_completion;

typedef struct {
  uint64_t submitted, admitted, completed, rejected;
  uint64_t batches, launches;
  uint64_t arena_bytes, arena_in_use, arena_peak;
} hpsp_server_info;

/* hbm_bytes: arena size; n_streams: 1..16 concurrent streams;
 * chunk_events: events per chunk (rounded up to a multiple of 4096; 0 = default). */
int hpsp_server_create(hpsp_server** server, size_t hbm_bytes, int n_streams,
                       uint64_t chunk_events);
/* Non-blocking.  HPSP_ERR_WORKSPACE if the request can never fit the arena. */
int hpsp_server_submit(hpsp_server* server, const hpsp_request* req, uint64_t* ticket);
/* Up to max completions; waits up to wait_ms for the first one. */
int hpsp_server_poll(hpsp_server* server, hpsp_completion* out, int max, int* n_out,
                     int wait_ms);
/* Frees a completed GENERATE request's arena space. */
int hpsp_server_release(hpsp_server* server, uint64_t ticket);
int hpsp_server_info_get(hpsp_server* server, hpsp_server_info* info);
/* Waits for in-flight work, then frees everything. */
int hpsp_server_destroy(hpsp_server* server);

/* ---- device memory helpers (for callers without a HIP toolchain) -------- */
int hpsp_malloc(void** ptr, size_t bytes);
int hpsp_free(void* ptr);
int hpsp_memcpy_to_device(void* dst, const void* src, size_t bytes);
int hpsp_memcpy_to_host(void* dst, const void* src, size_t bytes);
int hpsp_memcpy_device(void* dst, const void* src, size_t bytes, void* stream);
int hpsp_stream_synchronize(void* stream);
/* Strided copies (rows of SoA columns): height rows of width bytes. */
int hpsp_memcpy2d_to_host(void* dst, size_t dpitch, const void* src, size_t spitch,
                          size_t width, size_t height);
int hpsp_memcpy2d_to_device(void* dst, size_t dpitch, const void* src, size_t spitch,
                            size_t width, size_t height);

#ifdef __cplusplus
}
#endif
#endif /* HPSP_H */
