/*
 * CPU restatement of the Hydra 4.0.1 phase-space hot path, written from the
 * algorithm (see phsp_oracle.c for the per-function reference citations).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
 * load it.  The product path (hydra_amd/) never links or calls it.
 *
 * Conventions
 *   - engines: 0 squares3, 1 squares4, 2 philox (2x64, R=7),
 *              3 philox_long (4x64, R=7), 4 threefry (2x64, R=12),
 *              5 threefry_long (4x64, R=12),
 *              6 minstd_rand (thrust default_random_engine, phase_space_range)
 *   - SoA event columns: column (4*i + c) of particle i, component c
 *     (c = 0:E, 1:px, 2:py, 3:pz) starts at cols + (4*i + c) * n
 *   - AoS event rows: p[(k*N + i)*4 + c]
 *   - functors: id 0: f = 1
 *               id 1: f = |sum_k c_k / (M_k^2 - s_k - i M_k G_k)|^2,
 *                     s_k = (p_a + p_b)^2, par = {K, (M,G,cr,ci,a,b) x K}
 *               id 2: f = (p0 + p1)^2
 *               id 3: f = (p_a + p_b)^2, par = {a, b}
 *               id 4: helicity cosine of particle d in the pair (d, e) within
 *                     P = p_first + ... + p_(first+count-1),
 *                     par = {d, e, first, count}
 *               id 5: f = (p_a + p_b).mass(), par = {a, b}
 *               id 6: angle between the planes (d, e) and (d + e, h),
 *                     par = {d, e, h}
 *   - evaluate outputs: column 0 the weight, column 1+j functor j (ld = n);
 *     functor j's parameters at fpar + PO_FPAR*j
 */
#ifndef PHSP_ORACLE_H
#define PHSP_ORACLE_H
#include <stdint.h>

#define PO_MAXN 16
#define PO_FPAR 32

#ifdef __cplusplus
extern "C" {
#endif

uint64_t po_splitmix64(uint64_t* x);
uint32_t po_squares3(uint64_t key, uint64_t ctr);
uint32_t po_squares4(uint64_t key, uint64_t ctr);
void po_philox2x64(unsigned R, const uint64_t ctr[2], const uint64_t key[1], uint64_t out[2]);
void po_philox4x64(unsigned R, const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
void po_threefry2x64(unsigned R, const uint64_t ctr[2], const uint64_t key[2], uint64_t out[2]);
void po_threefry4x64(unsigned R, const uint64_t ctr[4], const uint64_t key[4], uint64_t out[4]);

int po_rng_words(int engine, uint64_t seed, uint64_t discard, uint64_t n, uint64_t* out);
double po_uniform(int engine, uint64_t word);

int po_params(int N, double M, const double* masses, double* ecm, double* max_weight);
void po_mother_beta(const double* mother4, double* beta3);
int po_range_params(int N, const double* mother4, const double* masses, double* ecm,
                    double* max_weight, double* beta3);

void po_decay_events(int N, int engine, const double* masses, double ecm, double max_weight,
                     const double* beta3, uint64_t seed, const uint64_t* evts, uint64_t nevt,
                     double* out_p, double* out_wt);
void po_generate(int N, int engine, const double* masses, double ecm, double max_weight,
                 const double* beta3, uint64_t seed, uint64_t first, uint64_t n, double* cols,
                 double* wt, int nthreads);
void po_generate_mothers(int N, int engine, const double* masses, double ecm,
                         double max_weight, uint64_t seed, uint64_t first, uint64_t n,
                         const double* mother_cols, double* cols, double* wt, int nthreads);

double po_functor(int fid, const double* fpar, int N, const double* p);
void po_event_weight(int N, const double* masses, double max_weight, int fid,
                     const double* fpar, const double* cols, uint64_t n, double* w_out,
                     int nthreads);
uint64_t po_unweight(int N, const double* masses, double max_weight, int fid,
                     const double* fpar, double max_value, const double* cols_in, uint64_t n,
                     uint64_t seed, double* cols_out, uint64_t* idx_out, uint8_t* flags_out,
                     int nthreads);
void po_average_on(int N, int engine, const double* masses, double ecm, double max_weight,
                   const double* beta3, uint64_t seed, uint64_t first, uint64_t n, int fid,
                   const double* fpar, int intervals, double* out3);
void po_average_on_mothers(int N, int engine, const double* masses, double ecm,
                           double max_weight, uint64_t seed, uint64_t n,
                           const double* mother_cols, int fid, const double* fpar,
                           int intervals, double* out3);
void po_stats_merge(const double* x3, const double* y3, double* out3);
uint64_t po_permutation_digest(const uint64_t* idx, uint64_t n);
uint64_t po_unweight_digest(int N, int engine, const double* masses, double ecm,
                            double max_weight, const double* beta3, uint64_t gen_seed,
                            uint64_t first, uint64_t n, uint64_t uw_seed, uint64_t* digest,
                            int nthreads);

void po_evaluate(int N, int engine, const double* masses, double ecm, double max_weight,
                 const double* beta3, uint64_t seed, uint64_t first, uint64_t n, int nf,
                 const int* fids, const double* fpar, double* out, int nthreads);
uint64_t po_pair_hash(uint64_t evt, uint64_t seed);
uint64_t po_unweight_sample(int ncols, const double* cols_in, uint64_t n, const double* w,
                            double max_pdf, uint64_t seed, uint64_t jump, int order,
                            double* cols_out, uint64_t* idx_out, uint8_t* flags_out);
uint64_t po_hist_key(int ndim, const uint64_t* grid, const double* lower, const double* upper,
                     const double* x);
void po_histogram_fill(int ndim, const uint64_t* grid, const double* lower, const double* upper,
                       const double* x, uint64_t ldx, uint64_t n, const double* w,
                       double* contents);
void po_evaluate_mothers(int N, int engine, const double* masses, uint64_t seed,
                         uint64_t first, uint64_t n, const double* mother_cols, int nf,
                         const int* fids, const double* fpar, double* out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
