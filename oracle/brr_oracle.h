/*
 * brr_oracle.h -- CPU ORACLE (test infrastructure only).
 *
 * A plain-C, FP64, dependency-free restatement of the reference BayesRRcmd
 * per-marker Gibbs sweep. It exists to CHECK the HIP product path: only the
 * test suite, the smoke check and the CPU-baseline leg of the benchmark may
 * load it. Product code never links or calls it.
 *
 * Every function cites the reference file:line it restates. Randomness is
 * injected: the reference's Boost mt19937 / glibc rand() streams are replaced
 * by a counter-based Philox4x32-10 stream whose exact definition is given in
 * oracle/rng.c (the product implements the same stream independently).
 *
 * The pinning status of each part is listed in README.md next to this file.
 */
#ifndef BRR_ORACLE_H
#define BRR_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- RNG ---- */

/* Philox4x32-10 (Salmon et al., SC'11). out[4] <- bijection(ctr[4], key[2]). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Stream tags (ctr word 2, low 8 bits). */
enum { ORC_STREAM_MARKER = 1, ORC_STREAM_SHUFFLE = 2, ORC_STREAM_SCALAR = 3, ORC_STREAM_FIXED = 5 };

/* Deterministic elementary functions built from + - * / only (bit-identical on
 * any IEEE-754 binary64 platform without contraction). */
double orc_det_log(double x);
double orc_det_exp(double x);
/* Inverse standard normal CDF, Wichura AS241 (PPND16). */
double orc_inv_norm_cdf(double p);

/* Per-marker (u, z) pair for (seed, iteration, marker id, chain). */
void orc_marker_uz(uint64_t seed, uint32_t iteration, uint32_t marker, uint32_t chain,
                   double* u, double* z);
void orc_fill_marker_uz(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t M,
                        double* u, double* z);

/* In-place Fisher-Yates in the rule of libstdc++ std::random_shuffle
 * (bits/stl_algo.h: j = r % (i + 1) for i = 1..M-1) with r drawn from the
 * Philox shuffle stream. */
void orc_shuffle_philox(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t M,
                        uint32_t* order);
/* The same rule with r from the given stream tag (ORC_STREAM_FIXED: the
 * covariate order of the fixed effects). */
void orc_shuffle_philox_stream(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t stream,
                               uint32_t M, uint32_t* order);
/* Same rule with r = rand() from the C library (the reference's stock source;
 * the caller seeds it with srand). */
void orc_shuffle_libc(uint32_t M, uint32_t* order);

/* Scalar draws (distributions_boost.cpp:22-65 parameterisations). draw_id
 * selects an independent substream. */
double orc_uniform01(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain);
double orc_norm_rng(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain,
                    double mean, double var);
double orc_rgamma(uint64_t seed, uint32_t iteration, uint32_t draw_id, uint32_t chain,
                  double shape, double scale);
double orc_inv_scaled_chisq_rng(uint64_t seed, uint32_t iteration, uint32_t draw_id,
                                uint32_t chain, double dof, double scale);
/* Dirichlet over K components; draw ids first_id .. first_id+K-1. */
void orc_dirichlet_rng(uint64_t seed, uint32_t iteration, uint32_t first_id, uint32_t chain,
                       int K, const double* alpha, double* out);

/* ------------------------------------------------------ markers (a1-a4) -- */

enum { ORC_DENSE = 1, ORC_RAGGED = 2, ORC_EIGEN = 3 };

/* a1: decode one SNP-major PLINK .bed column (ceil(N/4) bytes) into
 * (allele1, allele2) pairs, preprocessgraph.cpp:34-57. a1[i], a2[i] in {0,1}. */
void orc_decode_bed_column(const uint8_t* col, uint32_t N, uint8_t* a1, uint8_t* a2);
/* Returns 0 if the three magic bytes are 0x6c 0x1b 0x01 (preprocessgraph.cpp:195-200). */
int orc_check_bed_magic(const uint8_t* header3);

/* a3: DenseMarkerBuilder (densemarkerbuilder.cpp:23-60): Cx[N]. */
void orc_dense_build(const uint8_t* col, uint32_t N, double* Cx);
/* Dense column as a 4-entry table indexed by dosage 0,1,2 and 3 = missing; the
 * column value for individual i is lut[dosage_i]. Values equal orc_dense_build
 * element-for-element. */
void orc_dense_lut(const uint8_t* col, uint32_t N, double lut[4]);

/* a2: RaggedSparseMarkerBuilder + SparseMarker::updateStatistics
 * (raggedsparsemarkerbuilder.cpp:25-55, sparsemarker.cpp:8-19). Fills the
 * three ascending index lists (caller-sized to N) and returns counts.
 * Returns 1 if valid (raggedsparsemarker.cpp:68-77), 0 otherwise. */
int orc_ragged_build(const uint8_t* col, uint32_t N,
                     int32_t* ones, uint64_t* n_ones,
                     int32_t* twos, uint64_t* n_twos,
                     int32_t* miss, uint64_t* n_miss,
                     double stats[4] /* mean, sd, sqrdZ, Zsum */);

/* a4: serialised RaggedSparseMarker (raggedsparsemarker.cpp:43-66,
 * sparsemarker.cpp:43-58): 4 x f64 then, for ones/twos/missing, a u64 size
 * followed by int32[size]. Returns the byte count written (buf may be NULL to
 * query the size). */
size_t orc_ragged_serialise(const double stats[4],
                            const int32_t* ones, uint64_t n_ones,
                            const int32_t* twos, uint64_t n_twos,
                            const int32_t* miss, uint64_t n_miss, uint8_t* buf);

/* EigenSparseMarkerBuilder (eigensparsemarkerbuilder.cpp:22-78): the ragged
 * builder's lists, statistics with the missing calls imputed at the mean
 * (sqrdZ += mean^2 n_miss, Zsum += mean n_miss, then sd). Returns 1 if some
 * call is nonzero (eigensparsemarker.cpp:60-71), 0 otherwise. */
int orc_eigen_build(const uint8_t* col, uint32_t N,
                    int32_t* ones, uint64_t* n_ones,
                    int32_t* twos, uint64_t* n_twos,
                    int32_t* miss, uint64_t* n_miss,
                    double stats[4]);

/* Serialised EigenSparseMarker (eigensparsemarker.cpp:41-58): 4 x f64, an
 * int64 count, count f64 values (1, 2 or the mean for a missing call), count
 * int32 indices, in index order. buf may be NULL to query the size. */
size_t orc_eigen_serialise(const double stats[4],
                           const int32_t* ones, uint64_t n_ones,
                           const int32_t* twos, uint64_t n_twos,
                           const int32_t* miss, uint64_t n_miss, uint8_t* buf);

/* ---------------------------------------------------------- BayesR (a5-a10) */

typedef struct {
    uint32_t N, M;
    int layout;                 /* ORC_DENSE, ORC_RAGGED or ORC_EIGEN (the ragged
                                   lists with EigenSparse statistics) */
    /* dense: packed SNP-major columns (no magic), M x ceil(N/4) bytes, and
     * per-marker 4-entry tables (orc_dense_lut) */
    const uint8_t* bed;
    const double* lut;          /* M x 4 */
    /* ragged: CSR-style lists; offsets have M+1 entries */
    const int32_t* ones; const uint64_t* ones_off;
    const int32_t* twos; const uint64_t* twos_off;
    const int32_t* miss; const uint64_t* miss_off;
    const double* stats;        /* M x 4 {mean, sd, sqrdZ, Zsum} */
} orc_data;

typedef struct {
    int K;                      /* components including the zero one */
    int G;                      /* groups */
    const double* S;            /* G x (K-1) */
    const int32_t* group;       /* M labels; NULL = all group 0 */
    double v0E, s02E, v0G, s02G;
} orc_model;

typedef struct {
    /* chain state */
    double* eps;                /* N */
    double* beta;               /* M */
    int32_t* comp;              /* M */
    double* acum;               /* M */
    double mu, sigmaE, epsSum;
    double* sigmaG;             /* G */
    double* pi;                 /* G x K */
    double* v;                  /* G x K */
    double* betasqn;            /* G */
} orc_state;

/* Optional per-step trace of a sweep (indexed by position in the order). */
typedef struct {
    double* num; double* acum; int32_t* comp; double* beta;
} orc_trace;

/* a5: BayesRKernel::computeNum. Dense: densebayesrkernel.cpp:10-15.
 * Ragged: sparsebayesrkernel.cpp:24-27 with raggedbayesrkernel.cpp:32-35.
 * Eigen: sparsebayesrkernel.cpp:24-27 with eigenbayesrkernel.cpp:17-20. */
double orc_compute_num(const orc_data* d, uint32_t j, const double* eps,
                       double beta_old, double epsSum);

/* a7: residual update, applied only when not skipped. Dense:
 * densebayesrkernel.cpp:17-20; ragged raggedbayesrkernel.cpp:10-30 plus the
 * epsilonSum update sparsebayesrkernel.cpp:29-34 / SparseBayesRRG.cpp:84-92. */
void orc_apply_epsilon_change(const orc_data* d, uint32_t j, double* eps,
                              double* epsSum, double beta_old, double beta_new);

/* a6 + a5 + a7: BayesRBase::processColumn (BayesRBase.cpp:287-392). */
void orc_process_column(const orc_data* d, const orc_model* m, orc_state* s,
                        uint32_t j, double u, double z,
                        double* num_out, double* acum_out);

/* a10: Sequential::exec (sequential.cpp:23-26): processColumn over order.
 * u, z are indexed by marker id (BayesRBase.cpp:343-344). Clears v/betasqn
 * first as runGibbs does (BayesRBase.cpp:200-202). */
void orc_sweep(const orc_data* d, const orc_model* m, orc_state* s,
               const uint32_t* order, const double* u, const double* z,
               orc_trace* trace);

/* (f)-4: the asynchronous sweep of ParallelGraph (parallelgraph.cpp,
 * BayesRBase.cpp:394-523) in windows of `window` markers: each marker of a
 * window computes num from the window-start eps / epsSum and decides; the
 * window's residual changes are then applied in order. window = 1 is the
 * sequential sweep. */
void orc_sweep_async(const orc_data* d, const orc_model* m, orc_state* s,
                     const uint32_t* order, const double* u, const double* z, uint32_t window);

/* a8: BayesRBase::init (BayesRBase.cpp:39-99). y is the phenotype as stored
 * by the reference (already rounded through float, data.cpp:357-359). */
void orc_init(const orc_data* d, const orc_model* m, orc_state* s,
              const double* y, uint64_t seed, uint32_t chain);

/* a8: updateMu. Dense: DenseBayesRRmz.cpp:36-42; ragged SparseBayesRRG.cpp:106-113. */
void orc_update_mu(const orc_data* d, orc_state* s, uint64_t seed, uint32_t iteration,
                   uint32_t chain);

/* a8: the post-sweep draws of sigmaE, sigmaG, pi (BayesRBase.cpp:237-247). */
void orc_post_sweep(const orc_data* d, const orc_model* m, orc_state* s,
                    uint64_t seed, uint32_t iteration, uint32_t chain);

/* a8-a10: runGibbs (BayesRBase.cpp:130-285) for `iterations` iterations with
 * the Philox streams. order (M) carries the marker order across iterations
 * (std::iota on entry if order_init != 0). shuffle_mode 0 = Philox, 1 = libc
 * rand(). Per-iteration scalars go to hist[it * (4 + G)] =
 * {mu, sigmaE, m0_total, epsSqNorm, sigmaG[0..G-1]} when hist != NULL. */
void orc_run_gibbs(const orc_data* d, const orc_model* m, orc_state* s,
                   const double* y, uint64_t seed, uint32_t chain,
                   uint32_t iterations, uint32_t* order, int shuffle_mode,
                   double* hist);

/* Fixed effects of BayesRBase::runGibbs (BayesRBase.cpp:212-235), after the
 * sweep: the covariate order forder (F, carried across iterations like the
 * reference's xI) is shuffled on ORC_STREAM_FIXED, then for each covariate f in
 * that order num = sum_k X(k,f) (eps_k + gamma_old X(k,f)), denom = (N - 1) +
 * sigmaE / sigmaF, gamma_f ~ N(num / denom, sigmaE / denom) (scalar draw
 * ORC_DRAW_FIXED0(G, K) + f), eps += (gamma_old - gamma_f) X(., f). X is the
 * reference's float matrix (Data::X, column-major N x F). epsSum is left as is
 * (the reference does not update it here). sigmaF becomes s02F = 1. */
void orc_fixed_effects(const orc_data* d, const orc_model* m, orc_state* s, const float* X, uint32_t F,
                       double* gamma, uint32_t* forder, double* sigmaF, uint64_t seed, uint32_t iteration,
                       uint32_t chain);

/* orc_run_gibbs with fixed effects after every sweep (F = 0: none) and the sweep in
 * asynchronous windows (window 0: sequential). */
void orc_run_gibbs_fixed(const orc_data* d, const orc_model* m, orc_state* s,
                         const double* y, uint64_t seed, uint32_t chain,
                         uint32_t iterations, uint32_t* order, int shuffle_mode,
                         double* hist, uint32_t window, const float* X, uint32_t F,
                         double* gamma, uint32_t* forder, double* sigmaF);

/* orc_run_gibbs with the sweep in asynchronous windows (window 0: sequential). */
void orc_run_gibbs_window(const orc_data* d, const orc_model* m, orc_state* s,
                          const double* y, uint64_t seed, uint32_t chain,
                          uint32_t iterations, uint32_t* order, int shuffle_mode,
                          double* hist, uint32_t window);

/* Scalar-draw ids of the Philox scalar stream. */
enum { ORC_DRAW_MU = 0, ORC_DRAW_SIGMAE = 1, ORC_DRAW_SIGMAG0 = 2 };
#define ORC_DRAW_PI0(G) (2u + (uint32_t)(G))
#define ORC_DRAW_FIXED0(G, K) (2u + (uint32_t)(G) + (uint32_t)(G) * (uint32_t)(K))
#define ORC_INIT_ITERATION 0xFFFFFFFFu

/* ------------------------------------------------------ BayesW (a11-a13) -- */

/* ARMS uniforms: stream tag of ctr word 2 (see bayesw.c). */
enum { ORC_STREAM_ARMS = 4 };
#define ORC_ARMS_NPOINT 100
typedef double (*orc_logdens)(double x, void* ctx);
typedef struct {
    uint64_t seed;
    uint32_t site;        /* marker id, or ORC_BW_SITE_MU / _ALPHA */
    uint32_t iteration;
    uint32_t chain;
} orc_arms_rng;
#define ORC_BW_SITE_MU 0xFFFFFFF0u
#define ORC_BW_SITE_ALPHA 0xFFFFFFF1u
#define ORC_BW_SITE_THETA0 0xFFFFFF00u   /* + covariate index f < 128 */
/* scalar draws of a BayesW iteration: sigma_b, then the Dirichlet gammas */
enum { ORC_BW_DRAW_SIGMAB = 0, ORC_BW_DRAW_PI0 = 1 };
#define ORC_BW_MAXK 8

/* a13: one ARMS draw (BayesW_arms.cpp:133-240) with nsamp = 1 and no
 * Metropolis step. Returns 0 or the reference's error code (1001-1008, 2000). */
int orc_arms(const double* xinit, int ninit, double xl, double xr, orc_logdens f, void* ctx, int npoint,
             const orc_arms_rng* rng, double* xsamp, int* neval);

/* a11: the tabulated adaptive Gauss-Hermite rule of order n (3..15, odd):
 * n-1 off-centre nodes x / weights w in the reference's order and the centre
 * weight; returns n-1 (0 for an unsupported n). */
int orc_bw_gh_rule(int n, double* x, double* w, double* wc);
/* gauss_hermite_adaptive_integral for class C_k; vi = {vi_sum, vi_0, vi_1, vi_2}. */
double orc_bw_gh_integral(int n, double sigma, double C_k, double sigma_b, double alpha, double sum_failure,
                          double mean, double sd, const double vi[4]);
/* calculateSumFailure of a ragged marker over the failure indicator vector. */
double orc_bw_sum_failure(const orc_data* d, uint32_t j, const double* fail);

typedef struct {
    int K;                      /* classes including the zero one, <= ORC_BW_MAXK */
    const double* C;            /* K-1 mixture variances */
    int quad_points;            /* 3..15 odd */
    const double* sum_failure;  /* M, orc_bw_sum_failure per marker */
} orc_bw_model;

typedef struct {
    double* eps;                /* N */
    double* vi;                 /* N: exp(alpha eps - EulerGamma) */
    double* beta;               /* M */
    int32_t* comp;              /* M */
    double* v;                  /* K counts (1 + assignments) */
    double* pi;                 /* K */
    double mu, alpha, sigma_b, d;
} orc_bw_state;

typedef struct {
    double vi_sum, vi_1, vi_2, acum;
    double ml[ORC_BW_MAXK];
} orc_bw_trace;

/* a11: BayesWBase::processColumn on the ragged layout; p is the marker's
 * uniform. Returns 0 or an ARMS error code. */
int orc_bw_process_column(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, uint32_t j, double p,
                          uint64_t seed, uint32_t iteration, uint32_t chain, orc_bw_trace* tr);
/* processColumn over order with u indexed by marker id. */
int orc_bw_sweep(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const uint32_t* order,
                 const double* u, uint64_t seed, uint32_t iteration, uint32_t chain);
/* a12: BayesWBase::init, sampleMu, sampleAlpha, the sigma_b / pi_L draws and runGibbs. */
void orc_bw_init(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const double* y,
                 const double* fail);
int orc_bw_sample_mu(const orc_data* d, orc_bw_state* s, uint64_t seed, uint32_t iteration, uint32_t chain);
int orc_bw_sample_alpha(const orc_data* d, orc_bw_state* s, const double* fail, uint64_t seed,
                        uint32_t iteration, uint32_t chain);
void orc_bw_post_sweep(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, uint64_t seed,
                       uint32_t iteration, uint32_t chain);
/* hist: iterations x 5 {mu, alpha, sigma_b, nonzero markers, |beta|^2}. */
/* BayesW fixed effects: sum_failure_fix (bayeswbase.cpp:532-537) over the
 * reference's float covariates (column-major N x F), and sampleTheta
 * (bayeswbase.cpp:572-601) of covariate f: ARMS over theta_dens with the
 * starting points theta +- {0.01, 0, 0.005, 0.01}, bounds (-2, 2), the
 * residual shifted by X_f theta_f before and after (ARMS site
 * ORC_BW_SITE_THETA0 + f). */
void orc_bw_sum_failure_fixed(uint32_t N, const float* X, uint32_t F, const double* fail, double* sumf);
int orc_bw_sample_theta(const orc_data* d, orc_bw_state* s, const float* X, uint32_t f, double* theta,
                        const double* sumf, uint64_t seed, uint32_t iteration, uint32_t chain);
/* orc_bw_run_gibbs with the fixed effects after mu each iteration (F = 0: none). */
int orc_bw_run_gibbs_fixed(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const double* y,
                           const double* fail, uint64_t seed, uint32_t chain, uint32_t iterations,
                           uint32_t* order, double* hist, const float* X, uint32_t F, double* theta);
int orc_bw_run_gibbs(const orc_data* d, const orc_bw_model* m, orc_bw_state* s, const double* y,
                     const double* fail, uint64_t seed, uint32_t chain, uint32_t iterations, uint32_t* order,
                     double* hist);

#ifdef __cplusplus
}
#endif
#endif
