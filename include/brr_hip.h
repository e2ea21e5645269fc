/*
 * brr_hip.h -- C ABI of the MI355X-native BayesR marker sweep (libbrr_hip.so).
 *
 * This is the drop-in boundary a host C++ build of BayesRRcmd links against.
 * Plain C types only: pointers are host memory owned by the caller and are
 * only touched during the (synchronous) call. Every entry point returns a
 * status code (0 = ok) and never throws; brr_hip_last_error() holds the text
 * of the last failure of a context. One context is not re-entrant but may be
 * used from any single thread at a time (BayesRBase::processColumn is called
 * from one TBB serial node at a time, limitsequencegraph.cpp:46-55).
 *
 * Reference interfaces replaced (paths relative to the BayesRRcmd tree):
 *   - AnalysisGraph::exec(Analysis*, N, M, order)  src/analysisgraph.hpp:18-21
 *     as implemented by Sequential (src/sequential.cpp:10-27) and
 *     LimitSequenceGraph (src/limitsequencegraph.cpp:79-106), each driving
 *     BayesRBase::processColumn (src/BayesRBase.cpp:287-392)  -> brr_hip_sweep
 *   - BayesRKernel::computeNum / calculateEpsilonChange
 *     (src/bayesrkernel.h:11-15; dense src/densebayesrkernel.cpp:10-20,
 *     ragged src/sparsebayesrkernel.cpp:10-34 + src/raggedbayesrkernel.cpp:10-35)
 *                                      -> brr_hip_compute_num / brr_hip_epsilon_change
 *   - Marker storage of DenseMarker (src/densemarker.h:7-20, the column as
 *     a 4-entry table over the PLINK 2-bit codes) and RaggedSparseMarker
 *     (src/raggedsparsemarker.h:7-23)  -> brr_hip_upload_dense / _ragged
 *   - The per-iteration loop of BayesRBase::runGibbs (src/BayesRBase.cpp:186-253)
 *     with updateMu (src/DenseBayesRRmz.cpp:36-42, src/SparseBayesRRG.cpp:106-113)
 *     and the sigmaE / sigmaG / pi draws (src/BayesRBase.cpp:237-247), held on
 *     the device for a whole chain  -> brr_hip_run_chain
 *   - BayesW (Weibull survival) on ragged markers: AnalysisGraph::exec driving
 *     BayesWBase::processColumn (src/bayeswbase.cpp:604-704) with
 *     RaggedBayesWKernel (src/raggedbayeswkernel.cpp:10-77)  -> brr_hip_bw_sweep;
 *     BayesWKernel::setVi / calculateSumFailure / calculateResidualUpdate
 *     (src/bayeswkernel.h:18-26)  -> brr_hip_bw_set_vi / brr_hip_bw_sum_failure /
 *     brr_hip_bw_residual_update; BayesWBase::runGibbs (src/bayeswbase.cpp:737-830)
 *     with sampleMu / sampleAlpha (:542-569, :707-734) and the sigma_b / pi_L
 *     draws (:798-806)  -> brr_hip_bw_run_chain
 */
#ifndef BRR_HIP_H
#define BRR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRR_HIP_ABI_VERSION 7

enum brr_hip_status {
    BRR_HIP_OK = 0,
    BRR_HIP_EINVAL = 1,   /* bad argument or size */
    BRR_HIP_ENOMEM = 2,   /* device allocation failed */
    BRR_HIP_EDEVICE = 3,  /* HIP runtime error or no device */
    BRR_HIP_ESTATE = 4,   /* call out of order (e.g. sweep before upload) */
    BRR_HIP_ETIMEOUT = 5, /* a device-side wait exceeded its bound */
    BRR_HIP_ESAMPLER = 6  /* BayesW: the adaptive rejection sampler failed (its
                             error code is in brr_hip_last_error) */
};

/* Marker layouts (PreprocessDataType, src/common.h:27-32). */
enum brr_hip_layout {
    BRR_HIP_LAYOUT_DENSE_2BIT = 1, /* PreprocessDataType::Dense, columns kept as PLINK 2-bit */
    BRR_HIP_LAYOUT_RAGGED = 2,     /* PreprocessDataType::SparseRagged */
    BRR_HIP_LAYOUT_EIGEN_SPARSE = 3 /* PreprocessDataType::SparseEigen: EigenBayesRKernel
                                      (src/eigenbayesrkernel.cpp) on the ragged lists with
                                      EigenSparse statistics; held as 2-bit codes. BayesR only. */
};

typedef struct brr_hip_ctx brr_hip_ctx;

typedef struct {
    uint32_t num_inds;       /* N, Data::numInds */
    uint32_t num_snps;       /* M, Data::numSnps */
    uint32_t num_components; /* K = Options::S.cols() + 1 (zero component included), 2..8 */
    uint32_t num_groups;     /* G = Data::numGroups, 1..32 */
    uint32_t layout;         /* brr_hip_layout */
    uint32_t workgroups;     /* persistent workgroups (0 = choose) */
    uint32_t block_markers;  /* markers decided per speculation block, 1..64 (0 = choose) */
    uint32_t device;         /* HIP device ordinal */
    uint32_t ragged_storage; /* ragged layout only: device copy of the columns as
                                0 = whichever is smaller (index lists or 2-bit codes),
                                1 = index lists, 2 = 2-bit codes built on the device */
    uint32_t num_traits;     /* 0 = one chain. 1..64 = T chains (traits, one phenotype
                                each) over the same genotype columns and marker order,
                                swept together; needs num_groups == 1 and at most
                                256 individuals per workgroup (512 with the FP32
                                residual, BRR_HIP_PRECISION_FP32). The per-individual and
                                per-marker arrays of the state calls are then N x T and
                                M x T, trait fastest. */
} brr_hip_config;

/* Model constants: Options::S (G x (K-1), row-major) and the priors
 * v0E, s02E, v0G, s02G (src/options.hpp; BayesRBase.cpp:31-34). */
typedef struct {
    const double* cva;
    double v0E, s02E, v0G, s02G;
} brr_hip_model;

/* Chain scalars held by BayesRBase (src/BayesRBase.hpp:63-86). In set calls the
 * v / betasqn arrays are ignored (a sweep clears them, BayesRBase.cpp:200-202). */
typedef struct {
    double mu;           /* m_mu */
    double sigmaE;       /* m_sigmaE */
    double epsilon_sum;  /* m_epsilonSum (ragged layout only) */
    double* sigmaG;      /* m_sigmaG, G */
    double* pi;          /* m_pi, G x K */
    double* v;           /* m_v, G x K (out) */
    double* betasqn;     /* m_betasqnG, G (out) */
} brr_hip_scalars;

/* Counters of the last sweep / chain launch. */
typedef struct {
    uint64_t markers;      /* marker-updates performed */
    uint64_t rounds;       /* device exchange rounds */
    uint64_t changes;      /* markers whose beta changed (epsilon updated) */
    double kernel_ms;      /* device time of the launch, HIP events */
    uint32_t workgroups;   /* persistent workgroups used */
    uint32_t block_markers;
    uint32_t slice;        /* individuals per workgroup */
    uint32_t storage;      /* device copy of the columns: 1 index lists, 2 2-bit codes */
    uint64_t decision_rounds;  /* batched: decision rounds, summed over the waves of workgroup 0 */
} brr_hip_stats;

int brr_hip_abi_version(void);
int brr_hip_device_count(int* count);

/* Create a context and allocate the per-individual / per-marker device state. */
int brr_hip_create(const brr_hip_config* cfg, brr_hip_ctx** out);
void brr_hip_destroy(brr_hip_ctx* ctx);
const char* brr_hip_last_error(const brr_hip_ctx* ctx);

/* Dense columns: M SNP-major PLINK columns of ceil(N/4) bytes each, column j
 * starting at bed + j * column_stride (the .bed payload after the 3 magic
 * bytes has column_stride = ceil(N/4)). lut: M x 4 doubles, the standardised
 * column value for dosage 0, 1, 2 and for a missing call, as DenseMarkerBuilder
 * computes them (src/densemarkerbuilder.cpp:43-60). */
int brr_hip_upload_dense(brr_hip_ctx* ctx, const uint8_t* bed, uint64_t column_stride,
                         const double* lut);

/* Ragged columns: CSR lists of individuals with genotype 1, 2 and missing,
 * ascending within a column; *_off have M+1 entries. stats: M x 4 doubles
 * {mean, sd, sqrdZ, Zsum} (src/sparsemarker.h:8-11). On an EigenSparse context
 * the same lists (the nonzeros of EigenSparseMarker::Zg by value) with the
 * EigenSparse statistics (brr_eigen_fill / brr_eigen_read_fill). */
int brr_hip_upload_ragged(brr_hip_ctx* ctx,
                          const int32_t* ones, const uint64_t* ones_off,
                          const int32_t* twos, const uint64_t* twos_off,
                          const int32_t* miss, const uint64_t* miss_off,
                          const double* stats);

/* A whole PLINK .bed file image (magic 6c 1b 01, then M SNP-major columns of
 * ceil(N/4) bytes; file_bytes must be 3 + M * ceil(N/4)) decoded on the device
 * into the markers of `layout`, in place of the host preprocess pass
 * (PreprocessGraph, src/preprocessgraph.cpp:24-60, with the builders it runs)
 * and the upload calls above. Columns are held as 2-bit codes; the per-marker
 * table (dense: DenseMarkerBuilder::endColumn, src/densemarkerbuilder.cpp:43-60)
 * or statistics (ragged: SparseMarker::updateStatistics, src/sparsemarker.cpp:8-19)
 * are computed on the device and equal brr_dense_luts / brr_ragged_fill bit for
 * bit (EigenSparse statistics: brr_eigen_fill). No index lists are built. A
 * BayesR context takes its own layout only; a BayesW context takes dense or
 * ragged. BRR_HIP_EINVAL for a bad file, or (ragged) a
 * marker without genotype 1 or 2 calls, as RaggedSparseMarkerBuilder rejects it
 * (EigenSparse: a marker without any nonzero call, EigenSparseMarker::isValid).
 * brr_hip_last_stats().kernel_ms and brr_hip_last_phase_ms() give the device
 * time of the decode (phase 0: unpack and count, phase 1: dense tables). */
int brr_hip_load_bed(brr_hip_ctx* ctx, const uint8_t* bed_file, uint64_t file_bytes, int layout);
/* The M x 4 per-marker values held on the device: dense tables {dosage 0, 1, 2,
 * missing} or ragged statistics {mean, sd, sqrdZ, Zsum}. */
int brr_hip_get_marker_tables(brr_hip_ctx* ctx, double* out);

/* Arithmetic of the residual in brr_hip_sweep and brr_hip_run_chain. The
 * reference holds epsilon as an Eigen::VectorXd and updates it in FP64
 * (BayesRBase.cpp:370-378 with densebayesrkernel.cpp:17-20 /
 * raggedbayesrkernel.cpp:10-30). BRR_HIP_PRECISION_FP64 (the default) does the
 * same. BRR_HIP_PRECISION_FP32_REFRESH holds the residual slices in FP32 and
 * forms the x_j . epsilon dots in FP32 with compensated (TwoSum) wave
 * reductions; the residual changes are logged and replayed in FP64 on the
 * context's FP64 residual at least every refresh_markers (1..1024) markers and
 * at every iteration boundary, after which the FP32 slices are re-rounded from
 * it. So the residual returned by brr_hip_get_state, updateMu, the fixed
 * effects and |epsilon|^2 of the sigmaE draw are FP64, as are the decisions and
 * every scalar draw; only the dots see the FP32 residual. Single-chain
 * sequential sweeps whose columns are held as 2-bit codes (dense, EigenSparse,
 * or ragged with ragged_storage 0 or 2). */
enum brr_hip_precision {
    BRR_HIP_PRECISION_FP64 = 0,
    BRR_HIP_PRECISION_FP32_REFRESH = 1,
    /* Batched contexts: the residual tiles of every trait in FP32 on the f32
     * matrix-core instruction (slices of up to 512 individuals instead of 256,
     * so N up to 131,072 on 256 workgroups); each 16-individual tile's FP32 dot
     * and Gram products are summed over the tiles in FP64, decisions and draws
     * are FP64; the residual is not refreshed, its FP32 rounding accumulates. */
    BRR_HIP_PRECISION_FP32 = 2
};
int brr_hip_set_precision(brr_hip_ctx* ctx, int precision, uint32_t refresh_markers);

/* The per-marker BayesRKernel interface against a host residual, for a
 * HipBayesRKernel : BayesRKernel (src/bayesrkernel.h:11-15) that the reference's
 * Analysis::kernelForMarker hands to processColumn. Correctness path: every call
 * copies the N-vector across PCIe.
 *   brr_hip_kernel_compute_num   -> computeNum(const VectorXd &epsilon, double beta_old):
 *       dense markers num = Cx . epsilon + beta_old (N - 1) (densebayesrkernel.cpp:10-15);
 *       ragged / EigenSparse markers num = beta_old (N - 1) - mean epsilon_sum / sd +
 *       dot(epsilon) / sd (sparsebayesrkernel.cpp:24-27), epsilon_sum being the kernel's
 *       epsilonSum (ignored for dense markers).
 *   brr_hip_kernel_epsilon_change -> calculateEpsilonChange(beta_old, beta): delta (N)
 *       receives the residual change (densebayesrkernel.cpp:17-20, raggedbayesrkernel.cpp:10-30:
 *       the caller adds it to epsilon), *epsilon_sum_delta (if not NULL) the epsilonSum
 *       update of a sparse marker (sparsebayesrkernel.cpp:29-34; 0 for dense markers).
 * The context's own residual and state are untouched. */
int brr_hip_kernel_compute_num(brr_hip_ctx* ctx, uint32_t marker, const double* epsilon, double epsilon_sum,
                               double beta_old, double* num);
int brr_hip_kernel_epsilon_change(brr_hip_ctx* ctx, uint32_t marker, double beta_old, double beta,
                                  double* delta, double* epsilon_sum_delta);

/* Launch shape of a BayesR context (single-chain or batched): `workgroups`
 * persistent workgroups requested (0: the default rule of brr_hip_create; the
 * individuals are spread over them in slices that are a multiple of 16) and
 * `block_markers` markers per speculation block (1..64, 0: keep; single-chain
 * contexts). The reference's own tunables are --thread and the token counts of
 * its flow graphs (options.cpp; limitsequencegraph.cpp:41-44); these are the
 * device's. The state and the uploaded markers are kept. */
int brr_hip_set_launch(brr_hip_ctx* ctx, uint32_t workgroups, uint32_t block_markers);

/* The autotuner: times `iterations` (1..16) chain iterations from the current
 * state for each candidate launch shape -- lane runs per slice (1, 2, 4) with
 * the fewest workgroups that hold N, and all compute units, each with 32, 48
 * or 64 markers per block (batched contexts: slices of 256, 224, 192, 160 and
 * 128 individuals, and all compute units) -- restoring the state before every
 * trial, then keeps the fastest shape and the state as it was. `table` (if not
 * NULL) receives up to table_cap entries {workgroups, block markers, slice,
 * device ms per iteration; -1 if the shape could not run}. Tune from a state
 * like the one the chain will run from (the cost of a round depends on how
 * many markers change). */
typedef struct {
    uint32_t workgroups;
    uint32_t block_markers;
    uint32_t slice;
    double ms_per_iteration;
} brr_hip_tune_entry;
int brr_hip_autotune(brr_hip_ctx* ctx, uint64_t seed, uint32_t iterations, brr_hip_tune_entry* table,
                     uint32_t table_cap, uint32_t* entries);

/* Independent chains of several contexts -- cohorts of uneven N and M, the
 * reference's one runBayesRAnalysis per data set (analysisrunner.cpp:123-152)
 * -- run side by side: `iterations` whole-chain iterations of every context
 * (as brr_hip_run_chain; orders[q]: iterations x M_q, hists[q]: NULL or
 * iterations x (4 + G_q), chains[q]: NULL or each context's chain index). The
 * contexts are packed into persistent cooperative launches, each cohort on a
 * disjoint set of workgroups (its context's own count) with its own exchange,
 * at most one workgroup per compute unit and 64 cohorts per launch, longest
 * chains first (markers x (1 / block + the context's last change rate)); the
 * launches are captured into one hipGraph and run by one graph launch. All contexts: one device, single-chain sequential
 * BayesR without fixed effects, columns held as 2-bit codes, one residual
 * precision, slices of at most 1024 individuals (one or two lane runs; a context
 * with more is refused with BRR_HIP_EINVAL; brr_hip_set_launch sets a context's);
 * each launch runs the kernel of the most lane runs any of its cohorts has. report (if not NULL)
 * gets the number of launches, the device time of the graph and the launch of
 * each of the first 64 cohorts. */
typedef struct {
    uint32_t launches;
    double device_ms;
    uint32_t launch_of[64];
} brr_hip_cohort_report;
#define BRR_HIP_COHORTS_DIRECT 1u   /* flags: launch one after another instead of through a graph */
int brr_hip_run_cohorts(brr_hip_ctx* const* ctxs, uint32_t n, uint64_t seed, const uint32_t* chains,
                        uint32_t first_iteration, uint32_t iterations, const uint32_t* const* orders,
                        double* const* hists, uint32_t flags, brr_hip_cohort_report* report);

/* Measurement helpers (no counterpart in the reference):
 * brr_hip_probe_exchange runs `rounds` exchange rounds of a full speculation
 * block on the context's launch shape with no other work -- every workgroup
 * publishes its partials, the reduce-scatter and the all-gather -- and gives the
 * device time per round: the latency floor of the sweep's rounds (single-chain
 * BayesR contexts; the state is untouched).
 * brr_hip_calibration_copy copies `bytes` (a multiple of 16) between two device
 * buffers with one kernel and gives its time: a copy of known size whose HBM
 * counters scale the sweep's. */
int brr_hip_probe_exchange(brr_hip_ctx* ctx, uint32_t rounds, double* us_per_round);
int brr_hip_calibration_copy(uint32_t device, uint64_t bytes, double* ms);

/* Per-marker record of num, the value BayesRKernel::computeNum returns and
 * processColumn decides on (BayesRBase.cpp:300-305, densebayesrkernel.cpp:10-15,
 * sparsebayesrkernel.cpp:24-27): with enable != 0 every later sweep and chain
 * launch writes the num of each marker's decision (a chain: of its last
 * iteration), and brr_hip_get_num copies the M values out. Single-chain BayesR
 * contexts; the record costs one store per marker. It exists to state the
 * dot-product error of a precision mode against another run from the same
 * state: for dense markers num - beta_old (N - 1) is x_j . epsilon, for ragged
 * ones sd (num - beta_old (N - 1)) + mean epsilonSum is. */
int brr_hip_record_num(brr_hip_ctx* ctx, int enable);
int brr_hip_get_num(brr_hip_ctx* ctx, double* out);

/* Fixed effects (--fixed / Options::fixedEffectNumber, BayesRBase.cpp:212-235):
 * X is Data::X, the N x F float matrix in Eigen's column-major order (column f
 * = covariate f over the individuals), F <= 128. Resets gamma to 0, the
 * covariate order to 0..F-1 and sigmaF to s02F = 1 (the reference reads an
 * uninitialised m_sigmaF in iteration 0 and s02F after it). brr_hip_run_chain
 * then runs the fixed-effects block after every sweep; the reference's order
 * is kept: the covariate order (carried across iterations like xI) shuffled by
 * the random_shuffle rule on its own stream (tag 5), then per covariate f num =
 * sum_k X(k,f) (eps_k + gamma_old X(k,f)), denom = (N - 1) + sigmaE / sigmaF,
 * gamma_f ~ N(num / denom, sigmaE / denom) (scalar draw 2 + G + G K + f),
 * eps += (gamma_old - gamma_f) X(., f); epsilonSum is not updated, as in the
 * reference. F = 0 removes them. Single-chain BayesR contexts. */
int brr_hip_set_fixed(brr_hip_ctx* ctx, const float* X, uint32_t num_fixed);
/* gamma (F), the covariate order (F) and sigmaF; NULL arrays are skipped. */
int brr_hip_set_fixed_state(brr_hip_ctx* ctx, const double* gamma, const uint32_t* order, double sigma_f);
int brr_hip_get_fixed_state(brr_hip_ctx* ctx, double* gamma, uint32_t* order, double* sigma_f);
/* The fixed-effects block alone on the device state, for a host that runs
 * runGibbs itself and calls brr_hip_sweep for exec (BayesRBase.cpp:212-235). */
int brr_hip_fixed_effects(brr_hip_ctx* ctx, uint64_t seed, uint32_t iteration, uint32_t chain);

/* The asynchronous sweep of ParallelGraph (--analysis-type asyncppbayes,
 * src/parallelgraph.cpp; BayesRBase::processColumnAsync / updateGlobal,
 * src/BayesRBase.cpp:394-523) in its deterministic schedule: the order is cut
 * into windows of `window` markers (Options::analysisTokens, 1..64); every
 * marker of a window computes its num from the residual (and epsilonSum) at the
 * window start, draws its component and beta, and the window's residual
 * changes are then applied in order. window = 0 restores the exact sequential
 * sweep (with the configured block size). Single-chain BayesR contexts only;
 * applies to brr_hip_sweep and brr_hip_run_chain. Unlike processColumnAsync,
 * acum is still written per marker. */
int brr_hip_set_async(brr_hip_ctx* ctx, uint32_t window);

/* Group label per marker (Data::G, src/data.cpp:424-453); default all 0. */
int brr_hip_set_groups(brr_hip_ctx* ctx, const int32_t* group);
int brr_hip_set_model(brr_hip_ctx* ctx, const brr_hip_model* model);

/* State upload / download. Any pointer may be NULL to skip that array. */
int brr_hip_set_state(brr_hip_ctx* ctx, const double* epsilon, const double* beta,
                      const int32_t* components);
int brr_hip_set_scalars(brr_hip_ctx* ctx, const brr_hip_scalars* s);
int brr_hip_get_state(brr_hip_ctx* ctx, double* epsilon, double* beta,
                      int32_t* components, double* acum);
int brr_hip_get_scalars(brr_hip_ctx* ctx, brr_hip_scalars* s);
/* Scalars of one trait of a batched context (trait 0 is what set/get_scalars
 * address). One BayesRBase instance per trait holds these in the reference. */
int brr_hip_set_trait_scalars(brr_hip_ctx* ctx, uint32_t trait, const brr_hip_scalars* s);
int brr_hip_get_trait_scalars(brr_hip_ctx* ctx, uint32_t trait, brr_hip_scalars* s);

/* One AnalysisGraph::exec: processColumn over order[0..M) in sequence.
 * u, z: the per-marker uniform and standard normal of this iteration, indexed
 * by marker id (BayesRBase.cpp:343-344); M x T for a batched context. Uses and
 * updates the device state. */
int brr_hip_sweep(brr_hip_ctx* ctx, const uint32_t* order, const double* u, const double* z);

/* brr_hip_sweep without any device allocation. No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). The same sweep kernel on the
 * same context state, but order / u / z are device pointers (M uint32, M and
 * M doubles, caller-owned, read only) and the kernel's exchange words, control
 * words and counters, and a bitmap that checks the order on the device, live
 * in the library's file-scope device scratch (1.53 MiB per GPU, held until the
 * process ends). Limits: single-chain BayesR contexts, M <= 4,194,304 markers
 * (BRR_HIP_EINVAL beyond); an order that is not a permutation is refused with
 * BRR_HIP_EINVAL before the sweep touches the state. Calls that overlap on one
 * GPU race on its scratch; calls in sequence (each is synchronous) are safe.
 * Outputs: the context's state, read as after brr_hip_sweep. */
int brr_hip_sweep_noalloc(brr_hip_ctx* ctx, const uint32_t* d_order, const double* d_u, const double* d_z);

/* A device-resident chain: `iterations` full runGibbs iterations (updateMu,
 * sweep, sigmaE / sigmaG / pi draws) in one persistent launch, drawing u, z
 * and the scalars from the counter-based stream (seed, chain). orders holds
 * the marker order of each iteration (iterations x M). hist (may be NULL)
 * receives iterations x (4 + G) doubles {mu, sigmaE, m0, |eps|^2, sigmaG[G]}.
 * Batched context: trait t draws from stream (seed, chain + t) and hist is
 * iterations x T x 5. */
int brr_hip_run_chain(brr_hip_ctx* ctx, uint64_t seed, uint32_t chain,
                      uint32_t first_iteration, uint32_t iterations,
                      const uint32_t* orders, double* hist);

/* Per-marker kernel calls on the device epsilon of a single-chain context
 * (correctness path: one launch per call). compute_num: BayesRKernel::computeNum with the context's
 * epsilon_sum. epsilon_change: applies calculateEpsilonChange(beta_old, beta)
 * to the device epsilon (and epsilon_sum for ragged). */
int brr_hip_compute_num(brr_hip_ctx* ctx, uint32_t marker, double beta_old, double* num);
int brr_hip_epsilon_change(brr_hip_ctx* ctx, uint32_t marker, double beta_old, double beta);

int brr_hip_last_stats(const brr_hip_ctx* ctx, brr_hip_stats* stats);

/* Device time of workgroup 0 in each phase of the last launch, ms (n <= 12
 * values written). Single chain: block load, partial dots, exchange,
 * decisions, commit/update, per-iteration scalars. Batched: block commit, dot
 * tiles, exchange gather, decision draws, residual update, per-iteration
 * scalars, Gram tiles, u/z draws, exchange reduce, accept/correct, block-end
 * barrier, spare. */
int brr_hip_last_phase_ms(const brr_hip_ctx* ctx, double* out, uint32_t n);

/* ---- BayesW: Weibull survival model on ragged markers ---- */

/* A BayesW context holds N x {epsilon, failure indicator}, the columns as 2-bit
 * codes, beta and the mixture component per marker. Markers are uploaded with
 * brr_hip_upload_ragged (RaggedBayesWKernel, src/raggedbayeswkernel.cpp) or
 * brr_hip_upload_dense (DenseBayesWKernel, src/densebayeswkernel.cpp:10-42 with
 * DenseBayesW::beta_dens, src/densebayesw.cpp:25-33); the last upload fixes the
 * layout and must be followed by brr_hip_bw_set_model. brr_hip_set_state /
 * brr_hip_get_state move epsilon, beta and components (acum is not kept: pass NULL). */
typedef struct {
    uint32_t num_inds;       /* N */
    uint32_t num_snps;       /* M */
    uint32_t num_components; /* K = number of mixture classes + the zero class, 2..8 */
    uint32_t quad_points;    /* Gauss-Hermite points n_q: 3, 5, ..., 15 (Options::quad_points) */
    uint32_t workgroups;     /* persistent workgroups (0 = choose) */
    uint32_t device;         /* HIP device ordinal */
} brr_hip_bw_config;

/* Chain scalars of BayesWBase (src/bayeswbase.h:48-67). pi and v have K entries;
 * in set calls v is the class count a sweep starts from (runGibbs sets it to
 * ones before each sweep, src/bayeswbase.cpp:789) and betasq is ignored. */
typedef struct {
    double mu;        /* m_mu */
    double alpha;     /* m_alpha, the Weibull shape */
    double sigma_b;   /* m_sigma_b */
    double d;         /* d, the number of events (sum of the failure indicators) */
    double* pi;       /* m_pi_L, K */
    double* v;        /* m_v, K */
    double betasq;    /* out: sum of beta^2 drawn in the last sweep */
} brr_hip_bw_scalars;

int brr_hip_bw_create(const brr_hip_bw_config* cfg, brr_hip_ctx** out);

/* The mixture variances C_k (Options::S, K-1 values > 0) and the failure
 * indicator of each individual (N doubles, 0 or 1). Needs the markers: the
 * per-marker sum_failure (calculateSumFailure) is computed here. */
int brr_hip_bw_set_model(brr_hip_ctx* ctx, const double* mixture_C, const double* failure);
int brr_hip_bw_set_scalars(brr_hip_ctx* ctx, const brr_hip_bw_scalars* s);
int brr_hip_bw_get_scalars(brr_hip_ctx* ctx, brr_hip_bw_scalars* s);

/* One AnalysisGraph::exec of BayesW: processColumn over order[0..M) in
 * sequence with the uniform u[j] of marker j (indexed by marker id) choosing
 * the class; a nonzero class draws beta by adaptive rejection sampling from
 * the stream (seed, chain) at `iteration`. Uses and updates the device state. */
int brr_hip_bw_sweep(brr_hip_ctx* ctx, const uint32_t* order, const double* u,
                     uint64_t seed, uint32_t iteration, uint32_t chain);

/* `iterations` whole BayesWBase::runGibbs iterations in one persistent launch:
 * sampleMu, the sweep (u from the counter-based stream), sampleAlpha, then
 * sigma_b and pi_L. orders: iterations x M marker orders. hist (may be NULL):
 * iterations x 5 doubles {mu, alpha, sigma_b, M - v0 + 1, sum beta^2}. */
int brr_hip_bw_run_chain(brr_hip_ctx* ctx, uint64_t seed, uint32_t chain,
                         uint32_t first_iteration, uint32_t iterations,
                         const uint32_t* orders, double* hist);
/* BayesW fixed effects (Data::X, bayeswbase.cpp:532-537, 572-601, 783-788): X
 * is the N x F float matrix in Eigen's column-major order, F <= 128; theta
 * starts at 0 and sum_k X(k,f) failure_k is formed on the device (after
 * brr_hip_bw_set_model). brr_hip_bw_run_chain then draws theta_0 .. theta_F-1
 * in turn after mu each iteration: ARMS over theta_dens with the starting points
 * theta - 0.01, theta, theta + 0.005, theta + 0.01 and bounds (-2, 2), the
 * residual shifted by X_f theta before and X_f theta_new after (ARMS site
 * 0xFFFFFF00 + f). F = 0 removes them. */
int brr_hip_bw_set_fixed(brr_hip_ctx* ctx, const float* X, uint32_t num_fixed);
int brr_hip_bw_set_theta(brr_hip_ctx* ctx, const double* theta);
int brr_hip_bw_get_theta(brr_hip_ctx* ctx, double* theta);
/* The fixed effects alone (every sampleTheta of one iteration) on the device
 * state, for a host that runs runGibbs itself. */
int brr_hip_bw_fixed_effects(brr_hip_ctx* ctx, uint64_t seed, uint32_t iteration, uint32_t chain);


/* Per-marker kernel calls on the device epsilon (correctness path).
 * set_vi: RaggedBayesWKernel::setVi of vi = exp(alpha eps - EulerGamma):
 *   out[4] = {vi_sum, vi_0, vi_1, vi_2}.
 *   (the sums over the individuals with genotype 0 (incl. missing calls), 1 and 2)
 * residual_update: epsilon += sign * calculateResidualUpdate(beta), sign = +1 / -1.
 * sum_failure: calculateSumFailure of every marker (M doubles). */
int brr_hip_bw_set_vi(brr_hip_ctx* ctx, uint32_t marker, double alpha, double* out4);
int brr_hip_bw_residual_update(brr_hip_ctx* ctx, uint32_t marker, double beta, int sign);
int brr_hip_bw_sum_failure(brr_hip_ctx* ctx, double* out);

/* BayesWBase::init (src/bayeswbase.cpp:469-540, no fixed effects): from the
 * log survival times y (N, as the reference stores them, i.e. through float)
 * and failure indicators: eps = y - mu, mu = mean(y), alpha = 3.14159 /
 * sqrt(6 var(y)), sigma_b = 6.283185 / (6 alpha^2 M), pi = {0.99, 0.01/(K-1), ...},
 * v = ones, d = sum(failure). s->pi and s->v must hold K values. */
int brr_bayesw_init(const double* y, const double* failure, uint32_t num_inds, uint32_t num_snps,
                    uint32_t num_components, double* epsilon, brr_hip_bw_scalars* s);

/* ---- host-side marker building (MarkerBuilder semantics, no device) ---- */

/* PLINK .bed magic check (src/preprocessgraph.cpp:195-200): 0 if ok. */
int brr_bed_check_magic(const uint8_t* header3);

/* DenseMarkerBuilder (src/densemarkerbuilder.cpp:23-60) over a SNP-major
 * column: lut[4] for dosage 0,1,2,missing. */
int brr_dense_lut(const uint8_t* column, uint32_t num_inds, double* lut);

/* RaggedSparseMarkerBuilder (src/raggedsparsemarkerbuilder.cpp:25-55): counts
 * of genotype 1 / 2 / missing for each of M columns (pass 1). */
int brr_ragged_count(const uint8_t* bed, uint64_t column_stride, uint32_t num_inds,
                     uint32_t num_snps, uint64_t* ones_off, uint64_t* twos_off,
                     uint64_t* miss_off, uint32_t threads);
/* Pass 2: fill the lists (offsets from pass 1, prefix-summed by the call) and
 * stats (M x 4). Returns BRR_HIP_EINVAL if a column is invalid
 * (src/raggedsparsemarker.cpp:68-77). */
int brr_ragged_fill(const uint8_t* bed, uint64_t column_stride, uint32_t num_inds,
                    uint32_t num_snps, const uint64_t* ones_off, const uint64_t* twos_off,
                    const uint64_t* miss_off, int32_t* ones, int32_t* twos, int32_t* miss,
                    double* stats, uint32_t threads);
/* EigenSparseMarkerBuilder (src/eigensparsemarkerbuilder.cpp:22-78): pass 1 is
 * brr_ragged_count; pass 2 fills the same lists and the EigenSparse statistics
 * (missing calls imputed at the mean in sqrdZ and Zsum). BRR_HIP_EINVAL if a
 * column has no nonzero call (src/eigensparsemarker.cpp:60-71). */
int brr_eigen_fill(const uint8_t* bed, uint64_t column_stride, uint32_t num_inds,
                   uint32_t num_snps, const uint64_t* ones_off, const uint64_t* twos_off,
                   const uint64_t* miss_off, int32_t* ones, int32_t* twos, int32_t* miss,
                   double* stats, uint32_t threads);
/* All M dense tables at once (threads host workers). */
int brr_dense_luts(const uint8_t* bed, uint64_t column_stride, uint32_t num_inds,
                   uint32_t num_snps, double* lut, uint32_t threads);

/* The reference's preprocessed files (src/preprocessgraph.cpp:92-126, src/compression.cpp:137-154):
 * data = the file's bytes, index = its index file, M x {pos, compressedSize, originalSize} u64;
 * compressed = Options::compress (each marker one zlib stream). BRR_HIP_EINVAL on a malformed
 * entry. */
/* .ragged.sparsebed (RaggedSparseMarker::write, src/raggedsparsemarker.cpp:43-66) -> the CSR
 * arena of brr_hip_upload_ragged, in two passes like brr_ragged_count / brr_ragged_fill. */
int brr_ragged_read_counts(const uint8_t* data, uint64_t data_bytes, const uint64_t* index,
                           uint32_t num_snps, int compressed, uint64_t* ones_off, uint64_t* twos_off,
                           uint64_t* miss_off, uint32_t threads);
int brr_ragged_read_fill(const uint8_t* data, uint64_t data_bytes, const uint64_t* index,
                         uint32_t num_inds, uint32_t num_snps, int compressed, const uint64_t* ones_off,
                         const uint64_t* twos_off, const uint64_t* miss_off, int32_t* ones, int32_t* twos,
                         int32_t* miss, double* stats, uint32_t threads);
/* .eigen.sparsebed (EigenSparseMarker::write, src/eigensparsemarker.cpp:41-58) -> the same
 * CSR arena (values 1 and 2 give the genotype 1 / 2 lists, any other value is a missing call
 * imputed at the mean) with the file's statistics, for brr_hip_upload_ragged on an
 * EigenSparse context. */
int brr_eigen_read_counts(const uint8_t* data, uint64_t data_bytes, const uint64_t* index,
                          uint32_t num_snps, int compressed, uint64_t* ones_off, uint64_t* twos_off,
                          uint64_t* miss_off, uint32_t threads);
int brr_eigen_read_fill(const uint8_t* data, uint64_t data_bytes, const uint64_t* index,
                        uint32_t num_inds, uint32_t num_snps, int compressed, const uint64_t* ones_off,
                        const uint64_t* twos_off, const uint64_t* miss_off, int32_t* ones, int32_t* twos,
                        int32_t* miss, double* stats, uint32_t threads);
/* .ppbed (DenseMarker::write, src/densemarker.cpp:34-49, N doubles per column) -> 2-bit codes
 * (M x ceil(N/4) bytes) and tables (M x 4) for brr_hip_upload_dense that reproduce every column
 * value bit for bit; BRR_HIP_EINVAL if a column holds more than 4 distinct values. */
int brr_dense_read_ppbed(const uint8_t* data, uint64_t data_bytes, const uint64_t* index,
                         uint32_t num_inds, uint32_t num_snps, int compressed, uint8_t* bed, double* lut,
                         uint32_t threads);

/* ---- sample output (SampleWriter, src/samplewriter.cpp, src/writer.cpp) ---- */

/* The reference's CSV samples: the header of SampleWriter::openGroups (BayesR, kind 0) or
 * open_bayesW[_fixed] (BayesW, kind 1), then one row per kept iteration, values separated by
 * ", " and printed as a default std::ostream prints a double (6 significant digits).
 * Kinds 2 and 3 write the same rows as raw little-endian doubles (full precision, no text
 * for the N residuals of a BayesR row): the 8 bytes "BRRSAMP1", then u32 {kind, M, N, G, F,
 * doubles per row}, then the rows. */
enum brr_samples_kind { BRR_SAMPLES_BAYESR = 0, BRR_SAMPLES_BAYESW = 1, BRR_SAMPLES_BAYESR_BINARY = 2,
                        BRR_SAMPLES_BAYESW_BINARY = 3 };
typedef struct brr_samples brr_samples;
int brr_samples_open(const char* path, int kind, uint32_t num_snps, uint32_t num_inds, uint32_t num_groups,
                     uint32_t num_fixed, brr_samples** out);
/* BayesRBase.cpp:251: iteration, mu, beta[M], sigmaE, sigmaG[G], gamma[F], comp[M], acum[M], epsilon[N]. */
int brr_samples_write_bayesr(brr_samples* w, uint32_t iteration, double mu, const double* beta, double sigmaE,
                             const double* sigmaG, const double* gamma, const int32_t* comp, const double* acum,
                             const double* epsilon);
/* bayeswbase.cpp:810-815: iteration, alpha, mu, theta[F], beta[M], comp[M], sigma_b. */
int brr_samples_write_bayesw(brr_samples* w, uint32_t iteration, double alpha, double mu, const double* theta,
                             const double* beta, const int32_t* comp, double sigma_b);
int brr_samples_close(brr_samples* w);
/* 1 if iteration is written: iteration >= burn_in && iteration % thinning == 0 (BayesRBase.cpp:250). */
int brr_sample_due(uint32_t iteration, uint32_t burn_in, uint32_t thinning);

/* ---- host half of runGibbs over the counter-based stream ---- */

/* BayesRBase::init (src/BayesRBase.cpp:39-99): standardise y (N doubles, the
 * phenotype as the reference stores it, i.e. rounded through float), eps = y,
 * sigmaE = |eps|^2 / N / 2, epsilon_sum = sum(eps), pi = prior, sigmaG ~ U(0,1).
 * s->sigmaG (G) and s->pi (G x K) must be provided. */
int brr_bayesr_init(const double* y, uint32_t num_inds, uint32_t num_groups,
                    uint32_t num_components, uint64_t seed, uint32_t chain,
                    double* epsilon, brr_hip_scalars* s);

/* std::random_shuffle of order (src/BayesRBase.cpp:198; libstdc++ rule
 * j = r % (i + 1), i = 1..M-1) with r from the counter-based stream (mode 0)
 * or from the C library rand() (mode 1, the reference's stock source). */
int brr_shuffle(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t num_snps,
                uint32_t* order, int mode);

/* BayesRBase::prepareForAnylsis (src/BayesRBase.cpp:101-110): u ~ U[0,1) and
 * z ~ N(0,1) for every marker id of one iteration. */
int brr_marker_uz(uint64_t seed, uint32_t iteration, uint32_t chain, uint32_t num_snps,
                  double* u, double* z);

#ifdef __cplusplus
}
#endif
#endif
