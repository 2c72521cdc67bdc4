/* sdca.h -- C ABI of the MI355X plmDCA engine (libsdca.so).
 *
 * The entry points replace the per-locus solve of SuperDCA (paths relative to the
 * SuperDCA repository root):
 *   - run_plmDCA<double,triallelic_state_t>     SuperDCA/include/plmDCA.hpp:436-437
 *   - the TBB body plmDCA_solver                 SuperDCA/include/plmDCA.hpp:199-425
 *   - cppoptlib::Problem::value_and_gradient     SuperDCA/include/LBFGS_interface_CppNumericalSolvers.hpp:111-117
 *     -> plmDCA_objective_fval_and_gradient      SuperDCA/include/plmDCA_objective.hpp:38-249
 *   - calculate_weights                          SuperDCA/include/plmDCA_optimizer_parameters.hpp:39-98
 *   - B_eff / lambda initialisation              SuperDCA/include/plmDCA_optimizer_parameters.hpp:257-295
 *
 * Conventions: every call returns SDCA_OK (0) or a negative status; the message of
 * the last failure on the calling thread is sdca_last_error().  The caller owns all
 * host buffers; the library owns device memory and streams.  One context drives one
 * GPU and is not thread-safe (one host thread per context).  Calls block until their
 * results are in host memory.
 */
#ifndef SDCA_H
#define SDCA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDCA_OK 0
#define SDCA_E_ARG (-1)       /* invalid argument or shape */
#define SDCA_E_HIP (-2)       /* HIP runtime failure (allocation, launch) */
#define SDCA_E_NONFINITE (-3) /* objective became NaN/Inf */
#define SDCA_E_NOCONV (-4)    /* at least one locus stopped without convergence */
#define SDCA_E_STATE (-5)     /* call order violated (e.g. solve before weights) */
#define SDCA_E_IO (-6)        /* file could not be read or written */

/* Arithmetic of the energy and gradient contractions. */
#define SDCA_PREC_FP32 0       /* FP32 MFMA (v_mfma_f32_32x32x2f32), FP32 X and W */
#define SDCA_PREC_SPLIT_BF16 1 /* X exact in BF16, W and G as BF16 hi+lo, FP32 accumulation */

/* Per-locus solve status (sdca_locus_stats.status). */
#define SDCA_ST_CONVERGED 0
#define SDCA_ST_ITER_CAP 1
#define SDCA_ST_LS_FAIL 2
#define SDCA_ST_NONFINITE 3

typedef struct sdca_ctx sdca_ctx;

typedef struct sdca_opts {
    int32_t device;          /* HIP device ordinal (default 0) */
    int32_t precision;       /* SDCA_PREC_* (default SDCA_PREC_FP32) */
    int32_t batch_loci;      /* loci solved concurrently on the GPU (default 128; 0: 3072 for the
                                Gram formulation, 128 for the primal one) */
    int32_t history;         /* L-BFGS history m: 1, 3, 5, 8, 10 or 16; the Gram formulation also 24
                                and 32.  0 (default): the formulation's own, 32 for Gram, 10 primal */
    int32_t max_iter;        /* iteration cap per locus (default 2000, plmDCA.hpp:219) */
    int32_t max_ls;          /* function evaluations per line search (default 20) */
    double grad_threshold;   /* stop when ||g||_inf < this (default 1e-3, plmDCA_options.cpp:61) */
    int32_t verbose;         /* per-batch progress on stderr */
    int32_t autotune;        /* 1 (default): time the GEMM tile configurations once per shape */
    int32_t gemm_ksplit;     /* > 0: force the energy split-K slices (no tuning) */
    int32_t gemm_gm_energy;  /* > 0: force the energy XCD super-tile height, 1..64 (no tuning) */
    int32_t gemm_gm_grad;    /* > 0: force the gradient XCD super-tile height, 1..64 (no tuning) */
    int32_t gemm_tn_energy;  /* > 0: force the energy output tile width, 128 or 256 (no tuning) */
    int32_t gemm_tn_grad;    /* > 0: force the gradient output tile width, 128 or 256 (no tuning) */
    int32_t pipeline_groups; /* 2: two slot groups whose rounds overlap on two streams (4: four,
                                Gram formulation only); 1: one; 0: the formulation's default (Gram
                                two, primal one).  With two, the subspace warm start's short
                                restricted rounds run in four groups of half the size */
    int32_t coarse_factor;   /* split-BF16: a locus' GEMMs use the BF16 high parts alone (one MFMA
                                pass) while its ||g||_inf >= coarse_factor * grad_threshold, then
                                its iterate is evaluated again in full and it finishes at split
                                precision (default 4; 0: full precision throughout) */
    int32_t fine_factor;     /* with j_threshold > 0 under split-BF16: a locus moves to the fine
                                level (W and G as three BF16 parts, about 24 bits, i.e. FP32
                                contractions) once ||g||_inf < fine_factor * j_threshold, or when a
                                line search at split precision fails (default 10); 0: no fine
                                level, the coupling threshold is tested at split precision */
    double j_threshold;      /* > 0: a locus is converged only when, besides ||g||_inf <
                                grad_threshold, the couplings' gradient ||g_J||_inf < j_threshold,
                                evaluated at the highest level (the fine one unless fine_factor is
                                0).  The scores depend on J alone; the fields' own
                                gradient converges more slowly (the field column is the sum of every
                                locus' one-hot columns) without moving the scores.  0 (default):
                                SuperDCA's rule alone (plmDCA.hpp:218-221) */
    int32_t formulation;     /* SDCA_FORM_*: how a locus' couplings are represented during its
                                solve (default SDCA_FORM_AUTO) */
    int32_t graph_rounds;    /* 1: a Gram-formulation solve over one alignment captures each slot
                                group's round (command upload, kernels, result copy) as a HIP graph
                                once and replays it every round, with fixed grids over all of the
                                group's slots (idle slots return at once inside the kernels); the
                                round's device time is then reported whole, in lbfgs_ms.  0 (default):
                                the kernels are launched each round over the active slots */
    int32_t subspace_rank;   /* Gram formulation, one alignment, no scan: > 0 (128, the only rank
                                supported) warm-starts every locus from its solve restricted to
                                c = Q a, Q an orthonormal basis of the dominant 128-dimensional
                                eigen-subspace of the identical-locus counts (power iteration,
                                once per alignment); the restricted solve costs two B x 128 x q
                                products per evaluation instead of one B x B x q product.  The
                                optimum is unchanged (strictly convex); the full solve's iterations
                                drop.  Loci start from zero instead (nfeval_sub 0) when B < 256, when
                                the counts' restricted matrix Q^T K Q is singular (too few loci for
                                the rank) or when a call's start products would exceed 64 GiB.
                                0 (default): cold start from zero */
    double subspace_threshold; /* the restricted solve stops at ||g||_2 < this in its own metric
                                (default 3e-2), after at most max_iter iterations */
} sdca_opts;

/* Formulations of the per-locus solve (sdca_opts.formulation).
 *   PRIMAL  the parameter vector itself (q x K_x couplings and fields), energies and
 *           gradients as GEMMs against the one-hot alignment in opts->precision.
 *   GRAM    the couplings as J = X_r^T c with c the q x B coefficients of the sequences: the
 *           L2 term makes 2 lambda_J J = -X_r^T G at the optimum, and every L-BFGS iterate
 *           from zero lies in the same span, so the iterates are those of the primal solve
 *           while each evaluation costs a B x B x q product with the matrix of identical-locus
 *           counts instead of two B x K_x x q ones.  FP64 throughout; cold starts only; the
 *           stopping rule tests ||g||_2 and ||g_J||_2, which bound the infinity norms of the
 *           primal rule from above (sdca_locus_stats.gnorm and gnorm_J report them).  The
 *           couplings of a finished locus are formed once, by one contraction with FP32
 *           products (three BF16 parts of its coefficients).
 *   AUTO    GRAM when no warm start is given and B <= K_x for every alignment of the solve,
 *           else PRIMAL. */
#define SDCA_FORM_AUTO 0
#define SDCA_FORM_PRIMAL 1
#define SDCA_FORM_GRAM 2

/* How sdca_solve_loci_mlp uses the network's coupling blocks (flags). */
#define SDCA_WARM_DIRECTION 0 /* start at the fields; the blocks are the first search direction */
#define SDCA_WARM_POINT 1     /* start at the fields plus the blocks */
#define SDCA_WARM_NO_FIELDS 2 /* or-ed in: fields start at zero instead of the J = 0 optimum */

/* Defaults as documented above. */
void sdca_default_opts(sdca_opts *opts);

/* Per-locus result, plmDCA.hpp:306-315 (fval, gnorm, nfeval, iter) plus status. */
typedef struct sdca_locus_stats {
    double fval;
    double gnorm;
    int32_t nfeval;
    int32_t iter;
    int32_t status;
    int32_t nfeval_first;  /* evaluations up to the first accepted step (start point included) */
    double gnorm_J;        /* ||g||_inf over the coupling entries J_r (the fields h_r excluded);
                              under SDCA_FORM_GRAM gnorm and gnorm_J are 2-norms */
    int32_t nfeval_fine;   /* evaluations at the fine level (FP32-equivalent contractions) */
    int32_t nfeval_sub;    /* evaluations of the subspace warm start (sdca_opts.subspace_rank) */
} sdca_locus_stats;

/* Cumulative device timings of a context (milliseconds, HIP events). */
typedef struct sdca_timing {
    double gemm_energy_ms;   /* E = X W^T contractions */
    double gemm_grad_ms;     /* grad W = G^T X contractions */
    double softmax_ms;       /* LSE / softmax / node beliefs */
    double lbfgs_ms;         /* direction, trial point and dot-product passes */
    double score_ms;         /* gauge + Frobenius scoring */
    double reweight_ms;      /* bit-plane popcount kernel */
    double total_ms;         /* wall time inside sdca_solve_loci */
    int64_t n_evals;         /* batched objective evaluations (rounds) */
    int64_t n_energy_launches;
    int64_t n_grad_launches;
    double flops_energy;     /* algorithmic flops of all energy contractions (2*B*K_x*q per locus-eval) */
    double flops_grad;
    double bytes_lbfgs;      /* algorithmic HBM bytes of the L-BFGS vector passes */
    double update_ms;        /* trial point / new direction pass (part of lbfgs_ms) */
    double dots_ms;          /* gradient dot products + reduction (part of lbfgs_ms) */
    double bytes_update;     /* algorithmic HBM bytes of the update pass */
    double bytes_dots;       /* algorithmic HBM bytes of the dot-product pass */
    double subspace_ms;      /* rounds of the subspace warm start (sdca_opts.subspace_rank) */
    int64_t n_sub_evals;     /* their batched evaluations (rounds) */
    double flops_sub;        /* their algorithmic flops: 4 * B * 128 * q per locus-evaluation */
    double host_ms;          /* host time stepping the slots' L-BFGS state between rounds (Gram) */
    double gstep_ms;         /* the Gram step pass k_gstep alone (part of update_ms) */
    double bytes_gstep;      /* its algorithmic HBM bytes */
} sdca_timing;

/* Message of the calling thread's last failure (SuperDCA prints its errors to
 * plmDCA_options::err_stream() and exits, SuperDCA.cpp:90,178). */
const char *sdca_last_error(void);

/* Creates a context on opts->device over a B x L alignment of envelope states
 * states[b*L + j] in [0, q_env), q_env in {2..8}, gap_state in [0, q_env) (or -1 if
 * none).  The one-hot matrix X (B x K_x, K_x = observed states summed over loci)
 * is built on the device.  Replaces the set-up of run_plmDCA before its solve
 * (plmDCA.hpp:436-578: the alignment, block accounting and parameter dimensions of
 * plmDCA_optimizer_parameters.hpp:131-151); sdca_ctx_destroy ends it. */
int sdca_ctx_create(sdca_ctx **out, const uint8_t *states, int32_t B, int32_t L, int32_t q_env,
                    int32_t gap_state, const sdca_opts *opts);
void sdca_ctx_destroy(sdca_ctx *ctx);

/* Number of observed-state columns K_x of the context's one-hot X (no counterpart in
 * SuperDCA, whose parameter vector spans all L*q*q + q entries, plmDCA_optimizer_parameters.hpp:219-247). */
int64_t sdca_num_columns(const sdca_ctx *ctx);

/* calculate_weights (plmDCA_optimizer_parameters.hpp:39-98), bit-exact:
 * identity = double(L - d_ij)/double(L) >= theta; int_w[i] = m_i + sum of m_j over
 * such neighbours; w[i] = m_i / double(int_w[i]).  mult may be NULL (all 1);
 * int_w_out may be NULL.  Does not install the weights. */
int sdca_reweight(sdca_ctx *ctx, double theta, const uint32_t *mult, double *w_out,
                  uint64_t *int_w_out);

/* sdca_reweight without device allocation.  No counterpart in SuperDCA: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Inputs and outputs are device pointers: d_states
 * (B x L envelope states, row-major, values in [0, q)), d_mult (B multiplicities, may be
 * NULL), d_int_w (B neighbour counts, may be NULL) and d_w (B weights, FP64, the values
 * of sdca_reweight bit for bit).  The kernels run on `stream` (a hipStream_t, NULL for
 * the default stream) of the current device and the call returns without waiting; no
 * hipMalloc, hipFree or memset is issued, so the call may be captured in a HIP graph.
 * The bit-planes and counters live in the library's fixed 64 MiB device scratch array,
 * one per device: B * (8 + 8 * q * ceil(L/64)) bytes (rounded) must fit, else SDCA_E_ARG.
 * Every call on a device shares that device's scratch: calls that overlap on one device
 * (other streams, other threads) race. */
int sdca_reweight_noalloc(const uint8_t *d_states, int32_t B, int32_t L, int32_t q, double theta,
                          const uint32_t *d_mult, uint64_t *d_int_w, double *d_w, void *stream);

/* Installs sequence weights and derives B_eff and lambda_h, lambda_J as in
 * plmDCA_optimizer_parameters.hpp:257-284 (negative user value = automatic rule).
 * out3 (may be NULL) receives {B_eff, lambda_h, lambda_J}. */
int sdca_set_weights(sdca_ctx *ctx, const double *w, double lambda_h_user, double lambda_J_user,
                     double *out3);

/* Parity entry point for plmDCA_objective_fval_and_gradient (plmDCA_objective.hpp:38-249,
 * called through LBFGS_interface_CppNumericalSolvers.hpp:111-117): for each of the K loci
 * loci[k], evaluates f and g at x (K x n, reference layout: n = L*q*q + q, J_r[j](a,s)
 * at 16-locus blocks, Matrix_kernel_access_order.hpp:28-34, h_r at L*q*q), at the fine
 * level when opts.j_threshold > 0, else at the context's precision.  Rows of
 * states never observed at j contribute only their L2 term, which is added on the host
 * in FP64.  g (K x n) is overwritten; the block of locus r itself stays 0. */
int sdca_eval_batch(sdca_ctx *ctx, const int32_t *loci, int32_t K, const double *x, double *f_out,
                    double *g_out);

/* Replaces the tbb::parallel_reduce over plmDCA_solver::operator() (plmDCA.hpp:277-396,
 * 582-586).  Solves the K loci (any K; batched internally, converged slots are refilled) with the
 * batched L-BFGS and writes each locus' mean-of-norms score row (K x L floats,
 * plmDCA.hpp:366-394; row[k][loci[k]] = 0).  warm (K x n reference layout, may be NULL)
 * gives start points; NULL = cold start from zero (plmDCA.hpp:298).  x_out (K x n,
 * reference layout, may be NULL) receives the solutions.  st (K, may be NULL).
 * Returns SDCA_E_NOCONV if any locus hit the iteration cap or a line-search failure
 * (results are still written). */
int sdca_solve_loci(sdca_ctx *ctx, const int32_t *loci, int32_t K, const double *warm,
                    float *score_rows, double *x_out, sdca_locus_stats *st);

/* Two-alignment scan (SuperDCA.cpp:155-159, plmDCA.hpp:482-490): the context's alignment is
 * the first alignment's n_first loci followed by the second's, over the same sequences.
 * Parameters (x_out) keep the layout of that combined alignment, n = (n_first + L2) q q + q,
 * with the blocks over the first alignment zero; the reference's scan parameter vector
 * covers the L2 loci of the second alignment only (plmDCA.hpp:482-490), i.e. the blocks
 * from locus n_first on.
 * With n_first > 0 a solve takes loci of the first alignment only and fits each on the
 * second alignment's columns alone (plmDCA_objective.hpp:67: no block of the second
 * alignment is excluded, the first is not a predictor); its score row is zero over the
 * first alignment.  Install the second alignment's weights (plmDCA.hpp:538).  Cold start
 * only; sdca_eval_batch and sdca_solve_loci_mlp refuse a scanning context.  0 turns the
 * scan off. */
int sdca_set_scan_split(sdca_ctx *ctx, int32_t n_first);

/* ---- several alignments in one batched solve ----
 * SuperDCA solves the loci of one alignment per run (plmDCA.hpp:582-586).  A multi-solve
 * takes loci of several contexts (alignments of any B, L and q_env, each with its weights
 * installed, on one device) into the same slots: every round's kernels cover the loci of
 * all the alignments in flight, each GEMM launch with one group of tiles per alignment.
 * The contexts must outlive the multi-solve.  opts: the solver options (precision,
 * batch_loci, history, thresholds, coarse and fine levels); NULL = defaults. */
typedef struct sdca_multi sdca_multi;
int sdca_multi_create(sdca_multi **out, sdca_ctx *const *ctxs, int32_t n, const sdca_opts *opts);
void sdca_multi_destroy(sdca_multi *m);

/* Solves item k = locus loci[k] of context alns[k] (any order, alignments mixed) with the
 * batched L-BFGS of sdca_solve_loci, cold start.  Score row k (the L of its alignment)
 * goes to score_rows + row_off[k] (row_off NULL: rows packed in item order); score_rows
 * and st may be NULL.  SDCA_E_NOCONV as sdca_solve_loci. */
int sdca_multi_solve(sdca_multi *m, const int32_t *alns, const int32_t *loci, int32_t K, float *score_rows,
                     const int64_t *row_off, sdca_locus_stats *st);
int sdca_multi_get_timing(sdca_multi *m, sdca_timing *t, int32_t reset);

/* Timings accumulated since the context was created or last reset (no counterpart in
 * SuperDCA beyond its per-locus log, plmDCA.hpp:306-325). */
int sdca_get_timing(sdca_ctx *ctx, sdca_timing *t, int32_t reset);

/* GEMM configuration of the current workspace (no counterpart in SuperDCA):
 * cfg[0] = 1 if autotuned (or taken from the tuning cache), cfg[1] = energy split-K
 * slices, cfg[2] = columns per slice, cfg[3] / cfg[4] = XCD super-tile height of the
 * energy / gradient contraction, cfg[5] / cfg[6] = their output tile widths.
 * SDCA_E_STATE before the first evaluation. */
int sdca_gemm_config(const sdca_ctx *ctx, int32_t cfg[7]);

/* ---- pair statistics for the warm start (no counterpart in SuperDCA) ---- */

/* Exact co-occurrence counts N_ij(a,b) = #{sequences with state a at locus i and b at j}
 * for npairs listed pairs (pairs[2p], pairs[2p+1]); counts_out is npairs x q x q int32.
 * Popcounts of per-locus sequence bit-planes. */
int sdca_pair_counts(sdca_ctx *ctx, const int32_t *pairs, int64_t npairs, int32_t *counts_out);

/* Rows [i0*QP, (i0+nloci)*QP) of the full pair-count table X8 X8^T (QP = 4 for q <= 4,
 * else 8; row i*QP + a, column j*QP + b holds N_ij(a,b)), computed with int8 MFMA into
 * int32.  out is (nloci*QP) x (L*QP) int32.  For checking; the table is never stored whole. */
int sdca_pair_count_table(sdca_ctx *ctx, int32_t i0, int32_t nloci, int32_t *out);

/* Mutual information MI_ij = sum_ab f_ab log(f_ab / (f_a f_b)) of the unweighted pair
 * frequencies for all locus pairs (int8-MFMA count tiles reduced in the epilogue), and
 * for every locus the K partners of largest MI, ties to the lower index, listed in
 * ascending index order: partners_out is L x K; mi_out (L x K, may be NULL) their MI. */
int sdca_mi_topk(sdca_ctx *ctx, int32_t K, int32_t *partners_out, float *mi_out);

/* ---- warm starts: pair features and the learned network (no counterpart in SuperDCA) ---- */

/* Features of locus pairs (r, j) = (pairs[2p], pairs[2p+1]): the joint frequencies
 * f_rj(s, a) at [s*q + a], the marginals f_r(s), f_j(a) and the pair's mean-field coupling
 * estimate J_mf(s, a) at [q*q + 2q + s*q + a]: 2q*q + 2q floats per pair (unweighted
 * frequencies from exact counts).  J_mf = -(C^-1)_rj of the pair's connected correlations
 * over the first q-1 states of each locus, with pseudocount weight 1/2; 0 for the last
 * state. */
int sdca_pair_features(sdca_ctx *ctx, const int32_t *pairs, int64_t npairs, float *feats);

/* The warm-start network: features -> hidden -> hidden (ReLU) -> q*q outputs, where output
 * s*q + a is the start value of J_r[j](a, s).  hidden is a multiple of 32 in [32, 1024].
 * FP32 master parameters on the device; the layers run as BF16 products on MFMA with FP32
 * accumulation (weights, activations and backward output gradients rounded to BF16).
 * sdca_mlp_init draws He-uniform weights from seed; sdca_mlp_set_params installs given
 * weights (the layout W1 [H][Din], b1, W2 [H][H], b2, W3 [Dout][H], b3). */
int sdca_mlp_init(sdca_ctx *ctx, int32_t hidden, uint64_t seed);
int64_t sdca_mlp_num_params(const sdca_ctx *ctx);
int sdca_mlp_get_params(sdca_ctx *ctx, float *out, int64_t n);
int sdca_mlp_set_params(sdca_ctx *ctx, int32_t hidden, const float *params, int64_t n);
int sdca_mlp_forward(sdca_ctx *ctx, const float *x, int64_t n, float *y);

/* Masked mean squared error (1/sum(mask)) sum mask (y - t)^2 of n rows and its gradient
 * with respect to all parameters (hand-written backward kernels). mask may be NULL. */
int sdca_mlp_loss_grad(sdca_ctx *ctx, const float *x, const float *t, const float *mask, int64_t n, double *loss,
                       float *grad);

/* Adam training on the device: epochs over the n rows in seeded shuffled mini-batches;
 * epoch_loss (epochs entries, may be NULL) receives the mean masked squared error. */
int sdca_mlp_train(sdca_ctx *ctx, const float *x, const float *t, const float *mask, int64_t n, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double *epoch_loss);

/* sdca_solve_loci with warm starts built on the device: for each locus its npartners
 * largest-MI partners (sdca_mi_topk) receive the network's blocks, all other couplings
 * start at zero.  warm_mode (SDCA_WARM_* flags): the blocks are added to the start point
 * (POINT) or form the first search direction (DIRECTION); the fields start at the J = 0
 * optimum, or at zero with NO_FIELDS.  The optimum is the cold start's (the problem is
 * strictly convex); only the iteration count changes. */
int sdca_solve_loci_mlp(sdca_ctx *ctx, const int32_t *loci, int32_t K, int32_t npartners, int32_t warm_mode,
                        float *score_rows, double *x_out, sdca_locus_stats *st);

/* ---- host-side alignment pipeline (the apegrunt contract the CLI relies on) ---- */

/* MAF/gap filter and q_env-state transform (SuperDCA.cpp:274-277,354; README.md:32):
 * chars is B x L_in (row-major, FASTA letters).  A locus is kept if it has more than
 * one non-gap allele, minor-allele frequency (second most frequent allele over the
 * non-gap count) >= maf_threshold and gap fraction <= gap_threshold, unless no_filter;
 * alleles are then numbered by descending count (ties A<C<G<T), the gap is state
 * q_env-1, and loci with more than q_env-1 alleles are dropped.  states_out (B x L_in
 * capacity) and loci_out (L_in capacity, original column indices) receive the result;
 * *L_out the number of kept loci. */
int sdca_prepare_alignment(const char *chars, int32_t B, int32_t L_in, double maf_threshold,
                           double gap_threshold, int32_t no_filter, int32_t q_env,
                           uint8_t *states_out, int32_t *loci_out, int32_t *L_out);

/* Writes the symmetrised couplings "score i j" (plmDCA.hpp:667-693): for r over the
 * list in order and n before r: (float(S[r][n]) + float(S[n][r])) * 0.5 printed %.8f,
 * then translation[r]+base and translation[n]+base.  S is L x L (row = locus). */
int sdca_write_couplings(const char *path, const float *S, int32_t L, const int32_t *translation,
                         int32_t base);

/* The largest symmetrised scores only, in the line format of sdca_write_couplings: the
 * pairs with a value >= min_score (NaN never), the top_n largest of them (all when
 * top_n == 0), largest first and equal values in the order of the full file.  No
 * counterpart in SuperDCA, whose .all file lists every pair (plmDCA.hpp:667-693); its
 * users rank that file afterwards (the top pairs of the paper's S1 table). */
int sdca_write_couplings_top(const char *path, const float *S, int32_t L, const int32_t *translation,
                             int32_t base, int64_t top_n, double min_score);

/* The scan's output (plmDCA.hpp:625-638,672-681): for each of the K first-alignment loci
 * (row k of S, K x L2: its scores over the second alignment) and every second-alignment
 * locus n, the line "score r n" with the float score printed %.8f, r = translation1[k] +
 * base and n = translation2[n] + base; not symmetrised. */
int sdca_write_scan_couplings(const char *path, const float *S, int32_t K, int32_t L2,
                              const int32_t *translation1, const int32_t *translation2, int32_t base);

/* ---- file-backed stores and streaming writers (genome-scale output) ----
 * SuperDCA holds every score, or with --norm-of-mean-scoring every q x q coupling block,
 * of a run in memory and writes its output from there (plmDCA.hpp:86-190,589-697): 35 GB
 * of scores at 94,000 loci, and far more blocks.  A store keeps them in an unlinked file
 * instead, rows x cols entries of `entry` floats; the writers read it in tiles of rows, so
 * host memory stays at two bands of about 256 MB (or `tile` rows of each). */
typedef struct sdca_store sdca_store;

/* Creates the store's file at path (which must not exist; it is unlinked at once and
 * its space returned by sdca_store_destroy or the end of the process). */
int sdca_store_create(sdca_store **out, const char *path, int64_t rows, int64_t cols, int32_t entry);
void sdca_store_destroy(sdca_store *s);

/* Rows [r0, r0 + n) of the store: n x cols x entry floats, row-major. */
int sdca_store_put_rows(sdca_store *s, int64_t r0, int64_t n, const float *rows);

/* Row `row` of a store with entry q*q: for each column u the coupling block of locus r
 * with locus col_loci[u] from the solution x (reference layout, L*q*q + q doubles), as
 * float and transposed when col_loci[u] < r, so that both blocks of a pair run [state at
 * the larger locus][state at the smaller] (the storage of plmDCA.hpp:335-363, without the
 * row aliasing of its get_Ji_matrices at :133-137); the block of r itself is zero. */
int sdca_store_put_blocks(sdca_store *s, int64_t row, const double *x, int32_t L, int32_t q, int32_t r,
                          const int32_t *col_loci);

/* sdca_write_couplings and sdca_write_couplings_top over a square store of scores (entry
 * 1, row and column k = list position k), byte for byte the same files; either path may
 * be NULL.  tile: rows per tile (0: automatic). */
int sdca_store_write_couplings(const sdca_store *s, const char *path_all, const char *path_top,
                               const int32_t *translation, int32_t base, int64_t top_n, double min_score,
                               int32_t tile);

/* The --norm-of-mean-scoring output (plmDCA.hpp:617-666) from a square store of q*q
 * blocks: for r over the list and n before r, "J r n Jij Jji (Jij+Jji)/2" with Jij, Jji
 * the norms (gap row and column excluded) of the two gauged blocks and J the norm of their
 * mean, in the reference's float arithmetic; values %.8f. */
int sdca_store_write_norm_of_mean(const sdca_store *s, const char *path, const int32_t *translation, int32_t base,
                                  int32_t q, int32_t gap, int32_t tile);

/* The two-alignment scan's --norm-of-mean-scoring output (plmDCA.hpp:625-638): one line
 * "Jij r n" per block of the store (rows: first-alignment loci, columns: second). */
int sdca_store_write_norm_of_mean_scan(const sdca_store *s, const char *path, const int32_t *translation1,
                                       const int32_t *translation2, int32_t base, int32_t q, int32_t gap);

#ifdef __cplusplus
}
#endif
#endif
