/* plmdca_oracle.h -- FP64 CPU restatement of SuperDCA's per-locus plmDCA path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP engine in
 * superdca_amd/: tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it; the product (libsdca.so, the sdca CLI) never links it.
 *
 * Every function names the reference code it restates (paths relative to the
 * SuperDCA repository root).  Parity pinning per function:
 *   orc_ref_index        pinned: bijection test + layout of plmDCA_optimizer_parameters.hpp:245-247
 *   orc_hamming/weights  pinned: brute-force Python loops, closed-form special cases
 *   orc_lambda           pinned: closed form of plmDCA_optimizer_parameters.hpp:261-284
 *   orc_objective        pinned: central finite differences, L=2 multinomial-logistic
 *                        cross-check against SciPy (committed fixture), J=0 closed form,
 *                        sequence-permutation invariance
 *   orc_objective_blocked pinned: equality with orc_objective (same maths, other order)
 *   orc_lbfgs            pinned by the optimum only (strict convexity): gradient
 *                        certificate and agreement with SciPy L-BFGS-B; the trajectory
 *                        of CppNumericalSolvers (not vendored) is parity unpinned
 *   orc_score_row        pinned: hand-computed 4x4 blocks, textbook gauge for symmetric
 *                        blocks, transpose and constant-shift invariance
 *   orc_nom_*            pinned: a hand-computed block in exactly representable values,
 *                        agreement with the FP64 gauge to float rounding, the mean of a block
 *                        with itself, the triangle inequality, and the block orientation
 *                        against the reference layout
 */
#ifndef PLMDCA_ORACLE_H
#define PLMDCA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-locus solve statistics (plmDCA.hpp:306-315: fval, gnorm, nfeval, iter). */
typedef struct orc_stats {
    double fval;
    double gnorm; /* infinity norm of the final gradient */
    int32_t nfeval;
    int32_t iter;
    int32_t status; /* 0 converged, 1 iteration cap, 2 line search stalled, 3 non-finite */
    int32_t pad;
} orc_stats;

/* Number of FP64 parameters of one locus in the reference layout: L*q*q (J_r) + q (h_r). */
size_t orc_nparams(int32_t L, int32_t q);

/* Offset of J_r[j](a, s) in the reference parameter vector: 16-locus blocks,
 * STATES access order (Matrix_kernel_access_order.hpp:28-34,
 * Coupling_matrix_view.hpp:137-160).  h_r(s) sits at L*q*q + s. */
size_t orc_ref_index(int32_t L, int32_t q, int32_t j, int32_t a, int32_t s);

/* Mismatch count d_ij for all i>j, packed lower-triangular at i*(i-1)/2+j
 * (the distance_matrix() contract used at plmDCA_optimizer_parameters.hpp:56,68). */
void orc_hamming(const uint8_t *states, int32_t B, int32_t L, uint32_t *dmat, int32_t nthreads);

/* calculate_weights (plmDCA_optimizer_parameters.hpp:39-98). mult may be NULL (all 1).
 * w[i] = mult_i / int_w[i]; int_w may be NULL. */
void orc_weights(const uint8_t *states, int32_t B, int32_t L, double theta, const uint32_t *mult,
                 double *w, uint64_t *int_w, int32_t nthreads);

/* B_eff and the regularisation strengths (plmDCA_optimizer_parameters.hpp:257-284).
 * A negative user value selects the automatic rule. */
void orc_lambda(const double *w, int32_t B, double lambda_h_user, double lambda_J_user,
                double *B_eff, double *lambda_h, double *lambda_J);

/* Value and gradient of the asymmetric pseudolikelihood of column r
 * (plmDCA_objective.hpp:38-249), reference layout; g is overwritten. */
double orc_objective(const uint8_t *states, int32_t B, int32_t L, int32_t q, int32_t r,
                     const double *w, double lambda_h, double lambda_J, const double *x, double *g);

/* Block-compressed evaluation: unique 16-locus state patterns per block, the
 * organisation of plmDCA_objective.hpp:89-121,155-189.  Same value and gradient
 * as orc_objective up to summation order; used as the CPU baseline. */
typedef struct orc_blocks orc_blocks;
orc_blocks *orc_blocks_build(const uint8_t *states, int32_t B, int32_t L, int32_t q);
void orc_blocks_free(orc_blocks *blk);
int64_t orc_blocks_npatterns(const orc_blocks *blk);
double orc_objective_blocked(const orc_blocks *blk, int32_t r, const double *w, double lambda_h,
                             double lambda_J, const double *x, double *g, double *work);

/* Objective callback for orc_lbfgs. */
typedef double (*orc_objfn)(void *ctx, const double *x, double *g);

/* FP64 L-BFGS (history m, two-loop recursion) with the More-Thuente line search,
 * stopping when ||g||_inf < gtol or after max_iter iterations
 * (plmDCA.hpp:218-221,298,304).  Below the rounding level of f's FP64 sum a step is
 * accepted on the approximate Wolfe conditions of Hager and Zhang, so the solve
 * reaches gradients far below 1e-6.  x holds the start point on entry. */
int32_t orc_lbfgs(orc_objfn fn, void *ctx, int64_t n, double *x, int32_t m, double gtol,
                  int32_t max_iter, orc_stats *st);

/* Mean-of-norms score row of locus r (plmDCA.hpp:366-394): for every n != r the
 * Frobenius norm, excluding the gap row and column, of the Ising-gauged block
 * (plmDCA_utility.hpp:57-83, Matrix_math.hpp:134-178), stored as float.
 * row[r] is set to 0. */
void orc_score_row(const double *x, int32_t L, int32_t q, int32_t r, int32_t gap, float *row);

/* The gauged block itself (for tests): out[i*q+k]. */
void orc_ising_gauge(const double *M, int32_t q, int32_t transpose, double *out);
double orc_frobenius_excl(const double *G, int32_t q, int32_t exclude);

/* --norm-of-mean-scoring (plmDCA.hpp:335-363, 617-666): the stored float block of the pair
 * (r, n) from r's solution x (transposed when n < r); the float Ising gauge and norm; and
 * the line values {norm of the mean of the two gauged blocks, norm of G(Mrn), of G(Mnr)}. */
void orc_nom_block(const double *x, int32_t L, int32_t q, int32_t r, int32_t n, float *M);
void orc_ising_gauge_float(const float *M, int32_t q, float *out);
float orc_frobenius_excl_float(const float *G, int32_t q, int32_t exclude);
void orc_nom_pair(const float *Mrn, const float *Mnr, int32_t q, int32_t gap, float *out3);

/* Solve a list of loci, one locus per thread (the TBB grain-1 scheme of
 * plmDCA.hpp:582-586), cold start from zero unless x0 (K x n) is given.
 * score_rows (K x L, may be NULL), x_out (K x n, may be NULL), st (K). */
int32_t orc_solve_loci(const uint8_t *states, int32_t B, int32_t L, int32_t q, int32_t gap,
                       const double *w, double lambda_h, double lambda_J, const int32_t *loci,
                       int32_t K, double gtol, int32_t max_iter, int32_t m, int32_t nthreads,
                       int32_t blocked, const double *x0, float *score_rows, double *x_out,
                       orc_stats *st);

#ifdef __cplusplus
}
#endif
#endif
