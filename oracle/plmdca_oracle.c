/* plmdca_oracle.c -- FP64 CPU restatement of SuperDCA's per-locus plmDCA path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine, never linked into it.
 * See plmdca_oracle.h for the pinning of each function.  Reference citations are
 * relative to the SuperDCA repository root.
 */
#define _GNU_SOURCE
#include "plmdca_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define ORC_BLOCK 16 /* apegrunt::StateBlock_size (16 per the v0.1.1 changelog) */

/* ------------------------------------------------------------------ layout */

size_t orc_nparams(int32_t L, int32_t q) { return (size_t)L * q * q + q; }

/* plmDCA_optimizer_parameters.hpp:245 (Jr_size = L*N*N), Coupling_matrix_view.hpp:137-160
 * (block base bn*BlockSize*N*N, last block extent L % BlockSize),
 * Matrix_kernel_access_order.hpp:30-33 ((state*extent + pos)*N). */
size_t orc_ref_index(int32_t L, int32_t q, int32_t j, int32_t a, int32_t s)
{
    const size_t blk = (size_t)j / ORC_BLOCK;
    const size_t pos = (size_t)j % ORC_BLOCK;
    const size_t ext = ((blk + 1) * ORC_BLOCK > (size_t)L) ? (size_t)L % ORC_BLOCK : ORC_BLOCK;
    return blk * ORC_BLOCK * q * q + ((size_t)a * ext + pos) * q + (size_t)s;
}

/* ------------------------------------------------------------- reweighting */

typedef struct {
    const uint8_t *states;
    int32_t B, L;
    uint32_t *dmat;
    int32_t next; /* atomic row counter */
} ham_job;

static void *ham_worker(void *arg)
{
    ham_job *jb = (ham_job *)arg;
    for (;;) {
        const int32_t i = __atomic_fetch_add(&jb->next, 1, __ATOMIC_RELAXED);
        if (i >= jb->B) break;
        const uint8_t *si = jb->states + (size_t)i * jb->L;
        for (int32_t j = 0; j < i; ++j) {
            const uint8_t *sj = jb->states + (size_t)j * jb->L;
            uint32_t d = 0;
            for (int32_t l = 0; l < jb->L; ++l) d += (si[l] != sj[l]);
            jb->dmat[(size_t)i * (i - 1) / 2 + j] = d;
        }
    }
    return NULL;
}

/* apegrunt distance_matrix() as used at plmDCA_optimizer_parameters.hpp:56,68:
 * mismatching positions of the 4-state sequences, the gap being one more state. */
void orc_hamming(const uint8_t *states, int32_t B, int32_t L, uint32_t *dmat, int32_t nthreads)
{
    ham_job jb = {states, B, L, dmat, 0};
    if (nthreads < 1) nthreads = 1;
    pthread_t th[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, ham_worker, &jb);
    ham_worker(&jb);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* calculate_weights, plmDCA_optimizer_parameters.hpp:39-98. */
void orc_weights(const uint8_t *states, int32_t B, int32_t L, double theta, const uint32_t *mult,
                 double *w, uint64_t *int_w_out, int32_t nthreads)
{
    const size_t npairs = (size_t)B * (B > 0 ? B - 1 : 0) / 2;
    uint32_t *dmat = (uint32_t *)malloc((npairs ? npairs : 1) * sizeof(uint32_t));
    orc_hamming(states, B, L, dmat, nthreads);
    uint64_t *iw = (uint64_t *)malloc((B ? B : 1) * sizeof(uint64_t));
    for (int32_t i = 0; i < B; ++i) iw[i] = mult ? mult[i] : 1u; /* :49-50 */
    for (int32_t i = 0; i < B; ++i) {
        const uint64_t mi = mult ? mult[i] : 1u;
        for (int32_t j = 0; j < i; ++j) {
            const size_t n_elem = (size_t)L; /* min(len_i, len_j): equal lengths */
            const double identity =
                (double)(n_elem - dmat[(size_t)i * (i - 1) / 2 + j]) / (double)n_elem; /* :67-68 */
            if (identity >= theta) { /* :72-77 */
                iw[i] += mult ? mult[j] : 1u;
                iw[j] += mi;
            }
        }
    }
    for (int32_t i = 0; i < B; ++i) {
        const double mi = mult ? (double)mult[i] : 1.0;
        w[i] = mi / (double)iw[i]; /* :95 */
        if (int_w_out) int_w_out[i] = iw[i];
    }
    free(iw);
    free(dmat);
}

/* plmDCA_optimizer_parameters.hpp:257-284. */
void orc_lambda(const double *w, int32_t B, double lambda_h_user, double lambda_J_user,
                double *B_eff, double *lambda_h, double *lambda_J)
{
    double beff = 0.0;
    for (int32_t i = 0; i < B; ++i) beff += w[i]; /* std::accumulate from 0.0, :261 */
    const double auto_lambda = (beff > 500.0) ? 0.1 : 1.0 - ((1.0 - 0.1) * beff / 500.0); /* :263 */
    *lambda_J = (lambda_J_user < 0.0) ? auto_lambda * beff / 2.0 : lambda_J_user * beff;
    *lambda_h = (lambda_h_user < 0.0) ? auto_lambda * beff : lambda_h_user * beff;
    *B_eff = beff;
}

/* --------------------------------------------------------------- objective */

/* plmDCA_objective.hpp:38-249, one sequence at a time. */
double orc_objective(const uint8_t *states, int32_t B, int32_t L, int32_t q, int32_t r,
                     const double *w, double lambda_h, double lambda_J, const double *x, double *g)
{
    const size_t n = orc_nparams(L, q);
    const size_t hoff = (size_t)L * q * q;
    memset(g, 0, n * sizeof(double)); /* set_gradient(g, clear=true), plmDCA_optimizer_parameters.hpp:207 */
    double E[16], nb[16];
    double f = 0.0;
    for (int32_t b = 0; b < B; ++b) {
        const uint8_t *row = states + (size_t)b * L;
        for (int32_t s = 0; s < q; ++s) E[s] = x[hoff + s]; /* :86 */
        for (int32_t j = 0; j < L; ++j) {                  /* :88-122, j != r */
            if (j == r) continue;
            const double *J = x + orc_ref_index(L, q, j, row[j], 0);
            for (int32_t s = 0; s < q; ++s) E[s] += J[s];
        }
        const int32_t xr = row[r];
        double z = 0.0;
        for (int32_t s = 0; s < q; ++s) z += exp(E[s]);
        const double lz = log(z); /* :140, no max shift */
        f += w[b] * (lz - E[xr]); /* :143 */
        for (int32_t s = 0; s < q; ++s) nb[s] = w[b] * exp(E[s] - lz); /* :147 */
        nb[xr] -= w[b];                                                /* :148 */
        for (int32_t s = 0; s < q; ++s) g[hoff + s] += nb[s];           /* :149 */
        for (int32_t j = 0; j < L; ++j) {                              /* :153-190 */
            if (j == r) continue;
            double *G = g + orc_ref_index(L, q, j, row[j], 0);
            for (int32_t s = 0; s < q; ++s) G[s] += nb[s];
        }
    }
    /* L2 terms, :192-244 */
    for (int32_t s = 0; s < q; ++s) {
        f += lambda_h * x[hoff + s] * x[hoff + s];
        g[hoff + s] += 2.0 * lambda_h * x[hoff + s];
    }
    for (int32_t j = 0; j < L; ++j) {
        if (j == r) continue;
        for (int32_t a = 0; a < q; ++a) {
            const size_t o = orc_ref_index(L, q, j, a, 0);
            for (int32_t s = 0; s < q; ++s) {
                f += lambda_J * x[o + s] * x[o + s];
                g[o + s] += 2.0 * lambda_J * x[o + s];
            }
        }
    }
    return f;
}

/* ------------------------------------------------- block-compressed variant */

struct orc_blocks {
    int32_t B, L, q, nblk;
    int32_t *npat;       /* [nblk] */
    int64_t *pat_off;    /* [nblk] offset into pat_states (rows of 16) and pat_seq_off */
    uint8_t *pat_states; /* [total_pat][16] */
    int32_t *seq_off;    /* [total_pat + nblk] CSR offsets per block */
    int32_t *seqs;       /* [nblk][B] */
    const uint8_t *states;
};

typedef struct {
    uint64_t key;
    int32_t b;
} keyb;

static int keyb_cmp(const void *pa, const void *pb)
{
    const keyb *a = (const keyb *)pa, *b = (const keyb *)pb;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->b - b->b;
}

orc_blocks *orc_blocks_build(const uint8_t *states, int32_t B, int32_t L, int32_t q)
{
    orc_blocks *k = (orc_blocks *)calloc(1, sizeof(orc_blocks));
    k->B = B;
    k->L = L;
    k->q = q;
    k->states = states;
    k->nblk = (L + ORC_BLOCK - 1) / ORC_BLOCK;
    k->npat = (int32_t *)calloc(k->nblk, sizeof(int32_t));
    k->pat_off = (int64_t *)calloc(k->nblk + 1, sizeof(int64_t));
    k->seqs = (int32_t *)malloc((size_t)k->nblk * B * sizeof(int32_t));
    keyb *kb = (keyb *)malloc((size_t)B * sizeof(keyb));
    /* first pass: count patterns */
    int64_t total = 0;
    uint8_t **tmp_states = (uint8_t **)calloc(k->nblk, sizeof(uint8_t *));
    int32_t **tmp_off = (int32_t **)calloc(k->nblk, sizeof(int32_t *));
    for (int32_t blk = 0; blk < k->nblk; ++blk) {
        const int32_t j0 = blk * ORC_BLOCK;
        const int32_t ext = (j0 + ORC_BLOCK > L) ? L - j0 : ORC_BLOCK;
        for (int32_t b = 0; b < B; ++b) {
            uint64_t key = 0;
            for (int32_t p = 0; p < ext; ++p) key |= (uint64_t)states[(size_t)b * L + j0 + p] << (4 * p);
            kb[b].key = key;
            kb[b].b = b;
        }
        qsort(kb, B, sizeof(keyb), keyb_cmp);
        int32_t np = 0;
        for (int32_t b = 0; b < B; ++b)
            if (b == 0 || kb[b].key != kb[b - 1].key) ++np;
        k->npat[blk] = np;
        tmp_states[blk] = (uint8_t *)malloc((size_t)np * ORC_BLOCK);
        tmp_off[blk] = (int32_t *)malloc((size_t)(np + 1) * sizeof(int32_t));
        int32_t pi = -1;
        for (int32_t b = 0; b < B; ++b) {
            if (b == 0 || kb[b].key != kb[b - 1].key) {
                ++pi;
                tmp_off[blk][pi] = b;
                for (int32_t p = 0; p < ORC_BLOCK; ++p)
                    tmp_states[blk][(size_t)pi * ORC_BLOCK + p] = (uint8_t)((kb[b].key >> (4 * p)) & 15u);
            }
            k->seqs[(size_t)blk * B + b] = kb[b].b;
        }
        tmp_off[blk][np] = B;
        k->pat_off[blk] = total;
        total += np;
    }
    k->pat_off[k->nblk] = total;
    k->pat_states = (uint8_t *)malloc((size_t)(total ? total : 1) * ORC_BLOCK);
    k->seq_off = (int32_t *)malloc((size_t)(total + k->nblk) * sizeof(int32_t));
    for (int32_t blk = 0; blk < k->nblk; ++blk) {
        memcpy(k->pat_states + k->pat_off[blk] * ORC_BLOCK, tmp_states[blk], (size_t)k->npat[blk] * ORC_BLOCK);
        memcpy(k->seq_off + k->pat_off[blk] + blk, tmp_off[blk], (size_t)(k->npat[blk] + 1) * sizeof(int32_t));
        free(tmp_states[blk]);
        free(tmp_off[blk]);
    }
    free(tmp_states);
    free(tmp_off);
    free(kb);
    return k;
}

void orc_blocks_free(orc_blocks *k)
{
    if (!k) return;
    free(k->npat);
    free(k->pat_off);
    free(k->pat_states);
    free(k->seq_off);
    free(k->seqs);
    free(k);
}

int64_t orc_blocks_npatterns(const orc_blocks *k) { return k->pat_off[k->nblk]; }

/* plmDCA_objective.hpp:88-122 (energies per unique block pattern) and :153-190
 * (gradient scatter), with the per-pattern node beliefs summed before the scatter. */
double orc_objective_blocked(const orc_blocks *k, int32_t r, const double *w, double lambda_h,
                             double lambda_J, const double *x, double *g, double *work)
{
    const int32_t B = k->B, L = k->L, q = k->q;
    const size_t n = orc_nparams(L, q);
    const size_t hoff = (size_t)L * q * q;
    double *E = work;              /* B x q */
    double *nb = work + (size_t)B * q; /* B x q */
    memset(g, 0, n * sizeof(double));
    for (int32_t b = 0; b < B; ++b)
        for (int32_t s = 0; s < q; ++s) E[(size_t)b * q + s] = x[hoff + s];
    const int32_t rblk = r / ORC_BLOCK, rloc = r % ORC_BLOCK;
    double acc[16];
    for (int32_t blk = 0; blk < k->nblk; ++blk) {
        const int32_t j0 = blk * ORC_BLOCK;
        const int32_t ext = (j0 + ORC_BLOCK > L) ? L - j0 : ORC_BLOCK;
        const double *Jb = x + (size_t)blk * ORC_BLOCK * q * q;
        const int32_t *so = k->seq_off + k->pat_off[blk] + blk;
        const int32_t *sq = k->seqs + (size_t)blk * B;
        for (int32_t p = 0; p < k->npat[blk]; ++p) {
            const uint8_t *ps = k->pat_states + (k->pat_off[blk] + p) * ORC_BLOCK;
            for (int32_t s = 0; s < q; ++s) acc[s] = 0.0;
            for (int32_t pos = 0; pos < ext; ++pos) {
                if (blk == rblk && pos == rloc) continue;
                const double *J = Jb + ((size_t)ps[pos] * ext + pos) * q;
                for (int32_t s = 0; s < q; ++s) acc[s] += J[s];
            }
            for (int32_t t = so[p]; t < so[p + 1]; ++t) {
                double *Eb = E + (size_t)sq[t] * q;
                for (int32_t s = 0; s < q; ++s) Eb[s] += acc[s];
            }
        }
    }
    double f = 0.0;
    for (int32_t b = 0; b < B; ++b) {
        const double *Eb = E + (size_t)b * q;
        double *nbb = nb + (size_t)b * q;
        const int32_t xr = k->states[(size_t)b * L + r];
        double z = 0.0;
        for (int32_t s = 0; s < q; ++s) z += exp(Eb[s]);
        const double lz = log(z);
        f += w[b] * (lz - Eb[xr]);
        for (int32_t s = 0; s < q; ++s) nbb[s] = w[b] * exp(Eb[s] - lz);
        nbb[xr] -= w[b];
        for (int32_t s = 0; s < q; ++s) g[hoff + s] += nbb[s];
    }
    for (int32_t blk = 0; blk < k->nblk; ++blk) {
        const int32_t j0 = blk * ORC_BLOCK;
        const int32_t ext = (j0 + ORC_BLOCK > L) ? L - j0 : ORC_BLOCK;
        double *Gb = g + (size_t)blk * ORC_BLOCK * q * q;
        const int32_t *so = k->seq_off + k->pat_off[blk] + blk;
        const int32_t *sq = k->seqs + (size_t)blk * B;
        for (int32_t p = 0; p < k->npat[blk]; ++p) {
            const uint8_t *ps = k->pat_states + (k->pat_off[blk] + p) * ORC_BLOCK;
            for (int32_t s = 0; s < q; ++s) acc[s] = 0.0;
            for (int32_t t = so[p]; t < so[p + 1]; ++t) {
                const double *nbb = nb + (size_t)sq[t] * q;
                for (int32_t s = 0; s < q; ++s) acc[s] += nbb[s];
            }
            for (int32_t pos = 0; pos < ext; ++pos) {
                if (blk == rblk && pos == rloc) continue;
                double *G = Gb + ((size_t)ps[pos] * ext + pos) * q;
                for (int32_t s = 0; s < q; ++s) G[s] += acc[s];
            }
        }
    }
    for (int32_t s = 0; s < q; ++s) {
        f += lambda_h * x[hoff + s] * x[hoff + s];
        g[hoff + s] += 2.0 * lambda_h * x[hoff + s];
    }
    for (int32_t blk = 0; blk < k->nblk; ++blk) {
        const int32_t j0 = blk * ORC_BLOCK;
        const int32_t ext = (j0 + ORC_BLOCK > L) ? L - j0 : ORC_BLOCK;
        const size_t base = (size_t)blk * ORC_BLOCK * q * q;
        for (int32_t a = 0; a < q; ++a)
            for (int32_t pos = 0; pos < ext; ++pos) {
                if (blk == rblk && pos == rloc) continue;
                const size_t o = base + ((size_t)a * ext + pos) * q;
                for (int32_t s = 0; s < q; ++s) {
                    f += lambda_J * x[o + s] * x[o + s];
                    g[o + s] += 2.0 * lambda_J * x[o + s];
                }
            }
    }
    return f;
}

/* ---------------------------------------------------- More-Thuente search */

/* Line search of J. J. More and D. J. Thuente, "Line search algorithms with
 * guaranteed sufficient decrease", ACM TOMS 20 (1994) 286-307: the safeguarded
 * step (dcstep) and the two-stage driver (dcsrch), restated from the paper. */
typedef struct {
    int brackt, stage;
    double ftol, gtol, xtol, stpmin, stpmax;
    double finit, ginit, gtest, width, width1;
    double stx, fx, gx, sty, fy, gy, stmin, stmax;
} mt_state;

static void mt_dcstep(double *stx, double *fx, double *dx, double *sty, double *fy, double *dy,
                      double *stp, double fp, double dp, int *brackt, double stpmin, double stpmax)
{
    const double sgnd = dp * (*dx / fabs(*dx));
    double stpf, stpc, stpq, theta, s, gamma, p, qq, r;
    if (fp > *fx) {
        theta = 3.0 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = fmax(fabs(theta), fmax(fabs(*dx), fabs(dp)));
        gamma = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
        if (*stp < *stx) gamma = -gamma;
        p = (gamma - *dx) + theta;
        qq = ((gamma - *dx) + gamma) + dp;
        r = p / qq;
        stpc = *stx + r * (*stp - *stx);
        stpq = *stx + ((*dx / ((*fx - fp) / (*stp - *stx) + *dx)) / 2.0) * (*stp - *stx);
        if (fabs(stpc - *stx) < fabs(stpq - *stx)) stpf = stpc;
        else stpf = stpc + (stpq - stpc) / 2.0;
        *brackt = 1;
    } else if (sgnd < 0.0) {
        theta = 3.0 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = fmax(fabs(theta), fmax(fabs(*dx), fabs(dp)));
        gamma = s * sqrt((theta / s) * (theta / s) - (*dx / s) * (dp / s));
        if (*stp > *stx) gamma = -gamma;
        p = (gamma - dp) + theta;
        qq = ((gamma - dp) + gamma) + *dx;
        r = p / qq;
        stpc = *stp + r * (*stx - *stp);
        stpq = *stp + (dp / (dp - *dx)) * (*stx - *stp);
        stpf = (fabs(stpc - *stp) > fabs(stpq - *stp)) ? stpc : stpq;
        *brackt = 1;
    } else if (fabs(dp) < fabs(*dx)) {
        theta = 3.0 * (*fx - fp) / (*stp - *stx) + *dx + dp;
        s = fmax(fabs(theta), fmax(fabs(*dx), fabs(dp)));
        gamma = s * sqrt(fmax(0.0, (theta / s) * (theta / s) - (*dx / s) * (dp / s)));
        if (*stp > *stx) gamma = -gamma;
        p = (gamma - dp) + theta;
        qq = (gamma + (*dx - dp)) + gamma;
        r = p / qq;
        if (r < 0.0 && gamma != 0.0) stpc = *stp + r * (*stx - *stp);
        else if (*stp > *stx) stpc = stpmax;
        else stpc = stpmin;
        stpq = *stp + (dp / (dp - *dx)) * (*stx - *stp);
        if (*brackt) {
            stpf = (fabs(stpc - *stp) < fabs(stpq - *stp)) ? stpc : stpq;
            if (*stp > *stx) stpf = fmin(*stp + 0.66 * (*sty - *stp), stpf);
            else stpf = fmax(*stp + 0.66 * (*sty - *stp), stpf);
        } else {
            stpf = (fabs(stpc - *stp) > fabs(stpq - *stp)) ? stpc : stpq;
            stpf = fmin(stpmax, stpf);
            stpf = fmax(stpmin, stpf);
        }
    } else {
        if (*brackt) {
            theta = 3.0 * (fp - *fy) / (*sty - *stp) + *dy + dp;
            s = fmax(fabs(theta), fmax(fabs(*dy), fabs(dp)));
            gamma = s * sqrt((theta / s) * (theta / s) - (*dy / s) * (dp / s));
            if (*stp > *sty) gamma = -gamma;
            p = (gamma - dp) + theta;
            qq = ((gamma - dp) + gamma) + *dy;
            r = p / qq;
            stpc = *stp + r * (*sty - *stp);
            stpf = stpc;
        } else if (*stp > *stx) {
            stpf = stpmax;
        } else {
            stpf = stpmin;
        }
    }
    if (fp > *fx) {
        *sty = *stp;
        *fy = fp;
        *dy = dp;
    } else {
        if (sgnd < 0.0) {
            *sty = *stx;
            *fy = *fx;
            *dy = *dx;
        }
        *stx = *stp;
        *fx = fp;
        *dx = dp;
    }
    *stp = stpf;
}

static void mt_start(mt_state *m, double f, double g, double stp)
{
    m->brackt = 0;
    m->stage = 1;
    m->finit = f;
    m->ginit = g;
    m->gtest = m->ftol * g;
    m->width = m->stpmax - m->stpmin;
    m->width1 = 2.0 * m->width;
    m->stx = 0.0;
    m->fx = f;
    m->gx = g;
    m->sty = 0.0;
    m->fy = f;
    m->gy = g;
    m->stmin = 0.0;
    m->stmax = stp + 4.0 * stp;
}

/* Returns 0: evaluate at the new *stp; 1: converged; 2: warning (stop). */
static int mt_step(mt_state *m, double f, double g, double *stp)
{
    const double ftest = m->finit + *stp * m->gtest;
    if (m->stage == 1 && f <= ftest && g >= 0.0) m->stage = 2;
    if (m->brackt && (*stp <= m->stmin || *stp >= m->stmax)) return 2;
    if (m->brackt && m->stmax - m->stmin <= m->xtol * m->stmax) return 2;
    if (*stp == m->stpmax && f <= ftest && g <= m->gtest) return 2;
    if (*stp == m->stpmin && (f > ftest || g >= m->gtest)) return 2;
    if (f <= ftest && fabs(g) <= m->gtol * (-m->ginit)) return 1;
    if (m->stage == 1 && f <= m->fx && f > ftest) {
        double fm = f - *stp * m->gtest;
        double fxm = m->fx - m->stx * m->gtest, fym = m->fy - m->sty * m->gtest;
        double gm = g - m->gtest, gxm = m->gx - m->gtest, gym = m->gy - m->gtest;
        mt_dcstep(&m->stx, &fxm, &gxm, &m->sty, &fym, &gym, stp, fm, gm, &m->brackt, m->stmin, m->stmax);
        m->fx = fxm + m->stx * m->gtest;
        m->fy = fym + m->sty * m->gtest;
        m->gx = gxm + m->gtest;
        m->gy = gym + m->gtest;
    } else {
        mt_dcstep(&m->stx, &m->fx, &m->gx, &m->sty, &m->fy, &m->gy, stp, f, g, &m->brackt, m->stmin, m->stmax);
    }
    if (m->brackt) {
        if (fabs(m->sty - m->stx) >= 0.66 * m->width1) *stp = m->stx + 0.5 * (m->sty - m->stx);
        m->width1 = m->width;
        m->width = fabs(m->sty - m->stx);
    }
    if (m->brackt) {
        m->stmin = fmin(m->stx, m->sty);
        m->stmax = fmax(m->stx, m->sty);
    } else {
        m->stmin = *stp + 1.1 * (*stp - m->stx);
        m->stmax = *stp + 4.0 * (*stp - m->stx);
    }
    if (*stp < m->stpmin) *stp = m->stpmin;
    if (*stp > m->stpmax) *stp = m->stpmax;
    if ((m->brackt && (*stp <= m->stmin || *stp >= m->stmax)) ||
        (m->brackt && m->stmax - m->stmin <= m->xtol * m->stmax))
        *stp = m->stx;
    return 0;
}

/* ------------------------------------------------------------------ L-BFGS */

static double dotv(const double *a, const double *b, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double infnorm(const double *a, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s = fmax(s, fabs(a[i]));
    return s;
}

int32_t orc_lbfgs(orc_objfn fn, void *ctx, int64_t n, double *x, int32_t m, double gtol,
                  int32_t max_iter, orc_stats *st)
{
    double *g = (double *)malloc(n * sizeof(double));
    double *d = (double *)malloc(n * sizeof(double));
    double *xo = (double *)malloc(n * sizeof(double));
    double *go = (double *)malloc(n * sizeof(double));
    double *S = (double *)malloc((size_t)m * n * sizeof(double));
    double *Y = (double *)malloc((size_t)m * n * sizeof(double));
    double *rho = (double *)malloc(m * sizeof(double));
    double *alp = (double *)malloc(m * sizeof(double));
    int32_t nfev = 0, iter = 0, status = 0, hist = 0, head = 0;
    double f = fn(ctx, x, g);
    ++nfev;
    double gnorm = infnorm(g, n);
    if (!isfinite(f) || !isfinite(gnorm)) status = 3;
    while (status == 0 && !(gnorm < gtol)) {
        if (iter >= max_iter) {
            status = 1;
            break;
        }
        /* two-loop recursion (Nocedal & Wright, Algorithm 7.4) */
        memcpy(d, g, n * sizeof(double));
        for (int32_t t = 0; t < hist; ++t) {
            const int32_t i = (head - 1 - t + m) % m;
            alp[i] = rho[i] * dotv(S + (size_t)i * n, d, n);
            const double *yi = Y + (size_t)i * n;
            for (int64_t e = 0; e < n; ++e) d[e] -= alp[i] * yi[e];
        }
        if (hist > 0) {
            const int32_t i = (head - 1 + m) % m;
            const double gam = dotv(S + (size_t)i * n, Y + (size_t)i * n, n) /
                               dotv(Y + (size_t)i * n, Y + (size_t)i * n, n);
            for (int64_t e = 0; e < n; ++e) d[e] *= gam;
        }
        for (int32_t t = hist - 1; t >= 0; --t) {
            const int32_t i = (head - 1 - t + m) % m;
            const double be = rho[i] * dotv(Y + (size_t)i * n, d, n);
            const double *si = S + (size_t)i * n;
            for (int64_t e = 0; e < n; ++e) d[e] += (alp[i] - be) * si[e];
        }
        for (int64_t e = 0; e < n; ++e) d[e] = -d[e];
        double dg = dotv(g, d, n);
        if (!(dg < 0.0)) { /* not a descent direction: restart from steepest descent */
            hist = 0;
            for (int64_t e = 0; e < n; ++e) d[e] = -g[e];
            dg = dotv(g, d, n);
        }
        double stp = (hist == 0) ? 1.0 / sqrt(dotv(g, g, n)) : 1.0;
        memcpy(xo, x, n * sizeof(double));
        memcpy(go, g, n * sizeof(double));
        const double fo = f;
        mt_state ms;
        ms.ftol = 1e-4;
        ms.gtol = 0.9;
        ms.xtol = 1e-10;
        ms.stpmin = 1e-20;
        ms.stpmax = 1e20;
        mt_start(&ms, f, dg, stp);
        int ls = 0, approx = 0;
        for (int fe = 0; fe < 20; ++fe) {
            for (int64_t e = 0; e < n; ++e) x[e] = xo[e] + stp * d[e];
            f = fn(ctx, x, g);
            ++nfev;
            if (!isfinite(f)) {
                ls = 3;
                break;
            }
            const double dgt = dotv(g, d, n);
            /* Approximate Wolfe conditions (W. W. Hager and H. Zhang, SIAM J. Optim. 16
             * (2005) 170-192, eq. (23)): near the optimum a step lowers f by less than the
             * rounding of its FP64 sum over sequences, while the directional derivative is
             * still exact to many digits; the step is then accepted on the derivative, with
             * f allowed to exceed f(0) by that rounding level. */
            if (dgt >= ms.gtol * dg && dgt <= (2.0 * ms.ftol - 1.0) * dg && f <= fo + 1e-12 * fabs(fo)) {
                ls = 1;
                approx = 1;
                break;
            }
            /* within that rounding level the search runs on f(0) + stp (g0.d + g.d) / 2, the
             * trapezoid rule over the directional derivatives (exact for a quadratic) */
            const double f_ls = fabs(f - fo) <= 1e-12 * fabs(fo) ? fo + 0.5 * stp * (dg + dgt) : f;
            ls = mt_step(&ms, f_ls, dgt, &stp);
            if (ls != 0) break;
        }
        if (ls == 3) {
            status = 3;
            break;
        }
        if (!(f < fo) && !approx) { /* no progress: keep the previous iterate */
            memcpy(x, xo, n * sizeof(double));
            memcpy(g, go, n * sizeof(double));
            f = fo;
            status = 2;
            break;
        }
        double *si = S + (size_t)head * n, *yi = Y + (size_t)head * n;
        for (int64_t e = 0; e < n; ++e) {
            si[e] = x[e] - xo[e];
            yi[e] = g[e] - go[e];
        }
        const double ys = dotv(yi, si, n);
        if (ys > 0.0) {
            rho[head] = 1.0 / ys;
            head = (head + 1) % m;
            if (hist < m) ++hist;
        }
        ++iter;
        gnorm = infnorm(g, n);
    }
    if (st) {
        st->fval = f;
        st->gnorm = gnorm;
        st->nfeval = nfev;
        st->iter = iter;
        st->status = status;
        st->pad = 0;
    }
    free(g);
    free(d);
    free(xo);
    free(go);
    free(S);
    free(Y);
    free(rho);
    free(alp);
    return status;
}

/* ----------------------------------------------------------------- scoring */

/* ising_gauge, plmDCA_utility.hpp:57-83 with mean (row means, Matrix_math.hpp:364-374),
 * repmat (Matrix_math.hpp:51-76) and the elementwise operators (:214-330):
 * G[i][k] = ((M[i][k] - rowmean[k]) - colmean[i]) + mean(rowmean); the transposed
 * variant works on M^T with the roles of the two mean vectors exchanged. */
void orc_ising_gauge(const double *M, int32_t q, int32_t transpose, double *out)
{
    double A[16 * 16];
    for (int32_t i = 0; i < q; ++i)
        for (int32_t k = 0; k < q; ++k) A[i * q + k] = transpose ? M[k * q + i] : M[i * q + k];
    /* mean(A): row means of the (possibly transposed) matrix */
    double m1[16], m2[16];
    for (int32_t i = 0; i < q; ++i) {
        double a = 0.0;
        for (int32_t j = 0; j < q; ++j) a = a + A[i * q + j];
        m1[i] = a / (double)q;
    }
    double mm = 0.0;
    for (int32_t i = 0; i < q; ++i) mm += m1[i];
    mm = mm / (double)q;
    /* mean(A^T): column means of A */
    for (int32_t i = 0; i < q; ++i) {
        double a = 0.0;
        for (int32_t j = 0; j < q; ++j) a = a + A[j * q + i];
        m2[i] = a / (double)q;
    }
    for (int32_t i = 0; i < q; ++i)
        for (int32_t k = 0; k < q; ++k) out[i * q + k] = ((A[i * q + k] - m1[k]) - m2[i]) + mm;
}

/* frobenius_norm(a, exclude), Matrix_math.hpp:134-178. */
double orc_frobenius_excl(const double *G, int32_t q, int32_t exclude)
{
    double sum = 0.0;
    for (int32_t i = 0; i < q; ++i) {
        if (i == exclude) continue;
        for (int32_t k = 0; k < q; ++k) {
            if (k == exclude) continue;
            sum += G[i * q + k] * G[i * q + k];
        }
    }
    return sqrt(sum);
}

/* plmDCA.hpp:378-392: rows n<r use the transposed gauge, n>r the plain one;
 * the result is stored as float (CouplingStorage internal_real_t, :80). */
/* --norm-of-mean-scoring (plmDCA.hpp:335-363 storage, 617-666 output).  The stored block
 * of (r, n) is the float copy of J_r[n](a, s) (a: state at n, s: state at r), transposed
 * when n < r (plmDCA.hpp:345-357: copy(..., transpose = n < r)). */
void orc_nom_block(const double *x, int32_t L, int32_t q, int32_t r, int32_t n, float *M)
{
    for (int32_t a = 0; a < q; ++a)
        for (int32_t s = 0; s < q; ++s) {
            const float v = (float)x[orc_ref_index(L, q, n, a, s)];
            if (n < r) M[s * q + a] = v;
            else M[a * q + s] = v;
        }
}

/* ising_gauge on a float matrix view (plmDCA_utility.hpp:31-54) with the float mean,
 * repmat and element operators of Matrix_math.hpp: row means m1, column means m2 (mean of
 * the transpose), mm the mean of m1; G[i][k] = ((M[i][k] - m1[k]) - m2[i]) + mm in float. */
void orc_ising_gauge_float(const float *M, int32_t q, float *out)
{
    float m1[16], m2[16];
    for (int32_t i = 0; i < q; ++i) {
        float a = 0.0f;
        for (int32_t j = 0; j < q; ++j) a = a + M[i * q + j];
        m1[i] = a / (float)q;
    }
    float mm = 0.0f;
    for (int32_t i = 0; i < q; ++i) mm += m1[i];
    mm = mm / (float)q;
    for (int32_t i = 0; i < q; ++i) {
        float a = 0.0f;
        for (int32_t j = 0; j < q; ++j) a = a + M[j * q + i];
        m2[i] = a / (float)q;
    }
    for (int32_t i = 0; i < q; ++i)
        for (int32_t k = 0; k < q; ++k) out[i * q + k] = ((M[i * q + k] - m1[k]) - m2[i]) + mm;
}

/* frobenius_norm(a, exclude) on float (Matrix_math.hpp:132-178): sum is float and takes
 * std::pow(a, 2) (a double for a float argument) in each step: sum = float(sum + a*a). */
float orc_frobenius_excl_float(const float *G, int32_t q, int32_t exclude)
{
    float sum = 0.0f;
    for (int32_t i = 0; i < q; ++i) {
        if (i == exclude) continue;
        for (int32_t k = 0; k < q; ++k) {
            if (k == exclude) continue;
            const double a = (double)G[i * q + k];
            sum = (float)((double)sum + a * a);
        }
    }
    return sqrtf(sum);
}

/* One output line of plmDCA.hpp:646-663 for the stored blocks Mrn, Mnr:
 * out = {|G(Mrn) + G(Mnr)| * 0.5 as the norm of the float mean, |G(Mrn)|, |G(Mnr)|}. */
void orc_nom_pair(const float *Mrn, const float *Mnr, int32_t q, int32_t gap, float *out3)
{
    float Gij[256], Gji[256], Gm[256];
    orc_ising_gauge_float(Mrn, q, Gij);
    orc_ising_gauge_float(Mnr, q, Gji);
    for (int32_t e = 0; e < q * q; ++e) Gm[e] = (Gij[e] + Gji[e]) * 0.5f;
    out3[0] = orc_frobenius_excl_float(Gm, q, gap);
    out3[1] = orc_frobenius_excl_float(Gij, q, gap);
    out3[2] = orc_frobenius_excl_float(Gji, q, gap);
}

void orc_score_row(const double *x, int32_t L, int32_t q, int32_t r, int32_t gap, float *row)
{
    double M[16 * 16], G[16 * 16];
    for (int32_t nn = 0; nn < L; ++nn) {
        if (nn == r) {
            row[nn] = 0.0f;
            continue;
        }
        /* convert(Js, n): M[a][s] = J_r[n](a, s), Coupling_matrix_view.hpp:175-184 */
        for (int32_t a = 0; a < q; ++a)
            for (int32_t s = 0; s < q; ++s) M[a * q + s] = x[orc_ref_index(L, q, nn, a, s)];
        orc_ising_gauge(M, q, nn < r, G);
        row[nn] = (float)orc_frobenius_excl(G, q, gap);
    }
}

/* ------------------------------------------------------- threaded solving */

typedef struct {
    const uint8_t *states;
    int32_t B, L, q, r;
    const double *w;
    double lh, lJ;
    const orc_blocks *blk;
    double *work;
} obj_ctx;

static double obj_plain(void *c, const double *x, double *g)
{
    obj_ctx *o = (obj_ctx *)c;
    return orc_objective(o->states, o->B, o->L, o->q, o->r, o->w, o->lh, o->lJ, x, g);
}

static double obj_blocked(void *c, const double *x, double *g)
{
    obj_ctx *o = (obj_ctx *)c;
    return orc_objective_blocked(o->blk, o->r, o->w, o->lh, o->lJ, x, g, o->work);
}

typedef struct {
    const uint8_t *states;
    int32_t B, L, q, gap;
    const double *w;
    double lh, lJ;
    const int32_t *loci;
    int32_t K;
    double gtol;
    int32_t max_iter, m;
    const orc_blocks *blk;
    const double *x0;
    float *rows;
    double *x_out;
    orc_stats *st;
    int32_t next;
} solve_job;

static void *solve_worker(void *arg)
{
    solve_job *jb = (solve_job *)arg;
    const size_t n = orc_nparams(jb->L, jb->q);
    double *x = (double *)malloc(n * sizeof(double));
    double *work = (double *)malloc((size_t)2 * jb->B * jb->q * sizeof(double));
    for (;;) {
        const int32_t k = __atomic_fetch_add(&jb->next, 1, __ATOMIC_RELAXED);
        if (k >= jb->K) break;
        obj_ctx oc = {jb->states, jb->B, jb->L, jb->q, jb->loci[k], jb->w, jb->lh, jb->lJ, jb->blk, work};
        if (jb->x0) memcpy(x, jb->x0 + (size_t)k * n, n * sizeof(double));
        else memset(x, 0, n * sizeof(double)); /* cold start, plmDCA.hpp:298 */
        orc_lbfgs(jb->blk ? obj_blocked : obj_plain, &oc, (int64_t)n, x, jb->m, jb->gtol, jb->max_iter,
                  jb->st ? jb->st + k : NULL);
        if (jb->rows) orc_score_row(x, jb->L, jb->q, jb->loci[k], jb->gap, jb->rows + (size_t)k * jb->L);
        if (jb->x_out) memcpy(jb->x_out + (size_t)k * n, x, n * sizeof(double));
    }
    free(x);
    free(work);
    return NULL;
}

int32_t orc_solve_loci(const uint8_t *states, int32_t B, int32_t L, int32_t q, int32_t gap,
                       const double *w, double lambda_h, double lambda_J, const int32_t *loci,
                       int32_t K, double gtol, int32_t max_iter, int32_t m, int32_t nthreads,
                       int32_t blocked, const double *x0, float *score_rows, double *x_out,
                       orc_stats *st)
{
    orc_blocks *blk = blocked ? orc_blocks_build(states, B, L, q) : NULL;
    solve_job jb = {states, B, L, q, gap, w, lambda_h, lambda_J, loci, K, gtol, max_iter, m,
                    blk, x0, score_rows, x_out, st, 0};
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, solve_worker, &jb);
    solve_worker(&jb);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
    orc_blocks_free(blk);
    return 0;
}
