// sdca_internal.hpp -- shared definitions of the MI355X plmDCA engine (libsdca.so).
//
// Device data layout (one context = one alignment on one GPU):
//   states   B x L   uint8, envelope states (row-major, as given by the caller)
//   statesT  L x B   uint8, transposed copy: the column-r target states are contiguous
//   X        Bp x Kxp  one-hot alignment (FP32 or BF16), column col(j,a) = col_off[j] +
//            local index of state a among the states observed at locus j; column Kx is
//            the constant 1 ("bias" column) whose weights are the fields h_r; pad rows
//            b >= B and pad columns are zero
//   Xt       Kxp x Bp  its transpose (the gradient contraction reads X with b contiguous)
//   Per batch slot i (one locus r in flight) every L-BFGS vector is q x Kxp floats:
//   row s (state at r) and column k of X, i.e. W_r[s][col(j,a)] = J_r[j](a,s) and
//   W_r[s][Kx] = h_r(s).  Entries of the locus' own columns and of pad columns stay 0.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <mutex>
#include <vector>

#include "../../include/sdca.h"

namespace sdca {

constexpr int kTile = 128;        // GEMM output tile rows (and narrow tile columns)
constexpr int kTileWide = 256;    // wide GEMM tile columns; B and K_x are padded to this
constexpr int kMaxHistory = 32;   // largest L-BFGS history supported (the primal kernels: 16)
constexpr int kMaxGroups = 4;     // slot groups whose rounds are pipelined on separate streams

// BF16 <-> FP32 (round to nearest even; NaN stays NaN).
__device__ __forceinline__ uint16_t f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// Per-slot command of one batched evaluation round (host -> device).
//
// L-BFGS state of a slot on the device: iterate x_c, two gradient buffers (current g_k
// and trial g_t alternate), the search direction d, and m stored curvature pairs
// s_p = alpha d, y_p = g_new - g_old (BF16 under split precision, FP32 otherwise).
//
// A round's command list holds the group's active slots only, in slot order.  The slot
// vectors stay at their slot; the round's trial point x_t (and its BF16 copies), energies
// and node beliefs are packed at the command's rows, so that the GEMMs of the round cover
// active slots only.  x_t lives for one round: an accepted trial point is re-formed as
// x_c + pstep d_old by the next NEWDIR (the same fused multiply-add, hence the same value).
enum SlotMode : int32_t { MODE_IDLE = 0, MODE_INIT = 1, MODE_RETRY = 2, MODE_NEWDIR = 3 };

// Precision level of a command's contractions under split precision: the trial point W
// and the node beliefs G enter the MFMA as one BF16 part (coarse: the high part, ~8 bits),
// two (split: hi + lo, ~16 bits) or three (fine: hi + mid + lo, ~24 bits, the FP32
// contraction).  The fine level also takes exp and log at full accuracy in the softmax.
// Under FP32 precision the contractions are FP32 at every level; LV_FINE still selects
// the accurate exp and log.
enum Level : int32_t { LV_COARSE = 0, LV_SPLIT = 1, LV_FINE = 2 };

struct SlotCmd {
    int32_t mode;     // SlotMode
    int32_t locus;    // locus r of the slot, -1 if idle
    int32_t self_lo;  // columns [self_lo, self_hi) belong to locus r: gradient masked to 0
    int32_t self_hi;
    int32_t slot;     // slot of the command within its group
    int32_t row;      // first of its q rows in the round's packed layout
    int32_t level;    // contraction precision of the round (split precision): LV_COARSE, LV_SPLIT or LV_FINE
    int32_t aln;      // alignment of the locus (index into the solve's alignment table)
    int32_t q;        // its envelope: the command owns q packed rows, its slot vectors q x Kxp floats
    int32_t Kx, Kxp;  // its one-hot columns and their padded count (the slot vectors' row length)
    int32_t ksplit;   // split-K slices of its energy contraction in this round
    int32_t gt;       // gradient buffer (0 or 1) that receives the trial gradient
    int32_t gk;       // gradient buffer of the current iterate (after NEWDIR's accept); -1 at INIT
    int32_t accept;   // NEWDIR: 1 = the last trial point becomes the iterate first (x_c += pstep d_old)
    int32_t pnew;     // NEWDIR: pair slot receiving s = pstep d_old, y = g[gk] - g[1-gk]; -1: none
    uint32_t live;    // bit p: pair slot p holds a stored pair
    int32_t item;     // index of the locus in the solve's list (staged Gram warm starts), -1 if none
    float alpha;      // step of the trial point x_t = x_c + alpha d
    float pstep;      // NEWDIR with accept: step of the accepted trial
    float init_src;   // MODE_INIT: 1 = copy the staged start point, 0 = zero start
    float cd, cg;     // NEWDIR: d = cd d_old + cg g_k + sum_p (cs[p] s_p + cy[p] y_p)
    float cs[kMaxHistory], cy[kMaxHistory];
};

// Layout of the per-slot result row (device -> host), doubles, for history m.
struct ResLayout {
    int m;
    int ts(int p) const { return p; }          // g_t . s_p (live p)
    int ty(int p) const { return m + p; }      // g_t . y_p
    int tt() const { return 2 * m; }           // g_t . g_t
    int tg() const { return 2 * m + 1; }       // g_t . g_k
    int td() const { return 2 * m + 2; }       // g_t . d
    int ginf() const { return 2 * m + 3; }     // max |g_t|
    int gJinf() const { return 2 * m + 4; }    // max |g_t| over the coupling entries (column K_x excluded)
    int xsqJ() const { return 2 * m + 5; }     // sum of squares of x_t, J part
    int xsqh() const { return 2 * m + 6; }     // sum of squares of x_t, h part
    int dg() const { return 2 * m + 7; }       // NEWDIR: d . g_k
    int fdata() const { return 2 * m + 8; }    // sum_b w_b (LSE_b - E_b(x_br))
    int width() const { return 2 * m + 9; }
    int ndots() const { return 2 * m + 5; }    // entries produced by the dot-product pass
    int nupd() const { return 3; }             // entries produced by the trial-point pass
};

struct Timing {
    double energy = 0, grad = 0, softmax = 0, lbfgs = 0, score = 0, reweight = 0, total = 0;
    int64_t evals = 0, energy_launches = 0, grad_launches = 0;
    double flops_energy = 0, flops_grad = 0, bytes_lbfgs = 0;
    double update = 0, dots = 0, bytes_update = 0, bytes_dots = 0;
    double gstep = 0, bytes_gstep = 0;  // k_gstep alone (Gram full solve)
    double sub = 0, flops_sub = 0;
    int64_t sub_evals = 0;
    double host = 0;  // host time stepping the slots' L-BFGS state between rounds
};

// One alignment of a solve, as the kernels see it.  A solve covers one alignment (a
// context) or several (sdca_multi): every command names its alignment.
struct AlnDev {
    const uint32_t *X;         // Bp x Kxp one-hot bits
    const uint32_t *Xt;        // Kxp x Bp, its transpose
    const uint8_t *statesT;    // L x B envelope states, the target columns contiguous
    const float *w;            // Bp sequence weights (pads 0)
    const int32_t *state_col;  // L x q: column of (j, a) or -1
    int B, Bp, L, q, gap, Kx, Kxp;
    float lamJ2;               // 2 lambda_J
    double lamh2;              // 2 lambda_h
};

// One alignment of a Gram-formulation solve (sdca_gram.hip).
struct GramAln {
    const double *K;          // Gram matrix of identical-locus counts, Bp x Bp (pads 0)
    const uint8_t *statesT;   // L x B envelope states
    const double *w;          // Bp sequence weights in FP64 (pads 0)
    int B, Bp, q;
    int own;                  // 1: K_r = K minus the locus' own column; 0: K already excludes it (scan)
    double lamJ2, lamh2;      // 2 lambda_J, 2 lambda_h
};

// XCD-aware tile order of a GEMM launch.  Workgroups are dispatched to the eight XCDs
// round-robin (workgroup id % 8), so consecutive ids on one XCD are id, id+8, id+16, ...
// Tiles are grouped into super-tiles of GM x GN output tiles at one split-K index z; a
// super-tile runs on a single XCD, whose L2 then serves each A rows-tile to GN workgroups
// and each B rows-tile to GM workgroups.  Super-tiles are dealt to XCDs round-robin in
// the order N fastest, then split-K, then M: the row tiles start in increasing order, so
// the rows of a launch that take more MFMA passes (placed first) are not left for its end.
struct TileMap {
    int Mt, Nt, Z;   // output tiles along M and N, split-K slices
    int GM, GN;      // super-tile shape
    int MB, NB, S;   // super-tiles along M and N, total super-tiles
};

inline TileMap make_map(int Mt, int Nt, int Z, int gm)
{
    TileMap tm;
    tm.Mt = Mt;
    tm.Nt = Nt;
    tm.Z = Z;
    if (gm < 1) gm = 8;
    if (gm > 64) gm = 64;
    tm.GM = Mt < gm ? Mt : gm;
    tm.GN = Nt < 64 / tm.GM ? Nt : 64 / tm.GM;
    tm.MB = (Mt + tm.GM - 1) / tm.GM;
    tm.NB = (Nt + tm.GN - 1) / tm.GN;
    tm.S = Z * tm.MB * tm.NB;
    return tm;
}

// Workgroups of a tile map: whole rounds of eight super-tiles (one per XCD).
inline int map_grid(const TileMap &tm) { return ((tm.S + 7) / 8) * 8 * tm.GM * tm.GN; }

// The commands of a round are listed by level, finest first, each level's packed rows
// starting on a tile boundary: row[l] is the first packed row and cmd[l] the first command
// of the l-th part (fine, split, coarse), cmd[3] the number of commands.  Under split
// precision the fine part's GEMM tiles take three MFMA passes (hi, mid, lo), the split
// part's two and the coarse part's one.
struct Parts {
    int row[3];
    int cmd[4];
};

// One alignment's share of a round's GEMM launch: its commands' packed rows (parts, in
// absolute rows and command indices; the rows from parts.row[0] on, tile by tile), the
// B operand (its one-hot bits, ldbw words per row), the reduction length and split-K chunk,
// and its tile map; workgroups [tile0, tile0 + map_grid(tm)) of the launch compute it.
struct GemmGroup {
    int tile0;
    int aln, q;
    Parts parts;
    int K, kchunk;
    const uint32_t *Bbits;
    int64_t ldbw;
    TileMap tm;
};

// Largest mismatch count d with double(L - d) / double(L) >= theta, evaluated with the
// reference's own expression (plmDCA_optimizer_parameters.hpp:67-72: a size_t difference,
// then the FP64 division), or -1 if there is none; the reweighting kernels then compare
// integers.
inline int64_t identity_dmax(int L, double theta)
{
    int64_t dmax = -1;
    for (int64_t d = 0; d <= L; ++d) {
        const double identity = (double)((size_t)L - (size_t)d) / (double)(size_t)L;
        if (identity >= theta) dmax = d;
        else break;  // identity is non-increasing in d
    }
    return dmax;
}

// Kernel launchers (sdca_kernels.hip).
namespace kern {
void build_x(hipStream_t st, const uint8_t *d_states, int B, int L, int q, const int32_t *d_state_col,
             int Kx, int Bp, int Kxp, uint32_t *X, uint32_t *Xt);
void bitplanes(hipStream_t st, const uint8_t *d_states, int B, int L, int q, int W, uint64_t *planes);
void hamming_neighbours(hipStream_t st, const uint64_t *planes, int B, int q, int W, int L, int64_t dmax,
                        const uint32_t *d_mult, unsigned long long *d_iw);
// GEMMs over the groups of a round (d_groups on the device, ngroups of them, grid
// workgroups in all; tn the output tile width of every group).  Packed rows have ldx
// floats (trial points, W parts) or ldb (energies, beliefs); split-K slices of the energy
// partials are zstride floats apart.  Under split precision a group's fine part takes
// three MFMA passes, its split part two and its coarse part one.
void gemm_energy(hipStream_t st, int prec, const GemmGroup *d_groups, int ngroups, int grid, int tn, const void *W32,
                 const void *Whi, const void *Wlo, const void *Wl3, int64_t ldx, int64_t ldb, int64_t zstride,
                 float *Epart);
// hpart: per command and block of 256 sequences, the FP64 sums over b of G_b(s) (8 per
// block), from which the gradient epilogue forms the fields' gradient.
void softmax(hipStream_t st, int prec, const float *Epart, int64_t zstride, int64_t ldb, int ncmd, int nbchunk,
             const SlotCmd *cmds, const AlnDev *d_aln, float *G32, uint16_t *Ghi, uint16_t *Glo, uint16_t *Gl3,
             double *fpart, double *hpart);
void gemm_grad(hipStream_t st, int prec, const GemmGroup *d_groups, int ngroups, int grid, int tn, const void *G32,
               const void *Ghi, const void *Glo, const void *Gl3, int64_t ldb, const AlnDev *d_aln, const SlotCmd *cmds,
               const float *Xtrial, int64_t ldx, float *gbuf, int64_t n_slot, const double *hpart, int nbchunk);
// Slot vectors: gbuf [slot][2][n_slot], dbuf [slot][n_slot], S and Y [slot][m][n_slot]
// (dbuf, S and Y uint16_t BF16 when pair_bf16, else float); a command uses the first
// q x Kxp of its slot's.  Xtr, Whi, Wlo and Wl3 are packed rows of ldx.  One block row per
// command (ncmd commands).
void update(hipStream_t st, int m, bool pair_bf16, const SlotCmd *cmds, int ncmd, const float *gbuf, void *dbuf,
            void *S, void *Y, float *Xc, float *Xtr, uint16_t *Whi, uint16_t *Wlo, uint16_t *Wl3, const float *X0,
            int64_t n_slot, int64_t ldx, double *part, int nchunk);
void dots(hipStream_t st, int m, bool pair_bf16, const SlotCmd *cmds, int ncmd, const float *gbuf,
          const void *dbuf, const void *S, const void *Y, int64_t n_slot, double *part, int nchunk);
void reduce_results(hipStream_t st, int m, int ncmd, const double *upd_part, const double *dot_part,
                    const double *sm_part, int nchunk, int nbchunk, double *res);
// Score rows of nscore finished loci: item k reads its solution from xsrc + src_off[k]
// (rows src_ld[k] floats apart), locus loci[k] of alignment alns[k], and writes its row of
// that alignment's L entries at rows + k * ldrow.
void score(hipStream_t st, const float *xsrc, const int64_t *src_off, const int32_t *src_ld, const int32_t *loci,
           const int32_t *alns, int nscore, int Lmax, int64_t ldrow, const AlnDev *d_aln, float *rows, int coords);
// Couplings J = X_r^T c of finished Gram-formulation loci (the gradient contraction
// without its L2 term): item u's q x Kxp device vector at out + u * n_item, its field
// column from hpart[u][0..8).
void gemm_materialize(hipStream_t st, int prec, const GemmGroup *d_groups, int ngroups, int grid, int tn,
                      const void *G32, const void *Ghi, const void *Glo, const void *Gl3, int64_t ldb,
                      const AlnDev *d_aln, const SlotCmd *cmds, float *out, int64_t n_item, const double *hpart);
// sdca_pairs.hip
void build_x8(hipStream_t st, const uint8_t *states, int B, int L, int QP, int Bp, int8_t *X8);
void marginals(hipStream_t st, const uint8_t *states, int B, int L, int QP, int32_t *cnt);
void pair_gemm(hipStream_t st, int mode, int QP, const int8_t *Apan, const int8_t *Ball, int Bp, int Mrows,
               int Ncols, int i0, int L, int B, const int32_t *cnt, int32_t *Cout, float *MI);
void topk(hipStream_t st, const float *MI, int rows, int L, int K, int i0, int32_t *out_idx, float *out_val);
void seq_planes(hipStream_t st, const uint8_t *states, int B, int L, int q, int W, uint64_t *SP);
void pair_counts(hipStream_t st, const uint64_t *SP, int q, int W, const int32_t *pairs, int64_t npairs,
                 int32_t *counts);
// sdca_mlp.hip
// Epilogue of the network's BF16 products C = A B^T (mlp_gemm): bias, ReLU or the
// backward ReLU mask, and FP32 rows, BF16 rows and BF16 transposed rows as outputs.
struct MlpEpi {
    const float *bias = nullptr;     // [N]
    int relu = 0;
    const uint16_t *mask = nullptr;  // BF16 [M][ldm]: x = 0 where mask <= 0
    int64_t ldm = 0;
    float *f32 = nullptr;            // [M][ldf]
    int64_t ldf = 0;
    uint16_t *bf = nullptr;          // [Mt*32][ldbf]; rows >= M written as 0
    int64_t ldbf = 0;
    uint16_t *bt = nullptr;          // [N][ldbt]; columns >= M written as 0
    int64_t ldbt = 0;
    int64_t M = 0;                   // valid rows
    int N = 0;                       // valid columns
};
// Mt x Nt blocks of 32 x 32; A rows up to Mt*32 and B rows up to Nt*32 must be readable,
// K a multiple of 16, lda and ldb multiples of 8.
void mlp_gemm(hipStream_t st, const uint16_t *A, int64_t lda, const uint16_t *B, int64_t ldb, int K, int64_t Mt, int Nt,
              const MlpEpi &ep);
void rows_bf16(hipStream_t st, const float *X, int64_t n, int D, int64_t npad, int DP, uint16_t *Xb, uint16_t *Xt,
               int64_t ldt);
void pack_weights(hipStream_t st, const float *p, int din, int H, int dout, int DP, int DO, int64_t oW1, int64_t oW2,
                  int64_t oW3, uint16_t *W1b, uint16_t *W2b, uint16_t *W2t, uint16_t *W3b, uint16_t *W3t);
void colsum(hipStream_t st, const float *Z, int64_t ld, int64_t n, int O, float *out);
int mse(hipStream_t st, const float *Y, const float *T, const float *M, float *dY, uint16_t *dYb, uint16_t *dYt,
        int64_t n, int64_t npad, int dout, int DO, int64_t ldt, float scale, double *part, double *loss_acc);
void adam(hipStream_t st, float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
          float eps, float c1, float c2);
void gather_rows(hipStream_t st, const float *src, const int64_t *idx, int64_t nrows, int D, float *out);
void pair_features(hipStream_t st, const int32_t *counts, int64_t npairs, int q, int B, float *feats);
void scatter_warm(hipStream_t st, const int32_t *slot_of, const int32_t *item_of, const int32_t *locus_of,
                  const int32_t *partners, const float *J, const float *h, int nwarm, int P, int q, int64_t Kxp, int Kx,
                  const int32_t *state_col, int64_t n_slot, float *X0, void *XJ, bool xj_bf16, int64_t jstride);
}  // namespace kern

// Rank of the subspace warm start's basis (sdca_subspace.hip), and the sequences per
// split-K slice of its gradient product (kSubRank output rows: without slices a launch
// would hold one row of tiles).
constexpr int kSubRank = 128;
constexpr int kSubKc = 256;

// The subspace warm start (sdca_subspace.hip).
namespace sub {
void init(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, const double *Q, const double *Minv,
          int kp, double *aux, int64_t auxstride);
void pack(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, const double *V, int64_t vstride,
          int64_t nvs, int kp, const double *aux, int64_t auxstride, double *Ap, double *Wc);
void softmax(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, const double *V, int64_t vstride,
             int64_t nvs, int kp, const double *Ep, const double *Wc, double *Gp, int64_t ldG, double *sm_part,
             int nchunk);
// The restricted gradients from the nz split-K slices of (K Q)^T G at Z (row stride kp, zstride
// apart; slice 0 receives b, slice 1 u = M^-1 b over the round's ncol columns).
void grad(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride, int64_t nvs,
          int kp, const double *aux, int64_t auxstride, const double *sm_part, int nchunk, double *Z, int nz,
          int64_t zstride, const double *Minv, int ncol, double lamJ2);
// The start products of every staged restricted optimum: with a* of item i at rows
// [i R, (i + 1) R) of Wst (kp each), c = Q a* into Cst and K Q a* into Kst (R Bp per item,
// FP64 MFMA; nrows = K R rows).
void starts(hipStream_t st, const double *Q, const double *KQ, int kp, int Bp, const double *Wst, int64_t nrows,
            double *Cst, double *Kst);
// Staged starts of the round's new loci (INIT commands with init_src != 0):
// x_c = (c, K Q a* - U (U^T c), h*) from Cst, Kst and the fields Hst (8 per item).
void warm(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride, int64_t nvs,
          const double *Cst, const double *Kst, const double *Hst);
}  // namespace sub

// The Gram formulation (sdca_gram.hip): slot vectors x_c, x_t, g[0], g[1], d, then the m
// pairs s_p and the m pairs y_p, each nvs doubles apart (coefficients, image, fields).
namespace gram {
constexpr int kVecXc = 0, kVecXt = 1, kVecG = 2, kVecD = 4, kVecPairs = 5;

// Zero-sum contrasts.  Every coefficient column c(., b) of the Gram solve sums to zero over
// the q states: the gradient's coefficients G_b(s) = w_b (P_b(s) - [x_br == s]) do, every
// iterate started from zero is a combination of gradients, and the optimum is
// c* = -G*/(2 lambda_J).  The solve's vectors therefore hold the q - 1 coordinates
// w = U^T c of c in the orthonormal (Helmert) basis of the zero-sum vectors of R^q:
//   U[s][k] = h_k (s <= k),  -(k + 1) h_k (s = k + 1),  0 (s > k + 1),  h_k = 1/sqrt((k+1)(k+2)).
// With U^T U = I the primal products c_u . (K_r c_v) and all norms are unchanged and K_r
// acts on each coordinate row as on each state row: the GEMM, the vector passes and the
// pairs have q - 1 rows instead of q.  Energies and beliefs are formed per state
// (from_contrast), gradients return to coordinates (to_contrast).
inline __host__ __device__ int rows(int q) { return q - 1; }
// Slot vector: coefficients [0, R Bp), image [R Bp, 2 R Bp), fields [2 R Bp, + q), R = q - 1.
inline int64_t vec_len(int q, int Bp) { return 2 * (int64_t)rows(q) * Bp + 8; }

// w = U^T v for a q-vector v (entries s >= q ignored; w[k] = 0 for k >= q - 1).
__device__ __forceinline__ void to_contrast(const double (&v)[8], int q, double (&w)[8])
{
    constexpr double h[7] = {0.7071067811865475, 0.4082482904638631, 0.2886751345948129, 0.22360679774997896,
                             0.18257418583505536, 0.1543033499620919, 0.1336306209562122};
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
        run += v[k];
        w[k] = k < q - 1 ? h[k] * (run - (double)(k + 1) * v[k + 1]) : 0.0;
    }
    w[7] = 0.0;
}

// v = U w, the zero-sum q-vector of coordinates w[0, q - 1) (v[s] = 0 for s >= q).
__device__ __forceinline__ void from_contrast(const double (&w)[8], int q, double (&v)[8])
{
    constexpr double h[7] = {0.7071067811865475, 0.4082482904638631, 0.2886751345948129, 0.22360679774997896,
                             0.18257418583505536, 0.1543033499620919, 0.1336306209562122};
    double suf = 0.0;  // sum_{k >= s} h_k w[k]
#pragma unroll
    for (int s = 7; s >= 0; --s) {
        if (s < 7 && s < q - 1) suf += h[s] * w[s];
        double v_s = suf;
        if (s > 0 && s < q) v_s -= (double)s * h[s - 1] * w[s - 1];
        v[s] = s < q ? v_s : 0.0;
    }
}
// Identical-locus counts over loci [j0, L) from bit-planes (k_bitplanes layout).
void gram_matrix(hipStream_t st, const uint64_t *planes, int B, int q, int W, int j0, int64_t ldk, double *K);
// For u < n, with meta = [src offsets | dst offsets | len1 | src2 | dst2] (n entries each):
// meta[2n + u] doubles from src + meta[u] to dst + meta[n + u], then len2 doubles from
// src + meta[u] + meta[3n + u] to dst + meta[n + u] + meta[4n + u] (finished loci into the
// pending or staged-start buffers; maxlen1 bounds the first lengths).
void gather(hipStream_t st, const double *src, double *dst, const int64_t *meta, int n, int64_t maxlen1, int len2);
// Trial points and softmax of a round's commands: G into Gp (column row + s of ldG),
// partials upd_part [cmd][nchq + 1][3] (blocks of 256 coefficients) and sm_part
// [cmd][nchunk][65] (blocks of 256 sequences).
// Events ev_gstep0/1 (or null) are recorded around the k_gstep launch.
void step(hipStream_t st, int m, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride,
          int64_t nvs, double *Gp, int64_t ldG, double *upd_part, double *sm_part, int nchunk, int nchq,
          hipEvent_t ev_gstep0, hipEvent_t ev_gstep1);
// The trial-point part of step() alone (the subspace warm start's rounds: nchq blocks of 256
// over the q x k coefficients).
void step_vectors(hipStream_t st, int m, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V,
                  int64_t vstride, int64_t nvs, double *upd_part, int nchq);
// C[n][m] = sum_k Bm[n][k] A[m][k], FP64 MFMA: m < M (multiple of 128), k < Kd (multiple
// of 16), n < N stored; Bm rows up to round_up(N, kGemmCols) readable.
constexpr int kGemmCols = 128;
void gemm(hipStream_t st, const double *A, int64_t lda, const double *Bm, int64_t ldb, double *C, int64_t ldc, int M,
          int N, int Kd);
// The same as a launch of its own kernel instance (the subspace warm start's start points).
void gemm_warm(hipStream_t st, const double *A, int64_t lda, const double *Bm, int64_t ldb, double *C, int64_t ldc,
               int M, int N, int Kd);
// The same in ceil(Kd / kc) split-K slices (kc a multiple of 16): slice z sums k in
// [z kc, (z + 1) kc) into C + z zstride.  A kernel instance of its own (k_dgemm<1>): the
// subspace warm start's products.
void gemm_splitk(hipStream_t st, const double *A, int64_t lda, const double *Bm, int64_t ldb, double *C, int64_t ldc,
                 int M, int N, int Kd, int kc, int64_t zstride);
// Gradient images from Z = K G and the products: dot_part [cmd][nchd][2m + 5], nchd =
// dots_chunks(q Bp) blocks per command.
int dots_chunks(int64_t qB);
// part: 0 the full solve, 1 the subspace warm start (kernel instances of their own).
void dots(hipStream_t st, int m, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride,
          int64_t nvs, const double *Zp, int64_t ldG, const double *sm_part, int nchunk, double *dot_part, int nchd,
          int part);
void reduce(hipStream_t st, int m, int ncmd, const double *upd_part, const double *dot_part, const double *sm_part,
            int nchunk, int nchq, int nchd, double *res);
// Coefficients of n finished loci (V + src_off[u]) as contraction operands (rows[u]).
void split(hipStream_t st, const int64_t *src_off, const int32_t *rows, const int32_t *alns, int n, int nchunk,
           const GramAln *aln, const double *V, int64_t ldm, float *G32, uint16_t *Ghi, uint16_t *Glo,
           uint16_t *Gl3, double *hpart, int coords);
}  // namespace gram

// Workspace of Gram-formulation solves (sdca_api.cpp), sized for `slots` loci in flight,
// history m and the largest member alignment (q gq, padded depth gBp, gL loci, gKxp columns).
struct GramWs {
    int slots = 0, m = 0, gq = 0, gBp = 0, gL = 0, nmem = 0;
    int nchunk = 0, nchq = 0, nchd = 0;  // blocks per command: of 256 sequences, of 256 and of 1024 coefficients
    // slot groups whose rounds run on their own streams (c->gstream): gs slots each, colcap
    // columns of the packed GEMM operand each; a group's command lists, partials, result
    // rows and columns start at its first slot (or column) times the per-slot (per-column) size
    int ngroups = 1, gs = 0;
    int64_t colcap = 0;
    // the subspace warm start's rounds are short: its groups (rng of rgs slots and rcc
    // columns each, the same buffers) may be twice as many, so that more of them overlap
    int rng = 1, rgs = 0;
    int64_t rcc = 0;
    int64_t nvs = 0, vstride = 0, ldG = 0;
    double *V = nullptr;                  // slot vectors
    double *Gp = nullptr, *Zp = nullptr;  // packed GEMM operand G and product K G, ldG per column
    double *upd_part = nullptr, *sm_part = nullptr, *dot_part = nullptr, *res = nullptr;
    SlotCmd *d_cmds = nullptr, *h_cmds = nullptr;  // h_cmds, h_res: pinned
    double *h_res = nullptr;
    // staging of finished loci: per group, 5 gs entries of gram::gather metadata (pinned and
    // device), one gather kernel per round instead of two copies per locus
    int64_t *h_stage = nullptr, *d_stage = nullptr;
    GramAln *d_galn = nullptr;
    AlnDev *d_aln = nullptr;
    std::vector<int> cmd_slot[kMaxGroups];  // per group: slot of each listed command of its round in flight
    // finished loci wait in Pend (their coefficients and fields, at nvs doubles per locus,
    // up to pcap of them) until nmat have gathered; then, nmat per batch: contraction
    // operands (mrows packed rows of gBp), couplings (n_item floats each: q x Kxp), score rows
    // (gL each)
    struct Pending {
        int item, r, aln;
    };
    double *Pend = nullptr;  // two halves of pcap entries: new entries go to half phalf
    int pcap = 0, phalf = 0;
    std::vector<Pending> pend;
    hipEvent_t stage_ev[kMaxGroups] = {};
    // a flush reads one half of Pend on c->fstream while the rounds go on; pend_free[h]
    // follows the last flush of half h (gathers into h wait for it).  Its score rows (up to
    // nhrows, d_rows to pinned h_rows) are copied out at the next flush or at the end of the
    // solve: fout lists them, fev brackets the flush.
    struct RowOut {
        int item, aln;
        int64_t hrow;
    };
    hipEvent_t pend_free[2] = {}, fev[2] = {};
    std::vector<RowOut> fout;
    bool finflight = false;
    int nhrows = 0;
    int nmat = 0, mrows = 0;
    int64_t n_item = 0;
    float *M32 = nullptr, *Mout = nullptr;
    uint16_t *Mhi = nullptr, *Mlo = nullptr, *Ml3 = nullptr;
    double *mhpart = nullptr;
    SlotCmd *d_mcmds = nullptr;
    GemmGroup *d_mgroups = nullptr;
    int64_t *d_moff = nullptr, *d_soff = nullptr;
    int32_t *d_mrow = nullptr, *d_maln = nullptr, *d_sloc = nullptr, *d_sld = nullptr;
    float *d_rows = nullptr, *h_rows = nullptr;  // h_rows: pinned
    hipEvent_t ev[kMaxGroups][8] = {};
    // graph_rounds: per group, the captured round and the identical-locus counts it reads
    hipGraph_t graph[kMaxGroups] = {};
    hipGraphExec_t gexec[kMaxGroups] = {};
    const double *graph_K[kMaxGroups] = {};
    // subspace warm start: restricted slot vectors (nvs_r doubles, vstride_r per slot), per-slot
    // constants (aux, auxstride per slot), the packed restricted coefficients Ap and products
    // Zr (kSubRank per column), W = V_r a_t (64 per command), its descriptor, and the staged
    // restricted optima Wst (wcap items) with their start products Cst, Kst
    int64_t nvs_r = 0, vstride_r = 0, auxstride = 0, wcap = 0;
    int64_t wfld = 0;  // Wst: a* of item i at rows [i R, (i + 1) R) of kSubRank, its fields at wfld + 8 i
    int nchq_r = 0, nchd_r = 0;
    double *Vr = nullptr, *aux = nullptr, *Ap = nullptr, *Zr = nullptr, *Wc = nullptr, *Wst = nullptr;
    double *Cst = nullptr, *Kst = nullptr;  // per item: Q a* and K Q a* (R x Bp each)
    GramAln *d_galn_r = nullptr;
    bool warm_rounds = false;  // the current solve's full rounds stage warm starts (k_gwarm)
    // with four groups two host threads step two groups each: mu guards what they share
    // (the locus queue, the counters and timing, the pending buffer and coupling batches)
    std::mutex mu;
    const int *sub_nev = nullptr;  // the current solve's restricted evaluations per item
    ~GramWs();
};

// Warm starts of a solve call in compact device form: for item k (the k-th listed locus)
// P partner loci, the network's q x q blocks for them and the fields h.
struct CompactWarm {
    int P = 0;
    const int32_t *d_partners = nullptr;  // K x P
    const float *d_J = nullptr;           // K x P x q*q
    const float *d_h = nullptr;           // K x q
    // false: start at (J, h).  true: start at the fields h alone and take the blocks J as
    // the first search direction, so the line search scales them to the alignment at hand.
    bool as_direction = true;
};

// The warm-start network (sdca_mlp.hip) and its training state.
struct Mlp {
    int din = 0, dh = 0, dout = 0;
    int DP = 0, DO = 0;  // din and dout padded to 32: widths of the BF16 operands
    int64_t np = 0, step = 0;
    float *d_p = nullptr, *d_g = nullptr, *d_m = nullptr, *d_v = nullptr;  // FP32 masters, gradient, Adam moments
    // BF16 operands of the masters: W1 [dh][DP], W2 [dh][dh] and its transpose, W3 [DO][dh]
    // and its transpose [dh][DO]; packed says they match d_p
    uint16_t *d_W1b = nullptr, *d_W2b = nullptr, *d_W2t = nullptr, *d_W3b = nullptr, *d_W3t = nullptr;
    bool packed = false;
    // forward buffers, cap rows (a multiple of 32): FP32 input staging [cap][din], output
    // [cap][dout], BF16 input and activations
    int64_t cap = 0;
    float *d_x = nullptr, *d_y = nullptr;
    uint16_t *d_xb = nullptr, *d_h1b = nullptr, *d_h2b = nullptr;
    // training buffers, tcap rows: targets, mask, output gradients (FP32 [tcap][width]), BF16
    // copies and the transposes [width][tcap] that the weight-gradient products read
    int64_t tcap = 0;
    float *d_t = nullptr, *d_msk = nullptr, *d_dy = nullptr, *d_dz2 = nullptr, *d_dz1 = nullptr;
    uint16_t *d_xt = nullptr, *d_h1t = nullptr, *d_h2t = nullptr, *d_dyb = nullptr, *d_dyt = nullptr,
             *d_dz2b = nullptr, *d_dz2t = nullptr, *d_dz1t = nullptr;
    double *d_part = nullptr, *d_loss = nullptr;
    int64_t *d_idx = nullptr;
    ~Mlp();
};

}  // namespace sdca

// sdca_api.cpp: the batched solve with an optional warm-start source (dense reference
// layout on the host, or compact on the device).
int sdca_solve_impl(sdca_ctx *c, const int32_t *loci, int32_t K, const double *warm_dense,
                    const sdca::CompactWarm *cw, float *score_rows, double *x_out, sdca_locus_stats *st);
// The same over the member alignments of c (alns[k]: member of item k, NULL: all 0); the
// score row of item k goes to score_rows + row_off[k] (NULL: k * L).
int sdca_solve_batch(sdca_ctx *c, const int32_t *alns, const int32_t *loci, int32_t K, const double *warm_dense,
                     const sdca::CompactWarm *cw, float *score_rows, const int64_t *row_off, double *x_out,
                     sdca_locus_stats *st);

struct sdca_ctx {
    sdca_opts opts{};
    int device = 0;
    int B = 0, L = 0, q = 0, gap = -1;
    int Bp = 0;       // B padded to kTile
    int64_t Kx = 0;   // observed-state columns
    int64_t Kxp = 0;  // Kx + 1 padded to kTile
    std::vector<int32_t> col_off;    // L + 1
    int32_t scan_split = 0;          // > 0: two-alignment scan, loci [0, scan_split) are the first alignment's
    std::vector<int32_t> state_col;  // L x q: column of (j, a) or -1
    std::vector<uint8_t> h_states;   // B x L copy (host)
    hipStream_t stream = nullptr;

    uint8_t *d_states = nullptr, *d_statesT = nullptr;
    int32_t *d_state_col = nullptr;
    uint32_t *d_X = nullptr, *d_Xt = nullptr;  // one-hot X (Bp x Kxp) and X^T as bits
    float *d_w = nullptr;

    bool have_weights = false;
    double B_eff = 0, lambda_h = 0, lambda_J = 0;

    // batch workspace, sized for nslots slots and history m
    int ws_slots = 0, ws_m = 0;
    int64_t n_slot = 0;
    int Mpad = 0;
    int ngroups = 1, gslots = 0, Mg = 0;  // slot groups: slots and padded GEMM rows per group
    int Mgp = 0;        // packed rows per group: Mg plus two tiles (the three levels' parts pad separately)
    int64_t xc_off = 0;  // offset of the iterates x_c in the d_Xtr allocation
    int ws_groups = 0;
    int ksplit = 1, kchunk = 0;
    int ksplit_alloc = 1;            // split-K slices of Mgp rows the partial-energy buffer holds
    int gm_energy = 8, gm_grad = 8;  // XCD super-tile heights (autotuned)
    int tn_energy = 128, tn_grad = 128;  // output tile widths (autotuned)
    int tuned = 0;                   // 1 if the autotuner chose the GEMM configuration
    int nchunk = 0, nbchunk = 0;
    // The alignments the workspace serves: this context alone, or (for the internal
    // context of an sdca_multi) the member contexts of a multi-alignment solve.  The
    // workspace is sized for the largest: ws_q rows per command, packed rows of ws_ldx
    // (trial points) and ws_ldb (energies, beliefs) floats, score rows of ws_L, slot vectors
    // of n_slot floats.
    std::vector<sdca_ctx *> members;
    int ws_q = 0, ws_ldb = 0, ws_L = 0;
    int64_t ws_ldx = 0;
    sdca::AlnDev *d_aln = nullptr;                        // members' descriptors (device)
    sdca::GemmGroup *d_groups = nullptr, *h_groups = nullptr;  // per slot group: energy, then gradient groups
    int max_groups = 0;                                   // groups per launch (one per member)
    float *d_gbuf = nullptr;    // gradients: slots x 2 x n_slot (current and trial alternate)
    float *d_dbuf = nullptr;    // search directions: slots x n_slot (BF16 under split precision)
    void *d_S = nullptr, *d_Y = nullptr;  // curvature pairs: slots x m x n_slot (BF16 or float)
    float *d_Xc = nullptr, *d_Xtr = nullptr, *d_X0 = nullptr;
    uint16_t *d_Whi = nullptr, *d_Wlo = nullptr, *d_Wl3 = nullptr;  // BF16 parts of the trial points
    float *d_G32 = nullptr;
    uint16_t *d_Ghi = nullptr, *d_Glo = nullptr, *d_Gl3 = nullptr;  // BF16 parts of the node beliefs
    double *d_hpart = nullptr;  // FP64 field sums per command and block of 256 sequences
    float *d_Epart = nullptr;
    double *d_upd_part = nullptr, *d_dot_part = nullptr, *d_sm_part = nullptr, *d_res = nullptr;
    sdca::SlotCmd *d_cmds = nullptr;
    sdca::SlotCmd *h_cmds = nullptr;  // pinned host staging of the command lists
    std::vector<int> gcmd_slot[sdca::kMaxGroups];  // slot (in its group) of each listed command in flight
    std::vector<int64_t> h_xt_off;                 // per slot: d_Xtr offset of its last trial point
    std::vector<double> h_res_cmd;                 // result rows in command order (host scratch)
    double *h_res = nullptr;          // pinned host result rows
    float *d_rows = nullptr;
    int64_t *d_src_off = nullptr;
    int32_t *d_score_loci = nullptr;
    // pinned host staging, per group slices of gslots entries: score rows (L each), their
    // sources and loci, and the compact warm-start refill lists (3 per slot)
    float *h_rows = nullptr;
    int64_t *h_src_off = nullptr;
    int32_t *h_score_loci = nullptr;
    int32_t *h_score_ld = nullptr, *h_score_aln = nullptr;  // row length and alignment of each scored solution
    int32_t *d_score_ld = nullptr, *d_score_aln = nullptr;
    int32_t *h_wl = nullptr;

    // pair statistics (sdca_pairs.hip), built on first use
    int QP = 0;
    int64_t Kc_pad = 0, Kc_alloc = 0;
    int8_t *d_X8 = nullptr;
    int32_t *d_cnt = nullptr;
    uint64_t *d_SP = nullptr;  // sequence bit-planes L x q x ceil(B/64)

    // MI top-K partner table kept on the device (sdca_mi_topk)
    int32_t *d_topk = nullptr;
    int topk_K = 0;
    std::vector<double> h_w;  // installed weights (host copy)
    sdca::Mlp *mlp = nullptr;

    // Gram formulation: this alignment's identical-locus counts (computed for the scan
    // split gram_split) and FP64 weights, and the workspace of the solves it runs
    double *d_gram = nullptr;
    int gram_split = -1;
    // the dominant eigen-subspace of d_gram (subspace warm start): orthonormal basis Q
    // (Bp x kSubRank and its transpose), K Q and its transpose (kSubRank x Bp) and
    // M^-1 = (Q^T K Q)^-1; sub_rank 0 until built, -1 when these counts have none (Q^T K Q
    // singular: too few loci for the rank)
    int sub_rank = 0;
    double *d_subQ = nullptr, *d_subQt = nullptr, *d_subKQ = nullptr, *d_subKQt = nullptr, *d_subMinv = nullptr;
    double *d_w64 = nullptr;
    sdca::GramWs *gram = nullptr;

    hipEvent_t ev[8] = {};
    hipStream_t gstream[sdca::kMaxGroups] = {};  // gstream[0] is stream
    hipStream_t fstream = nullptr;  // the Gram solve's coupling batches of finished loci
    hipEvent_t gev[sdca::kMaxGroups][9] = {};  // [8]: a group's score rows are in h_rows
    sdca::Timing timing;

    ~sdca_ctx();
    void free_workspace();
};
