// sdca_subspace.hip -- the subspace warm start of the Gram formulation (kernels).
//
// The Gram solve holds a locus' couplings as J = X_r^T c with c the q x B coefficients of
// the sequences (sdca_gram.hip).  Its L-BFGS iterations are set by the large eigenvalues
// of the identical-locus counts K: the clade directions.  Before that solve, each locus is
// solved with c restricted to c = Q a, Q (B x k) an orthonormal basis of K's dominant
// eigen-subspace (power iteration), a (k x q).  The restricted problem is the same
// objective on fewer unknowns:
//   energies   E_b(s) = h(s) + (K_r Q a)(b, s),  K_r Q = K Q - U V_r,  V_r = U^T Q (q x k)
//              with U the B x q indicator of the locus' own states: (U V_r a)(b, s) =
//              (V_r a)(x_br, s)
//   penalty    lambda_J sum_s a_s^T M_r a_s,  M_r = Q^T K_r Q = M - V_r^T V_r,  M = Q^T K Q
//   gradient   dE/da_s = (K_r Q)^T G_s + 2 lambda_J M_r a_s = b_s + 2 lambda_J M_r a_s,
//              b_s = (K Q)^T G_s - V_r^T T(s, .),  T(s, a) = sum_{b: x_br = a} G_b(s)
// and its L-BFGS runs in the primal metric <u, v> = u_h v_h + a_u . M_r a_v, i.e. on the
// slot vectors of sdca_gram.hip with B replaced by k: coefficients a, image M_r a, fields.
// The gradient's coefficients are M_r^{-1} b_s + 2 lambda_J a_s, with
//   M_r^{-1} = M^{-1} + R_r Omega_r R_r^T,  R_r = M^{-1} V_r^T (k x q),
//   Omega_r = (I - V_r M^{-1} V_r^T)^{-1} (q x q)     (Woodbury: M_r = M - V_r^T V_r),
// and its image b_s + 2 lambda_J (M_r a)_s.  One restricted evaluation costs two B x k x q
// products (k_dgemm) where the full one costs a B x B x q product.  The restricted
// optimum (a*, h*) then starts the full solve at c = Q a*, K_r c = K Q a* - U V_r a*.
// Everything is FP64.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdca_internal.hpp"

namespace sdca {
namespace {

constexpr int kSmS = 65;  // softmax partials per block, as in sdca_gram.hip: T (8 x 8) and the data term

__device__ __forceinline__ double wsum(double x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;  // lane 0
}

// Per-slot constants of the restricted problem of an INIT command's locus, one block of
// 256 threads per command:
//   V_r[a][i] = sum_{b: x_br = a} Q[b][i]          (aux + 0,            q x kp)
//   R_r[i][a] = sum_j Minv[i][j] V_r[a][j]         (aux + 8 kp,         kp x 8)
//   Omega_r   = (I - V_r R_r)^{-1}                  (aux + 16 kp,        8 x 8)
// Q is Bp x kp (row-major), Minv kp x kp (symmetric); kp = kSubRank.
// V_r: wave w sums the sequences [w n4, (w + 1) n4) (n4 = ceil(B / 4)); lane l holds
// columns 2l and 2l + 1 of every state.  A sequence's state is the same for the whole wave,
// so its row goes to the registers of that state, picked by a wave-uniform index; eight
// rows are loaded ahead.  The four waves' sums are added in a fixed order.
// Omega_r: Gauss-Jordan with partial pivoting on [X | I] in LDS, one thread per entry.
__global__ __launch_bounds__(256) void k_sinit(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                               const double *__restrict__ Q, const double *__restrict__ Minv,
                                               double *__restrict__ aux, int64_t auxstride)
{
    constexpr int kp = kSubRank;
    static_assert(kp == 128, "V_r's lane layout covers 128 columns");
    const SlotCmd &c = cmds[blockIdx.x];
    if (c.mode != MODE_INIT) return;
    const GramAln &A = aln[c.aln];
    const int q = A.q, B = A.B;
    __shared__ double sP[4][8][kp];
    __shared__ double sX[8][8];
    __shared__ double sG[8][16];
    __shared__ int sPiv;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint8_t *y = A.statesT + (int64_t)c.locus * B;
    {
        const int n4 = (B + 3) >> 2;
        const int b0 = __builtin_amdgcn_readfirstlane(min(B, wave * n4));
        const int b1 = __builtin_amdgcn_readfirstlane(min(B, (wave + 1) * n4));
        double ax[8], ay[8];  // columns 2l and 2l + 1 (two arrays: each stays in registers)
#pragma unroll
        for (int a = 0; a < 8; ++a) ax[a] = ay[a] = 0.0;
        const double *qc = Q + 2 * lane;
        int b = b0;
        for (; b + 8 <= b1; b += 8) {
            double2 v[8];
            int ys[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                v[u] = *reinterpret_cast<const double2 *>(qc + (int64_t)(b + u) * kp);
                ys[u] = y[b + u];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int a = __builtin_amdgcn_readfirstlane(ys[u]) & 7;
                ax[a] += v[u].x;
                ay[a] += v[u].y;
            }
        }
        for (; b < b1; ++b) {
            const double2 v = *reinterpret_cast<const double2 *>(qc + (int64_t)b * kp);
            const int a = __builtin_amdgcn_readfirstlane((int)y[b]) & 7;
            ax[a] += v.x;
            ay[a] += v.y;
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            sP[wave][a][2 * lane] = ax[a];
            sP[wave][a][2 * lane + 1] = ay[a];
        }
    }
    __syncthreads();
    double *out = aux + (int64_t)c.slot * auxstride;
    if (tid < kp) {
        for (int a = 0; a < 8; ++a) {
            const double v =
                a < q ? (sP[0][a][tid] + sP[1][a][tid]) + (sP[2][a][tid] + sP[3][a][tid]) : 0.0;
            sP[0][a][tid] = v;
            out[(int64_t)a * kp + tid] = v;
        }
    }
    __syncthreads();
    // R_r[i][a] = sum_j Minv[i][j] V_r[a][j] (Minv symmetric: row i read as column i)
    if (tid < kp) {
        double r[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) r[a] = 0.0;
        for (int j = 0; j < kp; ++j) {
            const double mij = Minv[(int64_t)j * kp + tid];
#pragma unroll
            for (int a = 0; a < 8; ++a) r[a] += mij * sP[0][a][j];
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) out[8 * (int64_t)kp + (int64_t)tid * 8 + a] = r[a];
        // X[a][a'] = V_r[a] . R_r[., a'] : partial products of column tid
#pragma unroll
        for (int a = 0; a < 8; ++a) sP[1][a][tid] = r[a];
    }
    __syncthreads();
    if (tid < 64) {
        const int a = tid >> 3, a2 = tid & 7;
        double x = 0.0;
        if (a < q && a2 < q)
            for (int j = 0; j < kp; ++j) x += sP[0][a][j] * sP[1][a2][j];
        sX[a][a2] = (a == a2 ? 1.0 : 0.0) - x;
    }
    __syncthreads();
    // Gauss-Jordan on the q x q block (rows of unobserved states are those of the identity)
    const int rr = (tid >> 4) & 7, u = tid & 15;
    if (tid < 128) sG[rr][u] = u < 8 ? sX[rr][u] : (u - 8 == rr ? 1.0 : 0.0);
    __syncthreads();
    for (int col = 0; col < q; ++col) {
        if (tid == 0) {
            int piv = col;
            for (int r2 = col + 1; r2 < q; ++r2)
                if (fabs(sG[r2][col]) > fabs(sG[piv][col])) piv = r2;
            sPiv = piv;
        }
        __syncthreads();
        const int piv = sPiv;
        if (tid < 16 && piv != col) {
            const double t = sG[col][tid];
            sG[col][tid] = sG[piv][tid];
            sG[piv][tid] = t;
        }
        __syncthreads();
        double pr = 0.0, f = 0.0, inv = 0.0;
        if (tid < 128) {
            inv = 1.0 / sG[col][col];
            pr = sG[col][u] * inv;
            f = sG[rr][col];
        }
        __syncthreads();
        if (tid < 128 && rr < q) sG[rr][u] = rr == col ? pr : sG[rr][u] - f * pr;
        __syncthreads();
    }
    if (tid < 64) {
        const int a = tid >> 3, a2 = tid & 7;
        out[16 * (int64_t)kp + tid] = (a < q && a2 < q) ? sG[a][8 + a2] : 0.0;
    }
}

// The round's trial coordinates as the packed operand of the energy product,
// Ap[row + k][i] = a_t(k, i), and W[command][a][k] = (V_r a_t)(a, k) (8 x 8 per command).
// One block of 256 per command.
__global__ __launch_bounds__(256) void k_spack(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                               const double *__restrict__ V, int64_t vstride, int64_t nvs, int kp,
                                               const double *__restrict__ aux, int64_t auxstride,
                                               double *__restrict__ Ap, double *__restrict__ Wc)
{
    const int j = blockIdx.x;
    const SlotCmd &c = cmds[j];
    if (c.mode == MODE_IDLE) return;
    const int q = aln[c.aln].q, R = gram::rows(q);
    const double *xt = V + (int64_t)c.slot * vstride + gram::kVecXt * nvs;
    const double *Vr = aux + (int64_t)c.slot * auxstride;
    const int tid = threadIdx.x;
    for (int e = tid; e < R * kp; e += 256) Ap[(int64_t)c.row * kp + e] = xt[e];  // rows row + k
    // W[a][k] = sum_i V_r[a][i] a_t(k, i): one wave per (a, k) pair at a time
    const int lane = tid & 63, wave = tid >> 6;
    for (int p = wave; p < 64; p += 4) {
        const int a = p >> 3, k = p & 7;
        double v = 0.0;
        if (a < q && k < R)
            for (int i = lane; i < kp; i += 64) v += Vr[(int64_t)a * kp + i] * xt[(int64_t)k * kp + i];
        v = wsum(v);
        if (lane == 0) Wc[(int64_t)j * 64 + p] = v;
    }
}

// Softmax of the restricted trial points, one thread per sequence b (blocks of 256):
//   E_b(s) = h_t(s) + U (Ep[row + k][b] - W[a = x_br][k])_k,  Ep = (K Q) a_t from the
// product, U the coordinates' basis, then G, the T table and the data term exactly as
// k_gsoftmax forms them (G's coordinates into the packed operand of the gradient product).  sm_part[command][chunk] = {T(s, a) at 8 s + a, sum f_b}.
__global__ __launch_bounds__(256) void k_ssoftmax(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                                  const double *__restrict__ V, int64_t vstride, int64_t nvs, int kp,
                                                  const double *__restrict__ Ep, const double *__restrict__ Wc,
                                                  double *__restrict__ Gp, int64_t ldG, double *__restrict__ sm_part,
                                                  int nchunk)
{
    const int j = blockIdx.y;
    const SlotCmd &c = cmds[j];
    __shared__ double sW[4][kSmS];
    __shared__ double sGb[256][8];
    __shared__ int sYb[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *sp = sm_part + ((int64_t)j * nchunk + blockIdx.x) * kSmS;
    if (c.mode == MODE_IDLE) {
        if (tid < kSmS) sp[tid] = 0.0;
        return;
    }
    const GramAln &A = aln[c.aln];
    const int q = A.q, R = gram::rows(q), Bp = A.Bp, B = A.B;
    const double *xt = V + (int64_t)c.slot * vstride + gram::kVecXt * nvs;
    const double *W = Wc + (int64_t)j * 64;
    double hs[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) hs[s] = s < q ? xt[2 * (int64_t)R * kp + s] : 0.0;
    const int b = blockIdx.x * 256 + tid;
    double G[8];
    double fb = 0.0;
    int y = -1;
#pragma unroll
    for (int s = 0; s < 8; ++s) G[s] = 0.0;
    if (b < B) {
        y = A.statesT[(int64_t)c.locus * B + b];
        const double wb = A.w[b];
        double ec[8], es[8];  // (K_r Q a_t)(b, .): coordinates, then states
#pragma unroll
        for (int k = 0; k < 8; ++k) ec[k] = k < R ? Ep[(int64_t)(c.row + k) * ldG + b] - W[y * 8 + k] : 0.0;
        gram::from_contrast(ec, q, es);
        double E[8], mx = -INFINITY;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            E[s] = s < q ? hs[s] + es[s] : 0.0;
            if (s < q) mx = fmax(mx, E[s]);
        }
        double ex[8], z = 0.0, Ey = 0.0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            ex[s] = s < q ? exp(E[s] - mx) : 0.0;
            z += ex[s];
            if (s == y) Ey = E[s];
        }
        const double lz = mx + log(z);
        fb = wb * (lz - Ey);
        const double iz = 1.0 / z;
#pragma unroll
        for (int s = 0; s < 8; ++s) G[s] = s < q ? wb * (ex[s] * iz - (s == y ? 1.0 : 0.0)) : 0.0;
        double Gc[8];
        gram::to_contrast(G, q, Gc);
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < R) Gp[(int64_t)(c.row + k) * ldG + b] = Gc[k];
    } else if (b < Bp) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < R) Gp[(int64_t)(c.row + k) * ldG + b] = 0.0;  // pad sequences: zero
    }
    // T(s, a) = sum over the block's sequences with x_br = a of G_b(s): the block's beliefs
    // through LDS, then lane 8 s + a of wave w sums its 64 sequences in order; the data term
#pragma unroll
    for (int s = 0; s < 8; ++s) sGb[tid][s] = G[s];
    sYb[tid] = y;
    __syncthreads();
    {
        const int s = lane >> 3, u = lane & 7, b0 = wave * 64;
        double t = 0.0;
#pragma unroll 8
        for (int i = 0; i < 64; ++i) t += sYb[b0 + i] == u ? sGb[b0 + i][s] : 0.0;
        sW[wave][lane] = t;
    }
    fb = wsum(fb);
    if (lane == 0) sW[wave][64] = fb;
    __syncthreads();
    if (tid < kSmS) {
        const int s = tid >> 3, u = tid & 7;
        const bool ok = tid == 64 || (s < q && u < q);
        sp[tid] = ok ? (sW[0][tid] + sW[1][tid]) + (sW[2][tid] + sW[3][tid]) : 0.0;
    }
}

// The restricted gradient of every command, in two kernels of one block of kp (<= 128)
// threads per command around the product u = M^-1 b.  The split-K product gives, per
// packed row (row + k), the nz partials of (K Q)^T G_k (kp entries, zstride apart), so that
//   k_sgradb:  b_k = sum_z Z_z[row + k] - V_r^T T^(k, .)          (written back into Z_0)
//   (product)  u_k = M^-1 b_k                                      (into U)
//   k_sgradg:  g_k = u_k + R_r Omega_r (V_r u_k) + 2 lambda_J a_t(k) = M_r^-1 b_k + 2 lambda_J a_t(k)
// into the coordinates of g[gt], with T^ the coordinates of T over s; the products pass
// (k_gdots, reading b with row stride kp, image b + 2 lambda_J M_r a_t) completes it.
__device__ __forceinline__ void sgrad_tables(const SlotCmd &c, int j, int q, const double *sm_part, int nchunk,
                                             double *sT, double *sTc)
{
    const int tid = threadIdx.x;
    if (tid < 64) {
        const int s = tid >> 3, u = tid & 7;
        double t = 0.0;
        if (s < q && u < q)
            for (int ch = 0; ch < nchunk; ++ch) t += sm_part[((int64_t)j * nchunk + ch) * kSmS + tid];
        sT[tid] = t;
    }
    __syncthreads();
    if (tid < 8) {  // sTc[8 k + a]: the coordinates over s of T(s, a)
        double col[8], w[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) col[s] = sT[s * 8 + tid];
        gram::to_contrast(col, q, w);
#pragma unroll
        for (int k = 0; k < 8; ++k) sTc[k * 8 + tid] = w[k];
    }
    __syncthreads();
}

__global__ __launch_bounds__(128) void k_sgradb(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                                int kp, const double *__restrict__ aux, int64_t auxstride,
                                                const double *__restrict__ sm_part, int nchunk,
                                                double *__restrict__ Z, int nz, int64_t zstride)
{
    const int j = blockIdx.x;
    const SlotCmd &c = cmds[j];
    if (c.mode == MODE_IDLE || c.gt < 0 || c.gt > 1) return;
    const int q = aln[c.aln].q, R = gram::rows(q);
    __shared__ double sT[64], sTc[64];
    sgrad_tables(c, j, q, sm_part, nchunk, sT, sTc);
    const double *Vr = aux + (int64_t)c.slot * auxstride;
    const int i = threadIdx.x;
    if (i < kp)
        for (int k = 0; k < R; ++k) {
            double bv = 0.0;
            for (int z = 0; z < nz; ++z) bv += Z[z * zstride + (int64_t)(c.row + k) * kp + i];
            for (int a = 0; a < q; ++a) bv -= Vr[(int64_t)a * kp + i] * sTc[k * 8 + a];
            Z[(int64_t)(c.row + k) * kp + i] = bv;
        }
}

__global__ __launch_bounds__(128) void k_sgradg(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                                double *__restrict__ V, int64_t vstride, int64_t nvs, int kp,
                                                const double *__restrict__ aux, int64_t auxstride,
                                                const double *__restrict__ U, double lamJ2)
{
    const int j = blockIdx.x;
    const SlotCmd &c = cmds[j];
    if (c.mode == MODE_IDLE || c.gt < 0 || c.gt > 1) return;
    const int q = aln[c.aln].q, R = gram::rows(q);
    __shared__ double su[8][128];
    __shared__ double sv[8][8];
    const int tid = threadIdx.x;
    const double *Vr = aux + (int64_t)c.slot * auxstride;
    const double *Rr = Vr + 8 * (int64_t)kp;
    const double *Om = Vr + 16 * (int64_t)kp;
    const int i = tid;
    double u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) u[k] = 0.0;
    if (i < kp)
        for (int k = 0; k < R; ++k) {
            const double uv = U[(int64_t)(c.row + k) * kp + i];
            u[k] = uv;
            su[k][i] = uv;
        }
    __syncthreads();
    // v[a][k] = V_r[a] . u_k
    if (tid < 64) {
        const int a = tid >> 3, k = tid & 7;
        double v = 0.0;
        if (a < q && k < R)
            for (int t = 0; t < kp; ++t) v += Vr[(int64_t)a * kp + t] * su[k][t];
        sv[a][k] = v;
    }
    __syncthreads();
    if (i < kp) {
        double *gt = V + (int64_t)c.slot * vstride + (gram::kVecG + c.gt) * nvs;
        const double *xt = V + (int64_t)c.slot * vstride + gram::kVecXt * nvs;
        double r[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) r[a] = a < q ? Rr[(int64_t)i * 8 + a] : 0.0;
        for (int k = 0; k < R; ++k) {
            double g = u[k];
            for (int a = 0; a < q; ++a) {
                double w = 0.0;  // (Omega_r v)(a, k)
                for (int a2 = 0; a2 < q; ++a2) w += Om[a * 8 + a2] * sv[a2][k];
                g += r[a] * w;
            }
            gt[(int64_t)k * kp + i] = g + lamJ2 * xt[(int64_t)k * kp + i];
        }
    }
}

// The full solve's start of a new locus from its restricted optimum (a*, h*), given its
// start products c = Q a* (Cst) and K Q a* (Kst), R x Bp each:
//   coordinates  c
//   image        K_r c = K Q a* - U W,  W[a][k] = (V_r a*)(a, k) = sum_{b: x_br = a} c(k, b)
//   fields       h*
// into x_c.  One block of 256 per INIT command with a staged start; W from per-thread sums
// in LDS (thread t sums row k = t / 32 over the sequences b = t % 32 (mod 32)), reduced in a
// fixed order.
__global__ __launch_bounds__(256) void k_gwarm(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                               double *__restrict__ V, int64_t vstride, int64_t nvs,
                                               const double *__restrict__ Cst, const double *__restrict__ Kst,
                                               const double *__restrict__ Hst)
{
    const SlotCmd &c = cmds[blockIdx.x];
    if (c.mode != MODE_INIT || c.init_src == 0.f || c.item < 0) return;
    const GramAln &A = aln[c.aln];
    const int q = A.q, R = gram::rows(q), B = A.B, Bp = A.Bp;
    const int64_t qB = (int64_t)R * Bp;
    __shared__ double sAcc[256][9];
    __shared__ double sW[8][8];
    const int tid = threadIdx.x;
    const uint8_t *y = A.statesT + (int64_t)c.locus * B;
    const int64_t ldG = Bp;
    const double *cc = Cst + (int64_t)c.item * qB, *kc = Kst + (int64_t)c.item * qB;
#pragma unroll
    for (int a = 0; a < 9; ++a) sAcc[tid][a] = 0.0;
    {
        const int s = tid >> 5, g = tid & 31;
        if (s < R) {
            const double *cs = cc + (int64_t)s * ldG;
            int b = g;
            for (; b + 96 < B; b += 128) {  // four sequences' loads ahead of the sums
                double v[4];
                int ys[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    v[u] = cs[b + 32 * u];
                    ys[u] = y[b + 32 * u] & 7;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) sAcc[tid][ys[u]] += v[u];
            }
            for (; b < B; b += 32) sAcc[tid][y[b] & 7] += cs[b];
        }
    }
    __syncthreads();
    if (tid < 64) {
        const int a = tid >> 3, s = tid & 7;
        double w = 0.0;
        for (int g = 0; g < 32; ++g) w += sAcc[s * 32 + g][a];
        sW[a][s] = w;
    }
    __syncthreads();
    double *xc = V + (int64_t)c.slot * vstride + gram::kVecXc * nvs;
    for (int s = 0; s < R; ++s)
#pragma unroll 4
        for (int b = tid; b < Bp; b += 256) {
            const bool in = b < B;
            xc[(int64_t)s * Bp + b] = in ? cc[(int64_t)s * ldG + b] : 0.0;
            xc[qB + (int64_t)s * Bp + b] = in ? kc[(int64_t)s * ldG + b] - sW[y[b] & 7][s] : 0.0;
        }
    if (tid < 8) xc[2 * qB + tid] = tid < q ? Hst[(int64_t)c.item * 8 + tid] : 0.0;
}

}  // namespace

namespace sub {

void init(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, const double *Q, const double *Minv,
          int kp, double *aux, int64_t auxstride)
{
    if (ncmd <= 0) return;
    (void)kp;  // k_sinit is built for kp = kSubRank, the only rank the host allocates
    hipLaunchKernelGGL(k_sinit, dim3(ncmd), dim3(256), 0, st, cmds, aln, Q, Minv, aux, auxstride);
}

void pack(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, const double *V, int64_t vstride,
          int64_t nvs, int kp, const double *aux, int64_t auxstride, double *Ap, double *Wc)
{
    if (ncmd <= 0) return;
    hipLaunchKernelGGL(k_spack, dim3(ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, kp, aux, auxstride, Ap, Wc);
}

void softmax(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, const double *V, int64_t vstride,
             int64_t nvs, int kp, const double *Ep, const double *Wc, double *Gp, int64_t ldG, double *sm_part,
             int nchunk)
{
    if (ncmd <= 0) return;
    hipLaunchKernelGGL(k_ssoftmax, dim3(nchunk, ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, kp, Ep, Wc, Gp,
                       ldG, sm_part, nchunk);
}

void grad(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride, int64_t nvs,
          int kp, const double *aux, int64_t auxstride, const double *sm_part, int nchunk, double *Z, int nz,
          int64_t zstride, const double *Minv, int ncol, double lamJ2)
{
    if (ncmd <= 0) return;
    hipLaunchKernelGGL(k_sgradb, dim3(ncmd), dim3(128), 0, st, cmds, aln, kp, aux, auxstride, sm_part, nchunk, Z, nz,
                       zstride);
    // U = M^-1 b into the second slice (read by k_sgradb already)
    double *U = Z + zstride;
    gram::gemm_splitk(st, Minv, kp, Z, kp, U, kp, kp, ncol, kp, kp, 0);
    hipLaunchKernelGGL(k_sgradg, dim3(ncmd), dim3(128), 0, st, cmds, aln, V, vstride, nvs, kp, aux, auxstride, U,
                       lamJ2);
}

void starts(hipStream_t st, const double *Q, const double *KQ, int kp, int Bp, const double *Wst, int64_t nrows,
            double *Cst, double *Kst)
{
    if (nrows <= 0) return;
    gram::gemm_warm(st, Q, kp, Wst, kp, Cst, Bp, Bp, (int)nrows, kp);
    gram::gemm_warm(st, KQ, kp, Wst, kp, Kst, Bp, Bp, (int)nrows, kp);
}

void warm(hipStream_t st, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride, int64_t nvs,
          const double *Cst, const double *Kst, const double *Hst)
{
    if (ncmd <= 0) return;
    hipLaunchKernelGGL(k_gwarm, dim3(ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, Cst, Kst, Hst);
}

}  // namespace sub
}  // namespace sdca
