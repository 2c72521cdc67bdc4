// sdca_gram.hip -- the Gram formulation of the per-locus solve (kernels).
//
// The objective of locus r (plmDCA_objective.hpp:86-244) is a weighted multinomial
// log-loss plus lambda_h |h|^2 + lambda_J |J|^2.  Its stationarity condition in J reads
// 2 lambda_J J = -X_r^T G, with G_b(s) = w_b (P_b(s) - [x_br == s]) and X_r the one-hot
// alignment without locus r: the optimum lies in the span of the rows of X_r.  So does
// every L-BFGS iterate started from zero (plmDCA.hpp:298), since steps and gradients are
// combinations of gradients.  Each coupling vector v of the solve is therefore held as
// coefficients c_v (q x B, J_v = X_r^T c_v) together with its image K_r c_v, where
//   K_r[b][b'] = #{ j != r : x_bj == x_b'j } = K[b][b'] - [x_br == x_b'r]
// is the Gram matrix of X_r.  Then
//   energies     E_b(s) = h(s) + (K_r c_x)(s, b)                 (no product over K_x)
//   gradient     g = (sum_b G_b + 2 lambda_h h,  G + 2 lambda_J c_x)  in coefficients
//   its image    K_r G + 2 lambda_J K_r c_x                      (one B x B x q GEMM)
//   products     <u, v> = u_h . v_h + c_u . (K_r c_v)             (the primal dot product)
// so the L-BFGS recursion runs unchanged on the host, and a locus costs B^2 q instead of
// B K_x q per evaluation.  K is shared by every locus of the alignment; K_r differs from
// it by the locus' own column, T(s, a) = sum_{b: x_br = a} G_b(s), subtracted after the GEMM.
// All vectors and reductions are FP64.
//
// The coefficient columns sum to zero over the states; vectors hold their R = q - 1
// coordinates in an orthonormal basis of the zero-sum vectors (gram::to_contrast), so that
// the GEMM and every vector pass run over R rows instead of q.
// Slot vectors (FP64, a command uses the first 2 R Bp + 8 entries of each):
//   [0, R Bp)          coordinates c(k, b) at k * Bp + b (pads b >= B stay 0)
//   [R Bp, 2 R Bp)     the image (K_r c)(k, b), same layout
//   [2 R Bp, + q)      the fields h(s)
// in the order x_c, x_t, g[0], g[1], d, s_0..s_{m-1}, y_0..y_{m-1}.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdca_internal.hpp"

namespace sdca {
namespace {

using gram::kVecD;
using gram::kVecG;
using gram::kVecPairs;
using gram::kVecXc;
using gram::kVecXt;

constexpr int kSm = 65;  // softmax partials per block: T (8 x 8) and the data term

__device__ __forceinline__ double wave_sum(double x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;  // lane 0
}

// Sums NV values over a block of 256 threads into out[0..NV).
template <int NV>
__device__ __forceinline__ void block_sum_store(double (&v)[NV], double *out)
{
    __shared__ double red[4][NV];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double x = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) red[wave][k] = x;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < NV; k += blockDim.x) out[k] = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
}

// Sums 2 H values over a block of 256 threads: v[0, H) into outA[0, H), v[H, 2H) into
// outB[0, H).  The caller synchronises the block before the next call.
template <int H>
__device__ __forceinline__ void block_sum_store_split(double (&v)[2 * H], double *outA, double *outB)
{
    __shared__ double red[4][2 * H];
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 2 * H; ++k) {
        const double x = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) red[wave][k] = x;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * H; k += blockDim.x) {
        const double t = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
        if (k < H) outA[k] = t;
        else outB[k - H] = t;
    }
}

// ------------------------------------------------------------------ Gram matrix

// K[i][j] = sum_v popcount(P_i^v & P_j^v) over the loci [j0, L) (bits of loci below j0
// masked off): 64 x 64 lower-triangular tiles of sequence pairs, each thread 4 x 4 pairs,
// both triangles written.  planes[b][v][w] bit t set iff x_b(64 w + t) == v.
constexpr int GT = 64;
constexpr int GWC = 4;
__global__ __launch_bounds__(256) void k_gram(const uint64_t *__restrict__ planes, int B, int q, int W, int j0,
                                              int64_t ldk, double *__restrict__ K)
{
    __shared__ uint64_t sI[GT][8][GWC];
    __shared__ uint64_t sJ[GT][8][GWC];
    const int64_t t = blockIdx.x;
    int ti = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)(ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while ((int64_t)ti * (ti + 1) / 2 > t) --ti;
    const int tj = (int)(t - (int64_t)ti * (ti + 1) / 2);
    const int i0 = ti * GT, jb0 = tj * GT;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int wfirst = j0 >> 6;
    const uint64_t first_mask = ~0ull << (j0 & 63);
    int matches[4][4];
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) matches[a][c] = 0;
    for (int w0 = wfirst; w0 < W; w0 += GWC) {
        __syncthreads();
        for (int e = tid; e < GT * q * GWC; e += 256) {
            const int s = e / (q * GWC);
            const int rem = e - s * q * GWC;
            const int v = rem / GWC;
            const int ww = rem - v * GWC;
            const int w = w0 + ww;
            const uint64_t mask = w == wfirst ? first_mask : ~0ull;
            const int gi = i0 + s, gj = jb0 + s;
            sI[s][v][ww] = (gi < B && w < W) ? planes[((int64_t)gi * q + v) * W + w] & mask : 0ull;
            sJ[s][v][ww] = (gj < B && w < W) ? planes[((int64_t)gj * q + v) * W + w] & mask : 0ull;
        }
        __syncthreads();
        for (int v = 0; v < q; ++v)
            for (int ww = 0; ww < GWC; ++ww) {
                uint64_t pi[4], pj[4];
                for (int a = 0; a < 4; ++a) pi[a] = sI[ty * 4 + a][v][ww];
                for (int c = 0; c < 4; ++c) pj[c] = sJ[tx * 4 + c][v][ww];
                for (int a = 0; a < 4; ++a)
                    for (int c = 0; c < 4; ++c) matches[a][c] += __popcll(pi[a] & pj[c]);
            }
    }
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
            const int gi = i0 + ty * 4 + a, gj = jb0 + tx * 4 + c;
            if (gi < B && gj < B && gi >= gj) {
                const double v = (double)matches[a][c];
                K[(int64_t)gi * ldk + gj] = v;
                K[(int64_t)gj * ldk + gi] = v;
            }
        }
}

// ---------------------------------------------------------------- FP64 GEMM

// C[n][m] = sum_k Bm[n][k] A[m][k] (both operands K-contiguous; with A = K symmetric this
// is Z = K G per column n of the packed beliefs).  128 x 128 output tiles, four waves of
// 64 x 64 (4 x 4 tiles of v_mfma_f64_16x16x4f64), K in steps of 16 through LDS (k-major,
// double-buffered: one barrier per step).  One block per CU: the FP64 MFMA pipes stay busy
// with one wave per SIMD, and the other slot group's memory-bound passes share the CU (at
// C4 this gains more than the second block's latency hiding).  m < M (a multiple of 128),
// k < Kd (a multiple of 16); Bm rows [0, round_up(N, 128)) must be readable, only n < N is
// stored.
typedef double dx4 __attribute__((ext_vector_type(4)));
// DPAD 48: 90 KB of LDS per block, so one block (four waves) per CU
constexpr int DT = 128, DK = 16, DPAD = 48;
static_assert(DT == gram::kGemmCols, "the packed operand is padded for whole column tiles");

__device__ __forceinline__ void ld8d(const double *p, double (&r)[8])
{
    typedef double d2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const d2 v = *reinterpret_cast<const d2 *>(p + 2 * u);
        r[2 * u] = v.x;
        r[2 * u + 1] = v.y;
    }
}

// PART tells the launches apart in a profile: 0 the full solve's Z = K G, 1 the subspace
// warm start's products (split-K) and its basis, 2 its start points (Q a*, K Q a*).
template <int PART>
__global__ __launch_bounds__(256) void k_dgemm(const double *__restrict__ A, int64_t lda, const double *__restrict__ Bm,
                                               int64_t ldb, double *__restrict__ C, int64_t ldc, int N, int Kd,
                                               int kc, int64_t zstride)
{
    __shared__ double sA[2][DK][DT + DPAD];
    __shared__ double sB[2][DK][DT + DPAD];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * DT, n0 = blockIdx.x * DT;
    // split-K slice z: k in [z kc, min(Kd, (z + 1) kc)), partial product into C + z zstride
    const int kz = blockIdx.z * kc;
    Kd = min(Kd - kz, kc);
    A += kz;
    Bm += kz;
    C += blockIdx.z * zstride;
    const int wm = (wave & 1) * 64, wn = (wave >> 1) * 64;
    // staging: thread -> tile row r, k half kh (DK / 2 doubles of each operand per step)
    constexpr int KH = DK / 2;
    const int r = tid >> 1, kh = (tid & 1) * KH;
    const double *ga = A + (int64_t)(m0 + r) * lda + kh;
    const double *gb = Bm + (int64_t)(n0 + r) * ldb + kh;
    double ra[KH], rb[KH];
    auto load = [&]() {
#pragma unroll
        for (int h = 0; h < KH; h += 8) {
            double ta[8], tb[8];
            ld8d(ga + h, ta);
            ld8d(gb + h, tb);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                ra[h + u] = ta[u];
                rb[h + u] = tb[u];
            }
        }
    };
    load();
#pragma unroll
    for (int u = 0; u < KH; ++u) {
        sA[0][kh + u][r] = ra[u];
        sB[0][kh + u][r] = rb[u];
    }
    __syncthreads();
    dx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = dx4{0.0, 0.0, 0.0, 0.0};
    // D of a 16 x 16 tile holds row lk + 4 v of column li in lane (li + 16 lk), element v;
    // A rows enter permuted (lane li supplies row 4 (li % 4) + li / 4), so that the lane's
    // four values are rows 4 lk .. 4 lk + 3, contiguous in C
    const int li = lane & 15, lk = lane >> 4;
    const int ai = 4 * (li & 3) + (li >> 2);
    for (int k0 = 0; k0 < Kd; k0 += DK) {
        const int cur = (k0 / DK) & 1;
        const bool more = k0 + DK < Kd;
        if (more) {
            ga += DK;
            gb += DK;
            load();
        }
#pragma unroll
        for (int kk = 0; kk < DK; kk += 4) {
            double af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = sA[cur][kk + lk][wm + i * 16 + ai];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = sB[cur][kk + lk][wn + j * 16 + li];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
#pragma unroll
            for (int u = 0; u < KH; ++u) {
                sA[cur ^ 1][kh + u][r] = ra[u];
                sB[cur ^ 1][kh + u][r] = rb[u];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = n0 + wn + j * 16 + li;
        if (n >= N) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int m = m0 + wm + i * 16 + 4 * lk;
            *reinterpret_cast<dx4 *>(C + (int64_t)n * ldc + m) = acc[i][j];
        }
    }
}

// ------------------------------------------------------------- trial points

struct SlotPtrs {
    double *xc, *xt, *dv, *gt, *Sb, *Yb;
    const double *gk, *gold;
};

__device__ __forceinline__ SlotPtrs slot_ptrs(const SlotCmd &c, double *V, int64_t vstride, int64_t nvs, int M)
{
    double *base = V + (int64_t)c.slot * vstride;
    SlotPtrs p;
    p.xc = base + kVecXc * nvs;
    p.xt = base + kVecXt * nvs;
    p.dv = base + kVecD * nvs;
    p.gt = base + (kVecG + (c.gt == 1)) * nvs;
    p.gk = base + (kVecG + (c.gk == 1)) * nvs;    // NEWDIR: the current gradient
    p.gold = base + (kVecG + (c.gk != 1)) * nvs;  // NEWDIR: the previous one
    p.Sb = base + kVecPairs * nvs;
    p.Yb = p.Sb + (int64_t)M * nvs;
    return p;
}

struct StepArgs {
    int mode, pnew;
    bool accept, staged;
    double alpha, pstep, cd, cg;
};

template <int M>
__device__ __forceinline__ void load_step(const SlotCmd &c, StepArgs &a, double (&cs)[M], double (&cy)[M])
{
    a.mode = c.mode;
    a.pnew = c.pnew < M ? c.pnew : -1;
    a.accept = c.accept != 0;
    a.staged = c.init_src != 0.f;
    a.alpha = c.alpha;
    a.pstep = c.pstep;
    a.cd = c.cd;
    a.cg = c.cg;
#pragma unroll
    for (int p = 0; p < M; ++p) {
        cs[p] = c.cs[p];
        cy[p] = c.cy[p];
    }
}

// NEWDIR of step_elems for histories above 16: the same arithmetic in the same order, with
// the pairs streamed eight at a time through a loop that is not unrolled (their
// coefficients read from the command), so that a thread holds one chunk of pairs instead
// of 2m values and more waves fit a SIMD.
template <int M, int NE>
__device__ __forceinline__ void step_newdir_chunked(const StepArgs &a, const SlotCmd &c, const SlotPtrs &P,
                                                    int64_t nvs, const int64_t (&e)[NE], double (&xn)[NE],
                                                    double (&dn)[NE], double (&gv)[NE])
{
    constexpr int CH = 8;
    static_assert(M % CH == 0, "history a multiple of the chunk");
    double dold[NE], xv[NE], g[NE], go[NE], sn[NE], yn[NE], d[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        dold[k] = P.dv[e[k]];
        xv[k] = P.xc[e[k]];
        g[k] = P.gk[e[k]];
        go[k] = a.pnew >= 0 ? P.gold[e[k]] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        if (a.accept) {
            xv[k] = fma(a.pstep, dold[k], xv[k]);
            P.xc[e[k]] = xv[k];
        }
        sn[k] = yn[k] = 0.0;
        if (a.pnew >= 0) {
            sn[k] = a.pstep * dold[k];
            yn[k] = g[k] - go[k];
            P.Sb[(int64_t)a.pnew * nvs + e[k]] = sn[k];
            P.Yb[(int64_t)a.pnew * nvs + e[k]] = yn[k];
        }
        d[k] = a.cd * dold[k] + a.cg * g[k];
    }
#pragma unroll 1
    for (int p0 = 0; p0 < M; p0 += CH) {
        double sv[CH][NE], yv[CH][NE], csv[CH], cyv[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            csv[u] = c.cs[p0 + u];
            cyv[u] = c.cy[p0 + u];
        }
#pragma unroll
        for (int u = 0; u < CH; ++u)
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                const int p = p0 + u;
                sv[u][k] = (csv[u] != 0.0 && p != a.pnew) ? P.Sb[(int64_t)p * nvs + e[k]] : 0.0;
                yv[u][k] = (cyv[u] != 0.0 && p != a.pnew) ? P.Yb[(int64_t)p * nvs + e[k]] : 0.0;
            }
#pragma unroll
        for (int k = 0; k < NE; ++k)
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const bool mine = p0 + u == a.pnew;
                d[k] += csv[u] * (mine ? sn[k] : sv[u][k]);
                d[k] += cyv[u] * (mine ? yn[k] : yv[u][k]);
            }
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        P.dv[e[k]] = d[k];
        xn[k] = fma(a.alpha, d[k], xv[k]);
        P.xt[e[k]] = xn[k];
        dn[k] = d[k];
        gv[k] = g[k];
    }
}

// NE elements e[k] of a slot's vectors (coefficients, image entries or fields), as k_update
// does it for the primal vectors:
//   INIT    x_c = 0, x_t = 0; with a staged start (init_src, written into x_c by k_gwarm)
//           x_t = x_c
//   RETRY   x_t = x_c + alpha d
//   NEWDIR  x_c += pstep d_old (accept); s_pnew = pstep d_old, y_pnew = g[gk] - g[1-gk];
//           d = cd d_old + cg g_k + sum_p (cs[p] s_p + cy[p] y_p); x_t = x_c + alpha d
// xn receives x_t(e), dn and gv d(e) and g_k(e) (NEWDIR, else 0).  Every load comes before
// the arithmetic, so that a thread has all of its element's vectors in flight at once.
// The operations are linear, so an image entry follows its coefficients exactly.
template <int M, int NE>
__device__ __forceinline__ void step_elems(const StepArgs &a, const SlotCmd &c, const double (&cs)[M], const double (&cy)[M],
                                           const SlotPtrs &P, int64_t nvs, const int64_t (&e)[NE],
                                           double (&xn)[NE], double (&dn)[NE], double (&gv)[NE])
{
#pragma unroll
    for (int k = 0; k < NE; ++k) dn[k] = gv[k] = 0.0;
    if (a.mode == MODE_INIT) {
        if (a.staged) {
            double x[NE];
#pragma unroll
            for (int k = 0; k < NE; ++k) x[k] = P.xc[e[k]];
#pragma unroll
            for (int k = 0; k < NE; ++k) {
                xn[k] = x[k];
                P.xt[e[k]] = x[k];
            }
            return;
        }
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            xn[k] = 0.0;
            P.xc[e[k]] = 0.0;
            P.xt[e[k]] = 0.0;
        }
        return;
    }
    if (a.mode == MODE_RETRY) {
        double x[NE], d[NE];
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            x[k] = P.xc[e[k]];
            d[k] = P.dv[e[k]];
        }
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            xn[k] = fma(a.alpha, d[k], x[k]);
            P.xt[e[k]] = xn[k];
        }
        return;
    }
    if constexpr (M > 16) {
        step_newdir_chunked<M, NE>(a, c, P, nvs, e, xn, dn, gv);
        return;
    }
    constexpr int MS = M > 16 ? 1 : M;  // (the branch above serves longer histories)
    double dold[NE], xv[NE], g[NE], go[NE], sv[MS][NE], yv[MS][NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        dold[k] = P.dv[e[k]];
        xv[k] = P.xc[e[k]];
        g[k] = P.gk[e[k]];
        go[k] = a.pnew >= 0 ? P.gold[e[k]] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < MS; ++p)
#pragma unroll
        for (int k = 0; k < NE; ++k) {
            sv[p][k] = (cs[p] != 0.0 && p != a.pnew) ? P.Sb[(int64_t)p * nvs + e[k]] : 0.0;
            yv[p][k] = (cy[p] != 0.0 && p != a.pnew) ? P.Yb[(int64_t)p * nvs + e[k]] : 0.0;
        }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        if (a.accept) {
            xv[k] = fma(a.pstep, dold[k], xv[k]);
            P.xc[e[k]] = xv[k];
        }
        double sn = 0.0, yn = 0.0;
        if (a.pnew >= 0) {
            sn = a.pstep * dold[k];
            yn = g[k] - go[k];
            P.Sb[(int64_t)a.pnew * nvs + e[k]] = sn;
            P.Yb[(int64_t)a.pnew * nvs + e[k]] = yn;
        }
        double d = a.cd * dold[k] + a.cg * g[k];
#pragma unroll
        for (int p = 0; p < MS; ++p) {
            const bool mine = p == a.pnew;
            d += cs[p] * (mine ? sn : sv[p][k]);
            d += cy[p] * (mine ? yn : yv[p][k]);
        }
        P.dv[e[k]] = d;
        xn[k] = fma(a.alpha, d, xv[k]);
        P.xt[e[k]] = xn[k];
        dn[k] = d;
        gv[k] = g[k];
    }
}

// The fields of every command's trial point (one wave per command): the field entries
// step first, so that the softmax of k_gsoftmax reads the final h of x_t.
// upd_part[command][nchq] = {0, |h_t|^2, d_h . g_k,h}.  PART (here and in k_gstep, k_gdots)
// tells the full solve's launches (0) from the subspace warm start's (1) in a profile.
template <int M, int PART>
__global__ __launch_bounds__(64) void k_ghstep(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                               double *__restrict__ V, int64_t vstride, int64_t nvs,
                                               double *__restrict__ upd_part, int nchq)
{
    const int j = blockIdx.x;
    const SlotCmd &c = cmds[j];
    const int s = threadIdx.x;
    double xsqh = 0.0, dgh = 0.0;
    if (c.mode != MODE_IDLE) {
        const GramAln &A = aln[c.aln];
        if (s < A.q) {
            StepArgs a;
            double cs[M], cy[M];
            load_step<M>(c, a, cs, cy);
            const SlotPtrs P = slot_ptrs(c, V, vstride, nvs, M);
            const int64_t e[1] = {2 * (int64_t)gram::rows(A.q) * A.Bp + s};
            double xn[1], dn[1], g[1];
            step_elems<M, 1>(a, c, cs, cy, P, nvs, e, xn, dn, g);
            xsqh = xn[0] * xn[0];
            dgh = dn[0] * g[0];
        }
    }
    xsqh = wave_sum(xsqh);
    dgh = wave_sum(dgh);
    if (s == 0) {
        double *o = upd_part + ((int64_t)j * (nchq + 1) + nchq) * 3;
        o[0] = 0.0;
        o[1] = xsqh;
        o[2] = dgh;
    }
}

// Trial point of every command, one thread per coordinate (k, b) and its image entry
// (flat over the R Bp coordinates; blocks of 256).  upd_part[command][chunk] =
// {c_t . K_r c_t, 0, d . g_k} with d . g_k = c_d . (K_r c_gk).
template <int M, int PART>
__global__ __launch_bounds__(256) void k_gstep(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                               double *__restrict__ V, int64_t vstride, int64_t nvs,
                                               double *__restrict__ upd_part, int nchq)
{
    const int j = blockIdx.y;
    const SlotCmd &c = cmds[j];
    const int tid = threadIdx.x;
    double *up = upd_part + ((int64_t)j * (nchq + 1) + blockIdx.x) * 3;
    if (c.mode == MODE_IDLE) {
        if (tid < 3) up[tid] = 0.0;
        return;
    }
    const GramAln &A = aln[c.aln];
    const int64_t qB = (int64_t)gram::rows(A.q) * A.Bp;
    StepArgs a;
    double cs[M], cy[M];
    load_step<M>(c, a, cs, cy);
    const SlotPtrs P = slot_ptrs(c, V, vstride, nvs, M);
    const int64_t idx = (int64_t)blockIdx.x * 256 + tid;
    double red[3] = {0.0, 0.0, 0.0};
    if (idx < qB) {
        // the coefficient, then its image entry (one load phase each: fewer registers, so
        // that the pass can share a CU with the other slot group's GEMM)
        const int64_t ec[1] = {idx}, ei[1] = {idx + qB};
        double xc[1], dc[1], gc[1], xi[1], di[1], gi[1];
        step_elems<M, 1>(a, c, cs, cy, P, nvs, ec, xc, dc, gc);
        step_elems<M, 1>(a, c, cs, cy, P, nvs, ei, xi, di, gi);
        red[0] = xc[0] * xi[0];
        red[2] = dc[0] * gi[0];
    }
    block_sum_store<3>(red, up);
}

// Softmax of every command, one thread per sequence b (blocks of 256 sequences):
//   E_b(s) = h_t(s) + (K_r c_t)(s, b), lz = log sum_s exp E_b(s) (max-shifted),
//   f_b = w_b (lz - E_b(x_br)), G_b(s) = w_b (P_b(s) - [s == x_br])
// (plmDCA_objective.hpp:86-151), the image taken from its coordinates to states first.
// G's coordinates into the packed GEMM operand (column row + k, b contiguous) and the
// trial gradient's coordinates G + 2 lambda_J c_t into g[gt].
// sm_part[command][chunk] = {T(s, a) at 8 s + a, sum f_b}.
__global__ __launch_bounds__(256) void k_gsoftmax(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                                  double *__restrict__ V, int64_t vstride, int64_t nvs,
                                                  double *__restrict__ Gp, int64_t ldG, double *__restrict__ sm_part,
                                                  int nchunk)
{
    const int j = blockIdx.y;
    const SlotCmd &c = cmds[j];
    __shared__ double sW[4][kSm];
    __shared__ double sGb[256][8];
    __shared__ int sYb[256];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double *sp = sm_part + ((int64_t)j * nchunk + blockIdx.x) * kSm;
    if (c.mode == MODE_IDLE) {
        if (tid < kSm) sp[tid] = 0.0;
        return;
    }
    const GramAln &A = aln[c.aln];
    const int q = A.q, R = gram::rows(q), Bp = A.Bp, B = A.B;
    const int64_t qB = (int64_t)R * Bp;
    double *base = V + (int64_t)c.slot * vstride;
    const double *xt = base + kVecXt * nvs;
    double *gt = base + (kVecG + (c.gt == 1)) * nvs;
    double hs[8], xtc[8], xti[8];
    const int b = blockIdx.x * 256 + tid;
    const bool valid = b < Bp;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        hs[s] = s < q ? xt[2 * qB + s] : 0.0;  // from k_ghstep
        xtc[s] = (s < R && valid) ? xt[(int64_t)s * Bp + b] : 0.0;
        xti[s] = (s < R && valid) ? xt[qB + (int64_t)s * Bp + b] : 0.0;
    }
    double G[8];
    double fb = 0.0;
    int y = -1;
#pragma unroll
    for (int s = 0; s < 8; ++s) G[s] = 0.0;
    if (valid && b < B) {
        y = A.statesT[(int64_t)c.locus * B + b];
        const double wb = A.w[b];
        double img[8];
        gram::from_contrast(xti, q, img);  // (K_r c_t)(s, b) per state
        double E[8], mx = -INFINITY;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            E[s] = hs[s] + img[s];
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
        for (int k = 0; k < 8; ++k) {
            if (k >= R) continue;
            Gp[(int64_t)(c.row + k) * ldG + b] = Gc[k];
            gt[(int64_t)k * Bp + b] = Gc[k] + A.lamJ2 * xtc[k];
        }
    } else if (valid) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < R) {
                gt[(int64_t)k * Bp + b] = 0.0;  // pad sequences (the GEMM reads them: zero)
                Gp[(int64_t)(c.row + k) * ldG + b] = 0.0;
            }
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
    fb = wave_sum(fb);
    if (lane == 0) sW[wave][64] = fb;
    __syncthreads();
    if (tid < kSm) {
        const int s = tid >> 3, u = tid & 7;
        const bool ok = tid == 64 || (s < q && u < q);
        sp[tid] = ok ? (sW[0][tid] + sW[1][tid]) + (sW[2][tid] + sW[3][tid]) : 0.0;
    }
}

// The trial gradient's image and its products, flat over the R Bp coordinates (k, b) of a
// command, blocks of 1024 (four per thread, 256 apart):
//   (K_r g_t)(k, b) = Z(k, b) - T^(k, x_br) + 2 lambda_J (K_r c_t)(k, b), Z = K G (the GEMM),
// T^ the coordinates over s of T(s, a), stored into g[gt]; the fields' gradient
// sum_b G_b(s) + 2 lambda_h h_t(s) (block 0).
// part[command][chunk][2M + 5] = {g_t.s_p, g_t.y_p, g_t.g_t, g_t.g_k, g_t.d, 0, |g_t,J|^2}
// in the primal metric (coefficients of one against the image of the other).
constexpr int kDotsPerThread = 4;

// k_gdots for histories above 16: the same sums in the same order per thread, with the
// pairs' products reduced eight pairs at a time through a loop that is not unrolled, so
// that a thread holds 16 partial sums instead of 2m.
template <int M>
__device__ __forceinline__ void gdots_chunked(const SlotCmd &c, const GramAln &A, const double *xt, double *gt,
                                              const double *gk, const double *dv, const double *Sb,
                                              const double *Yb, int64_t nvs, const double *Zp, int64_t ldG,
                                              const double *sT, bool cur, uint32_t live, double *out)
{
    constexpr int CH = 8;
    static_assert(M % CH == 0, "history a multiple of the chunk");
    const int q = A.q, Bp = A.Bp, B = A.B;
    const int64_t qB = (int64_t)gram::rows(q) * Bp;
    const int tid = threadIdx.x;
    double tail[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // g.g, g.g_k, g.d, 0, |g_J|^2
    double gcv[kDotsPerThread];
    int64_t e1v[kDotsPerThread];
#pragma unroll
    for (int k = 0; k < kDotsPerThread; ++k) {
        gcv[k] = 0.0;
        e1v[k] = -1;
        const int64_t idx = (int64_t)blockIdx.x * (256 * kDotsPerThread) + k * 256 + tid;
        if (idx >= qB) continue;
        const int s = (int)(idx / Bp), b = (int)(idx - (int64_t)s * Bp);
        const int64_t e0 = idx, e1 = idx + qB;
        const bool seq = b < B;
        const double z = seq ? Zp[(int64_t)(c.row + s) * ldG + b] : 0.0;
        const int y = seq ? A.statesT[(int64_t)c.locus * B + b] : 0;
        const double xi = xt[e1];
        const double gc = gt[e0];
        const double gki = cur ? gk[e1] : 0.0, di = cur ? dv[e1] : 0.0;
        const double kg = seq ? z - (A.own ? sT[64 + s * 8 + y] : 0.0) + A.lamJ2 * xi : 0.0;
        gt[e1] = kg;
        tail[0] += gc * kg;
        tail[4] += gc * kg;
        tail[1] += gc * gki;
        tail[2] += gc * di;
        gcv[k] = gc;
        e1v[k] = e1;
    }
    const bool fields = blockIdx.x == 0 && tid < q;
    double gh = 0.0;
    int64_t eh = 0;
    if (fields) {
        // the fields: g_h(s) = sum_a T(s, a) + 2 lambda_h h_t(s)
        const int s = tid;
        for (int u = 0; u < q; ++u) gh += sT[s * 8 + u];
        eh = 2 * qB + s;
        gh += A.lamh2 * xt[eh];
        gt[eh] = gh;
        tail[0] += gh * gh;
        if (cur) {
            tail[1] += gh * gk[eh];
            tail[2] += gh * dv[eh];
        }
    }
#pragma unroll 1
    for (int p0 = 0; p0 < M; p0 += CH) {
        double a[2 * CH];
#pragma unroll
        for (int u = 0; u < 2 * CH; ++u) a[u] = 0.0;
#pragma unroll
        for (int k = 0; k < kDotsPerThread; ++k) {
            if (e1v[k] < 0) continue;
            double sp[CH], yp[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const bool on = (live >> (p0 + u)) & 1u;
                sp[u] = on ? Sb[(int64_t)(p0 + u) * nvs + e1v[k]] : 0.0;
                yp[u] = on ? Yb[(int64_t)(p0 + u) * nvs + e1v[k]] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                a[u] += gcv[k] * sp[u];
                a[CH + u] += gcv[k] * yp[u];
            }
        }
        if (fields && cur)
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                if (!((live >> (p0 + u)) & 1u)) continue;
                a[u] += gh * Sb[(int64_t)(p0 + u) * nvs + eh];
                a[CH + u] += gh * Yb[(int64_t)(p0 + u) * nvs + eh];
            }
        block_sum_store_split<CH>(a, out + p0, out + M + p0);
        __syncthreads();
    }
    block_sum_store<5>(tail, out + 2 * M);
}

template <int M, int PART>
__global__ __launch_bounds__(256) void k_gdots(const SlotCmd *__restrict__ cmds, const GramAln *__restrict__ aln,
                                               double *__restrict__ V, int64_t vstride, int64_t nvs,
                                               const double *__restrict__ Zp, int64_t ldG,
                                               const double *__restrict__ sm_part, int nchunk,
                                               double *__restrict__ part, int nchd)
{
    constexpr int NV = 2 * M + 5;
    const int j = blockIdx.y;
    const SlotCmd &c = cmds[j];
    __shared__ double sT[128];
    const int tid = threadIdx.x;
    double acc[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    double *out = part + ((int64_t)j * nchd + blockIdx.x) * NV;
    if (c.mode == MODE_IDLE || c.gt < 0 || c.gt > 1) {
        for (int k = tid; k < NV; k += 256) out[k] = 0.0;
        return;
    }
    const GramAln &A = aln[c.aln];
    const int q = A.q, Bp = A.Bp, B = A.B;
    const int64_t qB = (int64_t)gram::rows(q) * Bp;
    double *base = V + (int64_t)c.slot * vstride;
    const double *xt = base + kVecXt * nvs;
    double *gt = base + (kVecG + c.gt) * nvs;
    const double *gk = base + (kVecG + (1 - c.gt)) * nvs;
    const double *dv = base + kVecD * nvs;
    const double *Sb = base + kVecPairs * nvs;
    const double *Yb = Sb + (int64_t)M * nvs;
    // sT[8 s + a] = T(s, a); sT[64 + 8 k + a] = its coordinates over s (the own column's
    // term of the image, row k)
    if (tid < 64) {
        const int s = tid >> 3, u = tid & 7;
        double t = 0.0;
        if (s < q && u < q)
            for (int ch = 0; ch < nchunk; ++ch) t += sm_part[((int64_t)j * nchunk + ch) * kSm + tid];
        sT[tid] = t;
    }
    __syncthreads();
    if (tid < 8) {
        double col[8], w[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) col[s] = sT[s * 8 + tid];
        gram::to_contrast(col, q, w);
#pragma unroll
        for (int k = 0; k < 8; ++k) sT[64 + k * 8 + tid] = w[k];
    }
    __syncthreads();
    const bool cur = c.mode != MODE_INIT;  // a current gradient and a direction exist
    const uint32_t live = cur ? c.live : 0u;
    if constexpr (M > 16) {
        gdots_chunked<M>(c, A, xt, gt, gk, dv, Sb, Yb, nvs, Zp, ldG, sT, cur, live, out);
        return;
    }
#pragma unroll
    for (int k = 0; k < kDotsPerThread; ++k) {
        const int64_t idx = (int64_t)blockIdx.x * (256 * kDotsPerThread) + k * 256 + tid;
        if (idx >= qB) break;
        const int s = (int)(idx / Bp), b = (int)(idx - (int64_t)s * Bp);
        const int64_t e0 = idx, e1 = idx + qB;
        const bool seq = b < B;
        const double z = seq ? Zp[(int64_t)(c.row + s) * ldG + b] : 0.0;
        const int y = seq ? A.statesT[(int64_t)c.locus * B + b] : 0;
        const double xi = xt[e1];
        const double gc = gt[e0];
        const double gki = cur ? gk[e1] : 0.0, di = cur ? dv[e1] : 0.0;
        // the pairs' images, CH at a time (all of them for histories up to 16)
        constexpr int CH = M > 16 ? 8 : M;
        const double kg = seq ? z - (A.own ? sT[64 + s * 8 + y] : 0.0) + A.lamJ2 * xi : 0.0;
        gt[e1] = kg;
        acc[2 * M] += gc * kg;
        acc[2 * M + 4] += gc * kg;
        acc[2 * M + 1] += gc * gki;
        acc[2 * M + 2] += gc * di;
        for (int p0 = 0; p0 < M; p0 += CH) {
            double sp[CH], yp[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const bool on = (live >> (p0 + u)) & 1u;
                sp[u] = on ? Sb[(int64_t)(p0 + u) * nvs + e1] : 0.0;
                yp[u] = on ? Yb[(int64_t)(p0 + u) * nvs + e1] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                acc[p0 + u] += gc * sp[u];
                acc[M + p0 + u] += gc * yp[u];
            }
        }
    }
    if (blockIdx.x == 0 && tid < q) {
        // the fields: g_h(s) = sum_a T(s, a) + 2 lambda_h h_t(s)
        const int s = tid;
        double gh = 0.0;
        for (int u = 0; u < q; ++u) gh += sT[s * 8 + u];
        const int64_t eh = 2 * qB + s;
        gh += A.lamh2 * xt[eh];
        gt[eh] = gh;
        acc[2 * M] += gh * gh;
        if (cur) {
            acc[2 * M + 1] += gh * gk[eh];
            acc[2 * M + 2] += gh * dv[eh];
#pragma unroll
            for (int p = 0; p < M; ++p) {
                if (!((live >> p) & 1u)) continue;
                acc[p] += gh * Sb[(int64_t)p * nvs + eh];
                acc[M + p] += gh * Yb[(int64_t)p * nvs + eh];
            }
        }
    }
    block_sum_store<NV>(acc, out);
}

// Result rows (ResLayout): the products, {c_t . K_r c_t, |h_t|^2, d . g_k} and the data
// term.  The entry of max |g| holds |g_t|^2 (the host takes the square roots).  One block
// of 256 per command: eight threads per entry sum its chunk partials.
__global__ __launch_bounds__(256) void k_greduce(int m, const double *__restrict__ upd_part,
                                                 const double *__restrict__ dot_part,
                                                 const double *__restrict__ sm_part, int nchunk, int nchq, int nchd,
                                                 double *__restrict__ res)
{
    const int i = blockIdx.x;
    const int ND = 2 * m + 5, W = 2 * m + 9;
    double *out = res + (int64_t)i * W;
    const int sub = threadIdx.x & 7;
    for (int it = 0; it < (W + 31) / 32; ++it) {  // the same trip count for every lane
        const int t = it * 32 + (threadIdx.x >> 3);
        double a = 0.0;
        if (t < ND) {
            for (int ch = sub; ch < nchd; ch += 8) a += dot_part[((int64_t)i * nchd + ch) * ND + t];
        } else if (t < ND + 3) {
            for (int ch = sub; ch <= nchq; ch += 8) a += upd_part[((int64_t)i * (nchq + 1) + ch) * 3 + (t - ND)];
        } else if (t == W - 1) {
            for (int ch = sub; ch < nchunk; ch += 8) a += sm_part[((int64_t)i * nchunk + ch) * kSm + 64];
        }
        a += __shfl_xor(a, 1, 64);
        a += __shfl_xor(a, 2, 64);
        a += __shfl_xor(a, 4, 64);
        if (sub == 0 && t < W) out[t] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) out[2 * m + 3] = out[2 * m];
}

// Coefficients of finished loci as the A operand of the coupling contraction J = X_r^T c:
// item u's q state rows (from its R coordinate rows), or with coords its R coordinate rows
// as they are, from packed row rows[u] (FP32, or BF16 parts hi + mid, and lo unless Gl3 is
// null), and its fields into hpart[u][0..8) (the contraction's field column).
__global__ __launch_bounds__(256) void k_gsplit(const double *__restrict__ V, const int64_t *__restrict__ src_off,
                                                const int32_t *__restrict__ rows, const int32_t *__restrict__ alns,
                                                const GramAln *__restrict__ aln, int64_t ldm, float *__restrict__ G32,
                                                uint16_t *__restrict__ Ghi, uint16_t *__restrict__ Glo,
                                                uint16_t *__restrict__ Gl3, double *__restrict__ hpart, int coords)
{
    const int u = blockIdx.y;
    const GramAln &A = aln[alns[u]];
    const int q = A.q, R = gram::rows(q), Bp = A.Bp;
    const double *src = V + src_off[u];
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b < Bp) {
        double w[8], cs[8];  // the coordinates, then the coefficients per state
#pragma unroll
        for (int k = 0; k < 8; ++k) w[k] = k < R ? src[(int64_t)k * Bp + b] : 0.0;
        if (coords) {
#pragma unroll
            for (int k = 0; k < 8; ++k) cs[k] = w[k];
        } else {
            gram::from_contrast(w, q, cs);
        }
        const int nrow = coords ? R : q;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s >= nrow) continue;
            const double v = cs[s];
            const int64_t o = (int64_t)(rows[u] + s) * ldm + b;
            const float f = (float)v;
            if (G32) {
                G32[o] = f;
            } else {
                const uint16_t h = f2bf(f);
                const float r1 = f - bf2f(h);
                const uint16_t l = f2bf(r1);
                Ghi[o] = h;
                Glo[o] = l;
                if (Gl3) Gl3[o] = f2bf(r1 - bf2f(l));
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x < 8)
        hpart[(int64_t)u * 8 + threadIdx.x] = threadIdx.x < q ? src[2 * (int64_t)R * Bp + threadIdx.x] : 0.0;
}

// Copies of finished loci's vectors (gram::gather): block y is item u.
__global__ __launch_bounds__(256) void k_gather(const double *__restrict__ src, double *__restrict__ dst,
                                                const int64_t *__restrict__ meta, int n, int len2)
{
    const int u = blockIdx.y;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const double *s = src + meta[u];
    double *d = dst + meta[n + u];
    const int64_t len1 = meta[2 * n + u];
    if (t < len1) d[t] = s[t];
    else if (t < len1 + len2) d[meta[4 * n + u] + (t - len1)] = s[meta[3 * n + u] + (t - len1)];
}

}  // namespace

namespace gram {

void gather(hipStream_t st, const double *src, double *dst, const int64_t *meta, int n, int64_t maxlen1, int len2)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((maxlen1 + len2 + 255) / 256), n), dim3(256), 0, st, src, dst, meta,
                       n, len2);
}

void gram_matrix(hipStream_t st, const uint64_t *planes, int B, int q, int W, int j0, int64_t ldk, double *K)
{
    const int nt = (B + GT - 1) / GT;
    const int64_t tiles = (int64_t)nt * (nt + 1) / 2;
    hipLaunchKernelGGL(k_gram, dim3((unsigned)tiles), dim3(256), 0, st, planes, B, q, W, j0, ldk, K);
}

void gemm(hipStream_t st, const double *A, int64_t lda, const double *Bm, int64_t ldb, double *C, int64_t ldc, int M,
          int N, int Kd)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_dgemm<0>, dim3((N + DT - 1) / DT, M / DT, 1), dim3(256), 0, st, A, lda, Bm, ldb, C, ldc, N,
                       Kd, Kd, (int64_t)0);
}

void gemm_warm(hipStream_t st, const double *A, int64_t lda, const double *Bm, int64_t ldb, double *C, int64_t ldc,
               int M, int N, int Kd)
{
    if (N <= 0) return;
    hipLaunchKernelGGL(k_dgemm<2>, dim3((N + DT - 1) / DT, M / DT, 1), dim3(256), 0, st, A, lda, Bm, ldb, C, ldc, N,
                       Kd, Kd, (int64_t)0);
}

void gemm_splitk(hipStream_t st, const double *A, int64_t lda, const double *Bm, int64_t ldb, double *C, int64_t ldc,
                 int M, int N, int Kd, int kc, int64_t zstride)
{
    if (N <= 0) return;
    const int Z = (Kd + kc - 1) / kc;
    hipLaunchKernelGGL(k_dgemm<1>, dim3((N + DT - 1) / DT, M / DT, Z), dim3(256), 0, st, A, lda, Bm, ldb, C, ldc, N,
                       Kd, kc, zstride);
}

#define SDCA_GRAM_DISPATCH(m, CALL) \
    switch (m) {                    \
    case 1: CALL(1); break;         \
    case 3: CALL(3); break;         \
    case 5: CALL(5); break;         \
    case 8: CALL(8); break;         \
    case 10: CALL(10); break;       \
    case 16: CALL(16); break;       \
    case 24: CALL(24); break;       \
    case 32: CALL(32); break;       \
    default: break;                 \
    }

void step(hipStream_t st, int m, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride,
          int64_t nvs, double *Gp, int64_t ldG, double *upd_part, double *sm_part, int nchunk, int nchq,
          hipEvent_t ev_gstep0, hipEvent_t ev_gstep1)
{
#define SDCA_GSTEP(MM)                                                                                               \
    hipLaunchKernelGGL((k_ghstep<MM, 0>), dim3(ncmd), dim3(64), 0, st, cmds, aln, V, vstride, nvs, upd_part, nchq);  \
    if (ev_gstep0) (void)hipEventRecord(ev_gstep0, st);                                                              \
    hipLaunchKernelGGL((k_gstep<MM, 0>), dim3(nchq, ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, upd_part,   \
                       nchq);                                                                                        \
    if (ev_gstep1) (void)hipEventRecord(ev_gstep1, st)
    SDCA_GRAM_DISPATCH(m, SDCA_GSTEP)
#undef SDCA_GSTEP
    hipLaunchKernelGGL(k_gsoftmax, dim3(nchunk, ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, Gp, ldG, sm_part,
                       nchunk);
}

void step_vectors(hipStream_t st, int m, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V,
                  int64_t vstride, int64_t nvs, double *upd_part, int nchq)
{
#define SDCA_GSTEPV(MM)                                                                                              \
    hipLaunchKernelGGL((k_ghstep<MM, 1>), dim3(ncmd), dim3(64), 0, st, cmds, aln, V, vstride, nvs, upd_part, nchq);  \
    hipLaunchKernelGGL((k_gstep<MM, 1>), dim3(nchq, ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, upd_part,   \
                       nchq)
    SDCA_GRAM_DISPATCH(m, SDCA_GSTEPV)
#undef SDCA_GSTEPV
}

int dots_chunks(int64_t qB) { return (int)((qB + 256 * kDotsPerThread - 1) / (256 * kDotsPerThread)); }

void dots(hipStream_t st, int m, const SlotCmd *cmds, int ncmd, const GramAln *aln, double *V, int64_t vstride,
          int64_t nvs, const double *Zp, int64_t ldG, const double *sm_part, int nchunk, double *dot_part, int nchd,
          int part)
{
#define SDCA_GDOTS(MM)                                                                                               \
    if (part == 0)                                                                                                   \
        hipLaunchKernelGGL((k_gdots<MM, 0>), dim3(nchd, ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, Zp,     \
                           ldG, sm_part, nchunk, dot_part, nchd);                                                    \
    else                                                                                                             \
        hipLaunchKernelGGL((k_gdots<MM, 1>), dim3(nchd, ncmd), dim3(256), 0, st, cmds, aln, V, vstride, nvs, Zp,     \
                           ldG, sm_part, nchunk, dot_part, nchd)
    SDCA_GRAM_DISPATCH(m, SDCA_GDOTS)
#undef SDCA_GDOTS
}

void reduce(hipStream_t st, int m, int ncmd, const double *upd_part, const double *dot_part, const double *sm_part,
            int nchunk, int nchq, int nchd, double *res)
{
    hipLaunchKernelGGL(k_greduce, dim3(ncmd), dim3(256), 0, st, m, upd_part, dot_part, sm_part, nchunk, nchq, nchd,
                       res);
}

void split(hipStream_t st, const int64_t *src_off, const int32_t *rows, const int32_t *alns, int n, int nchunk,
           const GramAln *aln, const double *V, int64_t ldm, float *G32, uint16_t *Ghi, uint16_t *Glo,
           uint16_t *Gl3, double *hpart, int coords)
{
    hipLaunchKernelGGL(k_gsplit, dim3(nchunk, n), dim3(256), 0, st, V, src_off, rows, alns, aln, ldm, G32, Ghi, Glo,
                       Gl3, hpart, coords);
}

}  // namespace gram
}  // namespace sdca
