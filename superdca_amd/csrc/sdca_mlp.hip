// sdca_mlp.hip -- the warm-start network and its training kernels (no SuperDCA counterpart).
//
// A three-layer perceptron maps the statistics of a locus pair (r, j) to a start value of
// the coupling block J_r[j](a, s):
//   input   f_rj(s, a) (q*q joint frequencies, s = state at r, a = state at j),
//           f_r(s), f_j(a) (2q marginals) and the pair's mean-field coupling estimate
//           (q*q values, same layout as the joint frequencies): 2q^2 + 2q
//   hidden  two layers of H units (a multiple of 32), ReLU
//   output  q*q values, laid out like the joint frequencies: out[s*q + a] = J_r[j](a, s)
// Master parameters are FP32, stored contiguously: W1 [H][Din], b1 [H], W2 [H][H], b2 [H],
// W3 [Dout][H], b3 [Dout].  The layers run in BF16 on MFMA (v_mfma_f32_32x32x16_bf16):
// weights, activations and the backward pass's output gradients are rounded to BF16 (round
// to nearest even), products accumulate in FP32, biases, the loss and the bias gradients
// stay FP32.  Training minimises the masked mean squared error with Adam (D. P. Kingma and
// J. Ba, ICLR 2015) on the FP32 masters.  Every reduction runs in a fixed order, so results
// are deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdca_internal.hpp"

namespace sdca {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// C[m][n] = sum_k A[m][k] B[n][k]: BF16 operands (rows contiguous along k), FP32
// accumulation, one wave per 32 x 32 output block.  Rows of A are read up to Mt*32 and
// rows of B up to Nt*32; K is a multiple of 16 and pads of k hold zeros.  The epilogue
// (MlpEpi) adds the bias, applies ReLU or the ReLU mask of the backward pass and writes any
// of: FP32 rows, BF16 rows, BF16 transposed rows.  Output rows m in [M, Mt*32) are written
// as zeros to the BF16 outputs, so a later product over m (the batch) sees zero pads.

__global__ __launch_bounds__(256) void k_mlp_gemm(const uint16_t *__restrict__ A, int64_t lda,
                                                  const uint16_t *__restrict__ B, int64_t ldb, int K, int64_t Mt,
                                                  int Nt, kern::MlpEpi ep)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wid >= Mt * Nt) return;
    const int64_t mt = wid / Nt;
    const int nt = (int)(wid - mt * Nt);
    const int koff = (lane >> 5) * 8;
    const uint16_t *pa = A + (mt * 32 + (lane & 31)) * lda + koff;
    const uint16_t *pb = B + ((int64_t)nt * 32 + (lane & 31)) * ldb + koff;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        const bf16x8 a = *reinterpret_cast<const bf16x8 *>(pa + k0);
        const bf16x8 b = *reinterpret_cast<const bf16x8 *>(pb + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int n = nt * 32 + (lane & 31);
    if (n >= ep.N) return;
    const float bias = ep.bias ? ep.bias[n] : 0.f;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        // output layout of the 32x32 MFMA block: row 8*(v/4) + 4*(lane/32) + v%4, column lane%32
        const int64_t m = mt * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        if (m >= ep.M) {
            if (ep.bf) ep.bf[m * ep.ldbf + n] = 0;
            if (ep.bt) ep.bt[(int64_t)n * ep.ldbt + m] = 0;
            continue;
        }
        float x = acc[v] + bias;
        if (ep.relu) x = fmaxf(x, 0.f);
        if (ep.mask && !(bf2f(ep.mask[m * ep.ldm + n]) > 0.f)) x = 0.f;
        if (ep.f32) ep.f32[m * ep.ldf + n] = x;
        const uint16_t h = f2bf(x);
        if (ep.bf) ep.bf[m * ep.ldbf + n] = h;
        if (ep.bt) ep.bt[(int64_t)n * ep.ldbt + m] = h;
    }
}

// BF16 copies of n FP32 rows of width D: Xb [npad][DP] and (if Xt) Xt [DP][ldt], zero
// outside rows < n and columns < D.
__global__ void k_rows_bf16(const float *__restrict__ X, int64_t n, int D, int64_t npad, int DP,
                            uint16_t *__restrict__ Xb, uint16_t *__restrict__ Xt, int64_t ldt)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= npad * DP) return;
    const int64_t r = idx / DP;
    const int i = (int)(idx - r * DP);
    const uint16_t h = (r < n && i < D) ? f2bf(X[r * D + i]) : (uint16_t)0;
    Xb[idx] = h;
    if (Xt) Xt[(int64_t)i * ldt + r] = h;
}

// BF16 operands of the three layers from the FP32 masters: W1 [H][DP], W2 [H][H] and its
// transpose, W3 [DO][H] and its transpose [H][DO] (pads zero).
__global__ void k_pack_weights(const float *__restrict__ p, int din, int H, int dout, int DP, int DO,
                               int64_t oW1, int64_t oW2, int64_t oW3, uint16_t *__restrict__ W1b,
                               uint16_t *__restrict__ W2b, uint16_t *__restrict__ W2t, uint16_t *__restrict__ W3b,
                               uint16_t *__restrict__ W3t)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n1 = (int64_t)H * DP, n2 = (int64_t)H * H, n3 = (int64_t)DO * H;
    if (idx < n1) {
        const int h = (int)(idx / DP), i = (int)(idx - (int64_t)(idx / DP) * DP);
        W1b[idx] = i < din ? f2bf(p[oW1 + (int64_t)h * din + i]) : (uint16_t)0;
    } else if (idx < n1 + n2) {
        const int64_t u = idx - n1;
        const int a = (int)(u / H), b = (int)(u - (int64_t)a * H);
        const uint16_t v = f2bf(p[oW2 + u]);
        W2b[u] = v;
        W2t[(int64_t)b * H + a] = v;
    } else if (idx < n1 + n2 + n3) {
        const int64_t u = idx - n1 - n2;
        const int o = (int)(u / H), h = (int)(u - (int64_t)o * H);
        const uint16_t v = o < dout ? f2bf(p[oW3 + (int64_t)o * H + h]) : (uint16_t)0;
        W3b[u] = v;
        W3t[(int64_t)h * DO + o] = v;
    }
}

// out[o] = sum over rows r < n of Z[r][o] (FP64 in row order): the bias gradients.
__global__ void k_colsum(const float *__restrict__ Z, int64_t ld, int64_t n, int O, float *__restrict__ out)
{
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= O) return;
    double s = 0.0;
    for (int64_t r = 0; r < n; ++r) s += Z[r * ld + o];
    out[o] = (float)s;
}

// dY = scale (Y - T) M for rows < n, columns < dout (FP32 [n][dout]); its BF16 copy
// [npad][DO] and transpose [DO][ldt] (zero outside); per-block partial sums of
// M (Y - T)^2 in FP64.
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             const float *__restrict__ M, float *__restrict__ dY,
                                             uint16_t *__restrict__ dYb, uint16_t *__restrict__ dYt, int64_t n,
                                             int64_t npad, int dout, int DO, int64_t ldt, float scale,
                                             double *__restrict__ part)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (idx < npad * DO) {
        const int64_t r = idx / DO;
        const int o = (int)(idx - r * DO);
        uint16_t hb = 0;
        if (r < n && o < dout) {
            const int64_t e = r * dout + o;
            const float m = M ? M[e] : 1.f;
            const float d = (Y[e] - T[e]) * m;
            const float g = scale * d;
            dY[e] = g;
            hb = f2bf(g);
            l = (double)d * d;
        }
        dYb[idx] = hb;
        dYt[(int64_t)o * ldt + r] = hb;
    }
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_sum_parts(const double *__restrict__ part, int n, double *__restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < n; ++k) s += part[k];
        *out += s;
    }
}

__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const float gi = g[idx];
    const float mi = b1 * m[idx] + (1.f - b1) * gi;
    const float vi = b2 * v[idx] + (1.f - b2) * gi * gi;
    m[idx] = mi;
    v[idx] = vi;
    p[idx] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// rows out[k] = src[idx[k]] for a mini-batch
__global__ void k_gather_rows(const float *__restrict__ src, const int64_t *__restrict__ idx, int64_t nrows, int D,
                              float *__restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nrows * D) return;
    const int64_t k = t / D;
    const int d = (int)(t - k * D);
    out[t] = src[idx[k] * D + d];
}

// Pair features from exact q x q counts (q <= 8): joint and marginal frequencies, then the
// pair's mean-field coupling estimate.  The mean-field block comes from the connected
// correlations of the pair alone, with pseudocount weight 1/2 (f -> f/2 + 1/(2q) for the
// marginals, f/2 + 1/(2q^2) for the joint), over the first q-1 states of each locus (the
// last, the gap, is the reference state):
//   C = [[C_rr, C_rj], [C_jr, C_jj]],  J_mf = -(C^-1)_rj = T S^-1,
//   T = C_rr^-1 C_rj,  S = C_jj - C_jr T  (the Schur complement),
// with C_rr^-1 = diag(1/p_r) + 11^T / p_r(q-1) in closed form (Sherman-Morrison) and S
// inverted by Gauss-Jordan with partial pivoting, all in FP64.  Entries of the last state
// are 0.  Features: f_rj [q*q], f_r [q], f_j [q], J_mf [q*q].
__global__ void k_pair_features(const int32_t *__restrict__ counts, int64_t npairs, int q, int B,
                                float *__restrict__ feats)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const int D = 2 * q * q + 2 * q;
    const int32_t *c = counts + p * q * q;
    float *f = feats + p * D;
    double fr[8], fj[8];
    const double invB = 1.0 / (double)B;
    for (int s = 0; s < q; ++s) {
        fr[s] = 0.0;
        fj[s] = 0.0;
    }
    for (int s = 0; s < q; ++s)
        for (int a = 0; a < q; ++a) {
            const double v = (double)c[s * q + a] * invB;
            f[s * q + a] = (float)v;
            fr[s] += v;
            fj[a] += v;
        }
    for (int s = 0; s < q; ++s) {
        f[q * q + s] = (float)fr[s];
        f[q * q + q + s] = (float)fj[s];
    }
    // mean-field block
    const int n = q - 1;
    const double lam = 0.5;
    double pr[8], pj[8], T[7][7], S[7][7], Si[7][7];
    for (int s = 0; s < q; ++s) {
        pr[s] = (1.0 - lam) * fr[s] + lam / q;
        pj[s] = (1.0 - lam) * fj[s] + lam / q;
    }
    double crj[7][7], colsum[7];
    for (int a = 0; a < n; ++a) colsum[a] = 0.0;
    for (int s = 0; s < n; ++s)
        for (int a = 0; a < n; ++a) {
            const double pij = (1.0 - lam) * (double)c[s * q + a] * invB + lam / (q * q);
            crj[s][a] = pij - pr[s] * pj[a];
            colsum[a] += crj[s][a];
        }
    for (int s = 0; s < n; ++s)
        for (int a = 0; a < n; ++a) T[s][a] = crj[s][a] / pr[s] + colsum[a] / pr[n];
    for (int a = 0; a < n; ++a)
        for (int b2 = 0; b2 < n; ++b2) {
            double v = (a == b2 ? pj[a] : 0.0) - pj[a] * pj[b2];
            for (int s = 0; s < n; ++s) v -= crj[s][a] * T[s][b2];
            S[a][b2] = v;
            Si[a][b2] = a == b2 ? 1.0 : 0.0;
        }
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (fabs(S[r][col]) > fabs(S[piv][col])) piv = r;
        if (piv != col)
            for (int u = 0; u < n; ++u) {
                double t = S[col][u];
                S[col][u] = S[piv][u];
                S[piv][u] = t;
                t = Si[col][u];
                Si[col][u] = Si[piv][u];
                Si[piv][u] = t;
            }
        const double inv = 1.0 / S[col][col];
        for (int u = 0; u < n; ++u) {
            S[col][u] *= inv;
            Si[col][u] *= inv;
        }
        for (int r = 0; r < n; ++r) {
            if (r == col) continue;
            const double fct = S[r][col];
            for (int u = 0; u < n; ++u) {
                S[r][u] -= fct * S[col][u];
                Si[r][u] -= fct * Si[col][u];
            }
        }
    }
    float *mf = f + q * q + 2 * q;
    for (int s = 0; s < q; ++s)
        for (int a = 0; a < q; ++a) {
            double v = 0.0;
            if (s < n && a < n)
                for (int u = 0; u < n; ++u) v += T[s][u] * Si[u][a];
            mf[s * q + a] = (float)v;
        }
}

// Scatter compact warm starts into staged slot vectors.  Entry k refills slot slot_of[k]
// with item item_of[k] (locus locus_of[k]): partner list partners[item][0..P), network
// blocks J[item][p][s*q + a] and fields h[item][s]:
//   x[s][col(j, a)] = J  for observed a at partner j (j != r), x[s][Kx] = h[s].
// The caller zero-fills the staged vectors first.
// The blocks go to slot vector XJ + slot * jstride (the staged start, or the direction
// buffer as the first search direction, BF16 under split precision); the fields always go
// to X0.
__global__ void k_scatter_warm(const int32_t *__restrict__ slot_of, const int32_t *__restrict__ item_of,
                               const int32_t *__restrict__ locus_of, const int32_t *__restrict__ partners,
                               const float *__restrict__ J, const float *__restrict__ h, int nwarm, int P, int q,
                               int64_t Kxp, int Kx, const int32_t *__restrict__ state_col, int64_t n_slot,
                               float *__restrict__ X0, void *__restrict__ XJ, bool xj_bf16, int64_t jstride)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)P * q * q;
    if (t >= (int64_t)nwarm * (per + q)) return;
    const int k = (int)(t / (per + q));
    const int64_t u = t - (int64_t)k * (per + q);
    const int64_t item = item_of[k];
    float *x = X0 + (int64_t)slot_of[k] * n_slot;
    if (u >= per) {
        const int s = (int)(u - per);
        x[(int64_t)s * Kxp + Kx] = h[item * q + s];
        return;
    }
    const int p = (int)(u / (q * q));
    const int sa = (int)(u - (int64_t)p * q * q);
    const int s = sa / q, a = sa - (sa / q) * q;
    const int j = partners[item * P + p];
    if (j < 0 || j == locus_of[k]) return;
    const int col = state_col[(int64_t)j * q + a];
    if (col < 0) return;
    const int64_t o = (int64_t)slot_of[k] * jstride + (int64_t)s * Kxp + col;
    const float v = J[(item * P + p) * q * q + sa];
    if (xj_bf16) static_cast<uint16_t *>(XJ)[o] = f2bf(v);
    else static_cast<float *>(XJ)[o] = v;
}

}  // namespace

namespace kern {

static inline unsigned cd(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

void mlp_gemm(hipStream_t st, const uint16_t *A, int64_t lda, const uint16_t *B, int64_t ldb, int K, int64_t Mt, int Nt,
              const MlpEpi &ep)
{
    if (Mt <= 0 || Nt <= 0) return;
    hipLaunchKernelGGL(k_mlp_gemm, dim3(cd(Mt * Nt, 4)), dim3(256), 0, st, A, lda, B, ldb, K, Mt, Nt, ep);
}

void rows_bf16(hipStream_t st, const float *X, int64_t n, int D, int64_t npad, int DP, uint16_t *Xb, uint16_t *Xt,
               int64_t ldt)
{
    if (npad <= 0) return;
    hipLaunchKernelGGL(k_rows_bf16, dim3(cd(npad * DP, 256)), dim3(256), 0, st, X, n, D, npad, DP, Xb, Xt, ldt);
}

void pack_weights(hipStream_t st, const float *p, int din, int H, int dout, int DP, int DO, int64_t oW1, int64_t oW2,
                  int64_t oW3, uint16_t *W1b, uint16_t *W2b, uint16_t *W2t, uint16_t *W3b, uint16_t *W3t)
{
    const int64_t total = (int64_t)H * DP + (int64_t)H * H + (int64_t)DO * H;
    hipLaunchKernelGGL(k_pack_weights, dim3(cd(total, 256)), dim3(256), 0, st, p, din, H, dout, DP, DO, oW1, oW2, oW3,
                       W1b, W2b, W2t, W3b, W3t);
}

void colsum(hipStream_t st, const float *Z, int64_t ld, int64_t n, int O, float *out)
{
    hipLaunchKernelGGL(k_colsum, dim3(cd(O, 64)), dim3(64), 0, st, Z, ld, n, O, out);
}

int mse(hipStream_t st, const float *Y, const float *T, const float *M, float *dY, uint16_t *dYb, uint16_t *dYt,
        int64_t n, int64_t npad, int dout, int DO, int64_t ldt, float scale, double *part, double *loss_acc)
{
    const int nb = (int)cd(npad * DO, 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, Y, T, M, dY, dYb, dYt, n, npad, dout, DO, ldt, scale, part);
    hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, st, part, nb, loss_acc);
    return nb;
}

void adam(hipStream_t st, float *p, const float *g, float *m, float *v, int64_t n, float lr, float b1, float b2,
          float eps, float c1, float c2)
{
    hipLaunchKernelGGL(k_adam, dim3(cd(n, 256)), dim3(256), 0, st, p, g, m, v, n, lr, b1, b2, eps, c1, c2);
}

void gather_rows(hipStream_t st, const float *src, const int64_t *idx, int64_t nrows, int D, float *out)
{
    hipLaunchKernelGGL(k_gather_rows, dim3(cd(nrows * D, 256)), dim3(256), 0, st, src, idx, nrows, D, out);
}

void pair_features(hipStream_t st, const int32_t *counts, int64_t npairs, int q, int B, float *feats)
{
    if (npairs <= 0) return;
    hipLaunchKernelGGL(k_pair_features, dim3(cd(npairs, 128)), dim3(128), 0, st, counts, npairs, q, B, feats);
}

void scatter_warm(hipStream_t st, const int32_t *slot_of, const int32_t *item_of, const int32_t *locus_of,
                  const int32_t *partners, const float *J, const float *h, int nwarm, int P, int q, int64_t Kxp, int Kx,
                  const int32_t *state_col, int64_t n_slot, float *X0, void *XJ, bool xj_bf16, int64_t jstride)
{
    const int64_t total = (int64_t)nwarm * ((int64_t)P * q * q + q);
    if (total <= 0) return;
    hipLaunchKernelGGL(k_scatter_warm, dim3(cd(total, 256)), dim3(256), 0, st, slot_of, item_of, locus_of, partners, J,
                       h, nwarm, P, q, Kxp, Kx, state_col, n_slot, X0, XJ, xj_bf16, jstride);
}

}  // namespace kern
}  // namespace sdca
