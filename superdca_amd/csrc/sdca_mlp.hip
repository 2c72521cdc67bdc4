// sdca_mlp.hip -- the warm-start network and its training kernels (no SuperDCA counterpart).
//
// A three-layer perceptron maps the statistics of a locus pair (r, j) to a start value of
// the coupling block J_r[j](a, s):
//   input   f_rj(s, a) (q*q joint frequencies, s = state at r, a = state at j),
//           f_r(s), f_j(a) (2q marginals) and the pair's mutual information: q*q + 2q + 1
//   hidden  two layers of H units, ReLU
//   output  q*q values, laid out like the joint frequencies: out[s*q + a] = J_r[j](a, s)
// Parameters are FP32, stored contiguously: W1 [H][Din], b1 [H], W2 [H][H], b2 [H],
// W3 [Dout][H], b3 [Dout].  Training minimises the masked mean squared error with Adam
// (D. P. Kingma and J. Ba, ICLR 2015).  Every reduction runs in a fixed order, so results
// are deterministic.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdca_internal.hpp"

namespace sdca {
namespace {

// Y[n][o] = act(sum_i X[n][i] W[o][i] + b[o]); one thread per output element.
template <bool RELU>
__global__ void k_linear_fwd(const float *__restrict__ X, const float *__restrict__ W, const float *__restrict__ b,
                             float *__restrict__ Y, int64_t N, int I, int O)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * O) return;
    const int64_t n = idx / O;
    const int o = (int)(idx - n * O);
    const float *x = X + n * I;
    const float *w = W + (int64_t)o * I;
    float acc = b[o];
    for (int i = 0; i < I; ++i) acc += x[i] * w[i];
    Y[idx] = RELU ? fmaxf(acc, 0.f) : acc;
}

// dX[n][i] = (sum_o dY[n][o] W[o][i]) * [H[n][i] > 0]  (H = this layer's ReLU input's output)
__global__ void k_linear_bwd_dx(const float *__restrict__ dY, const float *__restrict__ W,
                                const float *__restrict__ H, float *__restrict__ dX, int64_t N, int I, int O)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * I) return;
    const int64_t n = idx / I;
    const int i = (int)(idx - n * I);
    const float *dy = dY + n * O;
    float acc = 0.f;
    for (int o = 0; o < O; ++o) acc += dy[o] * W[(int64_t)o * I + i];
    dX[idx] = H[idx] > 0.f ? acc : 0.f;
}

// dW[o][i] = sum_n dY[n][o] X[n][i]; db[o] = sum_n dY[n][o] (column i == I).
__global__ void k_linear_bwd_dw(const float *__restrict__ dY, const float *__restrict__ X, float *__restrict__ dW,
                                float *__restrict__ db, int64_t N, int I, int O)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)O * (I + 1)) return;
    const int o = (int)(idx / (I + 1));
    const int i = (int)(idx - (int64_t)o * (I + 1));
    float acc = 0.f;
    if (i < I) {
        for (int64_t n = 0; n < N; ++n) acc += dY[n * O + o] * X[n * I + i];
        dW[(int64_t)o * I + i] = acc;
    } else {
        for (int64_t n = 0; n < N; ++n) acc += dY[n * O + o];
        db[o] = acc;
    }
}

// dY = 2/(count) * (Y - T) * M and the per-block partial sums of M * (Y - T)^2 (FP64).
__global__ __launch_bounds__(256) void k_mse(const float *__restrict__ Y, const float *__restrict__ T,
                                             const float *__restrict__ M, float *__restrict__ dY, int64_t total,
                                             float scale, double *__restrict__ part)
{
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double l = 0.0;
    if (idx < total) {
        const float m = M ? M[idx] : 1.f;
        const float d = (Y[idx] - T[idx]) * m;
        dY[idx] = scale * d;
        l = (double)d * d;
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

// Pair features from exact q x q counts: joint and marginal frequencies and the mutual
// information (natural log) of the pair.
__global__ void k_pair_features(const int32_t *__restrict__ counts, int64_t npairs, int q, int B,
                                float *__restrict__ feats)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const int D = q * q + 2 * q + 1;
    const int32_t *c = counts + p * q * q;
    float *f = feats + p * D;
    float fr[8], fj[8];
    const float invB = 1.0f / (float)B;
    for (int s = 0; s < q; ++s) {
        fr[s] = 0.f;
        fj[s] = 0.f;
    }
    for (int s = 0; s < q; ++s)
        for (int a = 0; a < q; ++a) {
            const float v = (float)c[s * q + a] * invB;
            f[s * q + a] = v;
            fr[s] += v;
            fj[a] += v;
        }
    float mi = 0.f;
    for (int s = 0; s < q; ++s)
        for (int a = 0; a < q; ++a) {
            const float v = f[s * q + a];
            if (v > 0.f) mi += v * logf(v / (fr[s] * fj[a]));
        }
    for (int s = 0; s < q; ++s) {
        f[q * q + s] = fr[s];
        f[q * q + q + s] = fj[s];
    }
    f[q * q + 2 * q] = mi;
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

void linear_fwd(hipStream_t st, bool relu, const float *X, const float *W, const float *b, float *Y, int64_t N, int I,
                int O)
{
    if (relu) hipLaunchKernelGGL(k_linear_fwd<true>, dim3(cd(N * O, 256)), dim3(256), 0, st, X, W, b, Y, N, I, O);
    else hipLaunchKernelGGL(k_linear_fwd<false>, dim3(cd(N * O, 256)), dim3(256), 0, st, X, W, b, Y, N, I, O);
}

void linear_bwd_dx(hipStream_t st, const float *dY, const float *W, const float *H, float *dX, int64_t N, int I, int O)
{
    hipLaunchKernelGGL(k_linear_bwd_dx, dim3(cd(N * I, 256)), dim3(256), 0, st, dY, W, H, dX, N, I, O);
}

void linear_bwd_dw(hipStream_t st, const float *dY, const float *X, float *dW, float *db, int64_t N, int I, int O)
{
    hipLaunchKernelGGL(k_linear_bwd_dw, dim3(cd((int64_t)O * (I + 1), 128)), dim3(128), 0, st, dY, X, dW, db, N, I, O);
}

int mse(hipStream_t st, const float *Y, const float *T, const float *M, float *dY, int64_t total, float scale,
        double *part, double *loss_acc)
{
    const int nb = (int)cd(total, 256);
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, Y, T, M, dY, total, scale, part);
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
