// sdca_warm.cpp -- warm starts for the batched solve: pair features, the MLP (training on
// the device with the kernels of sdca_mlp.hip) and the compact warm-start tables that the
// solver scatters into its slots.  No counterpart in SuperDCA, whose solve always starts
// from zero (SuperDCA/include/plmDCA.hpp:298); a warm start changes only the starting
// point of a strictly convex problem, not its optimum.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "sdca_internal.hpp"

int sdca_fail(int code, const char *fmt, ...);  // sdca_api.cpp

namespace {

#define WTRY(expr)                                                                                                \
    do {                                                                                                          \
        hipError_t e_ = (expr);                                                                                   \
        if (e_ != hipSuccess) return sdca_fail(SDCA_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int walloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return sdca_fail(SDCA_E_HIP, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    e = hipMemset(*p, 0, count * sizeof(T));
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return sdca_fail(SDCA_E_HIP, "zero fill failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

template <typename T>
void wfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

struct Offs {
    int64_t W1, b1, W2, b2, W3, b3, n;
};

Offs offsets(int din, int dh, int dout)
{
    Offs o;
    o.W1 = 0;
    o.b1 = o.W1 + (int64_t)dh * din;
    o.W2 = o.b1 + dh;
    o.b2 = o.W2 + (int64_t)dh * dh;
    o.W3 = o.b2 + dh;
    o.b3 = o.W3 + (int64_t)dout * dh;
    o.n = o.b3 + dout;
    return o;
}

int64_t round32(int64_t n) { return (n + 31) / 32 * 32; }

void free_forward(sdca::Mlp *m)
{
    wfree(m->d_x);
    wfree(m->d_y);
    wfree(m->d_xb);
    wfree(m->d_h1b);
    wfree(m->d_h2b);
    m->cap = 0;
}

void free_training(sdca::Mlp *m)
{
    wfree(m->d_t);
    wfree(m->d_msk);
    wfree(m->d_dy);
    wfree(m->d_dz2);
    wfree(m->d_dz1);
    wfree(m->d_xt);
    wfree(m->d_h1t);
    wfree(m->d_h2t);
    wfree(m->d_dyb);
    wfree(m->d_dyt);
    wfree(m->d_dz2b);
    wfree(m->d_dz2t);
    wfree(m->d_dz1t);
    wfree(m->d_part);
    wfree(m->d_loss);
    wfree(m->d_idx);
    m->tcap = 0;
}

// forward buffers for at least `rows` rows
int ensure_scratch(sdca::Mlp *m, int64_t rows)
{
    rows = round32(std::max<int64_t>(rows, 1));
    if (m->cap >= rows) return SDCA_OK;
    free_forward(m);
    int rc;
    if ((rc = walloc(&m->d_x, rows * m->din)) || (rc = walloc(&m->d_y, rows * m->dout)) ||
        (rc = walloc(&m->d_xb, rows * m->DP)) || (rc = walloc(&m->d_h1b, rows * m->dh)) ||
        (rc = walloc(&m->d_h2b, rows * m->dh)))
        return rc;
    m->cap = rows;
    return SDCA_OK;
}

// forward and training buffers for mini-batches of at most `rows` rows
int ensure_training(sdca::Mlp *m, int64_t rows)
{
    rows = round32(std::max<int64_t>(rows, 1));
    int rc = ensure_scratch(m, rows);
    if (rc || m->tcap >= rows) return rc;
    free_training(m);
    const int64_t nblocks = rows * m->DO / 256 + 2;
    if ((rc = walloc(&m->d_t, rows * m->dout)) || (rc = walloc(&m->d_msk, rows * m->dout)) ||
        (rc = walloc(&m->d_dy, rows * m->dout)) || (rc = walloc(&m->d_dz2, rows * m->dh)) ||
        (rc = walloc(&m->d_dz1, rows * m->dh)) || (rc = walloc(&m->d_xt, rows * m->DP)) ||
        (rc = walloc(&m->d_h1t, rows * m->dh)) || (rc = walloc(&m->d_h2t, rows * m->dh)) ||
        (rc = walloc(&m->d_dyb, rows * m->DO)) || (rc = walloc(&m->d_dyt, rows * m->DO)) ||
        (rc = walloc(&m->d_dz2b, rows * m->dh)) || (rc = walloc(&m->d_dz2t, rows * m->dh)) ||
        (rc = walloc(&m->d_dz1t, rows * m->dh)) || (rc = walloc(&m->d_part, (size_t)nblocks)) ||
        (rc = walloc(&m->d_loss, 1)) || (rc = walloc(&m->d_idx, rows)))
        return rc;
    m->tcap = rows;
    return SDCA_OK;
}

// the BF16 operands from the FP32 masters (after a change of d_p)
void pack(hipStream_t st, sdca::Mlp *m)
{
    if (m->packed) return;
    const Offs o = offsets(m->din, m->dh, m->dout);
    sdca::kern::pack_weights(st, m->d_p, m->din, m->dh, m->dout, m->DP, m->DO, o.W1, o.W2, o.W3, m->d_W1b, m->d_W2b,
                             m->d_W2t, m->d_W3b, m->d_W3t);
    m->packed = true;
}

// Forward of n rows of FP32 features x (on the device, [n][din]) into m->d_y [n][dout];
// with train, the transposed BF16 input and activations that backward() reads as well
// (n <= tcap).  Three BF16 products: H1 = relu(X W1^T + b1), H2 = relu(H1 W2^T + b2),
// Y = H2 W3^T + b3.
void forward(hipStream_t st, sdca::Mlp *m, const float *x, int64_t n, bool train)
{
    if (n <= 0) return;
    pack(st, m);
    const Offs o = offsets(m->din, m->dh, m->dout);
    const int64_t npad = round32(n), Mt = npad / 32, ldt = m->tcap;
    const int Ht = m->dh / 32;
    sdca::kern::rows_bf16(st, x, n, m->din, npad, m->DP, m->d_xb, train ? m->d_xt : nullptr, ldt);
    sdca::kern::MlpEpi e1;
    e1.bias = m->d_p + o.b1;
    e1.relu = 1;
    e1.bf = m->d_h1b;
    e1.ldbf = m->dh;
    e1.bt = train ? m->d_h1t : nullptr;
    e1.ldbt = ldt;
    e1.M = n;
    e1.N = m->dh;
    sdca::kern::mlp_gemm(st, m->d_xb, m->DP, m->d_W1b, m->DP, m->DP, Mt, Ht, e1);
    sdca::kern::MlpEpi e2 = e1;
    e2.bias = m->d_p + o.b2;
    e2.bf = m->d_h2b;
    e2.bt = train ? m->d_h2t : nullptr;
    sdca::kern::mlp_gemm(st, m->d_h1b, m->dh, m->d_W2b, m->dh, m->dh, Mt, Ht, e2);
    sdca::kern::MlpEpi e3;
    e3.bias = m->d_p + o.b3;
    e3.f32 = m->d_y;
    e3.ldf = m->dout;
    e3.M = n;
    e3.N = m->dout;
    sdca::kern::mlp_gemm(st, m->d_h2b, m->dh, m->d_W3b, m->dh, m->dh, Mt, m->DO / 32, e3);
}

// Masked MSE gradient of the n rows of the last training forward (targets t and mask on
// the device) into m->d_g; the loss is accumulated in d_loss.  Weight gradients are BF16
// products over the batch (dW = dZ^T H); bias gradients FP32 column sums.
void backward(hipStream_t st, sdca::Mlp *m, const float *t, const float *mask, int64_t n, float count)
{
    const Offs o = offsets(m->din, m->dh, m->dout);
    const int64_t npad = round32(n), Mt = npad / 32, ldt = m->tcap;
    const int Ht = m->dh / 32;
    sdca::kern::mse(st, m->d_y, t, mask, m->d_dy, m->d_dyb, m->d_dyt, n, npad, m->dout, m->DO, ldt, 2.0f / count,
                    m->d_part, m->d_loss);
    sdca::kern::MlpEpi w3;  // dW3 [dout][dh] = dY^T H2
    w3.f32 = m->d_g + o.W3;
    w3.ldf = m->dh;
    w3.M = m->dout;
    w3.N = m->dh;
    sdca::kern::mlp_gemm(st, m->d_dyt, ldt, m->d_h2t, ldt, (int)npad, m->DO / 32, Ht, w3);
    sdca::kern::colsum(st, m->d_dy, m->dout, n, m->dout, m->d_g + o.b3);
    sdca::kern::MlpEpi z2;  // dZ2 = (dY W3) * [H2 > 0]
    z2.mask = m->d_h2b;
    z2.ldm = m->dh;
    z2.f32 = m->d_dz2;
    z2.ldf = m->dh;
    z2.bf = m->d_dz2b;
    z2.ldbf = m->dh;
    z2.bt = m->d_dz2t;
    z2.ldbt = ldt;
    z2.M = n;
    z2.N = m->dh;
    sdca::kern::mlp_gemm(st, m->d_dyb, m->DO, m->d_W3t, m->DO, m->DO, Mt, Ht, z2);
    sdca::kern::MlpEpi w2;  // dW2 [dh][dh] = dZ2^T H1
    w2.f32 = m->d_g + o.W2;
    w2.ldf = m->dh;
    w2.M = m->dh;
    w2.N = m->dh;
    sdca::kern::mlp_gemm(st, m->d_dz2t, ldt, m->d_h1t, ldt, (int)npad, Ht, Ht, w2);
    sdca::kern::colsum(st, m->d_dz2, m->dh, n, m->dh, m->d_g + o.b2);
    sdca::kern::MlpEpi z1;  // dZ1 = (dZ2 W2) * [H1 > 0]
    z1.mask = m->d_h1b;
    z1.ldm = m->dh;
    z1.f32 = m->d_dz1;
    z1.ldf = m->dh;
    z1.bt = m->d_dz1t;
    z1.ldbt = ldt;
    z1.M = n;
    z1.N = m->dh;
    sdca::kern::mlp_gemm(st, m->d_dz2b, m->dh, m->d_W2t, m->dh, m->dh, Mt, Ht, z1);
    sdca::kern::MlpEpi w1;  // dW1 [dh][din] = dZ1^T X
    w1.f32 = m->d_g + o.W1;
    w1.ldf = m->din;
    w1.M = m->dh;
    w1.N = m->din;
    sdca::kern::mlp_gemm(st, m->d_dz1t, ldt, m->d_xt, ldt, (int)npad, Ht, m->DP / 32, w1);
    sdca::kern::colsum(st, m->d_dz1, m->dh, n, m->dh, m->d_g + o.b1);
}

int check(sdca_ctx *c)
{
    if (!c) return sdca_fail(SDCA_E_ARG, "null context");
    WTRY(hipSetDevice(c->device));
    return SDCA_OK;
}

int need_mlp(sdca_ctx *c)
{
    if (!c->mlp) return sdca_fail(SDCA_E_STATE, "no network: call sdca_mlp_init or sdca_mlp_set_params first");
    return SDCA_OK;
}

int create_mlp(sdca_ctx *c, int hidden)
{
    if (hidden < 32 || hidden > 1024 || hidden % 32 != 0)
        return sdca_fail(SDCA_E_ARG, "hidden width %d: a multiple of 32 in [32, 1024] expected", hidden);
    delete c->mlp;
    c->mlp = new sdca::Mlp();
    sdca::Mlp *m = c->mlp;
    m->din = 2 * c->q * c->q + 2 * c->q;
    m->dh = hidden;
    m->dout = c->q * c->q;
    m->DP = (int)round32(m->din);
    m->DO = (int)round32(m->dout);
    m->np = offsets(m->din, m->dh, m->dout).n;
    int rc;
    if ((rc = walloc(&m->d_p, m->np)) || (rc = walloc(&m->d_g, m->np)) || (rc = walloc(&m->d_m, m->np)) ||
        (rc = walloc(&m->d_v, m->np)) || (rc = walloc(&m->d_W1b, (size_t)m->dh * m->DP)) ||
        (rc = walloc(&m->d_W2b, (size_t)m->dh * m->dh)) || (rc = walloc(&m->d_W2t, (size_t)m->dh * m->dh)) ||
        (rc = walloc(&m->d_W3b, (size_t)m->DO * m->dh)) || (rc = walloc(&m->d_W3t, (size_t)m->dh * m->DO)))
        return rc;
    return SDCA_OK;
}

// Closed-form fields of the J = 0 model: argmin_h  W lse(h) - c.h + lambda_h |h|^2 with
// c_s = sum_b w_b [x_br = s] (the h-part of plmDCA_objective.hpp:86-151,195-198), Newton.
void fields_j0(const sdca_ctx *c, int r, double *h)
{
    const int q = c->q;
    std::vector<double> cs(q, 0.0), p(q), g(q), H((size_t)q * q), d(q);
    double W = 0.0;
    for (int b = 0; b < c->B; ++b) {
        cs[c->h_states[(size_t)b * c->L + r]] += c->h_w[b];
        W += c->h_w[b];
    }
    std::fill(h, h + q, 0.0);
    for (int it = 0; it < 60; ++it) {
        double mx = h[0];
        for (int s = 1; s < q; ++s) mx = std::max(mx, h[s]);
        double z = 0.0;
        for (int s = 0; s < q; ++s) z += (p[s] = std::exp(h[s] - mx));
        double gmax = 0.0;
        for (int s = 0; s < q; ++s) {
            p[s] /= z;
            g[s] = W * p[s] - cs[s] + 2.0 * c->lambda_h * h[s];
            gmax = std::max(gmax, std::fabs(g[s]));
        }
        if (gmax < 1e-12 * std::max(1.0, W)) break;
        for (int s = 0; s < q; ++s)
            for (int u = 0; u < q; ++u) H[(size_t)s * q + u] = W * ((s == u ? p[s] : 0.0) - p[s] * p[u]) + (s == u ? 2.0 * c->lambda_h : 0.0);
        // Gaussian elimination with partial pivoting on the q x q system H d = g
        std::vector<double> A(H), bb(g);
        for (int col = 0; col < q; ++col) {
            int piv = col;
            for (int rr = col + 1; rr < q; ++rr)
                if (std::fabs(A[(size_t)rr * q + col]) > std::fabs(A[(size_t)piv * q + col])) piv = rr;
            for (int u = 0; u < q; ++u) std::swap(A[(size_t)col * q + u], A[(size_t)piv * q + u]);
            std::swap(bb[col], bb[piv]);
            for (int rr = col + 1; rr < q; ++rr) {
                const double f = A[(size_t)rr * q + col] / A[(size_t)col * q + col];
                for (int u = col; u < q; ++u) A[(size_t)rr * q + u] -= f * A[(size_t)col * q + u];
                bb[rr] -= f * bb[col];
            }
        }
        for (int rr = q - 1; rr >= 0; --rr) {
            double acc = bb[rr];
            for (int u = rr + 1; u < q; ++u) acc -= A[(size_t)rr * q + u] * d[u];
            d[rr] = acc / A[(size_t)rr * q + rr];
        }
        for (int s = 0; s < q; ++s) h[s] -= d[s];
    }
}

}  // namespace

sdca::Mlp::~Mlp()
{
    free_forward(this);
    free_training(this);
    wfree(d_p);
    wfree(d_g);
    wfree(d_m);
    wfree(d_v);
    wfree(d_W1b);
    wfree(d_W2b);
    wfree(d_W2t);
    wfree(d_W3b);
    wfree(d_W3t);
}

extern "C" {

int sdca_mlp_init(sdca_ctx *c, int32_t hidden, uint64_t seed)
{
    int rc = check(c);
    if (rc || (rc = create_mlp(c, hidden))) return rc;
    sdca::Mlp *m = c->mlp;
    // He-uniform weights, zero biases
    std::mt19937_64 gen(seed);
    std::vector<float> p(m->np, 0.f);
    const Offs o = offsets(m->din, m->dh, m->dout);
    auto fill = [&](int64_t off, int64_t count, int fan_in) {
        std::uniform_real_distribution<float> U(-std::sqrt(6.0f / fan_in), std::sqrt(6.0f / fan_in));
        for (int64_t k = 0; k < count; ++k) p[off + k] = U(gen);
    };
    fill(o.W1, (int64_t)m->dh * m->din, m->din);
    fill(o.W2, (int64_t)m->dh * m->dh, m->dh);
    fill(o.W3, (int64_t)m->dout * m->dh, m->dh);
    for (int64_t k = o.W3; k < o.b3; ++k) p[k] *= 0.1f;  // small initial outputs
    WTRY(hipMemcpy(m->d_p, p.data(), m->np * 4, hipMemcpyHostToDevice));
    return SDCA_OK;
}

int64_t sdca_mlp_num_params(const sdca_ctx *c) { return (c && c->mlp) ? c->mlp->np : -1; }

int sdca_mlp_get_params(sdca_ctx *c, float *out, int64_t n)
{
    int rc = check(c);
    if (rc || (rc = need_mlp(c))) return rc;
    if (!out || n != c->mlp->np) return sdca_fail(SDCA_E_ARG, "expected %lld parameters", (long long)c->mlp->np);
    WTRY(hipMemcpy(out, c->mlp->d_p, n * 4, hipMemcpyDeviceToHost));
    return SDCA_OK;
}

int sdca_mlp_set_params(sdca_ctx *c, int32_t hidden, const float *p, int64_t n)
{
    int rc = check(c);
    if (rc || (rc = create_mlp(c, hidden))) return rc;
    if (!p || n != c->mlp->np) return sdca_fail(SDCA_E_ARG, "expected %lld parameters", (long long)c->mlp->np);
    WTRY(hipMemcpy(c->mlp->d_p, p, n * 4, hipMemcpyHostToDevice));
    return SDCA_OK;
}

int sdca_mlp_forward(sdca_ctx *c, const float *x, int64_t n, float *y)
{
    int rc = check(c);
    if (rc || (rc = need_mlp(c))) return rc;
    if (n < 0 || (n > 0 && (!x || !y))) return sdca_fail(SDCA_E_ARG, "bad arguments");
    sdca::Mlp *m = c->mlp;
    const int64_t chunk = 1 << 16;
    if ((rc = ensure_scratch(m, std::min(n, chunk)))) return rc;
    for (int64_t r0 = 0; r0 < n; r0 += chunk) {
        const int64_t cnt = std::min(chunk, n - r0);
        WTRY(hipMemcpyAsync(m->d_x, x + r0 * m->din, cnt * m->din * 4, hipMemcpyHostToDevice, c->stream));
        forward(c->stream, m, m->d_x, cnt, false);
        WTRY(hipGetLastError());
        WTRY(hipMemcpyAsync(y + r0 * m->dout, m->d_y, cnt * m->dout * 4, hipMemcpyDeviceToHost, c->stream));
        WTRY(hipStreamSynchronize(c->stream));
    }
    return SDCA_OK;
}

int sdca_mlp_loss_grad(sdca_ctx *c, const float *x, const float *t, const float *mask, int64_t n, double *loss,
                       float *grad)
{
    int rc = check(c);
    if (rc || (rc = need_mlp(c))) return rc;
    if (n < 1 || !x || !t || !loss || !grad) return sdca_fail(SDCA_E_ARG, "bad arguments");
    sdca::Mlp *m = c->mlp;
    if ((rc = ensure_training(m, n))) return rc;
    WTRY(hipMemcpyAsync(m->d_x, x, n * m->din * 4, hipMemcpyHostToDevice, c->stream));
    WTRY(hipMemcpyAsync(m->d_t, t, n * m->dout * 4, hipMemcpyHostToDevice, c->stream));
    double count = (double)n * m->dout;
    if (mask) {
        WTRY(hipMemcpyAsync(m->d_msk, mask, n * m->dout * 4, hipMemcpyHostToDevice, c->stream));
        count = 0.0;
        for (int64_t k = 0; k < n * m->dout; ++k) count += mask[k];
    }
    if (count <= 0.0) return sdca_fail(SDCA_E_ARG, "mask selects no outputs");
    WTRY(hipMemsetAsync(m->d_loss, 0, 8, c->stream));
    forward(c->stream, m, m->d_x, n, true);
    backward(c->stream, m, m->d_t, mask ? m->d_msk : nullptr, n, (float)count);
    WTRY(hipGetLastError());
    double l = 0.0;
    WTRY(hipMemcpyAsync(&l, m->d_loss, 8, hipMemcpyDeviceToHost, c->stream));
    WTRY(hipMemcpyAsync(grad, m->d_g, m->np * 4, hipMemcpyDeviceToHost, c->stream));
    WTRY(hipStreamSynchronize(c->stream));
    *loss = l / count;
    return SDCA_OK;
}

int sdca_mlp_train(sdca_ctx *c, const float *x, const float *t, const float *mask, int64_t n, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double *epoch_loss)
{
    int rc = check(c);
    if (rc || (rc = need_mlp(c))) return rc;
    if (n < 1 || !x || !t || epochs < 1 || batch < 1 || !(lr > 0.f)) return sdca_fail(SDCA_E_ARG, "bad arguments");
    sdca::Mlp *m = c->mlp;
    const int64_t nb = std::min<int64_t>(batch, n);
    if ((rc = ensure_training(m, nb))) return rc;
    float *dX = nullptr, *dT = nullptr, *dM = nullptr;
    auto cleanup = [&]() {
        wfree(dX);
        wfree(dT);
        wfree(dM);
    };
    if ((rc = walloc(&dX, n * m->din)) || (rc = walloc(&dT, n * m->dout)) || (rc = walloc(&dM, n * m->dout))) {
        cleanup();
        return rc;
    }
    std::vector<float> ones;
    hipError_t e = hipMemcpy(dX, x, n * m->din * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dT, t, n * m->dout * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (mask) e = hipMemcpy(dM, mask, n * m->dout * 4, hipMemcpyHostToDevice);
        else {
            ones.assign(n * m->dout, 1.f);
            e = hipMemcpy(dM, ones.data(), n * m->dout * 4, hipMemcpyHostToDevice);
        }
    }
    if (e != hipSuccess) {
        cleanup();
        return sdca_fail(SDCA_E_HIP, "upload of the training set failed: %s", hipGetErrorString(e));
    }
    std::vector<int64_t> perm(n);
    std::mt19937_64 gen(seed);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs && e == hipSuccess; ++ep) {
        std::iota(perm.begin(), perm.end(), 0);
        std::shuffle(perm.begin(), perm.end(), gen);
        e = hipMemsetAsync(m->d_loss, 0, 8, c->stream);
        double total = 0.0;
        for (int64_t b0 = 0; b0 < n && e == hipSuccess; b0 += nb) {
            const int64_t cnt = std::min(nb, n - b0);
            e = hipMemcpyAsync(m->d_idx, perm.data() + b0, cnt * 8, hipMemcpyHostToDevice, c->stream);
            if (e != hipSuccess) break;
            sdca::kern::gather_rows(c->stream, dX, m->d_idx, cnt, m->din, m->d_x);
            sdca::kern::gather_rows(c->stream, dT, m->d_idx, cnt, m->dout, m->d_t);
            sdca::kern::gather_rows(c->stream, dM, m->d_idx, cnt, m->dout, m->d_msk);
            // the mask weight of this batch decides the gradient scale
            double count = 0.0;
            if (mask)
                for (int64_t k = 0; k < cnt; ++k)
                    for (int u = 0; u < m->dout; ++u) count += mask[perm[b0 + k] * m->dout + u];
            else count = (double)cnt * m->dout;
            if (count <= 0.0) continue;
            total += count;
            forward(c->stream, m, m->d_x, cnt, true);
            backward(c->stream, m, m->d_t, m->d_msk, cnt, (float)count);
            m->step += 1;
            const float c1 = 1.f - std::pow(b1, (float)m->step), c2 = 1.f - std::pow(b2, (float)m->step);
            sdca::kern::adam(c->stream, m->d_p, m->d_g, m->d_m, m->d_v, m->np, lr, b1, b2, eps, c1, c2);
            m->packed = false;
            e = hipGetLastError();
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // perm slice is read asynchronously
        }
        double l = 0.0;
        if (e == hipSuccess) e = hipMemcpy(&l, m->d_loss, 8, hipMemcpyDeviceToHost);
        if (epoch_loss) epoch_loss[ep] = total > 0.0 ? l / total : 0.0;
    }
    cleanup();
    if (e != hipSuccess) return sdca_fail(SDCA_E_HIP, "training failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

int sdca_pair_counts(sdca_ctx *ctx, const int32_t *pairs, int64_t npairs, int32_t *counts_out);

int sdca_pair_features(sdca_ctx *c, const int32_t *pairs, int64_t npairs, float *feats)
{
    int rc = check(c);
    if (rc) return rc;
    if (npairs < 0 || (npairs > 0 && (!pairs || !feats))) return sdca_fail(SDCA_E_ARG, "bad arguments");
    if (npairs == 0) return SDCA_OK;
    const int q = c->q, D = 2 * q * q + 2 * q;
    std::vector<int32_t> counts((size_t)npairs * q * q);
    if ((rc = sdca_pair_counts(c, pairs, npairs, counts.data()))) return rc;
    int32_t *d_c = nullptr;
    float *d_f = nullptr;
    if ((rc = walloc(&d_c, counts.size())) || (rc = walloc(&d_f, (size_t)npairs * D))) {
        wfree(d_c);
        return rc;
    }
    hipError_t e = hipMemcpy(d_c, counts.data(), counts.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        sdca::kern::pair_features(c->stream, d_c, npairs, q, c->B, d_f);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(feats, d_f, (size_t)npairs * D * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    wfree(d_c);
    wfree(d_f);
    if (e != hipSuccess) return sdca_fail(SDCA_E_HIP, "pair features failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

int sdca_mi_topk(sdca_ctx *ctx, int32_t K, int32_t *partners_out, float *mi_out);

int sdca_solve_loci_mlp(sdca_ctx *c, const int32_t *loci, int32_t K, int32_t npartners, int32_t warm_mode,
                        float *score_rows, double *x_out, sdca_locus_stats *st)
{
    int rc = check(c);
    if (rc || (rc = need_mlp(c))) return rc;
    if (warm_mode < 0 || warm_mode > 3) return sdca_fail(SDCA_E_ARG, "unknown warm_mode %d", warm_mode);
    if (!c->have_weights) return sdca_fail(SDCA_E_STATE, "sdca_set_weights must be called first");
    if (c->scan_split > 0) return sdca_fail(SDCA_E_STATE, "a two-alignment scan solves from a cold start only");
    if (K < 0 || (K > 0 && !loci)) return sdca_fail(SDCA_E_ARG, "bad arguments");
    if (npartners < 1 || npartners > c->L - 1) return sdca_fail(SDCA_E_ARG, "npartners must be in [1, L-1]");
    for (int k = 0; k < K; ++k)
        if (loci[k] < 0 || loci[k] >= c->L) return sdca_fail(SDCA_E_ARG, "locus %d out of range", loci[k]);
    if (K == 0) return SDCA_OK;
    const int q = c->q, P = npartners;
    sdca::Mlp *m = c->mlp;
    if (m->din != 2 * q * q + 2 * q || m->dout != q * q) return sdca_fail(SDCA_E_STATE, "network shape does not match q");
    // partners: the context's MI top-K table (computed here if missing or too short)
    if (!c->d_topk || c->topk_K < P) {
        std::vector<int32_t> tmp((size_t)c->L * P);
        if ((rc = sdca_mi_topk(c, P, tmp.data(), nullptr))) return rc;
    }
    const int KT = c->topk_K;
    int32_t *d_loci = nullptr, *d_part = nullptr, *d_pairs = nullptr, *d_cnt = nullptr;
    float *d_feat = nullptr, *d_J = nullptr, *d_h = nullptr;
    auto cleanup = [&]() {
        wfree(d_loci);
        wfree(d_part);
        wfree(d_pairs);
        wfree(d_cnt);
        wfree(d_feat);
        wfree(d_J);
        wfree(d_h);
    };
    const int64_t npairs = (int64_t)K * P;
    const int64_t chunk = std::min<int64_t>(npairs, 1 << 18);
    if ((rc = walloc(&d_part, (size_t)npairs)) || (rc = walloc(&d_J, (size_t)npairs * q * q)) ||
        (rc = walloc(&d_h, (size_t)K * q)) || (rc = walloc(&d_pairs, (size_t)chunk * 2)) ||
        (rc = walloc(&d_cnt, (size_t)chunk * q * q)) || (rc = walloc(&d_feat, (size_t)chunk * m->din)) ||
        (rc = ensure_scratch(m, chunk))) {
        cleanup();
        return rc;
    }
    // partner rows of the listed loci (first P of each stored list, in index order)
    std::vector<int32_t> part((size_t)npairs);
    {
        std::vector<int32_t> row(KT);
        for (int k = 0; k < K; ++k) {
            if (hipMemcpy(row.data(), c->d_topk + (size_t)loci[k] * KT, (size_t)KT * 4, hipMemcpyDeviceToHost) !=
                hipSuccess) {
                cleanup();
                return sdca_fail(SDCA_E_HIP, "reading the partner table failed");
            }
            std::copy(row.begin(), row.begin() + P, part.begin() + (size_t)k * P);
        }
    }
    hipError_t e = hipMemcpy(d_part, part.data(), part.size() * 4, hipMemcpyHostToDevice);
    const int W = (c->B + 63) / 64;
    if (e == hipSuccess && !c->d_SP) {
        if ((rc = walloc(&c->d_SP, (size_t)c->L * q * W))) {
            cleanup();
            return rc;
        }
        sdca::kern::seq_planes(c->stream, c->d_states, c->B, c->L, q, W, c->d_SP);
        e = hipGetLastError();
    }
    // features -> network -> J blocks, in chunks of pairs
    std::vector<int32_t> pairs((size_t)chunk * 2);
    for (int64_t p0 = 0; p0 < npairs && e == hipSuccess; p0 += chunk) {
        const int64_t cnt = std::min(chunk, npairs - p0);
        for (int64_t u = 0; u < cnt; ++u) {
            pairs[2 * u] = loci[(p0 + u) / P];
            pairs[2 * u + 1] = part[p0 + u];
        }
        e = hipMemcpyAsync(d_pairs, pairs.data(), cnt * 8, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) break;
        sdca::kern::pair_counts(c->stream, c->d_SP, q, W, d_pairs, cnt, d_cnt);
        sdca::kern::pair_features(c->stream, d_cnt, cnt, q, c->B, d_feat);
        forward(c->stream, m, d_feat, cnt, false);
        e = hipMemcpyAsync(d_J + p0 * q * q, m->d_y, cnt * q * q * 4, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);  // pairs is reused
    }
    // fields of the J = 0 model (or zero with SDCA_WARM_NO_FIELDS)
    std::vector<float> hf((size_t)K * q, 0.f);
    std::vector<double> hd(q);
    if (!(warm_mode & SDCA_WARM_NO_FIELDS))
        for (int k = 0; k < K; ++k) {
            fields_j0(c, loci[k], hd.data());
            for (int s = 0; s < q; ++s) hf[(size_t)k * q + s] = (float)hd[s];
        }
    if (e == hipSuccess) e = hipMemcpy(d_h, hf.data(), hf.size() * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        cleanup();
        return sdca_fail(SDCA_E_HIP, "building warm starts failed: %s", hipGetErrorString(e));
    }
    sdca::CompactWarm cw;
    cw.P = P;
    cw.d_partners = d_part;
    cw.d_J = d_J;
    cw.d_h = d_h;
    cw.as_direction = !(warm_mode & SDCA_WARM_POINT);
    rc = sdca_solve_impl(c, loci, K, nullptr, &cw, score_rows, x_out, st);
    cleanup();
    return rc;
}

}  // extern "C"
