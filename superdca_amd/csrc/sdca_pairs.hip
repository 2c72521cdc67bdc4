// sdca_pairs.hip -- pair statistics for the warm start (no counterpart in SuperDCA).
//
//   X8        Kc x Bp int8, Kc = L*QP: row j*QP + a is the indicator of state a at locus j
//             over the sequences (QP = q rounded up to 4 or 8; unused rows are zero)
//   counts    N_ij(a,b) = sum_b X8[i*QP+a][b] X8[j*QP+b][b]: an NT product of X8 with itself on
//             v_mfma_i32_32x32x32_i8, exact in int32 (B < 2^31)
//   MI        the epilogue turns every QP x QP locus-pair block of the count tile into the
//             mutual information sum_ab f_ab log(f_ab / (f_a f_b)), f = N / B, without ever
//             writing the counts
//   top-K     per locus, the K partners of largest MI (ties: lower index first), radix select
//   pair counts for listed pairs from sequence bit-planes (popcount), exact
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdca_internal.hpp"

namespace sdca {
namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x16 __attribute__((ext_vector_type(16)));
constexpr int PROWB = 144;  // LDS bytes per staged row of 128 int8

// X8[j*QP + a][b] = (states[b][j] == a); the caller zero-fills X8.
__global__ void k_build_x8(const uint8_t *__restrict__ states, int B, int L, int QP, int Bp, int8_t *__restrict__ X8)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)B * L) return;
    const int b = (int)(idx / L);
    const int j = (int)(idx - (int64_t)b * L);
    X8[((int64_t)j * QP + states[idx]) * Bp + b] = 1;
}

// Marginal counts cnt[j*QP + a] = number of sequences with state a at locus j.
__global__ void k_marginals(const uint8_t *__restrict__ states, int B, int L, int QP, int32_t *__restrict__ cnt)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= L) return;
    int c[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = 0; b < B; ++b) {
        const int a = states[(int64_t)b * L + j];
        for (int u = 0; u < 8; ++u) c[u] += (u == a);
    }
    for (int u = 0; u < QP; ++u) cnt[(int64_t)j * QP + u] = c[u];
}

// Count tile C[m][n] = sum_k A[m][k] B[n][k] for rows m of the panel (A = X8 + panel offset)
// and columns n of all of X8.  MODE 0 writes raw int32 counts (Cout[m][n]); MODE 1 writes
// the panel's MI rows MI[i_local][j] (i == j gets -1).
template <int MODE, int QP>
__global__ __launch_bounds__(256) void k_pair_gemm(const int8_t *__restrict__ Apan, const int8_t *__restrict__ Ball,
                                                   int Bp, int Mrows, int Ncols, int i0, int L, int B,
                                                   const int32_t *__restrict__ cnt, int32_t *__restrict__ Cout,
                                                   float *__restrict__ MI)
{
    __shared__ __attribute__((aligned(16))) uint8_t sA[128 * PROWB];
    __shared__ __attribute__((aligned(16))) uint8_t sB[128 * PROWB];
    // XCD-aware order: workgroup id % 8 selects the XCD; on each XCD the panel's row tiles
    // of one column tile run back to back, so that column tile is fetched into L2 once
    const int Mt = Mrows / 128, Nt = Ncols / 128;
    const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3;
    const int mt = local % Mt;
    const int nt = (local / Mt) * 8 + xcd;
    if (nt >= Nt) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m0 = (int64_t)mt * 128, n0 = (int64_t)nt * 128;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    i32x16 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0;
    const int lrow = tid >> 3, lcol = (tid & 7) * 16;
    i32x4 ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ra[i] = *reinterpret_cast<const i32x4 *>(Apan + (m0 + lrow + 32 * i) * Bp + k0 + lcol);
            rb[i] = *reinterpret_cast<const i32x4 *>(Ball + (n0 + lrow + 32 * i) * Bp + k0 + lcol);
        }
    };
    load(0);
    for (int k0 = 0; k0 < Bp; k0 += 128) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            *reinterpret_cast<i32x4 *>(&sA[(lrow + 32 * i) * PROWB + lcol]) = ra[i];
            *reinterpret_cast<i32x4 *>(&sB[(lrow + 32 * i) * PROWB + lcol]) = rb[i];
        }
        __syncthreads();
        if (k0 + 128 < Bp) load(k0 + 128);
#pragma unroll
        for (int kk = 0; kk < 128; kk += 32) {
            const int kb = kk + (lane >> 5) * 16;
            i32x4 a[2], b[2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
                a[mi] = *reinterpret_cast<const i32x4 *>(&sA[(wm + mi * 32 + (lane & 31)) * PROWB + kb]);
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
                b[ni] = *reinterpret_cast<const i32x4 *>(&sB[(wn + ni * 32 + (lane & 31)) * PROWB + kb]);
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni)
                    acc[mi][ni] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[mi], b[ni], acc[mi][ni], 0, 0, 0);
        }
    }
    // result layout of a 32x32 block: register v of lane l is row 8*(v/4) + 4*(l/32) + v%4,
    // column l%32
    if constexpr (MODE == 0) {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int64_t m = m0 + wm + mi * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    const int64_t n = n0 + wn + ni * 32 + (lane & 31);
                    Cout[m * Ncols + n] = acc[mi][ni][v];
                }
    } else {
        const float invB = 1.0f / (float)B;
        const float fB = (float)B;
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int64_t cg = n0 + wn + ni * 32 + (lane & 31);  // global column of this lane
                const int j = (int)(cg / QP);
                const int bst = (int)(cg % QP);
                const float nb = (float)cnt[cg];
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    // rows 8g + 4h + t of the block, h = lane / 32, t = 0..3
                    const int64_t rbase = m0 + wm + mi * 32 + 8 * g + 4 * (lane >> 5);
                    float part = 0.f;
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        const int64_t rg = rbase + t;  // panel row
                        const int n = acc[mi][ni][4 * g + t];
                        if (n > 0) {
                            const float na = (float)cnt[(int64_t)i0 * QP + rg];
                            part += (float)n * invB * logf((float)n * fB / (na * nb));
                        }
                    }
                    // reduce over the QP lanes (and both row halves if QP = 8) of one locus pair
                    part += __shfl_xor(part, 1, 64);
                    part += __shfl_xor(part, 2, 64);
                    if constexpr (QP == 8) {
                        part += __shfl_xor(part, 4, 64);
                        part += __shfl_xor(part, 32, 64);
                    }
                    const int il = (int)((rbase) / QP);  // panel-local locus of these rows
                    const int i = i0 + il;
                    const bool writer = (bst == 0) && (QP == 4 || (lane >> 5) == 0);
                    if (writer && i < L && j < L) MI[(int64_t)il * L + j] = (i == j) ? -1.f : part;
                }
            }
    }
}

// ----------------------------------------------------------------- top-K

__device__ __forceinline__ int block_excl_scan(int v, int *sh, int &total)
{
    // 256 threads, 4 waves
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int off = 0;
    for (int w = 0; w < wave; ++w) off += sh[w];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + x - v;
}

// One workgroup per row: the K largest non-negative values of MI[rr][0..L) (negative = the
// locus itself, excluded), by radix select on the float bits, then gathered in index order
// with ties at the threshold taken lowest index first.
__global__ __launch_bounds__(256) void k_topk(const float *__restrict__ MI, int L, int K, int i0,
                                              int32_t *__restrict__ out_idx, float *__restrict__ out_val)
{
    __shared__ int hist[256];
    __shared__ int sh[4];
    __shared__ uint32_t s_prefix;
    __shared__ int s_need;
    const int rr = blockIdx.x;
    const float *row = MI + (int64_t)rr * L;
    const uint32_t *bits = reinterpret_cast<const uint32_t *>(row);
    uint32_t prefix = 0, mask = 0;
    int need = K;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int u = threadIdx.x; u < 256; u += 256) hist[u] = 0;
        __syncthreads();
        for (int e = threadIdx.x; e < L; e += 256) {
            const uint32_t v = bits[e];
            if ((int32_t)v >= 0 && (v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255], 1);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int cum = 0, bsel = 0;
            for (int bk = 255; bk >= 0; --bk) {
                if (cum + hist[bk] >= need) {
                    bsel = bk;
                    break;
                }
                cum += hist[bk];
            }
            s_prefix = prefix | ((uint32_t)bsel << shift);
            s_need = need - cum;
        }
        __syncthreads();
        prefix = s_prefix;
        need = s_need;
        mask |= 255u << shift;
        __syncthreads();
    }
    // prefix = bits of the K-th largest value; take all above it and `need` ties
    int base = 0, ties = 0;
    int32_t *oi = out_idx + (int64_t)(i0 + rr) * K;
    float *ov = out_val ? out_val + (int64_t)(i0 + rr) * K : nullptr;
    for (int c0 = 0; c0 < L; c0 += 256) {
        const int e = c0 + threadIdx.x;
        uint32_t v = 0;
        bool valid = false;
        if (e < L) {
            v = bits[e];
            valid = (int32_t)v >= 0;
        }
        const int is_tie = (valid && v == prefix) ? 1 : 0;
        int ntie;
        const int tie_rank = block_excl_scan(is_tie, sh, ntie) + ties;
        const int take = (valid && (v > prefix || (is_tie && tie_rank < need))) ? 1 : 0;
        int ntake;
        const int pos = block_excl_scan(take, sh, ntake) + base;
        if (take && pos < K) {
            oi[pos] = e;
            if (ov) ov[pos] = row[e];
        }
        base += ntake;
        ties += ntie;
    }
}

// ------------------------------------------------- counts of listed pairs

// Sequence bit-planes: SP[j][a][w] bit t = (states[64w + t][j] == a).
__global__ void k_seq_planes(const uint8_t *__restrict__ states, int B, int L, int q, int W, uint64_t *__restrict__ SP)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)L * W) return;
    const int j = (int)(idx / W);
    const int w = (int)(idx - (int64_t)j * W);
    uint64_t word[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int b0 = w * 64, b1 = min(B, b0 + 64);
    for (int b = b0; b < b1; ++b) {
        const int a = states[(int64_t)b * L + j];
        for (int u = 0; u < 8; ++u)
            if (u == a) word[u] |= 1ull << (b - b0);
    }
    for (int a = 0; a < q; ++a) SP[((int64_t)j * q + a) * W + w] = word[a];
}

// counts[p][a][b] = popcount(SP[i_p][a] & SP[j_p][b]) summed over words; one wave per pair.
template <int Q>
__global__ __launch_bounds__(256) void k_pair_counts(const uint64_t *__restrict__ SP, int W,
                                                     const int32_t *__restrict__ pairs, int64_t npairs,
                                                     int32_t *__restrict__ counts)
{
    const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (p >= npairs) return;
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    int c[Q][Q];
#pragma unroll
    for (int a = 0; a < Q; ++a)
#pragma unroll
        for (int b = 0; b < Q; ++b) c[a][b] = 0;
    for (int w = lane; w < W; w += 64) {
        uint64_t pj[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b) pj[b] = SP[((int64_t)j * Q + b) * W + w];
#pragma unroll
        for (int a = 0; a < Q; ++a) {
            const uint64_t pi = SP[((int64_t)i * Q + a) * W + w];
#pragma unroll
            for (int b = 0; b < Q; ++b) c[a][b] += __popcll(pi & pj[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < Q; ++a)
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            int v = c[a][b];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
            if (lane == 0) counts[(p * Q + a) * Q + b] = v;
        }
}

}  // namespace

namespace kern {

static inline unsigned cdiv2(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

void build_x8(hipStream_t st, const uint8_t *states, int B, int L, int QP, int Bp, int8_t *X8)
{
    hipLaunchKernelGGL(k_build_x8, dim3(cdiv2((int64_t)B * L, 256)), dim3(256), 0, st, states, B, L, QP, Bp, X8);
}

void marginals(hipStream_t st, const uint8_t *states, int B, int L, int QP, int32_t *cnt)
{
    hipLaunchKernelGGL(k_marginals, dim3(cdiv2(L, 256)), dim3(256), 0, st, states, B, L, QP, cnt);
}

void pair_gemm(hipStream_t st, int mode, int QP, const int8_t *Apan, const int8_t *Ball, int Bp, int Mrows,
               int Ncols, int i0, int L, int B, const int32_t *cnt, int32_t *Cout, float *MI)
{
    const int Nt = Ncols / 128;
    const dim3 grid((unsigned)((Mrows / 128) * 8 * ((Nt + 7) / 8)));
#define SDCA_PG(MODE, Q)                                                                                          \
    hipLaunchKernelGGL((k_pair_gemm<MODE, Q>), grid, dim3(256), 0, st, Apan, Ball, Bp, Mrows, Ncols, i0, L, B, cnt, \
                       Cout, MI)
    if (mode == 0) {
        if (QP == 4) SDCA_PG(0, 4);
        else SDCA_PG(0, 8);
    } else {
        if (QP == 4) SDCA_PG(1, 4);
        else SDCA_PG(1, 8);
    }
#undef SDCA_PG
}

void topk(hipStream_t st, const float *MI, int rows, int L, int K, int i0, int32_t *out_idx, float *out_val)
{
    hipLaunchKernelGGL(k_topk, dim3(rows), dim3(256), 0, st, MI, L, K, i0, out_idx, out_val);
}

void seq_planes(hipStream_t st, const uint8_t *states, int B, int L, int q, int W, uint64_t *SP)
{
    hipLaunchKernelGGL(k_seq_planes, dim3(cdiv2((int64_t)L * W, 256)), dim3(256), 0, st, states, B, L, q, W, SP);
}

void pair_counts(hipStream_t st, const uint64_t *SP, int q, int W, const int32_t *pairs, int64_t npairs,
                 int32_t *counts)
{
    if (npairs <= 0) return;
    const dim3 grid(cdiv2(npairs, 4));
#define SDCA_PC(Q) hipLaunchKernelGGL(k_pair_counts<Q>, grid, dim3(256), 0, st, SP, W, pairs, npairs, counts)
    switch (q) {
    case 2: SDCA_PC(2); break;
    case 3: SDCA_PC(3); break;
    case 4: SDCA_PC(4); break;
    case 5: SDCA_PC(5); break;
    case 6: SDCA_PC(6); break;
    case 7: SDCA_PC(7); break;
    default: SDCA_PC(8); break;
    }
#undef SDCA_PC
}

}  // namespace kern
}  // namespace sdca
