// sdca_kernels.hip -- gfx950 kernels of the batched plmDCA solve.
//
// The per-locus asymmetric pseudolikelihood of SuperDCA
// (SuperDCA/include/plmDCA_objective.hpp:38-249) is evaluated for a batch of loci as
// two dense contractions over the shared one-hot alignment X:
//   energies   E^T[r*q+s][b] = sum_k W[r*q+s][k] X[b][k]          (gemm_energy)
//   gradient   dW[r*q+s][k]  = sum_b G^T[r*q+s][b] X^T[k][b] + 2 lambda W   (gemm_grad)
// with G = w (softmax(E) - onehot(x_r)) from the softmax kernel.  Both contractions are
// "NT" products (both operands contiguous along the reduction index) on 32x32 MFMA
// tiles: FP32 (v_mfma_f32_32x32x2f32) or BF16 (v_mfma_f32_32x32x16_bf16) with W or G
// split into hi + lo BF16 parts accumulated in FP32.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "sdca_internal.hpp"

namespace sdca {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));


// ---------------------------------------------------------------- one-hot X

// One-hot matrices as bits: X[b][col] set for col = state_col[j*q + x_bj] and the bias
// column Kx (b < B), rows of Kxp/32 words; Xt[col][b] likewise, rows of Bp/32 words.
// The caller zero-fills both first.
__global__ void k_build_x(const uint8_t *__restrict__ states, int B, int L, int q, const int32_t *__restrict__ state_col,
                          int Kx, int Bp, int64_t Kxp, uint32_t *__restrict__ X, uint32_t *__restrict__ Xt)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)B * (L + 1);
    if (idx >= total) return;
    const int b = (int)(idx / (L + 1));
    const int j = (int)(idx - (int64_t)b * (L + 1));
    int64_t col;
    if (j == L) {
        col = Kx;
    } else {
        const int a = states[(int64_t)b * L + j];
        col = state_col[(int64_t)j * q + a];
    }
    atomicOr(&X[(int64_t)b * (Kxp >> 5) + (col >> 5)], 1u << (col & 31));
    atomicOr(&Xt[col * (Bp >> 5) + (b >> 5)], 1u << (b & 31));
}

// ------------------------------------------------------------ reweighting

// planes[b][v][w] bit t set iff states[b][64*w + t] == v.
__global__ void k_bitplanes(const uint8_t *__restrict__ states, int B, int L, int q, int W, uint64_t *__restrict__ planes)
{
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)B * W) return;
    const int b = (int)(idx / W);
    const int w = (int)(idx - (int64_t)b * W);
    uint64_t word[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int l0 = w * 64;
    const int l1 = min(L, l0 + 64);
    for (int l = l0; l < l1; ++l) {
        const int v = states[(int64_t)b * L + l];
        for (int u = 0; u < 8; ++u)
            if (u == v) word[u] |= 1ull << (l - l0);
    }
    for (int v = 0; v < q; ++v) planes[((int64_t)b * q + v) * W + w] = word[v];
}

// Lower-triangular 64x64 tiles of sequence pairs; each thread owns 4x4 pairs.
// matches(i,j) = sum_v popcount(P_i^v & P_j^v); d = L - matches; the pair is a
// neighbour iff d <= dmax, where dmax is the largest d with
// double(L - d) / double(L) >= theta (computed on the host with that expression),
// so the comparison is exactly the one of plmDCA_optimizer_parameters.hpp:67-72.
constexpr int HT = 64;   // sequences per tile side
constexpr int HWC = 4;   // words per LDS chunk
__global__ __launch_bounds__(256) void k_hamming(const uint64_t *__restrict__ planes, int B, int q, int W, int L,
                                                 int64_t dmax, const uint32_t *__restrict__ mult,
                                                 unsigned long long *__restrict__ iw)
{
    __shared__ uint64_t sI[HT][8][HWC];
    __shared__ uint64_t sJ[HT][8][HWC];
    // blockIdx.x -> (ti, tj), ti >= tj
    const int64_t t = blockIdx.x;
    int ti = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((int64_t)(ti + 1) * (ti + 2) / 2 <= t) ++ti;
    while ((int64_t)ti * (ti + 1) / 2 > t) --ti;
    const int tj = (int)(t - (int64_t)ti * (ti + 1) / 2);
    const int i0 = ti * HT, j0 = tj * HT;
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    int matches[4][4];
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) matches[a][c] = 0;
    for (int w0 = 0; w0 < W; w0 += HWC) {
        __syncthreads();
        for (int e = tid; e < HT * q * HWC; e += 256) {
            const int s = e / (q * HWC);
            const int rem = e - s * q * HWC;
            const int v = rem / HWC;
            const int ww = rem - v * HWC;
            const int w = w0 + ww;
            const int gi = i0 + s, gj = j0 + s;
            sI[s][v][ww] = (gi < B && w < W) ? planes[((int64_t)gi * q + v) * W + w] : 0ull;
            sJ[s][v][ww] = (gj < B && w < W) ? planes[((int64_t)gj * q + v) * W + w] : 0ull;
        }
        __syncthreads();
        for (int v = 0; v < q; ++v)
            for (int ww = 0; ww < HWC; ++ww) {
                uint64_t pi[4], pj[4];
                for (int a = 0; a < 4; ++a) pi[a] = sI[ty * 4 + a][v][ww];
                for (int c = 0; c < 4; ++c) pj[c] = sJ[tx * 4 + c][v][ww];
                for (int a = 0; a < 4; ++a)
                    for (int c = 0; c < 4; ++c) matches[a][c] += __popcll(pi[a] & pj[c]);
            }
    }
    unsigned long long addI[4] = {0, 0, 0, 0}, addJ[4] = {0, 0, 0, 0};
    for (int a = 0; a < 4; ++a)
        for (int c = 0; c < 4; ++c) {
            const int gi = i0 + ty * 4 + a, gj = j0 + tx * 4 + c;
            if (gi < B && gj < B && gi > gj) {
                const int64_t d = (int64_t)L - matches[a][c];
                if (d <= dmax) {
                    addI[a] += mult ? mult[gj] : 1u;
                    addJ[c] += mult ? mult[gi] : 1u;
                }
            }
        }
    for (int a = 0; a < 4; ++a) {
        const int gi = i0 + ty * 4 + a;
        if (addI[a] && gi < B) atomicAdd(&iw[gi], addI[a]);
    }
    for (int c = 0; c < 4; ++c) {
        const int gj = j0 + tx * 4 + c;
        if (addJ[c] && gj < B) atomicAdd(&iw[gj], addJ[c]);
    }
}

// ------------------------------------------------------------------- GEMM

// C[m][n] = sum_k A[m][k] * B[n][k] on 256-thread workgroups of four waves (2 x 2).
// A tile is 128 rows of A by TN = 64 * NI rows of B; each wave computes 64 x 32*NI
// outputs as 2 x NI MFMA 32x32 blocks, with K staged through LDS in rows of 128 bytes.
// NA = 2 sums two A variants (hi and lo BF16 parts) against the same B fragment.  The
// wide tile (NI = 4) reads 0.5 LDS fragments per MFMA instead of 0.75 at NA = 2.
// B is a 0/1 matrix held as bits (one-hot X or its transpose); each thread loads the
// bits of its share of a K slice and expands them to 1.0 / 0.0 while staging to LDS,
// so B costs 1/16 (BF16) or 1/32 (FP32) of its element traffic.
constexpr int ROWB = 144;  // LDS bytes per staged row (128 data + 16 pad)

struct PolBF16 {
    using T = uint16_t;
    static constexpr int BK = 64;   // elements per staged row
    static constexpr int KPM = 16;  // K per MFMA
};
struct PolF32 {
    using T = float;
    static constexpr int BK = 32;
    static constexpr int KPM = 2;
};

template <class Pol, int NA, int NI>
__device__ __forceinline__ void mma_tile(const uint8_t *sA0, const uint8_t *sA1, const uint8_t *sB, int wm, int wn,
                                         int lane, floatx16 (&acc)[2][NI])
{
    if constexpr (sizeof(typename Pol::T) == 2) {
#pragma unroll
        for (int kk = 0; kk < Pol::BK; kk += Pol::KPM) {
            const int kb = (kk + (lane >> 5) * 8) * 2;
            bf16x8 a[NA][2], b[NI];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const int row = wm + mi * 32 + (lane & 31);
                a[0][mi] = *reinterpret_cast<const bf16x8 *>(sA0 + row * ROWB + kb);
                if constexpr (NA == 2) a[NA - 1][mi] = *reinterpret_cast<const bf16x8 *>(sA1 + row * ROWB + kb);
            }
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) {
                const int row = wn + ni * 32 + (lane & 31);
                b[ni] = *reinterpret_cast<const bf16x8 *>(sB + row * ROWB + kb);
            }
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                    for (int v = 0; v < NA; ++v)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[v][mi], b[ni], acc[mi][ni], 0, 0, 0);
        }
    } else {
#pragma unroll
        for (int kk = 0; kk < Pol::BK; kk += Pol::KPM) {
            const int kb = (kk + (lane >> 5)) * 4;
            float a[NA][2], b[NI];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const int row = wm + mi * 32 + (lane & 31);
                a[0][mi] = *reinterpret_cast<const float *>(sA0 + row * ROWB + kb);
                if constexpr (NA == 2) a[NA - 1][mi] = *reinterpret_cast<const float *>(sA1 + row * ROWB + kb);
            }
#pragma unroll
            for (int ni = 0; ni < NI; ++ni) {
                const int row = wn + ni * 32 + (lane & 31);
                b[ni] = *reinterpret_cast<const float *>(sB + row * ROWB + kb);
            }
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < NI; ++ni)
#pragma unroll
                    for (int v = 0; v < NA; ++v)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[v][mi], b[ni], acc[mi][ni], 0, 0, 0);
        }
    }
}

// Tile (mt, nt) and split-K slice z of workgroup bid of a tile map (TileMap, sdca_internal.hpp).
__device__ __forceinline__ bool map_tile(const TileMap &tm, int bid, int &mt, int &nt, int &z)
{
    const int xcd = bid & 7, local = bid >> 3;
    const int per = tm.GM * tm.GN;
    const int sidx = (local / per) * 8 + xcd;
    if (sidx >= tm.S) return false;
    const int t = local - (local / per) * per;
    const int nb = sidx % tm.NB;
    z = (sidx / tm.NB) % tm.Z;
    const int mb = sidx / (tm.NB * tm.Z);
    mt = mb * tm.GM + t % tm.GM;
    nt = nb * tm.GN + t / tm.GM;
    return mt < tm.Mt && nt < tm.Nt;
}

// Two workgroups per CU (eight waves): the register budget of a wave is 256 (VGPRs +
// AGPRs), which holds the wide tile's accumulators and one prefetch stage.
// Adds the 128 x 64 NI output tile (mt, nt) of split-K slice z of A B^T, A = A0 (+ A1),
// to acc: the K loop of k_gemm_nt for NA operand parts.
template <class Pol, int NA, int NI>
__device__ __forceinline__ void gemm_accum(uint8_t *sA0, uint8_t *sA1, uint8_t *sB,
                                           const typename Pol::T *__restrict__ A0,
                                           const typename Pol::T *__restrict__ A1, int64_t lda,
                                           const uint32_t *__restrict__ Bbits, int64_t ldbw, int K, int kchunk,
                                           int mt, int nt, int z, floatx16 (&acc)[2][NI])
{
    constexpr int TN = 64 * NI;  // B rows (output columns) per tile
    uint8_t *sA[2] = {sA0, sA1};
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t m0 = (int64_t)mt * kTile, n0 = (int64_t)nt * TN;
    const int kb = z * kchunk;
    const int ke = min(K, kb + kchunk);
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * (TN / 2);
    const int lrow = tid >> 3;         // 0..31
    const int lcol = (tid & 7) * 16;   // byte offset in the 128-byte row
    const uint8_t *pa0 = reinterpret_cast<const uint8_t *>(A0 + (m0 + lrow) * lda) + lcol;
    const uint8_t *pa1 = reinterpret_cast<const uint8_t *>(A1 + (m0 + lrow) * lda) + lcol;
    const int64_t sa = 32 * lda * (int64_t)sizeof(typename Pol::T);  // 32 rows, in bytes
    constexpr int EB = (int)sizeof(typename Pol::T);
    constexpr int BK = Pol::BK;
    // B bits: EPT elements of one row per thread and K slice (TPR threads per row)
    constexpr int EPT = TN * BK / 256, TPR = BK / EPT, NW = EPT >= 32 ? EPT / 32 : 1;
    static_assert(EPT == 16 || EPT == 32 || EPT == 64, "B bits per thread");
    const int brow = tid / TPR, bseg = tid - brow * TPR;
    const uint32_t *pbw = Bbits + (n0 + brow) * ldbw;
    uint8_t *sBrow = sB + brow * ROWB + bseg * EPT * EB;
    // software pipeline: the next K slice is in flight from memory while the current one
    // is multiplied out of LDS
    u32x4 ra[NA][4];
    uint32_t rbw[NW];
    auto load = [&](int k0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            ra[0][i] = *reinterpret_cast<const u32x4 *>(pa0 + i * sa + (int64_t)k0 * EB);
            if constexpr (NA == 2) ra[NA - 1][i] = *reinterpret_cast<const u32x4 *>(pa1 + i * sa + (int64_t)k0 * EB);
        }
        const int bit0 = k0 + bseg * EPT;
        if constexpr (NW == 2) {
            const uint2 w = *reinterpret_cast<const uint2 *>(pbw + (bit0 >> 5));
            rbw[0] = w.x;
            rbw[NW - 1] = w.y;
        } else {
            rbw[0] = pbw[bit0 >> 5] >> (bit0 & 31);
        }
    };
    if (kb < ke) load(kb);
    for (int k0 = kb; k0 < ke; k0 += BK) {
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int row = lrow + 32 * i;
#pragma unroll
            for (int v = 0; v < NA; ++v) *reinterpret_cast<u32x4 *>(sA[v] + row * ROWB + lcol) = ra[v][i];
        }
        // expand the B bits: 16-byte chunks of 8 BF16 or 4 FP32 ones and zeros
        {
            constexpr int EPC = 16 / EB;  // elements per chunk
#pragma unroll
            for (int ch = 0; ch < EPT / EPC; ++ch) {
                const int b0 = ch * EPC;
                const uint32_t w = rbw[b0 >> 5] >> (b0 & 31);
                u32x4 v;
                if constexpr (EB == 2) {
                    v.x = (w & 1u) * 0x3f80u + ((w >> 1) & 1u) * 0x3f800000u;
                    v.y = ((w >> 2) & 1u) * 0x3f80u + ((w >> 3) & 1u) * 0x3f800000u;
                    v.z = ((w >> 4) & 1u) * 0x3f80u + ((w >> 5) & 1u) * 0x3f800000u;
                    v.w = ((w >> 6) & 1u) * 0x3f80u + ((w >> 7) & 1u) * 0x3f800000u;
                } else {
                    v.x = (w & 1u) * 0x3f800000u;
                    v.y = ((w >> 1) & 1u) * 0x3f800000u;
                    v.z = ((w >> 2) & 1u) * 0x3f800000u;
                    v.w = ((w >> 3) & 1u) * 0x3f800000u;
                }
                *reinterpret_cast<u32x4 *>(sBrow + ch * 16) = v;
            }
        }
        __syncthreads();
        if (k0 + BK < ke) load(k0 + BK);
        // the MFMA block runs at raised wave priority, so that while one workgroup stages
        // its next slice the other's matrix work keeps issuing
        __builtin_amdgcn_s_setprio(1);
        mma_tile<Pol, NA, NI>(sA[0], sA[NA - 1], sB, wm, wn, lane, acc);
        __builtin_amdgcn_s_setprio(0);
    }
}

// A round's GEMM launch covers the groups of its commands (one per alignment, GemmGroup);
// workgroup bid belongs to the last group with tile0 <= bid.  Within a group the packed
// rows come in three parts by precision level (each starting on a tile boundary): tiles
// below parts.row[1] belong to fine commands and take three MFMA passes under split
// precision (hi and mid together, then lo), tiles below parts.row[2] two (hi and lo), the
// rest one pass over the BF16 high parts.  The dearer tiles have the lower row indices and
// are dispatched first, so the cheaper ones fill the end of the launch.
template <class Pol, int NA, int NI, class Epi>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_gemm_nt(const typename Pol::T *__restrict__ A0,
                                                    const typename Pol::T *__restrict__ A1,
                                                    const typename Pol::T *__restrict__ A2, int64_t lda,
                                                    const GemmGroup *__restrict__ groups, int ngroups, Epi epi)
{
    constexpr int TN = 64 * NI;
    __shared__ __attribute__((aligned(16))) uint8_t sA[NA][kTile * ROWB];
    __shared__ __attribute__((aligned(16))) uint8_t sB[TN * ROWB];
    const int bid = blockIdx.x;
    int g = 0;
    while (g + 1 < ngroups && groups[g + 1].tile0 <= bid) ++g;
    const GemmGroup &G = groups[g];
    int mt, nt, z;
    if (!map_tile(G.tm, bid - G.tile0, mt, nt, z)) return;
    mt += G.parts.row[0] / kTile;  // absolute packed row tile
    floatx16 acc[2][NI];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < NI; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0.f;
    const int m0 = mt * kTile;
    const uint32_t *Bbits = G.Bbits;
    const int64_t ldbw = G.ldbw;
    const int K = G.K, kchunk = G.kchunk;
    if constexpr (NA == 2) {
        if (m0 >= G.parts.row[2]) {
            gemm_accum<Pol, 1, NI>(sA[0], sA[0], sB, A0, A0, lda, Bbits, ldbw, K, kchunk, mt, nt, z, acc);
        } else {
            gemm_accum<Pol, 2, NI>(sA[0], sA[1], sB, A0, A1, lda, Bbits, ldbw, K, kchunk, mt, nt, z, acc);
            if (m0 < G.parts.row[1])
                gemm_accum<Pol, 1, NI>(sA[0], sA[0], sB, A2, A2, lda, Bbits, ldbw, K, kchunk, mt, nt, z, acc);
        }
    } else {
        gemm_accum<Pol, 1, NI>(sA[0], sA[0], sB, A0, A0, lda, Bbits, ldbw, K, kchunk, mt, nt, z, acc);
    }
    const int tid = threadIdx.x, wave = tid >> 6;
    epi.template tile<NI>(G, m0, (wave >> 1) * 64, (int64_t)nt * TN + (wave & 1) * (TN / 2), tid & 63, tid, acc, z);
}

// 32x32 MFMA result layout: register v of lane l holds row 8*(v/4) + 4*(l/32) + v%4,
// column l%32 of the block.
__device__ __forceinline__ int acc_row(int mi, int v, int lane) { return mi * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

// Epilogues receive a wave's 64 x 32*NI block of outputs: rows m0 + wm + acc_row(),
// columns nw + 32 ni + lane % 32.

// Energy epilogue: split-K partial sums, Epart[z][m][n].
struct EpiEnergy {
    float *E;
    int64_t ld, zstride;
    template <int NI>
    __device__ __forceinline__ void tile(const GemmGroup &, int m0, int wm, int64_t nw, int lane, int,
                                         floatx16 (&acc)[2][NI], int z) const
    {
        float *base = E + (int64_t)z * zstride + nw + (lane & 31);
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                float *row = base + (int64_t)(m0 + wm + acc_row(mi, v, lane)) * ld;
#pragma unroll
                for (int ni = 0; ni < NI; ++ni) row[ni * 32] = acc[mi][ni][v];
            }
    }
};

// Gradient epilogue: packed row m = row_of(k) + s of command k, column n.  Adds 2*lambda*x
// (plmDCA_objective.hpp:192-244), zeroes the locus' own columns (the j != r exclusions of
// :95,111,161,176) and pad columns, and writes the trial gradient into gradient buffer gt
// of the command's slot.  The field column K_x takes its data term from the FP64 sums of
// the softmax (sum_b G_b(s), hpart) instead of the FP32 contraction, and adds
// 2 lambda_h h in FP64: that sum runs over every sequence, and its FP32 partial sums grow
// to about B_eff while the result is small.  The commands under the tile's 128 rows are
// staged in LDS once; each group of four rows first loads its trial-point values (packed
// rows, like the beliefs), then stores.
// Without Xtrial (null) the epilogue stores the contraction itself: the couplings
// J = X_r^T c of coefficients c (the Gram formulation's solutions), with column K_x from hpart.
struct EpiGrad {
    float *gring;
    const float *Xtrial;
    const SlotCmd *cmds;
    const double *hpart;
    const AlnDev *aln;
    int64_t n_slot, ldx;
    int ng, g0, nbchunk;  // gradients per slot, ring index of the first; softmax blocks per command
    // command owning packed row m of group G (G's end command for a pad row) and its first row
    __device__ __forceinline__ static int cmd_of(const GemmGroup &G, int m)
    {
        const Parts &P = G.parts;
        const int p = m >= P.row[2] ? 2 : (m >= P.row[1] ? 1 : 0);
        const int i = (m - P.row[p]) / G.q;
        return i < P.cmd[p + 1] - P.cmd[p] ? P.cmd[p] + i : P.cmd[3];
    }
    __device__ __forceinline__ static int row_of(const GemmGroup &G, int i)
    {
        const Parts &P = G.parts;
        const int p = i >= P.cmd[2] ? 2 : (i >= P.cmd[1] ? 1 : 0);
        return P.row[p] + (i - P.cmd[p]) * G.q;
    }
    template <int NI>
    __device__ __forceinline__ void tile(const GemmGroup &G, int m0, int wm, int64_t nw, int lane, int tid,
                                         floatx16 (&acc)[2][NI], int) const
    {
        // s_gt < 0: no gradient to write
        __shared__ int s_gt[kTile + 2], s_lo[kTile + 2], s_hi[kTile + 2], s_slot[kTile + 2];
        const int cend = G.parts.cmd[3];
        const int i0 = cmd_of(G, m0);
        const int ns = min(cend - i0, (kTile - 1) / G.q + 2);  // commands under rows [m0, m0 + 128)
        if (tid < ns) {
            const SlotCmd &c = cmds[i0 + tid];
            s_gt[tid] = (c.mode == MODE_IDLE || c.gt < g0 || c.gt >= g0 + ng) ? -1 : c.gt;
            s_lo[tid] = c.self_lo;
            s_hi[tid] = c.self_hi;
            s_slot[tid] = c.slot;
        }
        __syncthreads();
        const AlnDev &A = aln[G.aln];
        const int64_t Kx = A.Kx, Kxp = A.Kxp;
        const float lamJ2 = A.lamJ2;
        const double lamh2 = A.lamh2;
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int vq = 0; vq < 16; vq += 4) {  // four rows at a time
                int gt[4], lo[4], hi[4], cm[4], srow[4];
                int64_t roff[4];
                float xv[4][NI];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int m = m0 + wm + acc_row(mi, vq + u, lane);
                    const int i = cmd_of(G, m);
                    const int il = i - i0;
                    const bool ok = il < ns;
                    gt[u] = ok ? s_gt[il] : -1;
                    lo[u] = ok ? s_lo[il] : 0;
                    hi[u] = ok ? s_hi[il] : 0;
                    cm[u] = i;
                    srow[u] = ok ? m - row_of(G, i) : 0;
                    const int slot = ok ? s_slot[il] : 0;
                    roff[u] = ((int64_t)slot * ng + (gt[u] < 0 ? 0 : gt[u] - g0)) * n_slot + (int64_t)srow[u] * Kxp;
                    const float *xr = Xtrial ? Xtrial + (int64_t)m * ldx : nullptr;
#pragma unroll
                    for (int ni = 0; ni < NI; ++ni) {
                        const int64_t n = nw + ni * 32 + (lane & 31);
                        const bool use = Xtrial && gt[u] >= 0 && n <= Kx && !(n >= lo[u] && n < hi[u]);
                        xv[u][ni] = use ? __builtin_nontemporal_load(xr + n) : 0.f;
                    }
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (gt[u] < 0) continue;
#pragma unroll
                    for (int ni = 0; ni < NI; ++ni) {
                        const int64_t n = nw + ni * 32 + (lane & 31);
                        if (n >= Kxp) continue;
                        float val = 0.f;
                        if (n == Kx) {
                            const double *hp = hpart + (int64_t)cm[u] * nbchunk * 8 + srow[u];
                            double hs = 0.0;
                            for (int ch = 0; ch < nbchunk; ++ch) hs += hp[(int64_t)ch * 8];
                            val = (float)(hs + lamh2 * (double)xv[u][ni]);
                        } else if (n < Kx && !(n >= lo[u] && n < hi[u])) {
                            val = acc[mi][ni][vq + u] + lamJ2 * xv[u][ni];
                        }
                        __builtin_nontemporal_store(val, gring + roff[u] + n);
                    }
                }
            }
    }
};

// ---------------------------------------------------------------- softmax

// One thread per (command, sequence b): sums the split-K partial energies, forms the
// log-partition lz = log sum_s exp E_b(s) (max-shifted; the same value as the
// unshifted plmDCA_objective.hpp:140), the data term w_b (lz - E_b(x_br)) (:143) and the
// node beliefs G_b(s) = w_b P_b(s) - w_b [s == x_br] (:147-148), with 1 - P(x_br)
// formed as the sum of the other probabilities.  Coarse and split commands take the
// hardware exp2/log2 approximations; fine ones the accurate expf/logf and P = e_s / z.
// The data term and, per state, sum_b G_b(s) (the fields' data gradient, :151) are
// reduced in FP64 per block of 256 sequences.
__global__ __launch_bounds__(256) void k_softmax(const float *__restrict__ Epart, int64_t zstride, int64_t ldb,
                                                 const SlotCmd *__restrict__ cmds, const AlnDev *__restrict__ aln,
                                                 float *__restrict__ G32, uint16_t *__restrict__ Ghi,
                                                 uint16_t *__restrict__ Glo, uint16_t *__restrict__ Gl3,
                                                 double *__restrict__ fpart, double *__restrict__ hpart, int nbchunk)
{
    const int i = blockIdx.y;
    const int b = blockIdx.x * 256 + threadIdx.x;
    const SlotCmd &c = cmds[i];
    const int r = c.mode == MODE_IDLE ? -1 : c.locus;
    const int64_t row0 = c.row;
    const int level = c.level, q = c.q, ksplit = c.ksplit;
    const AlnDev &A = aln[c.aln];
    const int B = A.B, Bp = A.Bp;
    const uint8_t *statesT = A.statesT;
    const float *wts = A.w;
    double fb = 0.0;
    double hs[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (b < Bp) {
        float e[8], g[8];
        if (r >= 0) {
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                if (s >= q) break;
                float acc = 0.f;
                for (int z = 0; z < ksplit; ++z) acc += Epart[(int64_t)z * zstride + (row0 + s) * ldb + b];
                e[s] = acc;
            }
            const float w = wts[b];
            const int xr = b < B ? statesT[(int64_t)r * B + b] : 0;
            float mx = e[0], exr = e[0];
#pragma unroll
            for (int s = 1; s < 8; ++s) {
                if (s >= q) break;
                mx = fmaxf(mx, e[s]);
                if (s == xr) exr = e[s];
            }
            float p[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            float lz;
            if (level == LV_FINE) {
                float z = 0.f;
#pragma unroll
                for (int s = 0; s < 8; ++s) {
                    if (s >= q) break;
                    p[s] = expf(e[s] - mx);
                    z += p[s];
                }
                lz = mx + logf(z);
                const float iz = 1.f / z;
#pragma unroll
                for (int s = 0; s < 8; ++s) p[s] *= iz;
            } else {
                float z = 0.f;
#pragma unroll
                for (int s = 0; s < 8; ++s)
                    if (s < q) z += __expf(e[s] - mx);
                lz = mx + __logf(z);
#pragma unroll
                for (int s = 0; s < 8; ++s) p[s] = s < q ? __expf(e[s] - lz) : 0.f;
            }
            fb = (double)w * ((double)lz - (double)exr);
            float other = 0.f;
#pragma unroll
            for (int s = 0; s < 8; ++s)
                if (s < q && s != xr) other += p[s];
#pragma unroll
            for (int s = 0; s < 8; ++s) {
                g[s] = s == xr ? -w * other : w * p[s];
                hs[s] = (double)g[s];
            }
        } else {
#pragma unroll
            for (int s = 0; s < 8; ++s) g[s] = 0.f;
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            if (s >= q) break;
            const int64_t o = (row0 + s) * ldb + b;
            if (G32) {
                G32[o] = g[s];
            } else {
                const uint16_t h = f2bf(g[s]);
                Ghi[o] = h;
                if (level != LV_COARSE) {
                    const float rest = g[s] - bf2f(h);
                    const uint16_t l = f2bf(rest);
                    Glo[o] = l;
                    if (level == LV_FINE) Gl3[o] = f2bf(rest - bf2f(l));
                }
            }
        }
    }
    // block reductions of fb and of the field sums
    __shared__ double red[4][9];
    for (int off = 32; off > 0; off >>= 1) {
        fb += __shfl_down(fb, off, 64);
#pragma unroll
        for (int s = 0; s < 8; ++s)
            if (s < q) hs[s] += __shfl_down(hs[s], off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[threadIdx.x >> 6][8] = fb;
#pragma unroll
        for (int s = 0; s < 8; ++s) red[threadIdx.x >> 6][s] = hs[s];
    }
    __syncthreads();
    if (threadIdx.x == 0) fpart[(int64_t)i * nbchunk + blockIdx.x] = red[0][8] + red[1][8] + red[2][8] + red[3][8];
    if (threadIdx.x < 8) {
        const int s = threadIdx.x;
        hpart[((int64_t)i * nbchunk + blockIdx.x) * 8 + s] = red[0][s] + red[1][s] + red[2][s] + red[3][s];
    }
}

// ------------------------------------------------------ L-BFGS vector passes

template <int NV>
__device__ __forceinline__ void block_reduce_store(double (&v)[NV], double *out, const bool (&is_max)[NV])
{
    __shared__ double red[4][NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double x = v[k];
        for (int off = 32; off > 0; off >>= 1) {
            const double y = __shfl_down(x, off, 64);
            x = is_max[k] ? fmax(x, y) : x + y;
        }
        v[k] = x;
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[wave][k] = v[k];
    __syncthreads();
    for (int k = threadIdx.x; k < NV; k += blockDim.x)
        out[k] = is_max[k] ? fmax(fmax(red[0][k], red[1][k]), fmax(red[2][k], red[3][k]))
                           : (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
}

__device__ __forceinline__ double dot4(float4 a, float4 b)
{
    return (double)a.x * b.x + (double)a.y * b.y + (double)a.z * b.z + (double)a.w * b.w;
}

// The vector passes move 8 consecutive elements per thread and iteration: two float4 of
// an FP32 vector, or one 16-byte load of a BF16 curvature pair.  Pair loads widen to
// float; stores round and hand back the rounded value (every product uses the stored
// vector).  Every slot vector is streamed once per pass and is far larger than the
// caches, so loads and stores are non-temporal.
struct f8 {
    float4 a, b;
};
__device__ __forceinline__ f8 zero8() { return {make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)}; }
typedef float fx4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ f8 ld8(const float *v, int64_t e)
{
    const fx4 a = __builtin_nontemporal_load(reinterpret_cast<const fx4 *>(v + e));
    const fx4 b = __builtin_nontemporal_load(reinterpret_cast<const fx4 *>(v + e + 4));
    return {make_float4(a.x, a.y, a.z, a.w), make_float4(b.x, b.y, b.z, b.w)};
}
__device__ __forceinline__ f8 unpack8(const u32x4 &u)
{
    auto lo = [](uint32_t w) { return __uint_as_float(w << 16); };
    auto hi = [](uint32_t w) { return __uint_as_float(w & 0xffff0000u); };
    return {make_float4(lo(u.x), hi(u.x), lo(u.y), hi(u.y)), make_float4(lo(u.z), hi(u.z), lo(u.w), hi(u.w))};
}
__device__ __forceinline__ f8 ld8(const uint16_t *v, int64_t e)
{
    return unpack8(__builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(v + e)));
}
__device__ __forceinline__ void st8(float *v, int64_t e, const f8 &x)
{
    const fx4 a = {x.a.x, x.a.y, x.a.z, x.a.w}, b = {x.b.x, x.b.y, x.b.z, x.b.w};
    __builtin_nontemporal_store(a, reinterpret_cast<fx4 *>(v + e));
    __builtin_nontemporal_store(b, reinterpret_cast<fx4 *>(v + e + 4));
}
__device__ __forceinline__ void st8(float *v, int64_t e, f8 &x, int) { st8(v, e, x); }
__device__ __forceinline__ void st8(uint16_t *v, int64_t e, f8 &x, int)
{
    const float f[8] = {x.a.x, x.a.y, x.a.z, x.a.w, x.b.x, x.b.y, x.b.z, x.b.w};
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) w[k] = (uint32_t)f2bf(f[2 * k]) | ((uint32_t)f2bf(f[2 * k + 1]) << 16);
    u32x4 u;
    u.x = w[0];
    u.y = w[1];
    u.z = w[2];
    u.w = w[3];
    __builtin_nontemporal_store(u, reinterpret_cast<u32x4 *>(v + e));
    x = unpack8(u);
}
__device__ __forceinline__ f8 lin(float a, const f8 &x, float b, const f8 &y)
{
    return {make_float4(a * x.a.x + b * y.a.x, a * x.a.y + b * y.a.y, a * x.a.z + b * y.a.z, a * x.a.w + b * y.a.w),
            make_float4(a * x.b.x + b * y.b.x, a * x.b.y + b * y.b.y, a * x.b.z + b * y.b.z, a * x.b.w + b * y.b.w)};
}
__device__ __forceinline__ void axpy8(f8 &d, float a, const f8 &v)
{
    d.a.x += a * v.a.x;
    d.a.y += a * v.a.y;
    d.a.z += a * v.a.z;
    d.a.w += a * v.a.w;
    d.b.x += a * v.b.x;
    d.b.y += a * v.b.y;
    d.b.z += a * v.b.z;
    d.b.w += a * v.b.w;
}
// x + a d with one rounding per element: an accepted trial point is formed again from
// x_c and d by the same operation, which reproduces it exactly.
__device__ __forceinline__ f8 step8(const f8 &x, float a, const f8 &d)
{
    return {make_float4(__builtin_fmaf(a, d.a.x, x.a.x), __builtin_fmaf(a, d.a.y, x.a.y),
                        __builtin_fmaf(a, d.a.z, x.a.z), __builtin_fmaf(a, d.a.w, x.a.w)),
            make_float4(__builtin_fmaf(a, d.b.x, x.b.x), __builtin_fmaf(a, d.b.y, x.b.y),
                        __builtin_fmaf(a, d.b.z, x.b.z), __builtin_fmaf(a, d.b.w, x.b.w))};
}
__device__ __forceinline__ double dot8(const f8 &x, const f8 &y) { return dot4(x.a, y.a) + dot4(x.b, y.b); }
__device__ __forceinline__ float amax8(const f8 &x)
{
    return fmaxf(fmaxf(fmaxf(fabsf(x.a.x), fabsf(x.a.y)), fmaxf(fabsf(x.a.z), fabsf(x.a.w))),
                 fmaxf(fmaxf(fabsf(x.b.x), fabsf(x.b.y)), fmaxf(fabsf(x.b.z), fabsf(x.b.w))));
}

// Trial point of every command's slot (one launch per round):
//   INIT    x_c = x0 (staged) or 0,           x_t = x_c
//   RETRY   x_t = x_c + alpha d
//   NEWDIR  x_c = x_c + pstep d_old (if accept: the last trial point); the new pair
//           s = pstep d, y = g[gk] - g[1-gk] into pair slot pnew (if any);
//           d <- cd d + cg g_k + sum_p (cs[p] s_p + cy[p] y_p);
//           x_t = x_c + alpha d; and the product d . g_k of the stored d.  Pairs with a
//           zero coefficient are not read.  (The products d . y_p follow on the host from
//           the coefficients and the pair products.)
// x_t (and its BF16 parts: hi for a coarse command, hi and lo for a split one, hi, mid and
// lo for a fine one) is written at the
// command's packed rows.  Plus the squared norms of x_t for the L2 terms:
// part[command][chunk][3] = {sum x^2 over the whole vector, sum over the h entries
// (column Kx of each row), d . g_k}.
template <int M, class PT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(M > 10 ? 1 : 4))) void k_update(const SlotCmd *__restrict__ cmds, const float *__restrict__ gbuf,
                                                PT *__restrict__ dbuf, PT *__restrict__ S, PT *__restrict__ Y,
                                                float *__restrict__ Xc, float *__restrict__ Xtr,
                                                uint16_t *__restrict__ Whi, uint16_t *__restrict__ Wlo,
                                                uint16_t *__restrict__ Wl3,
                                                const float *__restrict__ X0, int64_t n_slot, int64_t ldx,
                                                double *__restrict__ part, int nchunk)
{
    constexpr int NV = 3;
    const int j = blockIdx.y;
    const SlotCmd &c = cmds[j];
    double acc[NV];
    bool is_max[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        acc[k] = 0.0;
        is_max[k] = false;
    }
    const int mode = c.mode;
    const bool ok = mode == MODE_INIT || mode == MODE_RETRY ||
                    (mode == MODE_NEWDIR && (c.gk == 0 || c.gk == 1) && c.pnew < M);
    if (mode != MODE_IDLE && ok) {
        const int64_t i = c.slot, xo = (int64_t)c.row * ldx;
        const uint32_t Kxp = (uint32_t)c.Kxp, Kx = (uint32_t)c.Kx;
        float *xc = Xc + i * n_slot;
        float *xt = Xtr + xo;
        PT *dv = dbuf + (int64_t)i * n_slot;
        const float *x0 = X0 ? X0 + (int64_t)i * n_slot : nullptr;
        const float *gk = gbuf + ((int64_t)i * 2 + (c.gk == 1)) * n_slot;
        const float *gold = gbuf + ((int64_t)i * 2 + (c.gk != 1)) * n_slot;
        PT *Sb = S + (int64_t)i * M * n_slot;
        PT *Yb = Y + (int64_t)i * M * n_slot;
        const float alpha = c.alpha, pstep = c.pstep, cd = c.cd, cg = c.cg;
        const int pnew = c.pnew;
        const int64_t n8 = ((int64_t)c.q * Kxp) >> 3;  // the command's q x Kxp elements
        const bool accept = c.accept != 0;
        const int level = c.level;  // BF16 parts the GEMMs read: 1 (coarse), 2 (split) or 3 (fine)
        const bool from_x0 = c.init_src != 0.f && x0;
        float cs[M], cy[M];
#pragma unroll
        for (int p = 0; p < M; ++p) {
            cs[p] = c.cs[p];
            cy[p] = c.cy[p];
        }
        for (int64_t e8 = (int64_t)blockIdx.x * 256 + threadIdx.x; e8 < n8; e8 += (int64_t)nchunk * 256) {
            const int64_t e = e8 * 8;
            // row s and column of the element in the command's q x Kxp vector, and its place
            // in the packed trial-point rows (K_x is padded to a multiple of 8, so an
            // 8-element chunk lies in one row; a slot vector has fewer than 2^32 elements)
            const uint32_t srow = (uint32_t)e / Kxp, col = (uint32_t)e - srow * Kxp;
            const int64_t pe = (int64_t)srow * ldx + col;
            f8 xn;
            if (mode == MODE_INIT) {
                xn = from_x0 ? ld8(x0, e) : zero8();
                st8(xc, e, xn);
            } else if (mode == MODE_RETRY) {
                xn = step8(ld8(xc, e), alpha, ld8(dv, e));
            } else {  // NEWDIR
                const f8 dold = ld8(dv, e);
                f8 xv = ld8(xc, e);
                if (accept) {
                    xv = step8(xv, pstep, dold);
                    st8(xc, e, xv);
                }
                const f8 g = ld8(gk, e);
                f8 sn = zero8(), yn = zero8();
                if (pnew >= 0) {
                    sn = lin(pstep, dold, 0.f, dold);
                    yn = lin(1.f, g, -1.f, ld8(gold, e));
                    st8(Sb + (int64_t)pnew * n_slot, e, sn, 0);
                    st8(Yb + (int64_t)pnew * n_slot, e, yn, 0);
                }
                f8 d = lin(cd, dold, cg, g);
#pragma unroll
                for (int p = 0; p < M; ++p) {
                    const bool mine = p == pnew;
                    if (cs[p] != 0.f) {
                        f8 sv = mine ? zero8() : ld8(Sb + (int64_t)p * n_slot, e);
                        if (mine) sv = sn;
                        axpy8(d, cs[p], sv);
                    }
                    if (cy[p] != 0.f) {
                        f8 yv = mine ? zero8() : ld8(Yb + (int64_t)p * n_slot, e);
                        if (mine) yv = yn;
                        axpy8(d, cy[p], yv);
                    }
                }
                st8(dv, e, d, 0);  // rounded to the stored precision: the products use it
                acc[2] += dot8(d, g);
                xn = step8(xv, alpha, d);
            }
            st8(xt, pe, xn);
            const float xs[8] = {xn.a.x, xn.a.y, xn.a.z, xn.a.w, xn.b.x, xn.b.y, xn.b.z, xn.b.w};
            if (Whi) {
                // x = hi + lo (+ l3): each part is the rounded remainder of the previous ones,
                // and x - hi and (x - hi) - lo are exact in FP32
                uint32_t h[4], l[4], t[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint16_t h0 = f2bf(xs[2 * k]), h1 = f2bf(xs[2 * k + 1]);
                    const float r0 = xs[2 * k] - bf2f(h0), r1 = xs[2 * k + 1] - bf2f(h1);
                    const uint16_t l0 = f2bf(r0), l1 = f2bf(r1);
                    h[k] = (uint32_t)h0 | ((uint32_t)h1 << 16);
                    l[k] = (uint32_t)l0 | ((uint32_t)l1 << 16);
                    t[k] = (uint32_t)f2bf(r0 - bf2f(l0)) | ((uint32_t)f2bf(r1 - bf2f(l1)) << 16);
                }
                u32x4 hv, lv, tv;
                hv.x = h[0], hv.y = h[1], hv.z = h[2], hv.w = h[3];
                lv.x = l[0], lv.y = l[1], lv.z = l[2], lv.w = l[3];
                tv.x = t[0], tv.y = t[1], tv.z = t[2], tv.w = t[3];
                __builtin_nontemporal_store(hv, reinterpret_cast<u32x4 *>(Whi + xo + pe));
                if (level != LV_COARSE) __builtin_nontemporal_store(lv, reinterpret_cast<u32x4 *>(Wlo + xo + pe));
                if (level == LV_FINE) __builtin_nontemporal_store(tv, reinterpret_cast<u32x4 *>(Wl3 + xo + pe));
            }
            acc[0] += dot8(xn, xn);
            // the h entry of row s is column Kx
            if (Kx >= col && Kx < col + 8) {
                const float hv = xs[Kx - col];
                acc[1] += (double)hv * hv;
            }
        }
    }
    block_reduce_store<NV>(acc, part + ((int64_t)j * nchunk + blockIdx.x) * NV, is_max);
}

// Products of each command slot's trial gradient g_t with the stored pairs, itself, the
// current gradient and the direction, and its infinity norms over all entries and over
// the coupling entries (the field column K_x of each row excluded):
// part[command][chunk][2M + 5] = {g_t.s_p, g_t.y_p, g_t.g_t, g_t.g_k, g_t.d, max |g_t|, max |g_t,J|}.
template <int M, class PT>
__global__ __launch_bounds__(256) void k_dots(const SlotCmd *__restrict__ cmds, const float *__restrict__ gbuf,
                                              const PT *__restrict__ dbuf, const PT *__restrict__ S,
                                              const PT *__restrict__ Y, int64_t n_slot, double *__restrict__ part,
                                              int nchunk)
{
    constexpr int NV = 2 * M + 5;
    const int j = blockIdx.y;
    const SlotCmd &c = cmds[j];
    double acc[NV];
    bool is_max[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        acc[k] = 0.0;
        is_max[k] = (k >= 2 * M + 3);
    }
    if (c.mode != MODE_IDLE && (c.gt == 0 || c.gt == 1)) {
        const int64_t i = c.slot;
        const float *g = gbuf + (i * 2 + c.gt) * n_slot;
        const bool cur = c.mode != MODE_INIT;  // a current gradient and a direction exist
        const float *gk = gbuf + (i * 2 + (1 - c.gt)) * n_slot;
        const PT *dv = dbuf + i * n_slot;
        const PT *Sb = S + i * M * n_slot;
        const PT *Yb = Y + i * M * n_slot;
        const uint32_t live = c.live;
        const uint32_t Kxp = (uint32_t)c.Kxp, Kx = (uint32_t)c.Kx;
        const int64_t n8 = ((int64_t)c.q * Kxp) >> 3;  // the command's q x Kxp elements
        for (int64_t e8 = (int64_t)blockIdx.x * 256 + threadIdx.x; e8 < n8; e8 += (int64_t)nchunk * 256) {
            const int64_t e = e8 * 8;
            const f8 gv = ld8(g, e);
            const float ga = amax8(gv);
            acc[2 * M + 3] = fmax(acc[2 * M + 3], (double)ga);
            // the coupling entries: the chunk that holds a row's field entry leaves it out
            const uint32_t col = (uint32_t)e % Kxp;
            float gJ = ga;
            if (Kx >= col && Kx < col + 8) {
                const float gs[8] = {gv.a.x, gv.a.y, gv.a.z, gv.a.w, gv.b.x, gv.b.y, gv.b.z, gv.b.w};
                const uint32_t hpos = Kx - col;
                gJ = 0.f;
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if ((uint32_t)u != hpos) gJ = fmaxf(gJ, fabsf(gs[u]));
            }
            acc[2 * M + 4] = fmax(acc[2 * M + 4], (double)gJ);
            acc[2 * M] += dot8(gv, gv);
            if (cur) {
                acc[2 * M + 1] += dot8(gv, ld8(gk, e));
                acc[2 * M + 2] += dot8(gv, ld8(dv, e));
            }
#pragma unroll
            for (int p = 0; p < M; ++p) {
                if (!((live >> p) & 1u)) continue;
                acc[p] += dot8(gv, ld8(Sb + (int64_t)p * n_slot, e));
                acc[M + p] += dot8(gv, ld8(Yb + (int64_t)p * n_slot, e));
            }
        }
    }
    block_reduce_store<NV>(acc, part + ((int64_t)j * nchunk + blockIdx.x) * NV, is_max);
}

// Per-command reduction of the chunk partials into the result row (ResLayout): the dots
// pass fills [0, 2m+5) (its last two entries are maxima), the trial-point pass
// [2m+5, 2m+8) (two squares and d.g_k; the first square is the whole vector, stored as
// the J part), the softmax the data term.
__global__ void k_reduce(int m, const double *__restrict__ upd_part, const double *__restrict__ dot_part,
                         const double *__restrict__ sm_part, int nchunk, int nbchunk, double *__restrict__ res)
{
    const int i = blockIdx.x;
    const int t = threadIdx.x;
    const int ND = 2 * m + 5, NU = 3, W = 2 * m + 9;
    double *out = res + (int64_t)i * W;
    if (t < ND) {
        const bool mx = t >= ND - 2;
        double a = 0.0;
        for (int ch = 0; ch < nchunk; ++ch) {
            const double v = dot_part[((int64_t)i * nchunk + ch) * ND + t];
            a = mx ? fmax(a, v) : a + v;
        }
        out[t] = a;
    } else if (t < ND + NU) {
        const int u = t - ND;
        double a = 0.0;
        for (int ch = 0; ch < nchunk; ++ch) a += upd_part[((int64_t)i * nchunk + ch) * NU + u];
        out[t] = a;
    } else if (t == W - 1) {
        double a = 0.0;
        for (int ch = 0; ch < nbchunk; ++ch) a += sm_part[(int64_t)i * nbchunk + ch];
        out[t] = a;
    }
    __syncthreads();
    if (t == 0) out[ND] -= out[ND + 1];  // J part = all - h
}

// ---------------------------------------------------------------- scoring

// Score row of a solved locus r (plmDCA.hpp:378-392): for n != r the block
// M[a][s] = J_r[n](a,s) (zero for states a not observed at n; with coords, the solution
// holds the q - 1 zero-sum coordinates of J_r[n](a, .) per row, gram::to_contrast), the gauge
// G[i][k] = ((M[i][k] - rowmean[k]) - colmean[i]) + mean(rowmean) of
// plmDCA_utility.hpp:57-83 (on M^T for n < r), and the Frobenius norm over
// i, k != gap (Matrix_math.hpp:134-178), accumulated in FP64 and stored as float.
__global__ __launch_bounds__(256) void k_score(const float *__restrict__ xsrc, const int64_t *__restrict__ src_off,
                                               const int32_t *__restrict__ src_ld, const int32_t *__restrict__ loci,
                                               const int32_t *__restrict__ alns, int64_t ldrow,
                                               const AlnDev *__restrict__ aln, float *__restrict__ rows, int coords)
{
    const int k = blockIdx.y;
    const int n = blockIdx.x * 256 + threadIdx.x;
    const AlnDev &Ad = aln[alns[k]];
    const int L = Ad.L, q = Ad.q, gap = Ad.gap;
    const int32_t *state_col = Ad.state_col;
    const int64_t Kxp = src_ld[k];  // row length of the solution's q rows
    if (n >= L) return;
    const int r = loci[k];
    float *row = rows + (int64_t)k * ldrow;
    if (n == r) {
        row[n] = 0.f;
        return;
    }
    const float *x = xsrc + src_off[k];
    double A[8][8];
    const bool tr = n < r;
    for (int a = 0; a < q; ++a) {
        const int col = state_col[(int64_t)n * q + a];
        double js[8];  // J_r[n](a, s) over s
        if (coords) {
            double w[8];
#pragma unroll
            for (int kk = 0; kk < 8; ++kk)
                w[kk] = (col >= 0 && kk < gram::rows(q)) ? (double)x[(int64_t)kk * Kxp + col] : 0.0;
            gram::from_contrast(w, q, js);
        } else {
#pragma unroll
            for (int s = 0; s < 8; ++s) js[s] = (col >= 0 && s < q) ? (double)x[(int64_t)s * Kxp + col] : 0.0;
        }
        for (int s = 0; s < q; ++s) {
            if (tr) A[s][a] = js[s];
            else A[a][s] = js[s];
        }
    }
    double m1[8], m2[8];
    double mm = 0.0;
    for (int a = 0; a < q; ++a) {
        double t = 0.0;
        for (int b2 = 0; b2 < q; ++b2) t = t + A[a][b2];
        m1[a] = t / (double)q;
    }
    for (int a = 0; a < q; ++a) mm += m1[a];
    mm = mm / (double)q;
    for (int a = 0; a < q; ++a) {
        double t = 0.0;
        for (int b2 = 0; b2 < q; ++b2) t = t + A[b2][a];
        m2[a] = t / (double)q;
    }
    double sum = 0.0;
    for (int a = 0; a < q; ++a) {
        if (a == gap) continue;
        for (int s = 0; s < q; ++s) {
            if (s == gap) continue;
            const double g = ((A[a][s] - m1[s]) - m2[a]) + mm;
            sum += g * g;
        }
    }
    row[n] = (float)sqrt(sum);
}

}  // namespace

// ============================================================== launchers

namespace kern {

static inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

void build_x(hipStream_t st, const uint8_t *d_states, int B, int L, int q, const int32_t *d_state_col, int Kx,
             int Bp, int Kxp, uint32_t *X, uint32_t *Xt)
{
    const int64_t total = (int64_t)B * (L + 1);
    hipLaunchKernelGGL(k_build_x, dim3(cdiv(total, 256)), dim3(256), 0, st, d_states, B, L, q, d_state_col, Kx, Bp,
                       (int64_t)Kxp, X, Xt);
}

void bitplanes(hipStream_t st, const uint8_t *d_states, int B, int L, int q, int W, uint64_t *planes)
{
    hipLaunchKernelGGL(k_bitplanes, dim3(cdiv((int64_t)B * W, 256)), dim3(256), 0, st, d_states, B, L, q, W, planes);
}

void hamming_neighbours(hipStream_t st, const uint64_t *planes, int B, int q, int W, int L, int64_t dmax,
                        const uint32_t *d_mult, unsigned long long *d_iw)
{
    const int64_t T = (B + HT - 1) / HT;
    const int64_t ntiles = T * (T + 1) / 2;
    hipLaunchKernelGGL(k_hamming, dim3((unsigned)ntiles), dim3(256), 0, st, planes, B, q, W, L, dmax, d_mult, d_iw);
}

// Launches k_gemm_nt with output tiles 128 x tn (tn = 128 or 256; every group's N a multiple).
template <class Pol, int NA, class Epi>
static void launch_gemm(hipStream_t st, int tn, int grid, const void *A0, const void *A1, const void *A2, int64_t lda,
                        const GemmGroup *groups, int ngroups, const Epi &epi)
{
    using T = typename Pol::T;
    if (grid <= 0) return;
    if (tn == 256)
        hipLaunchKernelGGL((k_gemm_nt<Pol, NA, 4, Epi>), dim3(grid), dim3(256), 0, st, (const T *)A0, (const T *)A1,
                           (const T *)A2, lda, groups, ngroups, epi);
    else
        hipLaunchKernelGGL((k_gemm_nt<Pol, NA, 2, Epi>), dim3(grid), dim3(256), 0, st, (const T *)A0, (const T *)A1,
                           (const T *)A2, lda, groups, ngroups, epi);
}

void gemm_energy(hipStream_t st, int prec, const GemmGroup *d_groups, int ngroups, int grid, int tn, const void *W32,
                 const void *Whi, const void *Wlo, const void *Wl3, int64_t ldx, int64_t ldb, int64_t zstride,
                 float *Epart)
{
    EpiEnergy epi{Epart, ldb, zstride};
    if (prec == SDCA_PREC_FP32)
        launch_gemm<PolF32, 1>(st, tn, grid, W32, W32, W32, ldx, d_groups, ngroups, epi);
    else
        launch_gemm<PolBF16, 2>(st, tn, grid, Whi, Wlo, Wl3, ldx, d_groups, ngroups, epi);
}

void softmax(hipStream_t st, int prec, const float *Epart, int64_t zstride, int64_t ldb, int ncmd, int nbchunk,
             const SlotCmd *cmds, const AlnDev *d_aln, float *G32, uint16_t *Ghi, uint16_t *Glo, uint16_t *Gl3,
             double *fpart, double *hpart)
{
    hipLaunchKernelGGL(k_softmax, dim3(nbchunk, ncmd), dim3(256), 0, st, Epart, zstride, ldb, cmds, d_aln,
                       prec == SDCA_PREC_FP32 ? G32 : nullptr, Ghi, Glo, Gl3, fpart, hpart, nbchunk);
}

void gemm_grad(hipStream_t st, int prec, const GemmGroup *d_groups, int ngroups, int grid, int tn, const void *G32,
               const void *Ghi, const void *Glo, const void *Gl3, int64_t ldb, const AlnDev *d_aln, const SlotCmd *cmds,
               const float *Xtrial, int64_t ldx, float *gbuf, int64_t n_slot, const double *hpart, int nbchunk)
{
    EpiGrad epi{gbuf, Xtrial, cmds, hpart, d_aln, n_slot, ldx, 2, 0, nbchunk};
    if (prec == SDCA_PREC_FP32)
        launch_gemm<PolF32, 1>(st, tn, grid, G32, G32, G32, ldb, d_groups, ngroups, epi);
    else
        launch_gemm<PolBF16, 2>(st, tn, grid, Ghi, Glo, Gl3, ldb, d_groups, ngroups, epi);
}

void gemm_materialize(hipStream_t st, int prec, const GemmGroup *d_groups, int ngroups, int grid, int tn,
                      const void *G32, const void *Ghi, const void *Glo, const void *Gl3, int64_t ldb,
                      const AlnDev *d_aln, const SlotCmd *cmds, float *out, int64_t n_item, const double *hpart)
{
    EpiGrad epi{out, nullptr, cmds, hpart, d_aln, n_item, 0, 1, 0, 1};
    if (prec == SDCA_PREC_FP32)
        launch_gemm<PolF32, 1>(st, tn, grid, G32, G32, G32, ldb, d_groups, ngroups, epi);
    else
        launch_gemm<PolBF16, 2>(st, tn, grid, Ghi, Glo, Gl3, ldb, d_groups, ngroups, epi);
}

#define SDCA_DISPATCH_M(m, CALL) \
    switch (m) {                 \
    case 1: CALL(1); break;      \
    case 3: CALL(3); break;      \
    case 5: CALL(5); break;      \
    case 8: CALL(8); break;      \
    case 10: CALL(10); break;    \
    case 16: CALL(16); break;    \
    default: break;              \
    }

void update(hipStream_t st, int m, bool pair_bf16, const SlotCmd *cmds, int ncmd, const float *gbuf, void *dbuf,
            void *S, void *Y, float *Xc, float *Xtr, uint16_t *Whi, uint16_t *Wlo, uint16_t *Wl3, const float *X0,
            int64_t n_slot, int64_t ldx, double *part, int nchunk)
{
#define SDCA_UPD(MM)                                                                                                \
    if (pair_bf16)                                                                                                  \
        hipLaunchKernelGGL((k_update<MM, uint16_t>), dim3(nchunk, ncmd), dim3(256), 0, st, cmds, gbuf,          \
                           (uint16_t *)dbuf, (uint16_t *)S, (uint16_t *)Y, Xc, Xtr, Whi, Wlo, Wl3, X0, n_slot, ldx, \
                           part, nchunk);                                                                           \
    else                                                                                                            \
        hipLaunchKernelGGL((k_update<MM, float>), dim3(nchunk, ncmd), dim3(256), 0, st, cmds, gbuf, (float *)dbuf, \
                           (float *)S, (float *)Y, Xc, Xtr, Whi, Wlo, Wl3, X0, n_slot, ldx, part, nchunk)
    SDCA_DISPATCH_M(m, SDCA_UPD)
#undef SDCA_UPD
}

void dots(hipStream_t st, int m, bool pair_bf16, const SlotCmd *cmds, int ncmd, const float *gbuf,
          const void *dbuf, const void *S, const void *Y, int64_t n_slot, double *part, int nchunk)
{
#define SDCA_DOT(MM)                                                                                                \
    if (pair_bf16)                                                                                                  \
        hipLaunchKernelGGL((k_dots<MM, uint16_t>), dim3(nchunk, ncmd), dim3(256), 0, st, cmds, gbuf,            \
                           (const uint16_t *)dbuf, (const uint16_t *)S, (const uint16_t *)Y, n_slot, part, nchunk); \
    else                                                                                                            \
        hipLaunchKernelGGL((k_dots<MM, float>), dim3(nchunk, ncmd), dim3(256), 0, st, cmds, gbuf,               \
                           (const float *)dbuf, (const float *)S, (const float *)Y, n_slot, part, nchunk)
    SDCA_DISPATCH_M(m, SDCA_DOT)
#undef SDCA_DOT
}

void reduce_results(hipStream_t st, int m, int ncmd, const double *upd_part, const double *dot_part,
                    const double *sm_part, int nchunk, int nbchunk, double *res)
{
    const int W = ResLayout{m}.width();
    hipLaunchKernelGGL(k_reduce, dim3(ncmd), dim3((W + 63) / 64 * 64), 0, st, m, upd_part, dot_part, sm_part, nchunk,
                       nbchunk, res);
}

void score(hipStream_t st, const float *xsrc, const int64_t *src_off, const int32_t *src_ld, const int32_t *loci,
           const int32_t *alns, int nscore, int Lmax, int64_t ldrow, const AlnDev *d_aln, float *rows, int coords)
{
    hipLaunchKernelGGL(k_score, dim3(cdiv(Lmax, 256), nscore), dim3(256), 0, st, xsrc, src_off, src_ld, loci, alns,
                       ldrow, d_aln, rows, coords);
}

}  // namespace kern
}  // namespace sdca
