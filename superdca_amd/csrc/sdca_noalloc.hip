// sdca_noalloc.hip -- Hamming reweighting that allocates no device memory.
//
// sdca_reweight (sdca_api.cpp) allocates its bit-planes and neighbour counters per call.
// Callers that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc) use sdca_reweight_noalloc instead: the same kernels,
// with the planes and counters in the file-scope device array `scratch` below.  Each GPU
// holds its own copy of the array until the process ends; its address on a device is
// taken once with hipGetSymbolAddress (with that device current) and kept per device.
// Every call on a GPU shares that GPU's array, so calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "sdca_internal.hpp"

int sdca_fail(int code, const char *fmt, ...);  // sdca_api.cpp

namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kMaxDevices = 64;

__device__ __align__(256) uint8_t scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's `scratch` (looked up on the first call per device).
int scratch_on(int dev, void **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return sdca_fail(SDCA_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                             hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SDCA_OK;
}

// Neighbour counters start at each sequence's own multiplicity (plmDCA_optimizer_parameters.hpp:60-83
// counts m_i itself); a kernel rather than a memset, so the call can be captured in a graph.
__global__ void k_init_counts(unsigned long long *__restrict__ iw, const uint32_t *__restrict__ mult, int B)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B) iw[i] = mult ? mult[i] : 1u;
}

// w_i = m_i / double(int_w_i) (plmDCA_optimizer_parameters.hpp:95), IEEE division in FP64.
__global__ void k_finish_weights(const unsigned long long *__restrict__ iw, const uint32_t *__restrict__ mult, int B,
                                 double *__restrict__ w, uint64_t *__restrict__ int_w)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    const unsigned long long c = iw[i];
    w[i] = (mult ? (double)mult[i] : 1.0) / (double)c;
    if (int_w) int_w[i] = c;
}

}  // namespace

extern "C" int sdca_reweight_noalloc(const uint8_t *d_states, int32_t B, int32_t L, int32_t q, double theta,
                                     const uint32_t *d_mult, uint64_t *d_int_w, double *d_w, void *stream)
{
    if (!d_states || !d_w) return sdca_fail(SDCA_E_ARG, "d_states and d_w must be device pointers");
    if (B < 1 || L < 1) return sdca_fail(SDCA_E_ARG, "need B >= 1 and L >= 1 (got B=%d L=%d)", B, L);
    if (q < 2 || q > 8) return sdca_fail(SDCA_E_ARG, "q_env must be in [2, 8] (got %d)", q);
    const int W = (L + 63) / 64;
    const size_t counts = (size_t)B * sizeof(unsigned long long);
    const size_t planes = (size_t)B * q * W * sizeof(uint64_t);
    const size_t need = ((counts + 255) / 256) * 256 + planes;
    if (need > kScratchBytes)
        return sdca_fail(SDCA_E_ARG,
                         "alignment too large for the %zu-byte scratch of sdca_reweight_noalloc (B x q x ceil(L/64) "
                         "= %zu words need %zu bytes); use sdca_reweight",
                         kScratchBytes, (size_t)B * q * W, need);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return sdca_fail(SDCA_E_HIP, "no current HIP device");
    void *base = nullptr;
    int rc = scratch_on(dev, &base);
    if (rc) return rc;
    auto *iw = reinterpret_cast<unsigned long long *>(base);
    auto *pl = reinterpret_cast<uint64_t *>(reinterpret_cast<uint8_t *>(base) + ((counts + 255) / 256) * 256);
    const int64_t dmax = sdca::identity_dmax(L, theta);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const unsigned nb = (unsigned)((B + 255) / 256);
    hipLaunchKernelGGL(k_init_counts, dim3(nb), dim3(256), 0, st, iw, d_mult, B);
    sdca::kern::bitplanes(st, d_states, B, L, q, W, pl);
    sdca::kern::hamming_neighbours(st, pl, B, q, W, L, dmax, d_mult, iw);
    hipLaunchKernelGGL(k_finish_weights, dim3(nb), dim3(256), 0, st, iw, d_mult, B, d_w, d_int_w);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return sdca_fail(SDCA_E_HIP, "reweighting kernels failed to launch: %s", hipGetErrorString(e));
    return SDCA_OK;
}
