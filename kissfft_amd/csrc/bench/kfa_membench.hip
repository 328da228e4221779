// kfa_membench.hip -- streaming ceilings of this device, for the benchmark's roofline.
//
// Not part of the FFT library (built as its own libkfa_membench.so).  Three hand-written
// streaming kernels with 16-byte lanes, each moving a known number of HBM bytes:
//   read   every 16-byte word of a buffer is loaded once (XOR-reduced, one word per
//          workgroup stored so the loads are not removed)
//   copy   a buffer copied to another one
//   rows   the real-forward row pattern: rows of n floats read, rows of n/2 + 1 complex
//          (8 bytes each, so every other row starts 8 bytes off a 16-byte boundary) written;
//          the bytes a real-forward transform of the same batch must move; in 16-byte
//          lanes (rows) and in 8-byte lanes, one complex value per lane (rows8: faster on
//          this part, 5.58 against 5.43 TB/s)
// bench.py reports their rates beside the FFT kernel's as the part's practical ceilings.
#include <hip/hip_runtime.h>

#include <stdint.h>

namespace {

constexpr int kThreads = 256;

template <int kUnroll>
__global__ void __launch_bounds__(kThreads) k_read(const uint4 *__restrict__ src, size_t nvec, uint4 *sink)
{
    uint4 acc = make_uint4(0, 0, 0, 0);
    const size_t stride = (size_t)gridDim.x * kThreads;
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < nvec; i += kUnroll * stride) {
        uint4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            acc.x ^= v[u].x;
            acc.y ^= v[u].y;
            acc.z ^= v[u].z;
            acc.w ^= v[u].w;
        }
    }
    for (; i < nvec; i += stride) {
        const uint4 v = src[i];
        acc.x ^= v.x;
        acc.y ^= v.y;
        acc.z ^= v.z;
        acc.w ^= v.w;
    }
    __shared__ uint4 red[kThreads];
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint4 r = make_uint4(0, 0, 0, 0);
        for (int t = 0; t < kThreads; ++t) {
            r.x ^= red[t].x;
            r.y ^= red[t].y;
            r.z ^= red[t].z;
            r.w ^= red[t].w;
        }
        sink[blockIdx.x] = r;
    }
}

template <int kUnroll>
__global__ void __launch_bounds__(kThreads) k_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t nvec)
{
    const size_t stride = (size_t)gridDim.x * kThreads;
    size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    for (; i + (kUnroll - 1) * stride < nvec; i += kUnroll * stride) {
        uint4 v[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u)
            dst[i + u * stride] = v[u];
    }
    for (; i < nvec; i += stride)
        dst[i] = src[i];
}

// One workgroup per row: n floats in (n % 1024 == 0: n/4 16-byte words, n/1024 per thread),
// (n/2 + 1) * 8 bytes out.
template <int WORDS_PER_THREAD>
__global__ void __launch_bounds__(kThreads) k_rows(const uint4 *__restrict__ src, char *__restrict__ dst, int n)
{
    const long long r = blockIdx.x;
    const uint4 *in = src + r * (n / 4);
    const size_t out_row = (size_t)(n / 2 + 1) * 8;
    char *out = dst + (size_t)r * out_row;
    uint4 v[WORDS_PER_THREAD];
#pragma unroll
    for (int u = 0; u < WORDS_PER_THREAD; ++u)
        v[u] = in[threadIdx.x + u * kThreads];
    // rows start 8 bytes off a 16-byte boundary every other row: 8-byte head there,
    // 8-byte tail otherwise
    const int head = (reinterpret_cast<uintptr_t>(out) & 15) ? 8 : 0;
    uint4 *o = reinterpret_cast<uint4 *>(out + head);
#pragma unroll
    for (int u = 0; u < WORDS_PER_THREAD; ++u)
        o[threadIdx.x + u * kThreads] = v[u];
    if (threadIdx.x == 0) {
        uint2 *edge = reinterpret_cast<uint2 *>(head ? out : out + out_row - 8);
        *edge = make_uint2(v[0].x, v[0].y);
    }
}

// The same row pattern in 8-byte lanes: n/2 8-byte words in, n/2 + 1 out (n % 512 == 0).
template <int WORDS_PER_THREAD>
__global__ void __launch_bounds__(kThreads) k_rows8(const uint2 *__restrict__ src, uint2 *__restrict__ dst, int n)
{
    const long long r = blockIdx.x;
    const uint2 *in = src + r * (n / 2);
    uint2 *out = dst + r * (n / 2 + 1);
    uint2 v[WORDS_PER_THREAD];
#pragma unroll
    for (int u = 0; u < WORDS_PER_THREAD; ++u)
        v[u] = in[threadIdx.x + u * kThreads];
#pragma unroll
    for (int u = 0; u < WORDS_PER_THREAD; ++u)
        out[threadIdx.x + u * kThreads] = v[u];
    if (threadIdx.x == 0)
        out[n / 2] = v[0];
}

int grid_for(size_t nvec, int blocks_per_cu)
{
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    const size_t want = (nvec + kThreads - 1) / kThreads;
    const size_t cap = (size_t)cus * (blocks_per_cu > 0 ? blocks_per_cu : 16);
    return (int)(want < cap ? (want ? want : 1) : cap);
}

}  // namespace

extern "C" {

// sink: at least 16 bytes per workgroup (16 * blocks_per_cu * CUs); blocks_per_cu <= 0: 16;
// unroll: 16-byte loads in flight per thread, 4 or 8
__attribute__((visibility("default"))) int kfa_mb_read(const void *src, size_t bytes, void *sink, void *stream,
                                                       int blocks_per_cu, int unroll)
{
    const size_t nvec = bytes / 16;
    const dim3 g(grid_for(nvec, blocks_per_cu));
    if (unroll == 8)
        hipLaunchKernelGGL(k_read<8>, g, dim3(kThreads), 0, (hipStream_t)stream, (const uint4 *)src, nvec,
                           (uint4 *)sink);
    else
        hipLaunchKernelGGL(k_read<4>, g, dim3(kThreads), 0, (hipStream_t)stream, (const uint4 *)src, nvec,
                           (uint4 *)sink);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

__attribute__((visibility("default"))) int kfa_mb_copy(const void *src, void *dst, size_t bytes, void *stream,
                                                       int blocks_per_cu, int unroll)
{
    const size_t nvec = bytes / 16;
    const dim3 g(grid_for(nvec, blocks_per_cu));
    if (unroll == 8)
        hipLaunchKernelGGL(k_copy<8>, g, dim3(kThreads), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst,
                           nvec);
    else
        hipLaunchKernelGGL(k_copy<4>, g, dim3(kThreads), 0, (hipStream_t)stream, (const uint4 *)src, (uint4 *)dst,
                           nvec);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// n must be 4096 or 1024 (one or four 16-byte words per thread); src 16-byte aligned
__attribute__((visibility("default"))) int kfa_mb_rows(const void *src, void *dst, long long rows, int n,
                                                       void *stream)
{
    if (rows <= 0 || rows > 0x7fffffffLL || (reinterpret_cast<uintptr_t>(src) & 15) ||
        (reinterpret_cast<uintptr_t>(dst) & 7))
        return -2;
    if (n == 4096)
        hipLaunchKernelGGL(k_rows<4>, dim3((unsigned)rows), dim3(kThreads), 0, (hipStream_t)stream,
                           (const uint4 *)src, (char *)dst, n);
    else if (n == 1024)
        hipLaunchKernelGGL(k_rows<1>, dim3((unsigned)rows), dim3(kThreads), 0, (hipStream_t)stream,
                           (const uint4 *)src, (char *)dst, n);
    else
        return -2;
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// the row pattern in 8-byte lanes; n must be 4096 or 1024, src and dst 8-byte aligned
__attribute__((visibility("default"))) int kfa_mb_rows8(const void *src, void *dst, long long rows, int n,
                                                        void *stream)
{
    if (rows <= 0 || rows > 0x7fffffffLL || (reinterpret_cast<uintptr_t>(src) & 7) ||
        (reinterpret_cast<uintptr_t>(dst) & 7))
        return -2;
    if (n == 4096)
        hipLaunchKernelGGL(k_rows8<8>, dim3((unsigned)rows), dim3(kThreads), 0, (hipStream_t)stream,
                           (const uint2 *)src, (uint2 *)dst, n);
    else if (n == 1024)
        hipLaunchKernelGGL(k_rows8<2>, dim3((unsigned)rows), dim3(kThreads), 0, (hipStream_t)stream,
                           (const uint2 *)src, (uint2 *)dst, n);
    else
        return -2;
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // extern "C"
