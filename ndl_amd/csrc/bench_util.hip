// bench_util.hip -- measurement helpers for bench.py (libndl_benchutil.so).
// Not part of the product ABI.
//
// ndl_bench_flush reads `bytes` of device memory with wide loads and writes
// one word per workgroup, so the 256 MB Infinity Cache ends up holding clean
// lines of the flush buffer instead of the benchmark's working set (a cold
// cache with no dirty write-backs pending for the timed kernel).
#include <hip/hip_runtime.h>
#include <stdint.h>

__global__ __launch_bounds__(256) void flush_read_kernel(const uint4* __restrict__ p, size_t n16,
                                                         uint32_t* __restrict__ sink)
{
    uint32_t acc = 0;
    const size_t step = (size_t)gridDim.x * 256;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += step) {
        uint4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    __shared__ uint32_t red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] ^= red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sink[blockIdx.x] = red[0];
}

// streaming copy of n16 16-byte words, one per thread: the measured-copy
// ceiling (on MI355X this form moves ~6.4 TB/s for 1 GiB, a grid-stride loop
// over 4096 workgroups ~4.9 TB/s)
__global__ __launch_bounds__(256) void copy_kernel(const uint4* __restrict__ a, uint4* __restrict__ b, size_t n16)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n16) b[i] = a[i];
}

// one wave that waits `ticks` of the 100 MHz real-time counter: queued
// before a timed region it keeps the GPU busy while the host enqueues the
// region's commands, so the events around them time the GPU work alone
// (without it a short kernel's event pair also spans the host's launch
// overhead).  Touches no memory but one word.
__global__ __launch_bounds__(64) void spin_kernel(uint64_t ticks, uint32_t* sink)
{
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    uint32_t n = 0;
    while (__builtin_amdgcn_s_memrealtime() - t0 < ticks) {
        __builtin_amdgcn_s_sleep(8);
        n++;
    }
    if (threadIdx.x == 0 && n == 0xffffffffu) sink[0] = n;
}

extern "C" __attribute__((visibility("default"))) int ndl_bench_spin(double microseconds, void* sink, void* stream)
{
    const uint64_t ticks = (uint64_t)(microseconds * 100.0);
    hipLaunchKernelGGL(spin_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream, ticks, (uint32_t*)sink);
    return hipGetLastError() == hipSuccess ? 0 : -6;
}

extern "C" __attribute__((visibility("default"))) int ndl_bench_flush(const void* buf, size_t bytes, void* sink,
                                                                      void* stream)
{
    hipLaunchKernelGGL(flush_read_kernel, dim3(2048), dim3(256), 0, (hipStream_t)stream, (const uint4*)buf,
                       bytes / 16, (uint32_t*)sink);
    return hipGetLastError() == hipSuccess ? 0 : -6;
}

extern "C" __attribute__((visibility("default"))) int ndl_bench_copy(const void* src, void* dst, size_t bytes,
                                                                     void* stream)
{
    const size_t n16 = bytes / 16;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)((n16 + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)src, (uint4*)dst, n16);
    return hipGetLastError() == hipSuccess ? 0 : -6;
}
