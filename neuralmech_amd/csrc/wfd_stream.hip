// HBM streaming ceilings for the roofline harness: the rates a plain streaming kernel reaches for
// the access mixes the step kernels use.  One 16-byte chunk per lane (the fastest shape measured:
// more chunks per lane lower the rate by 5-10 %).
#include <hip/hip_runtime.h>

#include "wfd_internal.h"

namespace wfd {
namespace {

typedef float sb_v4f __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void stream_copy(const sb_v4f* __restrict__ a, sb_v4f* __restrict__ o, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n4) o[i] = a[i];
}

template <bool NT>
__global__ __launch_bounds__(256) void stream_r3w1(const sb_v4f* __restrict__ a, const sb_v4f* __restrict__ b,
                                                   const sb_v4f* __restrict__ c, sb_v4f* __restrict__ o, int64_t n4)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n4) return;
    if constexpr (NT) {
        const sb_v4f x = __builtin_nontemporal_load(a + i), y = __builtin_nontemporal_load(b + i),
                     z = __builtin_nontemporal_load(c + i);
        __builtin_nontemporal_store(x + y + z, o + i);
    } else {
        o[i] = a[i] + b[i] + c[i];
    }
}

// read-only: each block sums its 4096 floats into o[block] (0.1 % of the bytes read)
__global__ __launch_bounds__(256) void stream_read(const sb_v4f* __restrict__ a, float* __restrict__ o, int64_t n4)
{
    const int64_t i0 = (int64_t)blockIdx.x * 1024 + threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int64_t i = i0 + 256 * u;
        if (i < n4) {
            const sb_v4f v = a[i];
            s += (v.x + v.y) + (v.z + v.w);
        }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
    __shared__ float part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) o[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" int wfd_stream_bench(const float* a, const float* b, const float* c, float* o, int64_t n, int32_t mode,
                                wfd_stream_t stream)
{
    if (mode < 0 || mode > 3) return set_error(WFD_EINVAL, "wfd_stream_bench: mode must be 0..3");
    if (!a || !o || (mode == 1 || mode == 2) && (!b || !c))
        return set_error(WFD_EINVAL, "wfd_stream_bench: NULL array");
    if (n < 0 || n % 4) return set_error(WFD_EINVAL, "wfd_stream_bench: n must be a non-negative multiple of 4");
    const uintptr_t mis = ((uintptr_t)a | (uintptr_t)o | (uintptr_t)(b ? b : a) | (uintptr_t)(c ? c : a)) & 15u;
    if (mis) return set_error(WFD_EINVAL, "wfd_stream_bench: arrays must be 16-byte aligned");
    if (n == 0) return WFD_OK;
    const int64_t n4 = n / 4;
    hipStream_t s = (hipStream_t)stream;
    const sb_v4f* A = (const sb_v4f*)a;
    const dim3 blk(256);
    const dim3 grid((unsigned)((n4 + 255) / 256));
    switch (mode) {
        case 0: hipLaunchKernelGGL(stream_copy, grid, blk, 0, s, A, (sb_v4f*)o, n4); break;
        case 1:
            hipLaunchKernelGGL((stream_r3w1<false>), grid, blk, 0, s, A, (const sb_v4f*)b, (const sb_v4f*)c,
                               (sb_v4f*)o, n4);
            break;
        case 2:
            hipLaunchKernelGGL((stream_r3w1<true>), grid, blk, 0, s, A, (const sb_v4f*)b, (const sb_v4f*)c,
                               (sb_v4f*)o, n4);
            break;
        default: hipLaunchKernelGGL(stream_read, dim3((unsigned)((n4 + 1023) / 1024)), blk, 0, s, A, o, n4); break;
    }
    if (hipGetLastError() != hipSuccess) return set_error(WFD_EHIP, "wfd_stream_bench: launch failed");
    return WFD_OK;
}
