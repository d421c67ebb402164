// Kernels behind the reference-layout entry points (wfd_ref_*): the constant-speed 2nd-order step
// of projects/18_simple_wave_fd in the layout its drivers use (N rows of M_padded elements, the
// boundary ring stored in the array and never written), in fp32, fp16 storage and int16
// block-scaled storage, plus the 1D step and an LDS temporal-blocked multi-step.
//
// Arithmetic is the slice form of 2D_wave_cupy.py:54-62 evaluated in fp32 with IEEE division
// and no contraction:
//   lx = ((u1[i-1] - 2 u1) + u1[i+1]) / dx2;   ly = ((u1[j-1] - 2 u1) + u1[j+1]) / dy2
//   u2 = ((-u0) + 2 u1) + coef * (lx + ly)       coef = f32(dt^2 c^2), dx2 = f32(dx^2), dy2 = f32(dy^2)
// which is what NumPy computes for those slices with float32 arrays and Python-float scalars.
#include <hip/hip_fp16.h>

#include "wfd_internal.h"

namespace wfd {

struct RefCoef {
    float coef, dx2, dy2;
};

__device__ __forceinline__ float ref_update(float u0, float c, float n, float s, float w, float e, const RefCoef& k)
{
    const float lx = ((n - 2.f * c) + s) / k.dx2;
    const float ly = ((w - 2.f * c) + e) / k.dy2;
    return ((-u0) + 2.f * c) + k.coef * (lx + ly);
}

// ---------------------------------------------------------------- one step, global / cache
// block 64 x 4: a wave covers 64 consecutive columns of one row (coalesced), the neighbours
// along the row come from the same cache lines.
__global__ __launch_bounds__(256) void ref_step2d_kernel(const float* u0, const float* __restrict__ u1, float* u2,
                                                         RefCoef k, int N, int M, int64_t ld)
{
    const int j = blockIdx.x * 64 + (threadIdx.x & 63);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i <= 0 || i >= N - 1 || j <= 0 || j >= M - 1) return;
    const int64_t c = (int64_t)i * ld + j;
    u2[c] = ref_update(u0[c], u1[c], u1[c - ld], u1[c + ld], u1[c - 1], u1[c + 1], k);
}

// ---------------------------------------------------------------- one step, 4-node chunks
// For rows whose starts are 16-byte aligned (ld % 4 == 0): a lane owns 4 consecutive nodes, a
// wave 256 columns of RI consecutive rows, the RI + 2 rows of u1 it needs are loaded once into
// registers; the row neighbours across chunk boundaries come from lanes l - 1 / l + 1 (one
// scalar load at each end of the wave).  Nodes outside the interior are never stored: chunks
// that straddle the boundary column or the padding store element by element.
template <int RI>
__global__ __launch_bounds__(256) void ref_step2d_vec_kernel(const float* u0, const float* __restrict__ u1, float* u2,
                                                             RefCoef k, int N, int M, int64_t ld)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t base = (int64_t)blockIdx.x * 256;
    const int64_t c = base + 4 * lane;
    const int64_t r0 = 1 + ((int64_t)blockIdx.y * 4 + wave) * RI;
    if (r0 >= N - 1) return;  // whole wave
    const bool in = c < M;
    float4 w[RI + 2];
#pragma unroll
    for (int t = 0; t < RI + 2; t++) {
        const int64_t r = r0 - 1 + t;
        w[t] = (in && r < N) ? *reinterpret_cast<const float4*>(u1 + r * ld + c) : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < RI; q++) {
        const int64_t r = r0 + q;
        if (r >= N - 1) break;  // whole wave
        const float4 C = w[q + 1], Nn = w[q], S = w[q + 2];
        float wl = __shfl_up(C.w, 1), er = __shfl_down(C.x, 1);
        if (lane == 0) wl = base > 0 ? u1[r * ld + base - 1] : 0.f;
        if (lane == 63) er = base + 256 < ld ? u1[r * ld + base + 256] : 0.f;
        if (!in) continue;
        const float4 P = *reinterpret_cast<const float4*>(u0 + r * ld + c);
        float o[4];
        o[0] = ref_update(P.x, C.x, Nn.x, S.x, wl, C.y, k);
        o[1] = ref_update(P.y, C.y, Nn.y, S.y, C.x, C.z, k);
        o[2] = ref_update(P.z, C.z, Nn.z, S.z, C.y, C.w, k);
        o[3] = ref_update(P.w, C.w, Nn.w, S.w, C.z, er, k);
        float* dst = u2 + r * ld + c;
        if (c >= 1 && c + 3 <= M - 2) {
            *reinterpret_cast<float4*>(dst) = make_float4(o[0], o[1], o[2], o[3]);
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++)
                if (c + e >= 1 && c + e <= M - 2) dst[e] = o[e];
        }
    }
}

// ---------------------------------------------------------------- one step, LDS tile
// 64 x 4 output tile staged with a one-node halo in LDS, as the shared-memory variant of the
// reference does (step2D_wave.cu:187-240); identical results.
__global__ __launch_bounds__(256) void ref_step2d_lds_kernel(const float* u0, const float* __restrict__ u1, float* u2,
                                                             RefCoef k, int N, int M, int64_t ld)
{
    __shared__ float t[6][66];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int j = blockIdx.x * 64 + tx;
    const int i = blockIdx.y * 4 + ty;
    auto at = [&](int ii, int jj) -> float {
        return (ii >= 0 && ii < N && jj >= 0 && jj < M) ? u1[(int64_t)ii * ld + jj] : 0.f;
    };
    t[ty + 1][tx + 1] = at(i, j);
    if (ty == 0) t[0][tx + 1] = at(i - 1, j);
    if (ty == 3) t[5][tx + 1] = at(i + 1, j);
    if (tx == 0) t[ty + 1][0] = at(i, j - 1);
    if (tx == 63) t[ty + 1][65] = at(i, j + 1);
    __syncthreads();
    if (i <= 0 || i >= N - 1 || j <= 0 || j >= M - 1) return;
    const int64_t c = (int64_t)i * ld + j;
    u2[c] = ref_update(u0[c], t[ty + 1][tx + 1], t[ty][tx + 1], t[ty + 2][tx + 1], t[ty + 1][tx], t[ty + 1][tx + 2], k);
}

// ---------------------------------------------------------------- 1D
__global__ __launch_bounds__(256) void ref_step1d_kernel(const float* u0, const float* __restrict__ u1, float* u2,
                                                         float coef, float dx2, int N)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i <= 0 || i >= N - 1) return;
    const float c = u1[i];
    const float lx = ((u1[i - 1] - 2.f * c) + u1[i + 1]) / dx2;
    u2[i] = ((-u0[i]) + 2.f * c) + coef * lx;
}

// ---------------------------------------------------------------- temporal blocking
// S steps per launch on an LDS tile of TI x TJ nodes whose outer S nodes are halo: the valid
// region shrinks by one node per step, the inner (TI-2S) x (TJ-2S) nodes are written to (v0, v1)
// (older, newer).  S is even so the two LDS levels end in place.  Bit-identical to S single
// steps: every value is produced by the same expression from the same inputs.
template <int S, int TI, int TJ>
__global__ __launch_bounds__(256) void ref_temporal_kernel(const float* __restrict__ u0, const float* __restrict__ u1,
                                                           float* v0, float* v1, RefCoef k, int N, int M, int64_t ld)
{
    __shared__ float lv[2][TI][TJ];
    const int i0 = blockIdx.y * (TI - 2 * S) - S;
    const int j0 = blockIdx.x * (TJ - 2 * S) - S;
    for (int x = threadIdx.x; x < TI * TJ; x += 256) {
        const int ti = x / TJ, tj = x % TJ;
        const int i = i0 + ti, j = j0 + tj;
        const bool in = i >= 0 && i < N && j >= 0 && j < M;
        const int64_t c = (int64_t)i * ld + j;
        lv[0][ti][tj] = in ? u0[c] : 0.f;
        lv[1][ti][tj] = in ? u1[c] : 0.f;
    }
    __syncthreads();
#pragma unroll 1
    for (int s = 0; s < S; s++) {
        float(*a)[TJ] = lv[s & 1];        // older level, overwritten by the new one
        float(*b)[TJ] = lv[1 - (s & 1)];  // current level
        for (int x = threadIdx.x; x < TI * TJ; x += 256) {
            const int ti = x / TJ, tj = x % TJ;
            const int i = i0 + ti, j = j0 + tj;
            if (ti > s && ti < TI - 1 - s && tj > s && tj < TJ - 1 - s && i > 0 && i < N - 1 && j > 0 && j < M - 1)
                a[ti][tj] = ref_update(a[ti][tj], b[ti][tj], b[ti - 1][tj], b[ti + 1][tj], b[ti][tj - 1], b[ti][tj + 1], k);
        }
        __syncthreads();
    }
    for (int x = threadIdx.x; x < TI * TJ; x += 256) {
        const int ti = x / TJ, tj = x % TJ;
        if (ti < S || ti >= TI - S || tj < S || tj >= TJ - S) continue;
        const int i = i0 + ti, j = j0 + tj;
        if (i >= N || j >= M) continue;
        const int64_t c = (int64_t)i * ld + j;
        v0[c] = lv[0][ti][tj];
        v1[c] = lv[1][ti][tj];
    }
}

// ---------------------------------------------------------------- fp16 storage
// fp16 fields, fp32 arithmetic, round-to-nearest-even on store; a lane updates two adjacent
// nodes with 4-byte loads and stores.
__global__ __launch_bounds__(256) void ref_mixed_kernel(const __half* u0, const __half* __restrict__ u1, __half* u2,
                                                        RefCoef k, int N, int M, int64_t ld)
{
    const int j = 2 * (blockIdx.x * 64 + (threadIdx.x & 63));
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i <= 0 || i >= N - 1 || j >= M) return;
    const int64_t c = (int64_t)i * ld + j;
    const float2 p0 = __half22float2(*reinterpret_cast<const __half2*>(u0 + c));
    const float2 pc = __half22float2(*reinterpret_cast<const __half2*>(u1 + c));
    const float2 pn = __half22float2(*reinterpret_cast<const __half2*>(u1 + c - ld));
    const float2 ps = __half22float2(*reinterpret_cast<const __half2*>(u1 + c + ld));
    const float pw = j > 0 ? __half2float(u1[c - 1]) : 0.f;
    const float pe = __half2float(u1[c + 2]);  // inside the row padding at the last pair
    const float a = ref_update(p0.x, pc.x, pn.x, ps.x, pw, pc.y, k);
    const float b = ref_update(p0.y, pc.y, pn.y, ps.y, pc.x, pe, k);
    if (j > 0 && j + 1 < M - 1) {
        *reinterpret_cast<__half2*>(u2 + c) = __floats2half2_rn(a, b);
    } else {
        if (j > 0 && j < M - 1) u2[c] = __float2half_rn(a);
        if (j + 1 < M - 1) u2[c + 1] = __float2half_rn(b);
    }
}

// ---------------------------------------------------------------- int16 block-scaled storage
// u = s[block] * q with one fp32 scale per 32 consecutive nodes of a row (16 lanes of a pair
// each).  Decode, update as above, then the block's largest magnitude m sets the new scale:
// q = rint(u * (32767/m)), s = m / 32767; q = 0 where 32767/m is not finite.  Rows 0 and N-1
// are never written; nodes j = 0 and j >= M-1 of interior rows encode zero.
__global__ __launch_bounds__(256) void ref_scaled_kernel(const int16_t* q0, const float* s0, const int16_t* q1,
                                                         const float* s1, int16_t* q2, float* s2, RefCoef k, int N,
                                                         int M, int64_t ld)
{
    const int lane = threadIdx.x & 63;
    const int j = 2 * (blockIdx.x * 64 + lane);
    const int i = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i <= 0 || i >= N - 1) return;  // whole waves: ld % 128 == 0 keeps j < ld for every lane
    const int64_t c = (int64_t)i * ld + j;
    const int64_t S = ld / 32;
    const int64_t b = (int64_t)i * S + j / 32;
    auto dec = [&](const float* s, const int16_t* q, int64_t blk, int64_t x) { return s[blk] * (float)q[x]; };
    const short2 a0 = *reinterpret_cast<const short2*>(q0 + c);
    const short2 a1 = *reinterpret_cast<const short2*>(q1 + c);
    const short2 an = *reinterpret_cast<const short2*>(q1 + c - ld);
    const short2 as = *reinterpret_cast<const short2*>(q1 + c + ld);
    const float u0x = s0[b] * (float)a0.x, u0y = s0[b] * (float)a0.y;
    const float ucx = s1[b] * (float)a1.x, ucy = s1[b] * (float)a1.y;
    const float unx = s1[b - S] * (float)an.x, uny = s1[b - S] * (float)an.y;
    const float usx = s1[b + S] * (float)as.x, usy = s1[b + S] * (float)as.y;
    const float uw = j > 0 ? dec(s1, q1, b - (j % 32 == 0 ? 1 : 0), c - 1) : 0.f;
    const float ue = j + 2 < ld ? dec(s1, q1, b + (j % 32 == 30 ? 1 : 0), c + 2) : 0.f;
    float x = ref_update(u0x, ucx, unx, usx, uw, ucy, k);
    float y = ref_update(u0y, ucy, uny, usy, ucx, ue, k);
    if (j == 0 || j >= M - 1) x = 0.f;
    if (j + 1 >= M - 1) y = 0.f;
    float m = fmaxf(fabsf(x), fabsf(y));
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) m = fmaxf(m, __shfl_xor(m, off, 16));
    // 32767/m overflows for an all-zero block and for a block maximum below ~1e-34: such a
    // block encodes as zero (the division by a block maximum that small is undefined in the
    // reference's encoding)
    float inv = 32767.f / m;
    if (!(inv <= 3.402823466e38f)) inv = 0.f;
    short2 out;
    out.x = (short)__float2int_rn(x * inv);
    out.y = (short)__float2int_rn(y * inv);
    *reinterpret_cast<short2*>(q2 + c) = out;
    if ((lane & 15) == 0) s2[b] = m / 32767.f;
}

}  // namespace wfd

// =============================================================================== C ABI
using namespace wfd;

namespace {
int ref_args(int N, int M, int M_padded, double c, double dt, double dx, double dy, RefCoef& k)
{
    if (N < 3 || M < 3 || M_padded < M) return set_error(WFD_EINVAL, "wfd_ref: need N, M >= 3 and M_padded >= M");
    if (!(dt > 0) || !(dx > 0) || !(dy > 0) || !(c > 0))
        return set_error(WFD_EINVAL, "wfd_ref: wavespeed, dt, dx, dy must be > 0");
    k.coef = (float)((dt * dt) * (c * c));
    k.dx2 = (float)(dx * dx);
    k.dy2 = (float)(dy * dy);
    return WFD_OK;
}
// Peak fp32 FMA rate probe: per thread a chain of fused multiply-adds held in registers, 4 per
// iteration (8 flop).  The chain is dependent, so throughput comes from the other waves of the SIMD.
__global__ __launch_bounds__(256) void fma_chain_kernel(float* out, float s, int n)
{
    float x = (float)threadIdx.x, y = s;
    for (int i = 0; i < n; i++) {
        x = __builtin_fmaf(x, y, s);
        y = __builtin_fmaf(y, x, s);
        x = __builtin_fmaf(x, y, s);
        y = __builtin_fmaf(y, x, s);
    }
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = x + y;
}

int launched(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        return set_error(WFD_EHIP, m.c_str());
    }
    return WFD_OK;
}
}  // namespace

extern "C" {

int wfd_ref_step2d(const float* u0, const float* u1, float* u2, double wavespeed, double dt, double dx, double dy,
                   int32_t N, int32_t M, int32_t M_padded, int32_t variant, wfd_stream_t stream)
{
    RefCoef k;
    int rc = ref_args(N, M, M_padded, wavespeed, dt, dx, dy, k);
    if (rc) return rc;
    if (!u0 || !u1 || !u2 || u1 == u2) return set_error(WFD_EINVAL, "wfd_ref_step2d: NULL or aliased u1/u2");
    constexpr int RI = 8;
    const bool vec = variant == 0 && M_padded % 4 == 0 && ((uintptr_t)u0 | (uintptr_t)u1 | (uintptr_t)u2) % 16 == 0;
    if (vec) {
        const dim3 g4((M + 255) / 256, (N - 2 + 4 * RI - 1) / (4 * RI));
        hipLaunchKernelGGL((ref_step2d_vec_kernel<RI>), g4, dim3(256), 0, (hipStream_t)stream, u0, u1, u2, k, N, M,
                           (int64_t)M_padded);
        return launched("wfd_ref_step2d");
    }
    const dim3 grid((M + 63) / 64, (N + 3) / 4);
    if (variant == 1)
        hipLaunchKernelGGL(ref_step2d_lds_kernel, grid, dim3(256), 0, (hipStream_t)stream, u0, u1, u2, k, N, M,
                           (int64_t)M_padded);
    else
        hipLaunchKernelGGL(ref_step2d_kernel, grid, dim3(256), 0, (hipStream_t)stream, u0, u1, u2, k, N, M,
                           (int64_t)M_padded);
    return launched("wfd_ref_step2d");
}

int wfd_ref_step1d(const float* u0, const float* u1, float* u2, double wavespeed, double dt, double dx, int32_t N,
                   wfd_stream_t stream)
{
    RefCoef k;
    int rc = ref_args(N, 3, 3, wavespeed, dt, dx, 1.0, k);
    if (rc) return rc;
    if (!u0 || !u1 || !u2 || u1 == u2) return set_error(WFD_EINVAL, "wfd_ref_step1d: NULL or aliased u1/u2");
    hipLaunchKernelGGL(ref_step1d_kernel, dim3((N + 255) / 256), dim3(256), 0, (hipStream_t)stream, u0, u1, u2, k.coef,
                       k.dx2, N);
    return launched("wfd_ref_step1d");
}

int wfd_ref_temporal(const float* u0, const float* u1, float* v0, float* v1, double wavespeed, double dt, double dx,
                     double dy, int32_t N, int32_t M, int32_t M_padded, int32_t steps, wfd_stream_t stream)
{
    RefCoef k;
    int rc = ref_args(N, M, M_padded, wavespeed, dt, dx, dy, k);
    if (rc) return rc;
    if (!u0 || !u1 || !v0 || !v1) return set_error(WFD_EINVAL, "wfd_ref_temporal: NULL field");
    if (v0 == u0 || v0 == u1 || v1 == u0 || v1 == u1 || v0 == v1)
        return set_error(WFD_EINVAL, "wfd_ref_temporal: (v0, v1) must be a second pair of fields");
    constexpr int TI = 48, TJ = 128;
    if (steps == 8) {
        const dim3 grid((M + TJ - 16 - 1) / (TJ - 16), (N + TI - 16 - 1) / (TI - 16));
        hipLaunchKernelGGL((ref_temporal_kernel<8, TI, TJ>), grid, dim3(256), 0, (hipStream_t)stream, u0, u1, v0, v1, k,
                           N, M, (int64_t)M_padded);
    } else if (steps == 4) {
        const dim3 grid((M + TJ - 8 - 1) / (TJ - 8), (N + TI - 8 - 1) / (TI - 8));
        hipLaunchKernelGGL((ref_temporal_kernel<4, TI, TJ>), grid, dim3(256), 0, (hipStream_t)stream, u0, u1, v0, v1, k,
                           N, M, (int64_t)M_padded);
    } else if (steps == 2) {
        const dim3 grid((M + TJ - 4 - 1) / (TJ - 4), (N + TI - 4 - 1) / (TI - 4));
        hipLaunchKernelGGL((ref_temporal_kernel<2, TI, TJ>), grid, dim3(256), 0, (hipStream_t)stream, u0, u1, v0, v1, k,
                           N, M, (int64_t)M_padded);
    } else {
        return set_error(WFD_EINVAL, "wfd_ref_temporal: steps must be 2, 4 or 8");
    }
    return launched("wfd_ref_temporal");
}

int wfd_ref_mixed_step(const void* u0, const void* u1, void* u2, double wavespeed, double dt, double dx, double dy,
                       int32_t N, int32_t M, int32_t M_padded, wfd_stream_t stream)
{
    RefCoef k;
    int rc = ref_args(N, M, M_padded, wavespeed, dt, dx, dy, k);
    if (rc) return rc;
    if (!u0 || !u1 || !u2 || u1 == u2) return set_error(WFD_EINVAL, "wfd_ref_mixed_step: NULL or aliased u1/u2");
    if (M_padded % 2 || M_padded < M + 2)
        return set_error(WFD_EINVAL, "wfd_ref_mixed_step: M_padded must be even and >= M + 2");
    const dim3 grid((M / 2 + 63) / 64 + 1, (N + 3) / 4);
    hipLaunchKernelGGL(ref_mixed_kernel, grid, dim3(256), 0, (hipStream_t)stream, (const __half*)u0, (const __half*)u1,
                       (__half*)u2, k, N, M, (int64_t)M_padded);
    return launched("wfd_ref_mixed_step");
}

int wfd_fma_peak(float* out, int32_t blocks, int32_t n, wfd_stream_t stream)
{
    if (!out || blocks < 1 || n < 0) return set_error(WFD_EINVAL, "wfd_fma_peak: bad arguments");
    hipLaunchKernelGGL(fma_chain_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, out, 1e-9f, n);
    return launched("wfd_fma_peak");
}

int wfd_ref_scaled_step(const int16_t* q0, const float* s0, const int16_t* q1, const float* s1, int16_t* q2,
                        float* s2, double wavespeed, double dt, double dx, double dy, int32_t N, int32_t M,
                        int32_t M_padded, wfd_stream_t stream)
{
    RefCoef k;
    int rc = ref_args(N, M, M_padded, wavespeed, dt, dx, dy, k);
    if (rc) return rc;
    if (!q0 || !s0 || !q1 || !s1 || !q2 || !s2 || q1 == q2 || s1 == s2)
        return set_error(WFD_EINVAL, "wfd_ref_scaled_step: NULL or aliased fields");
    if (M_padded % 128) return set_error(WFD_EINVAL, "wfd_ref_scaled_step: M_padded must be a multiple of 128");
    const dim3 grid(M_padded / 128, (N + 3) / 4);
    hipLaunchKernelGGL(ref_scaled_kernel, grid, dim3(256), 0, (hipStream_t)stream, q0, s0, q1, s1, q2, s2, k, N, M,
                       (int64_t)M_padded);
    return launched("wfd_ref_scaled_step");
}

}  // extern "C"
