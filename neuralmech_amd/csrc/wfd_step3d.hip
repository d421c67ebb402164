// 3D leapfrog step kernel for gfx950, plus the small utility kernels (gradient flush, medium
// check, device step counter).
//
// 2.5D scheme: a block of 256 threads (4 waves) owns a 16-row x 64-column tile of the two fast
// axes and marches along the slowest axis over RK planes.  Each lane owns a 4-node chunk of
// one row; the 2K+1 planes of its own chunk that the stencil needs along the marching axis
// stay in registers (one new plane loaded per step).  The centre plane of the tile, with a
// 4-node halo on the four in-plane sides, is staged in LDS (double-buffered, one barrier per
// plane) and supplies the in-plane neighbours.  Arithmetic follows wavefd.h exactly.
#include "wfd_internal.h"
#include "wfd_points_dev.h"

namespace wfd {

template <typename T>
struct V4 {
    T e[4];
};

template <typename T>
__device__ __forceinline__ V4<T> gld4(const T* p);
template <>
__device__ __forceinline__ V4<float> gld4<float>(const float* p)
{
    const float4 v = *reinterpret_cast<const float4*>(p);
    return {{v.x, v.y, v.z, v.w}};
}
template <>
__device__ __forceinline__ V4<double> gld4<double>(const double* p)
{
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    return {{a.x, a.y, b.x, b.y}};
}
template <typename T>
__device__ __forceinline__ void gst4(T* p, const T (&v)[4]);
template <>
__device__ __forceinline__ void gst4<float>(float* p, const float (&v)[4])
{
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void gst4<double>(double* p, const double (&v)[4])
{
    *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
}

// Streaming access to the fields a step touches once (u_prev, kc, u_next): with NT the loads and
// stores carry the nontemporal hint so they do not displace the planes of u_cur that neighbouring
// tiles still read.  It pays off when the fields exceed the Infinity Cache and costs when they fit,
// so it is a launch option.
typedef float s3_v4f __attribute__((ext_vector_type(4)));
typedef double s3_v2d __attribute__((ext_vector_type(2)));
template <bool NT, typename T>
__device__ __forceinline__ V4<T> gld4s(const T* p)
{
    if constexpr (NT && sizeof(T) == 4) {
        const s3_v4f v = __builtin_nontemporal_load(reinterpret_cast<const s3_v4f*>(p));
        return {{v.x, v.y, v.z, v.w}};
    } else if constexpr (NT) {
        const s3_v2d a = __builtin_nontemporal_load(reinterpret_cast<const s3_v2d*>(p));
        const s3_v2d b = __builtin_nontemporal_load(reinterpret_cast<const s3_v2d*>(p + 2));
        return {{a.x, a.y, b.x, b.y}};
    } else {
        return gld4<T>(p);
    }
}
template <bool NT, typename T>
__device__ __forceinline__ void gst4s(T* p, const T (&v)[4])
{
    if constexpr (NT && sizeof(T) == 4) {
        const s3_v4f w = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(w, reinterpret_cast<s3_v4f*>(p));
    } else if constexpr (NT) {
        const s3_v2d a = {v[0], v[1]}, b = {v[2], v[3]};
        __builtin_nontemporal_store(a, reinterpret_cast<s3_v2d*>(p));
        __builtin_nontemporal_store(b, reinterpret_cast<s3_v2d*>(p + 2));
    } else {
        gst4<T>(p, v);
    }
}

__device__ __forceinline__ float fma3(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma3(double a, double b, double c) { return __builtin_fma(a, b, c); }

constexpr int T3_TI = 16;  // tile rows: a block has 16 lanes per row (4-node chunks), 4 rows per wave
constexpr int T3_THREADS = 16 * T3_TI;
constexpr int T3_TJ = 64;  // tile columns
constexpr int T3_LH = T3_TI + 2 * HALO;
constexpr int T3_LW = T3_TJ + 2 * HALO;

// Stage the centre plane chunk (already in registers) and the in-plane halo of plane `k`
// of field `u` into LDS tile `t`.
template <typename T>
__device__ __forceinline__ void stage_plane(T (*t)[T3_LW], const T* u, const V4<T>& centre, int64_t k,
                                            int64_t i0, int64_t j0, int ty, int tx, int64_t nr,
                                            int64_t ld, int64_t ps)
{
#pragma unroll
    for (int e = 0; e < 4; e++) t[ty + HALO][HALO + tx * 4 + e] = centre.e[e];
    const int tid = threadIdx.x;
    if (tid < 128) {
        const int hr = tid >> 4, hc = tid & 15;
        const int64_t row = hr < 4 ? i0 - 4 + hr : i0 + T3_TI + (hr - 4);
        const int lr = hr < 4 ? hr : T3_TI + HALO + (hr - 4);
        V4<T> v = {{T(0), T(0), T(0), T(0)}};
        if (row <= nr + 3) v = gld4(u + k * ps + row * ld + j0 + hc * 4);
#pragma unroll
        for (int e = 0; e < 4; e++) t[lr][HALO + hc * 4 + e] = v.e[e];
    } else if (tid < 128 + 2 * T3_TI) {
        const int s = tid - 128, side = s / T3_TI, rr = s % T3_TI;
        const int64_t row = i0 + rr;
        V4<T> v = {{T(0), T(0), T(0), T(0)}};
        if (row <= nr + 3) v = gld4(u + k * ps + row * ld + (side ? j0 + T3_TJ : j0 - 4));
        const int lc = side ? HALO + T3_TJ : 0;
#pragma unroll
        for (int e = 0; e < 4; e++) t[HALO + rr][lc + e] = v.e[e];
    }
}

// The in-plane halo chunk this thread loads for plane k of field u (threads 0..127: the 4 rows above
// and below the tile, 16 chunks each; threads 128..159: the 4-column chunks left and right of each
// tile row; zero for the others and for rows past the layout), and where it goes in the LDS tile.
constexpr int T3_HALO_THREADS = 128 + 2 * T3_TI;
template <typename T>
__device__ __forceinline__ V4<T> halo_fetch(const T* u, int64_t k, int64_t i0, int64_t j0, int64_t nr, int64_t ld,
                                            int64_t ps)
{
    const int tid = threadIdx.x;
    V4<T> v = {{T(0), T(0), T(0), T(0)}};
    if (tid < 128) {
        const int hr = tid >> 4, hc = tid & 15;
        const int64_t row = hr < 4 ? i0 - 4 + hr : i0 + T3_TI + (hr - 4);
        if (row <= nr + 3) v = gld4(u + k * ps + row * ld + j0 + hc * 4);
    } else if (tid < T3_HALO_THREADS) {
        const int s = tid - 128, side = s / T3_TI, rr = s % T3_TI;
        const int64_t row = i0 + rr;
        if (row <= nr + 3) v = gld4(u + k * ps + row * ld + (side ? j0 + T3_TJ : j0 - 4));
    }
    return v;
}
template <typename T>
__device__ __forceinline__ void halo_put(T (*t)[T3_LW], const V4<T>& v)
{
    const int tid = threadIdx.x;
    if (tid < 128) {
        const int hr = tid >> 4, hc = tid & 15;
        const int lr = hr < 4 ? hr : T3_TI + HALO + (hr - 4);
#pragma unroll
        for (int e = 0; e < 4; e++) t[lr][HALO + hc * 4 + e] = v.e[e];
    } else if (tid < T3_HALO_THREADS) {
        const int s = tid - 128, side = s / T3_TI, rr = s % T3_TI;
        const int lc = side ? HALO + T3_TJ : 0;
#pragma unroll
        for (int e = 0; e < 4; e++) t[HALO + rr][lc + e] = v.e[e];
    }
}

template <typename T, int K>
__device__ __forceinline__ T lap3(const T wc, const T* wp, const T* wr, const T* wcol, const V4<T>* win,
                                  const T (*t)[T3_LW], int ty, int tx, int e)
{
    const int lr = ty + HALO, lc = HALO + tx * 4 + e;
    T lap = wc * win[K].e[e];
#pragma unroll
    for (int m = 1; m <= K; m++) lap = fma3(wp[m - 1], win[K - m].e[e] + win[K + m].e[e], lap);
#pragma unroll
    for (int m = 1; m <= K; m++) lap = fma3(wr[m - 1], t[lr - m][lc] + t[lr + m][lc], lap);
#pragma unroll
    for (int m = 1; m <= K; m++) lap = fma3(wcol[m - 1], t[lr][lc - m] + t[lr][lc + m], lap);
    return lap;
}

// Forward launches load the in-plane halo of plane k + K -- the plane whose centre chunks the
// neighbouring tiles load into their z-windows in the same iteration, so the lines are fresh in
// the L2 -- and keep it K iterations in an LDS ring of K + 1 halo slots, from which the padded
// tile of plane k is completed.  Written at iteration q - K and read at iteration q (both before
// that iteration's barrier), a slot's previous plane q - K - 1 was read one barrier earlier.
// Imaging launches (two staged fields, more registers) load the halo of plane k directly.
template <typename T, int K, bool VARC, bool SPONGE, bool IMG, bool NT>
__global__ __launch_bounds__(T3_THREADS) void step3d_kernel(const StepArgs<T> a, int rk, int nkb)
{
    constexpr bool RING = !IMG;
    __shared__ T tile[IMG ? 2 : 1][2][T3_LH][T3_LW];
    __shared__ V4<T> ring[RING ? K + 1 : 1][RING ? T3_HALO_THREADS : 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const StepBlock bp = step_block(a.ngb, a.gx, a.gy);
    const int64_t step = a.step_dev ? *a.step_dev + a.step : a.step;
    if (bp.gather) {
        if (threadIdx.x < 256) gather_points(a, bp.g, step);  // 256 points per gather block
        return;
    }
    const int shot = bp.z / nkb, kb = bp.z % nkb;
    const int64_t j0 = (int64_t)bp.x * T3_TJ;
    const int64_t c = j0 + tx * 4;
    const int64_t i0 = 1 + (int64_t)bp.y * T3_TI;
    const int64_t i = i0 + ty;
    const int64_t k0 = 1 + (int64_t)kb * rk;
    const int64_t k1 = min(k0 + rk, a.np - 1);
    const bool load_row = i <= a.nr + 3;
    const bool out = (i <= a.nr - 2) && (c < a.nc);

    const int64_t sb = (int64_t)shot * a.shot_stride;
    const T* __restrict__ uc = a.uc + sb;
    const T* up = a.up + sb;
    T* un = a.un + sb;
    const T* __restrict__ coef = VARC ? a.coef + (int64_t)shot * a.coef_shot_stride : nullptr;
    const T* __restrict__ uf = IMG ? a.uf + sb : nullptr;
    T* gp = IMG ? a.g + sb : nullptr;
    const int64_t ld = a.ld, ps = a.ps;
    const int64_t rowoff = i * ld + c;
    const V4<T> zero = {{T(0), T(0), T(0), T(0)}};

    V4<T> win[2 * K + 1];
    V4<T> wf[IMG ? 2 * K + 1 : 1];
#pragma unroll
    for (int t = 1; t <= 2 * K; t++) {
        const int64_t k = k0 - K - 1 + t;
        win[t] = load_row ? gld4(uc + k * ps + rowoff) : zero;
        if constexpr (IMG) wf[t] = load_row ? gld4(uf + k * ps + rowoff) : zero;
    }

    bool valid[4];
    T bc[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        valid[e] = out && (c + e >= 1) && (c + e <= a.nc - 2);
        bc[e] = T(0);
        if constexpr (SPONGE) bc[e] = (c + e < a.nc) ? a.beta_c[c + e] : T(0);
    }
    T br = T(0);
    if constexpr (SPONGE) br = (i < a.nr) ? a.beta_r[i] : T(0);

    const int64_t wr0 = min(i0 + 4 * wave, a.nr - 1), wr1 = min(i0 + 4 * wave + 4, a.nr - 1);
    uint64_t src_mask = 0;
    int64_t mask_k = k0 - 64;
    int buf = 0;
    if constexpr (RING) {  // halos of the first K planes
        if (threadIdx.x < T3_HALO_THREADS)
            for (int t = 0; t < K; t++)
                if (k0 + t < k1) ring[(k0 + t) % (K + 1)][threadIdx.x] = halo_fetch(uc, k0 + t, i0, j0, a.nr, ld, ps);
        __syncthreads();
    }
    for (int64_t k = k0; k < k1; k++, buf ^= 1) {
#pragma unroll
        for (int t = 0; t < 2 * K; t++) win[t] = win[t + 1];
        win[2 * K] = load_row ? gld4(uc + (k + K) * ps + rowoff) : zero;
        if constexpr (IMG) {
#pragma unroll
            for (int t = 0; t < 2 * K; t++) wf[t] = wf[t + 1];
            wf[2 * K] = load_row ? gld4(uf + (k + K) * ps + rowoff) : zero;
        }
        if constexpr (RING) {
#pragma unroll
            for (int e = 0; e < 4; e++) tile[0][buf][ty + HALO][HALO + tx * 4 + e] = win[K].e[e];
            if (threadIdx.x < T3_HALO_THREADS) {
                const V4<T> next = k + K < k1 ? halo_fetch(uc, k + K, i0, j0, a.nr, ld, ps) : zero;
                halo_put(tile[0][buf], ring[k % (K + 1)][threadIdx.x]);
                if (k + K < k1) ring[(k + K) % (K + 1)][threadIdx.x] = next;
            }
        } else {
            stage_plane<T>(tile[0][buf], uc, win[K], k, i0, j0, ty, tx, a.nr, ld, ps);
            if constexpr (IMG) stage_plane<T>(tile[IMG ? 1 : 0][buf], uf, wf[K], k, i0, j0, ty, tx, a.nr, ld, ps);
        }
        __syncthreads();
        const int64_t o = k * ps + rowoff;
        const V4<T> P = out ? gld4s<NT>(up + o) : zero;
        T kc[4];
        if constexpr (VARC) {
            const V4<T> kv = out ? gld4s<NT>(coef + o) : zero;
#pragma unroll
            for (int e = 0; e < 4; e++) kc[e] = kv.e[e];
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) kc[e] = a.kconst;
        }
        T b[4] = {T(0), T(0), T(0), T(0)};
        bool damp = false;
        if constexpr (SPONGE) {
            const T bpr = a.beta_p[k] + br;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b[e] = bpr + bc[e];
                damp |= b[e] != T(0);
            }
        }
        T val[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T ctr = win[K].e[e];
            const T lap = lap3<T, K>(a.wc, a.wp, a.wr, a.wcol, win, tile[0][buf], ty, tx, e);
            const T t = damp ? T(2) * ctr - (T(1) - b[e]) * P.e[e] : T(2) * ctr - P.e[e];
            val[e] = fma3(kc[e], lap, t);
        }
        // source pass: the wave holds 4 rows of the tile, 16 lanes each.  Which of the next 64
        // planes have sources in these rows is found once per 64 planes (lane l checks plane
        // k + l, one ballot), so no per-plane load latency sits in the loop.
        if (a.src.count && k >= mask_k + 64) {
            mask_k = k;
            const int64_t kk = k + lane;
            bool has = false;
            if (kk < k1) {
                const int32_t* rp = a.src.row_ptr + (int64_t)shot * (a.src.nrows + 1) + kk * a.nr;
                has = rp[wr0] != rp[wr1];
            }
            src_mask = __ballot(has);
        }
        const bool src_here = a.src.count && ((src_mask >> (k - mask_k)) & 1);
        if (src_here) {
            for (int rr = 0; rr < 4; rr++) {
                const int64_t ir = i0 + 4 * wave + rr;
                if (ir <= a.nr - 2)
                    wave_inject(a.src, shot, k * a.nr + ir, j0 - 64 * rr, 16 * rr, 16 * rr + 15, step, lane, kc, val);
            }
        }
        if (damp) {
#pragma unroll
            for (int e = 0; e < 4; e++) val[e] = val[e] / (T(1) + b[e]);
        }
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (!valid[e]) val[e] = T(0);
        if (out) gst4s<NT>(un + o, val);
        if constexpr (IMG) {
            T gv[4];
#pragma unroll
            for (int e = 0; e < 4; e++) gv[e] = lap3<T, K>(a.wc, a.wp, a.wr, a.wcol, wf, tile[IMG ? 1 : 0][buf], ty, tx, e);
            if (out) {
                const V4<T> G = gld4(gp + o);
#pragma unroll
                for (int e = 0; e < 4; e++) gv[e] = valid[e] ? fma3(val[e], gv[e], G.e[e]) : T(0);
                gst4(gp + o, gv);
            }
        }
    }
}

template <typename T, int K, bool VARC, bool SPONGE, bool IMG, bool NT>
static hipError_t launch3d_kn(const StepArgs<T>& a, int rk, int nshots, hipStream_t s)
{
    StepArgs<T> b = a;
    b.gx = (uint32_t)((a.nc + T3_TJ - 1) / T3_TJ);
    b.gy = (uint32_t)((a.nr - 2 + T3_TI - 1) / T3_TI);
    const int64_t nkb = (a.np - 2 + rk - 1) / rk;
    // 1D launch: b.ngb gather blocks, then gx * gy * (nkb * nshots) tiles
    const int64_t nblocks = step_grid((uint32_t)((int64_t)b.gx * b.gy * nkb * nshots), b.ngb);
    hipLaunchKernelGGL((step3d_kernel<T, K, VARC, SPONGE, IMG, NT>), dim3((unsigned)nblocks), dim3(T3_THREADS), 0, s, b, rk,
                       (int)nkb);
    return hipGetLastError();
}

template <typename T, int K, bool VARC, bool SPONGE, bool IMG>
static hipError_t launch3d_k(const StepArgs<T>& a, int rk, int nshots, hipStream_t s, bool nt)
{
    if (nt) return launch3d_kn<T, K, VARC, SPONGE, IMG, true>(a, rk, nshots, s);
    return launch3d_kn<T, K, VARC, SPONGE, IMG, false>(a, rk, nshots, s);
}

template <typename T, int K>
static hipError_t launch3d_flags(const StepArgs<T>& a, bool varc, bool sponge, bool img, int rk, bool nt, int nshots,
                                 hipStream_t s)
{
    const int f = (varc ? 4 : 0) | (sponge ? 2 : 0) | (img ? 1 : 0);
    switch (f) {
        case 0: return launch3d_k<T, K, false, false, false>(a, rk, nshots, s, nt);
        case 1: return launch3d_k<T, K, false, false, true>(a, rk, nshots, s, nt);
        case 2: return launch3d_k<T, K, false, true, false>(a, rk, nshots, s, nt);
        case 3: return launch3d_k<T, K, false, true, true>(a, rk, nshots, s, nt);
        case 4: return launch3d_k<T, K, true, false, false>(a, rk, nshots, s, nt);
        case 5: return launch3d_k<T, K, true, false, true>(a, rk, nshots, s, nt);
        case 6: return launch3d_k<T, K, true, true, false>(a, rk, nshots, s, nt);
        default: return launch3d_k<T, K, true, true, true>(a, rk, nshots, s, nt);
    }
}

template <typename T>
hipError_t launch_step3d(const StepArgs<T>& a, int order, bool varc, bool sponge, bool img, int rk, bool nt, int nshots,
                         hipStream_t s)
{
    switch (order) {
        case 2: return launch3d_flags<T, 1>(a, varc, sponge, img, rk, nt, nshots, s);
        case 4: return launch3d_flags<T, 2>(a, varc, sponge, img, rk, nt, nshots, s);
        case 6: return launch3d_flags<T, 3>(a, varc, sponge, img, rk, nt, nshots, s);
        default: return launch3d_flags<T, 4>(a, varc, sponge, img, rk, nt, nshots, s);
    }
}

template hipError_t launch_step3d<float>(const StepArgs<float>&, int, bool, bool, bool, int, bool, int, hipStream_t);
template hipError_t launch_step3d<double>(const StepArgs<double>&, int, bool, bool, bool, int, bool, int, hipStream_t);

// ------------------------------------------------------------------------- small kernels
// Gather-only launch: copy a field at the receiver nodes into one trace row.
template <typename T>
__global__ __launch_bounds__(256) void gather_kernel(const StepArgs<T> a)
{
    gather_points(a, blockIdx.x, a.step);
}

template <typename T>
hipError_t launch_gather(const StepArgs<T>& a, hipStream_t s)
{
    if (a.ngb == 0) return hipSuccess;
    hipLaunchKernelGGL((gather_kernel<T>), dim3(a.ngb), dim3(256), 0, s, a);
    return hipGetLastError();
}
template hipError_t launch_gather<float>(const StepArgs<float>&, hipStream_t);
template hipError_t launch_gather<double>(const StepArgs<double>&, hipStream_t);

template <typename T>
__global__ void flush_kernel(T* g, double* G, int64_t elems, int nshots)
{
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < elems;
         x += (int64_t)gridDim.x * blockDim.x) {
        double acc = G[x];
        for (int s = 0; s < nshots; s++) {
            acc += (double)g[(int64_t)s * elems + x];
            g[(int64_t)s * elems + x] = T(0);
        }
        G[x] = acc;
    }
}

template <typename T>
hipError_t launch_flush(T* g, double* G, int64_t elems, int nshots, hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>((elems + 255) / 256, 8192);
    hipLaunchKernelGGL((flush_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, g, G, elems, nshots);
    return hipGetLastError();
}
template hipError_t launch_flush<float>(float*, double*, int64_t, int, hipStream_t);
template hipError_t launch_flush<double>(double*, double*, int64_t, int, hipStream_t);

// flag |= 1 if a medium coefficient at an interior node is NaN/Inf or not > 0, flag |= 2 if one
// exceeds kmax (the CFL limit; kmax <= 0 checks nothing)
template <typename T>
__global__ void check_coef_kernel(const T* origin, int64_t np, int64_t nr, int64_t nc, int64_t ld, int64_t ps,
                                  T kmax, int32_t* flag)
{
    const int64_t n = np * nr * nc;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = x % nc, r = x / nc, i = r % nr, k = r / nr;
        const bool inner = (j >= 1 && j <= nc - 2) && (nr == 1 || (i >= 1 && i <= nr - 2)) &&
                           (np == 1 || (k >= 1 && k <= np - 2));
        if (!inner) continue;
        const T y = origin[k * ps + i * ld + j];
        if (!(y > T(0)) || y - y != T(0))
            atomicOr(flag, 1);
        else if (kmax > T(0) && y > kmax)
            atomicOr(flag, 2);
    }
}

// flag |= 1 if any of the n damping entries is NaN/Inf or negative
template <typename T>
__global__ void check_beta_kernel(const T* v, int64_t n, int32_t* flag)
{
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const T y = v[x];
        if (!(y >= T(0)) || y - y != T(0)) atomicOr(flag, 1);
    }
}

template <typename T>
hipError_t launch_check(const T* origin, int64_t np, int64_t nr, int64_t nc, int64_t ld, int64_t ps, T kmax,
                        const T* beta, int64_t nbeta, int32_t* flag, hipStream_t s)
{
    if (origin) {
        const int64_t n = np * nr * nc;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL((check_coef_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, origin, np, nr, nc, ld,
                           ps, kmax, flag);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (beta && nbeta > 0) {
        const int64_t blocks = std::min<int64_t>((nbeta + 255) / 256, 256);
        hipLaunchKernelGGL((check_beta_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s, beta, nbeta, flag);
        return hipGetLastError();
    }
    return hipSuccess;
}
template hipError_t launch_check<float>(const float*, int64_t, int64_t, int64_t, int64_t, int64_t, float,
                                        const float*, int64_t, int32_t*, hipStream_t);
template hipError_t launch_check<double>(const double*, int64_t, int64_t, int64_t, int64_t, int64_t, double,
                                         const double*, int64_t, int32_t*, hipStream_t);

__global__ void add_i64_kernel(int64_t* p, int64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t s)
{
    hipLaunchKernelGGL(add_i64_kernel, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

__global__ void set_i64_kernel(int64_t* p, int64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

hipError_t launch_set_i64(int64_t* p, int64_t v, hipStream_t s)
{
    hipLaunchKernelGGL(set_i64_kernel, dim3(1), dim3(64), 0, s, p, v);
    return hipGetLastError();
}

// fill n elements of type T with zero (a kernel, so that it may sit inside a captured graph)
template <typename T>
__global__ void zero_kernel(T* p, int64_t n)
{
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x)
        p[x] = T(0);
}

hipError_t launch_zero(void* p, int64_t bytes, hipStream_t s)
{
    if (bytes <= 0) return hipSuccess;
    if (bytes % 16 == 0 && (uintptr_t)p % 16 == 0) {
        const int64_t n = bytes / 16;
        const int64_t blocks = std::min<int64_t>((n + 255) / 256, 16384);
        hipLaunchKernelGGL(zero_kernel<float4>, dim3((unsigned)blocks), dim3(256), 0, s, (float4*)p, n);
    } else {
        const int64_t blocks = std::min<int64_t>((bytes + 255) / 256, 16384);
        hipLaunchKernelGGL(zero_kernel<unsigned char>, dim3((unsigned)blocks), dim3(256), 0, s, (unsigned char*)p,
                           bytes);
    }
    return hipGetLastError();
}

}  // namespace wfd
