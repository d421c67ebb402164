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

// The lane's 4-node chunk of LDS tile row r starting at column c (16-byte aligned: rows of T3_LW
// elements, c a multiple of 4), read in one 16-byte access per 4 bytes of element.  Reading the
// chunk element by element instead puts the 64 lanes of a wave on 8 of the 32 LDS banks (lanes
// tx and tx + 8 of a row, and rows 2 apart, share banks), 4 conflicting passes per read.
template <typename T>
__device__ __forceinline__ V4<T> lds4(const T (*t)[T3_LW], int r, int c)
{
    if constexpr (sizeof(T) == 4) {
        const s3_v4f v = *reinterpret_cast<const s3_v4f*>(&t[r][c]);
        return {{v.x, v.y, v.z, v.w}};
    } else {
        const s3_v2d a = *reinterpret_cast<const s3_v2d*>(&t[r][c]);
        const s3_v2d b = *reinterpret_cast<const s3_v2d*>(&t[r][c + 2]);
        return {{a.x, a.y, b.x, b.y}};
    }
}

// value at column offset j in [-4, 8) of a chunk, from the chunks left of, at and right of it
template <typename T>
__device__ __forceinline__ T pick3(const V4<T>& L, const V4<T>& C, const V4<T>& R, int j)
{
    return j < 0 ? L.e[4 + j] : (j < 4 ? C.e[j] : R.e[j - 4]);
}

// Laplacian of the lane's 4 nodes of plane k: z-window planes k-K..k+K-1 in win[0..2K-1], plane
// k+K in top, in-plane neighbours from the staged tile t.  Per node the terms are summed in the
// order of wavefd.h (centre, marching axis, row axis, column axis, m = 1..K each).
template <typename T, int K>
__device__ __forceinline__ void lap3x4(const T wc, const T* wp, const T* wr, const T* wcol, const V4<T>* win,
                                       const V4<T>& top, const T (*t)[T3_LW], int ty, int tx, T (&lap)[4])
{
    const int lr = ty + HALO, c0 = HALO + tx * 4;
#pragma unroll
    for (int e = 0; e < 4; e++) lap[e] = wc * win[K].e[e];
#pragma unroll
    for (int m = 1; m <= K; m++)
#pragma unroll
        for (int e = 0; e < 4; e++) lap[e] = fma3(wp[m - 1], win[K - m].e[e] + (m < K ? win[K + m].e[e] : top.e[e]), lap[e]);
#pragma unroll
    for (int m = 1; m <= K; m++) {
        const V4<T> up = lds4(t, lr - m, c0), dn = lds4(t, lr + m, c0);
#pragma unroll
        for (int e = 0; e < 4; e++) lap[e] = fma3(wr[m - 1], up.e[e] + dn.e[e], lap[e]);
    }
    const V4<T> L = lds4(t, lr, c0 - 4), C = lds4(t, lr, c0), R = lds4(t, lr, c0 + 4);
#pragma unroll
    for (int m = 1; m <= K; m++)
#pragma unroll
        for (int e = 0; e < 4; e++) lap[e] = fma3(wcol[m - 1], pick3(L, C, R, e - m) + pick3(L, C, R, e + m), lap[e]);
}

// Per-block constants of a 3D tile and its lane's chunk.
template <typename T>
struct Tile3 {
    int64_t i0, j0, i, c, k0, k1, rowoff, ld, ps, sb;
    int shot, wave, lane, tx, ty;
    bool load_row, out, halo_thread;
    bool valid[4];
    T bc[4];
    T br;
    int64_t wr0, wr1;  // rows of this wave (source mask)
};

// Source bookkeeping of the march: which of the next 64 planes hold sources in this wave's rows.
struct SrcMask {
    uint64_t bits;
    int64_t k;
};

// Update of the lane's 4-node chunk of plane k from the z-window (`win`: planes k-K..k+K-1, `top`:
// plane k+K), the staged in-plane tile `t`, u_prev P and coefficient kv (the update documented in
// wavefd.h, in exactly that order), into val (zero where the node is not interior); the caller
// stores it.
template <typename T, int K, bool VARC, bool SPONGE, bool NT>
__device__ __forceinline__ void plane_update(const StepArgs<T>& a, const Tile3<T>& b, const V4<T>* win,
                                             const V4<T>& top, const T (*t)[T3_LW], const V4<T>& P, const V4<T>& kv, T bp, int64_t k,
                                             int64_t step, SrcMask& sm, T (&val)[4])
{
    T kc[4];
#pragma unroll
    for (int e = 0; e < 4; e++) kc[e] = VARC ? kv.e[e] : a.kconst;
    T bb[4] = {T(0), T(0), T(0), T(0)};
    bool damp = false;
    if constexpr (SPONGE) {
        const T bpr = bp + b.br;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            bb[e] = bpr + b.bc[e];
            damp |= bb[e] != T(0);
        }
    }
    T lap[4];
    lap3x4<T, K>(a.wc, a.wp, a.wr, a.wcol, win, top, t, b.ty, b.tx, lap);
#pragma unroll
    for (int e = 0; e < 4; e++) {
        const T ctr = win[K].e[e];
        const T tt = damp ? T(2) * ctr - (T(1) - bb[e]) * P.e[e] : T(2) * ctr - P.e[e];
        val[e] = fma3(kc[e], lap[e], tt);
    }
    // source pass: the wave holds 4 rows of the tile, 16 lanes each.  Which of the next 64 planes
    // have sources in these rows is found once per 64 planes (lane l checks plane k + l, one
    // ballot), so no per-plane load latency sits in the march.
    if (a.src.count && k >= sm.k + 64) {
        sm.k = k;
        const int64_t kk = k + b.lane;
        bool has = false;
        if (kk < b.k1) {
            const int32_t* rp = a.src.row_ptr + (int64_t)b.shot * (a.src.nrows + 1) + kk * a.nr;
            has = rp[b.wr0] != rp[b.wr1];
        }
        sm.bits = __ballot(has);
    }
    const bool src_here = a.src.count && ((sm.bits >> (k - sm.k)) & 1);
    if (src_here) {
        for (int rr = 0; rr < 4; rr++) {
            const int64_t ir = b.i0 + 4 * b.wave + rr;
            if (ir <= a.nr - 2)
                wave_inject(a.src, b.shot, k * a.nr + ir, b.j0 - 64 * rr, 16 * rr, 16 * rr + 15, step, b.lane, kc, val);
        }
    }
    if (damp) {
#pragma unroll
        for (int e = 0; e < 4; e++) val[e] = val[e] / (T(1) + bb[e]);
    }
#pragma unroll
    for (int e = 0; e < 4; e++)
        if (!b.valid[e]) val[e] = T(0);
}

// Forward launches: each iteration issues the loads of plane k + 1 (its u_prev, kc and sponge
// value, the z-window chunk of plane k + K + 1 and that plane's in-plane halo chunk), updates plane
// k from operands that have already landed, then stages plane k + 1's tile and meets the others at
// a barrier that orders LDS only -- so the look-ahead loads are in flight through the stencil,
// and no load is outstanding across the loop's back edge but the store.  The halo of plane
// k + K + 1 is loaded with the neighbouring tiles' loads of that plane's centre chunks (fresh in
// the L2) and kept in an LDS ring of K + 1 slots until the tile of that plane is staged from it;
// the ring is private to each thread (a thread reads back only the entries it wrote), and slot
// (k + K + 1) % (K + 1), written at iteration k, was read (for plane k) at iteration k - 1.
// Imaging launches march the same way with two staged fields (the adjoint field and the forward
// field whose Laplacian the imaging condition takes), each with its own z-window and halo ring,
// and also load the imaging partial of plane k + 1 ahead.
template <typename T, int K, bool VARC, bool SPONGE, bool IMG, bool NT>
__global__ __launch_bounds__(T3_THREADS, sizeof(T) == 4 ? (IMG ? 2 : 3) : 1) void step3d_kernel(const StepArgs<T> a, int rk,
                                                                                             int nkb)
{
    constexpr bool RING = !IMG;
    constexpr int NF = IMG ? 2 : 1;  // staged fields
    __shared__ __align__(16) T tile[NF][2][T3_LH][T3_LW];  // 16-byte chunk reads (lds4)
    __shared__ V4<T> ring[NF][K + 1][T3_HALO_THREADS];
    const StepBlock bp = step_block(a.ngb, a.gx, a.gy);
    const int64_t step = a.step_dev ? *a.step_dev + a.step : a.step;
    if (bp.gather) {
        if (threadIdx.x < 256) gather_points(a, bp.g, step);  // 256 points per gather block
        return;
    }
    Tile3<T> b;
    b.tx = threadIdx.x & 15;
    b.ty = threadIdx.x >> 4;
    b.lane = threadIdx.x & 63;
    b.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    b.shot = bp.z / nkb;
    const int kb = bp.z % nkb;
    b.j0 = (int64_t)bp.x * T3_TJ;
    b.c = b.j0 + b.tx * 4;
    b.i0 = 1 + (int64_t)bp.y * T3_TI;
    b.i = b.i0 + b.ty;
    b.k0 = 1 + (int64_t)kb * rk;
    b.k1 = min(b.k0 + rk, a.np - 1);
    b.load_row = b.i <= a.nr + 3;
    b.out = (b.i <= a.nr - 2) && (b.c < a.nc);
    b.halo_thread = threadIdx.x < T3_HALO_THREADS;
    b.ld = a.ld;
    b.ps = a.ps;
    b.rowoff = b.i * b.ld + b.c;
    b.sb = (int64_t)b.shot * a.shot_stride;
#pragma unroll
    for (int e = 0; e < 4; e++) {
        b.valid[e] = b.out && (b.c + e >= 1) && (b.c + e <= a.nc - 2);
        b.bc[e] = T(0);
        if constexpr (SPONGE) b.bc[e] = (b.c + e < a.nc) ? a.beta_c[b.c + e] : T(0);
    }
    b.br = T(0);
    if constexpr (SPONGE) b.br = (b.i < a.nr) ? a.beta_r[b.i] : T(0);
    b.wr0 = min(b.i0 + 4 * b.wave, a.nr - 1);
    b.wr1 = min(b.i0 + 4 * b.wave + 4, a.nr - 1);

    const T* __restrict__ uc = a.uc + b.sb;
    const T* up = a.up + b.sb;
    T* un = a.un + b.sb;
    const T* __restrict__ coef = VARC ? a.coef + (int64_t)b.shot * a.coef_shot_stride : nullptr;
    const int64_t k0 = b.k0, k1 = b.k1, ps = b.ps, rowoff = b.rowoff;
    const V4<T> zero = {{T(0), T(0), T(0), T(0)}};
    SrcMask sm{0, k0 - 64};

    V4<T> win[2 * K + 1];
#pragma unroll
    for (int m = 0; m < 2 * K; m++) win[m] = b.load_row ? gld4(uc + (k0 - K + m) * ps + rowoff) : zero;

    if constexpr (RING) {
        V4<T> wc = b.load_row ? gld4(uc + (k0 + K) * ps + rowoff) : zero;  // plane k + K
        V4<T> P = b.out ? gld4s<NT>(up + k0 * ps + rowoff) : zero;
        V4<T> kv = zero;
        if constexpr (VARC) kv = b.out ? gld4s<NT>(coef + k0 * ps + rowoff) : zero;
        T bpk = T(0);
        if constexpr (SPONGE) bpk = a.beta_p[k0];
        if (b.halo_thread) {  // halos of planes k0 .. k0 + K
            for (int m = 0; m <= K; m++)
                if (k0 + m < k1) ring[0][(k0 + m) % (K + 1)][threadIdx.x] = halo_fetch(uc, k0 + m, b.i0, b.j0, a.nr, b.ld, ps);
        }
#pragma unroll
        for (int e = 0; e < 4; e++) tile[0][0][b.ty + HALO][HALO + b.tx * 4 + e] = win[K].e[e];
        if (b.halo_thread) halo_put(tile[0][0], ring[0][k0 % (K + 1)][threadIdx.x]);
        __syncthreads();
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        int buf = 0;
        for (int64_t k = k0; k < k1; k++, buf ^= 1) {
            // plane k + 1 <= k1 <= np - 1 and plane k + K + 1 <= np - 1 + K lie in the layout
            // (K <= 4 zero halo planes), so the last plane's look-ahead loads stay in bounds
            const V4<T> wn = b.load_row ? gld4(uc + (k + K + 1) * ps + rowoff) : zero;
            const V4<T> Pn = b.out ? gld4s<NT>(up + (k + 1) * ps + rowoff) : zero;
            V4<T> kn = zero;
            if constexpr (VARC) kn = b.out ? gld4s<NT>(coef + (k + 1) * ps + rowoff) : zero;
            T bpn = T(0);
            if constexpr (SPONGE) bpn = a.beta_p[k + 1];
            const V4<T> hn = b.halo_thread && k + K + 1 < k1 ? halo_fetch(uc, k + K + 1, b.i0, b.j0, a.nr, b.ld, ps) : zero;
            T val[4];
            plane_update<T, K, VARC, SPONGE, NT>(a, b, win, wc, tile[0][buf], P, kv, bpk, k, step, sm, val);
            // Take over the look-ahead operands before storing plane k: a wait for a load issued
            // before an outstanding store also waits for the store.  The explicit wait lets the
            // compiler see every look-ahead register as landed on every path to the back edge.
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
#pragma unroll
            for (int m = 0; m < 2 * K - 1; m++) win[m] = win[m + 1];
            win[2 * K - 1] = wc;
            wc = wn;
            P = Pn;
            kv = kn;
            bpk = bpn;
            if (b.halo_thread && k + K + 1 < k1) ring[0][(k + K + 1) % (K + 1)][threadIdx.x] = hn;
            if (b.out) gst4s<NT>(un + k * ps + rowoff, val);
            if (k + 1 < k1) {
#pragma unroll
                for (int e = 0; e < 4; e++) tile[0][buf ^ 1][b.ty + HALO][HALO + b.tx * 4 + e] = win[K].e[e];
                if (b.halo_thread) halo_put(tile[0][buf ^ 1], ring[0][(k + 1) % (K + 1)][threadIdx.x]);
                lds_barrier();
            }
        }
    } else {
        const T* __restrict__ uf = a.uf + b.sb;
        T* gp = a.g + b.sb;
        V4<T> wf[2 * K + 1];
#pragma unroll
        for (int m = 0; m < 2 * K; m++) wf[m] = b.load_row ? gld4(uf + (k0 - K + m) * ps + rowoff) : zero;
        V4<T> wc = b.load_row ? gld4(uc + (k0 + K) * ps + rowoff) : zero;   // plane k + K of both fields
        V4<T> wcf = b.load_row ? gld4(uf + (k0 + K) * ps + rowoff) : zero;
        V4<T> P = b.out ? gld4s<NT>(up + k0 * ps + rowoff) : zero;
        V4<T> kv = zero;
        if constexpr (VARC) kv = b.out ? gld4s<NT>(coef + k0 * ps + rowoff) : zero;
        V4<T> G = b.out ? gld4(gp + k0 * ps + rowoff) : zero;
        T bpk = T(0);
        if constexpr (SPONGE) bpk = a.beta_p[k0];
        if (b.halo_thread) {  // halos of planes k0 .. k0 + K
            for (int m = 0; m <= K; m++)
                if (k0 + m < k1) {
                    ring[0][(k0 + m) % (K + 1)][threadIdx.x] = halo_fetch(uc, k0 + m, b.i0, b.j0, a.nr, b.ld, ps);
                    ring[NF - 1][(k0 + m) % (K + 1)][threadIdx.x] = halo_fetch(uf, k0 + m, b.i0, b.j0, a.nr, b.ld, ps);
                }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) {
            tile[0][0][b.ty + HALO][HALO + b.tx * 4 + e] = win[K].e[e];
            tile[NF - 1][0][b.ty + HALO][HALO + b.tx * 4 + e] = wf[K].e[e];
        }
        if (b.halo_thread) {
            halo_put(tile[0][0], ring[0][k0 % (K + 1)][threadIdx.x]);
            halo_put(tile[NF - 1][0], ring[NF - 1][k0 % (K + 1)][threadIdx.x]);
        }
        __syncthreads();
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        int buf = 0;
        for (int64_t k = k0; k < k1; k++, buf ^= 1) {
            // the look-ahead loads of plane k + 1 and of plane k + K + 1 stay in the layout (as in
            // the forward march); the partial of plane k + 1 is not the one stored below (plane k)
            const V4<T> wn = b.load_row ? gld4(uc + (k + K + 1) * ps + rowoff) : zero;
            const V4<T> wfn = b.load_row ? gld4(uf + (k + K + 1) * ps + rowoff) : zero;
            const V4<T> Pn = b.out ? gld4s<NT>(up + (k + 1) * ps + rowoff) : zero;
            V4<T> kn = zero;
            if constexpr (VARC) kn = b.out ? gld4s<NT>(coef + (k + 1) * ps + rowoff) : zero;
            const V4<T> Gn = b.out ? gld4(gp + (k + 1) * ps + rowoff) : zero;
            T bpn = T(0);
            if constexpr (SPONGE) bpn = a.beta_p[k + 1];
            const bool hk = b.halo_thread && k + K + 1 < k1;
            const V4<T> hn = hk ? halo_fetch(uc, k + K + 1, b.i0, b.j0, a.nr, b.ld, ps) : zero;
            const V4<T> hfn = hk ? halo_fetch(uf, k + K + 1, b.i0, b.j0, a.nr, b.ld, ps) : zero;
            T val[4];
            plane_update<T, K, VARC, SPONGE, NT>(a, b, win, wc, tile[0][buf], P, kv, bpk, k, step, sm, val);
            T gv[4];
            lap3x4<T, K>(a.wc, a.wp, a.wr, a.wcol, wf, wcf, tile[NF - 1][buf], b.ty, b.tx, gv);
#pragma unroll
            for (int e = 0; e < 4; e++) gv[e] = b.valid[e] ? fma3(val[e], gv[e], G.e[e]) : T(0);
            // take over the look-ahead operands before the stores (see the forward march)
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
#pragma unroll
            for (int m = 0; m < 2 * K - 1; m++) {
                win[m] = win[m + 1];
                wf[m] = wf[m + 1];
            }
            win[2 * K - 1] = wc;
            wf[2 * K - 1] = wcf;
            wc = wn;
            wcf = wfn;
            P = Pn;
            kv = kn;
            G = Gn;
            bpk = bpn;
            if (hk) {
                ring[0][(k + K + 1) % (K + 1)][threadIdx.x] = hn;
                ring[NF - 1][(k + K + 1) % (K + 1)][threadIdx.x] = hfn;
            }
            const int64_t o = k * ps + rowoff;
            if (b.out) {
                gst4s<NT>(un + o, val);
                gst4(gp + o, gv);
            }
            if (k + 1 < k1) {
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    tile[0][buf ^ 1][b.ty + HALO][HALO + b.tx * 4 + e] = win[K].e[e];
                    tile[NF - 1][buf ^ 1][b.ty + HALO][HALO + b.tx * 4 + e] = wf[K].e[e];
                }
                if (b.halo_thread) {
                    halo_put(tile[0][buf ^ 1], ring[0][(k + 1) % (K + 1)][threadIdx.x]);
                    halo_put(tile[NF - 1][buf ^ 1], ring[NF - 1][(k + 1) % (K + 1)][threadIdx.x]);
                }
                lds_barrier();
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
