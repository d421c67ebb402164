// Two leapfrog steps per launch (3D), bit-identical to two single steps.
//
// A block of 768 threads (12 waves) owns an output tile of 40 rows x 56 columns of the two fast
// axes and marches along the slowest axis over RK planes.  Step 1 (u^n -> u^{n+1}) is evaluated
// on the tile extended by 4 nodes on every in-plane side -- 48 rows x 64 columns, one 4-node chunk
// per lane -- and K planes before and after the block's planes; step 2 (u^{n+1} -> u^{n+2}) on the
// output tile, K planes behind step 1.  Per lane:
//   * a window of the 2K + 1 planes of u^n of its chunk (one plane loaded per iteration), the
//     centre plane staged with a 4-node in-plane halo in LDS (double-buffered) for the in-plane
//     neighbours of step 1;
//   * a window of the 2K + 1 planes of u^{n+1} of its chunk (its own step-1 results), the
//     in-plane neighbours of step 2 read from an LDS ring of K + 1 step-1 planes.
// One barrier per plane.  Step-1 values are zero outside the interior, exactly as a single step
// stores them, so step 2 is exact on the output tile.  Per output node the launch reads u^{n-1},
// u^n and kc once from HBM (the overlap of neighbouring tiles, 1.37x, is served by the caches) and
// writes u^{n+1} and u^{n+2}: 20 bytes per two steps instead of 32.  The four fields are distinct.
// The price is arithmetic: 1.37 step-1 evaluations per output node, two register windows (164
// VGPRs, 3 waves per SIMD, one block per CU for its 94 KB of LDS).  Without packed FP32 the kernel
// is bound by the vector ALU (about 480 instructions per lane and plane), so it is an option, not
// the default.  float32 only.
//
// Sources are injected in both steps (step n on the extended tile, n + 1 on the output tile);
// the gather blocks record u^{n-1} and u^n (the launch's inputs) at the receivers.
#include "wfd_step2d_impl.h"

namespace wfd {

constexpr int T32_ER = 48;                 // extended tile rows (one lane row each)
constexpr int T32_EC = 64;                 // extended tile columns (16 chunks)
constexpr int T32_OR = T32_ER - 2 * HALO;  // output rows
constexpr int T32_OC = T32_EC - 2 * HALO;  // output columns
constexpr int T32_SH = T32_ER + 2 * HALO;  // staged u^n plane with its in-plane halo
constexpr int T32_SW = T32_EC + 2 * HALO;
constexpr int T32_THREADS = 16 * T32_ER;
constexpr int T32_HALO_THREADS = 8 * 16 + 2 * T32_ER;

// A 4-node chunk of an LDS tile (16-byte aligned) in one 16-byte access per 4 bytes of element:
// element-by-element reads of the lanes' chunks put a wave on 8 of the 32 LDS banks.
template <typename T>
__device__ __forceinline__ C4<T> lds_chunk(const T* p)
{
    if constexpr (sizeof(T) == 4) {
        const wfd_v4f v = *reinterpret_cast<const wfd_v4f*>(p);
        return {{v.x, v.y, v.z, v.w}};
    } else {
        const wfd_v2d x = *reinterpret_cast<const wfd_v2d*>(p);
        const wfd_v2d y = *reinterpret_cast<const wfd_v2d*>(p + 2);
        return {{x.x, x.y, y.x, y.y}};
    }
}

// Laplacian of the lane's 4 nodes: t points at the lane's chunk in an LDS row-major tile with row
// stride lw; per node the terms are summed in the order of wavefd.h.
template <typename T, int K>
__device__ __forceinline__ void t32_lap4(const StepArgs<T>& a, const C4<T>* win, const T* t, int lw, T (&lap)[4])
{
#pragma unroll
    for (int e = 0; e < 4; e++) lap[e] = a.wc * win[K].e[e];
#pragma unroll
    for (int m = 1; m <= K; m++)
#pragma unroll
        for (int e = 0; e < 4; e++) lap[e] = fmaT(a.wp[m - 1], win[K - m].e[e] + win[K + m].e[e], lap[e]);
#pragma unroll
    for (int m = 1; m <= K; m++) {
        const C4<T> up = lds_chunk(t - m * lw), dn = lds_chunk(t + m * lw);
#pragma unroll
        for (int e = 0; e < 4; e++) lap[e] = fmaT(a.wr[m - 1], up.e[e] + dn.e[e], lap[e]);
    }
    const C4<T> L = lds_chunk(t - 4), C = lds_chunk(t), R = lds_chunk(t + 4);
#pragma unroll
    for (int m = 1; m <= K; m++)
#pragma unroll
        for (int e = 0; e < 4; e++) lap[e] = fmaT(a.wcol[m - 1], pick(L, C, R, e - m) + pick(L, C, R, e + m), lap[e]);
}

// Which of the planes pk .. pk + 63 (lane l: plane pk + l, within [lo, hi)) have sources in CSR
// rows [r0, r1) of their plane; one ballot.
template <typename T>
__device__ __forceinline__ uint64_t t32_src_mask(const PtsDev<T>& p, int shot, int64_t nr, int64_t pk, int64_t lo,
                                                 int64_t hi, int64_t r0, int64_t r1, int lane)
{
    const int64_t kk = pk + lane;
    bool has = false;
    if (r0 < r1 && kk >= lo && kk < hi) {
        const int32_t* rp = p.row_ptr + (int64_t)shot * (p.nrows + 1) + kk * nr;
        has = rp[r0] != rp[r1];
    }
    return __ballot(has);
}

template <typename T, int K, bool VARC, bool SPONGE>
__global__ __launch_bounds__(T32_THREADS) __attribute__((amdgpu_waves_per_eu(3))) void step3d_t2_kernel(const StepArgs<T> a, int rk, int nkb)
{
    constexpr int W = 2 * K + 1;
    __shared__ __align__(16) T stage[2][T32_SH][T32_SW];  // 16-byte chunk reads (lds_chunk)
    __shared__ __align__(16) T ring[K + 1][T32_ER][T32_EC];
    const StepBlock bp = step_block(a.ngb, a.gx, a.gy);
    const int64_t step = a.step_dev ? *a.step_dev + a.step : a.step;
    if (bp.gather) {  // first half: u^n into row step + gather_off; second half: u^{n-1} into the row before
        const uint32_t half = a.ngb / 2;
        if (bp.g < half) {
            gather_points(a, bp.g, step);
        } else {
            StepArgs<T> b = a;
            b.uc = a.up;
            b.gather_off = a.gather_off - 1;
            gather_points(b, bp.g - half, step);
        }
        return;
    }
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int shot = bp.z / nkb, kb = bp.z % nkb;
    const int64_t np = a.np, nr = a.nr, nc = a.nc, ld = a.ld, ps = a.ps;
    const int64_t j0 = (int64_t)bp.x * T32_OC, i0 = 1 + (int64_t)bp.y * T32_OR;
    const int64_t J0 = j0 - HALO, I0 = i0 - HALO;  // origin of the extended tile
    const int64_t r = I0 + ty, c = J0 + 4 * tx;    // this lane's chunk
    const int64_t k0 = 1 + (int64_t)kb * rk;
    const int64_t k1 = min(k0 + rk, np - 1);

    const int64_t sb = (int64_t)shot * a.shot_stride;
    const T* __restrict__ B = a.uc + sb;
    const T* A = a.up + sb;
    T* C = a.un + sb;
    T* D = a.un2 + sb;
    const T* __restrict__ coef = VARC ? a.coef + (int64_t)shot * a.coef_shot_stride : nullptr;

    // u^n is loaded where it can reach an interior step-1 node: planes / rows within 4 of the
    // node box, chunks overlapping columns [-3, nc + 2] (all in-bounds elements of the layout)
    auto plane_ok = [&](int64_t k) { return k >= -3 && k <= np + 2; };
    auto chunk_ok = [&](int64_t rr, int64_t cc) { return rr >= -3 && rr <= nr + 2 && cc >= -4 && cc <= nc + 2; };
    const bool my_ok = chunk_ok(r, c);
    const bool row_in = r >= 1 && r <= nr - 2;
    const bool chunk_in = row_in && c + 3 >= 1 && c <= nc - 2;  // some element is interior
    const bool is_out = tx >= 1 && tx <= 14 && ty >= HALO && ty < HALO + T32_OR;
    const bool store = is_out && r <= nr - 2 && c < nc;
    bool valid[4];
    T bc[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        valid[e] = row_in && c + e >= 1 && c + e <= nc - 2;
        bc[e] = T(0);
        if constexpr (SPONGE) bc[e] = (c + e >= 0 && c + e < nc) ? a.beta_c[c + e] : T(0);
    }
    T br = T(0);
    if constexpr (SPONGE) br = (r >= 0 && r < nr) ? a.beta_r[r] : T(0);

    // in-plane halo chunk of the staged u^n plane this thread loads (threads 0..191)
    int64_t hrow = 0, hcol = 0;
    int hl = -1;  // LDS element index of the chunk, -1: none
    if (tid < 128) {
        const int hr = tid >> 4, hc = tid & 15;
        hrow = hr < 4 ? I0 - 4 + hr : I0 + T32_ER + (hr - 4);
        hcol = J0 + 4 * hc;
        hl = (hr < 4 ? hr : T32_ER + HALO + (hr - 4)) * T32_SW + HALO + 4 * hc;
    } else if (tid < T32_HALO_THREADS) {
        const int s = tid - 128, side = s / T32_ER, rr = s % T32_ER;
        hrow = I0 + rr;
        hcol = side ? J0 + T32_EC : J0 - 4;
        hl = (HALO + rr) * T32_SW + (side ? HALO + T32_EC : 0);
    }
    const bool halo_ok = hl >= 0 && chunk_ok(hrow, hcol);
    auto fetch_halo = [&](int64_t k) {
        return halo_ok && plane_ok(k) ? ld4(B + k * ps + hrow * ld + hcol) : zero4<T>();
    };
    auto fetch_u = [&](int64_t k) { return my_ok && plane_ok(k) ? ld4(B + k * ps + r * ld + c) : zero4<T>(); };

    // source rows of this wave (rows I0 + 4 wave .. + 3), clipped to the interior
    const int64_t wr0 = min(max(I0 + 4 * wave, (int64_t)1), nr - 1);
    const int64_t wr1 = min(max(I0 + 4 * wave + 4, (int64_t)1), nr - 1);
    const bool out_wave = wave >= 1 && wave <= T32_ER / 4 - 2;  // waves holding output rows
    uint64_t m1 = 0, m2 = 0;
    int64_t m1k = INT64_MIN / 2, m2k = INT64_MIN / 2;

    const int64_t qs = k0 - K;  // first step-1 plane
    C4<T> w1[W], w2[W];
#pragma unroll
    for (int t = 0; t < W; t++) w2[t] = zero4<T>();
#pragma unroll
    for (int t = 1; t < W; t++) w1[t] = fetch_u(qs - K - 1 + t);
    C4<T> nb = fetch_u(qs + K);
    C4<T> hn = fetch_halo(qs);
    int buf = 0;
    for (int64_t q = qs; q < k1 + K; q++, buf ^= 1) {
#pragma unroll
        for (int t = 0; t + 1 < W; t++) w1[t] = w1[t + 1];
        w1[W - 1] = nb;
        if (q + 1 < k1 + K) nb = fetch_u(q + K + 1);
        T* st = &stage[buf][0][0];
        {
            T* cp = st + (ty + HALO) * T32_SW + HALO + 4 * tx;
#pragma unroll
            for (int e = 0; e < 4; e++) cp[e] = w1[K].e[e];
        }
        if (hl >= 0) {
#pragma unroll
            for (int e = 0; e < 4; e++) st[hl + e] = hn.e[e];
        }
        if (q + 1 < k1 + K) hn = fetch_halo(q + 1);
        const bool plane1 = q >= 1 && q <= np - 2;
        const int64_t o1 = q * ps + r * ld + c;
        const C4<T> P = plane1 && chunk_in ? ld4(A + o1) : zero4<T>();
        C4<T> kv = zero4<T>();
        if constexpr (VARC) kv = plane1 && chunk_in ? ld4(coef + o1) : zero4<T>();
        const int64_t p = q - K;  // step-2 plane
        const bool do2 = p >= k0;
        const int64_t o2 = p * ps + r * ld + c;
        C4<T> kv2 = zero4<T>();
        if constexpr (VARC) kv2 = do2 && store ? ld4(coef + o2) : zero4<T>();
        __syncthreads();

        // ---- step 1 at plane q on the extended tile
        T kc[4];
#pragma unroll
        for (int e = 0; e < 4; e++) kc[e] = VARC ? kv.e[e] : a.kconst;
        T b[4] = {T(0), T(0), T(0), T(0)};
        bool damp = false;
        if constexpr (SPONGE) {
            const T bpr = (plane1 ? a.beta_p[q] : T(0)) + br;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b[e] = bpr + bc[e];
                damp |= b[e] != T(0);
            }
        }
        T val[4];
        const T* tc = st + (ty + HALO) * T32_SW + HALO + 4 * tx;
        T lap1[4];
        t32_lap4<T, K>(a, w1, tc, T32_SW, lap1);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T t = damp ? T(2) * w1[K].e[e] - (T(1) - b[e]) * P.e[e] : T(2) * w1[K].e[e] - P.e[e];
            val[e] = fmaT(kc[e], lap1[e], t);
        }
        if (a.src.count && plane1) {
            if (q >= m1k + 64) {
                m1k = q;
                m1 = t32_src_mask(a.src, shot, nr, q, 1, min(k1 + K, np - 1), wr0, wr1, lane);
            }
            if ((m1 >> (q - m1k)) & 1) {
                for (int rr = 0; rr < 4; rr++) {
                    const int64_t ir = I0 + 4 * wave + rr;
                    if (ir >= 1 && ir <= nr - 2)
                        wave_inject(a.src, shot, q * nr + ir, J0 - 64 * rr, 16 * rr, 16 * rr + 15, step, lane, kc, val);
                }
            }
        }
#pragma unroll
        for (int t = 0; t + 1 < W; t++) w2[t] = w2[t + 1];
        {
            T* rw = &ring[q % (K + 1)][ty][4 * tx];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const T x = damp ? val[e] / (T(1) + b[e]) : val[e];
                w2[W - 1].e[e] = (plane1 && valid[e]) ? x : T(0);
                rw[e] = w2[W - 1].e[e];
            }
        }
        if (!do2 || !out_wave) continue;  // waves 0 and 11 hold no output rows

        // ---- step 2 at plane p = q - K on the output tile (ring slot of p written K barriers ago)
        T kc2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) kc2[e] = VARC ? kv2.e[e] : a.kconst;
        T b2[4] = {T(0), T(0), T(0), T(0)};
        bool damp2 = false;
        if constexpr (SPONGE) {
            const T bpr = a.beta_p[p] + br;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b2[e] = bpr + bc[e];
                damp2 |= b2[e] != T(0);
            }
        }
        T val2[4];
        const T* rc = &ring[p % (K + 1)][ty][4 * tx];
        T lap2[4];
        t32_lap4<T, K>(a, w2, rc, T32_EC, lap2);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            // u_prev of step 2 is u^n at plane p: the oldest plane of the u^n window
            const T t = damp2 ? T(2) * w2[K].e[e] - (T(1) - b2[e]) * w1[0].e[e] : T(2) * w2[K].e[e] - w1[0].e[e];
            val2[e] = fmaT(kc2[e], lap2[e], t);
        }
        if (a.src.count) {
            if (p >= m2k + 64) {
                m2k = p;
                m2 = t32_src_mask(a.src, shot, nr, p, k0, k1, wr0, wr1, lane);
            }
            if ((m2 >> (p - m2k)) & 1) {
                for (int rr = 0; rr < 4; rr++) {
                    const int64_t ir = I0 + 4 * wave + rr;
                    if (ir <= nr - 2)
                        wave_inject(a.src, shot, p * nr + ir, J0 - 64 * rr, 16 * rr + 1, 16 * rr + 14, step + 1, lane,
                                    kc2, val2);
                }
            }
        }
        if (store) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const T x = damp2 ? val2[e] / (T(1) + b2[e]) : val2[e];
                val2[e] = valid[e] ? x : T(0);
            }
            const T s1v[4] = {w2[K].e[0], w2[K].e[1], w2[K].e[2], w2[K].e[3]};
            st4(C + o2, s1v);
            st4(D + o2, val2);
        }
    }
}

template <typename T, int K, bool VARC, bool SPONGE>
static hipError_t launch3d_t2_k(const StepArgs<T>& a0, int rk, int nshots, hipStream_t s)
{
    StepArgs<T> a = a0;
    a.gx = (uint32_t)((a.nc + T32_OC - 1) / T32_OC);
    a.gy = (uint32_t)((a.nr - 2 + T32_OR - 1) / T32_OR);
    const int64_t nkb = (a.np - 2 + rk - 1) / rk;
    const uint32_t nblocks = step_grid((uint32_t)((int64_t)a.gx * a.gy * nkb * nshots), a.ngb);
    hipLaunchKernelGGL((step3d_t2_kernel<T, K, VARC, SPONGE>), dim3(nblocks), dim3(T32_THREADS), 0, s, a, rk,
                       (int)nkb);
    return hipGetLastError();
}

template <typename T, int K>
static hipError_t launch3d_t2_flags(const StepArgs<T>& a, bool varc, bool sponge, int rk, int nshots, hipStream_t s)
{
    if (varc) {
        if (sponge) return launch3d_t2_k<T, K, true, true>(a, rk, nshots, s);
        return launch3d_t2_k<T, K, true, false>(a, rk, nshots, s);
    }
    if (sponge) return launch3d_t2_k<T, K, false, true>(a, rk, nshots, s);
    return launch3d_t2_k<T, K, false, false>(a, rk, nshots, s);
}

template <typename T>
hipError_t launch_step3d_t2(const StepArgs<T>& a, int order, bool varc, bool sponge, int rk, int nshots,
                            hipStream_t s)
{
    switch (order) {
        case 2: return launch3d_t2_flags<T, 1>(a, varc, sponge, rk, nshots, s);
        case 4: return launch3d_t2_flags<T, 2>(a, varc, sponge, rk, nshots, s);
        case 6: return launch3d_t2_flags<T, 3>(a, varc, sponge, rk, nshots, s);
        default: return launch3d_t2_flags<T, 4>(a, varc, sponge, rk, nshots, s);
    }
}

template hipError_t launch_step3d_t2<float>(const StepArgs<float>&, int, bool, bool, int, int, hipStream_t);
// float64 would need 188 KB of LDS per block with this tile: not provided (wfd_forward2 rejects it)
template <>
hipError_t launch_step3d_t2<double>(const StepArgs<double>&, int, bool, bool, int, int, hipStream_t)
{
    return hipErrorInvalidValue;
}

}  // namespace wfd
