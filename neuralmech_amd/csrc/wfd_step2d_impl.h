// 1D / 2D leapfrog step kernels for gfx950 (wave64).
//
// Mapping: a lane owns a chunk of 4 consecutive nodes along the fastest axis (one 16-byte vector
// load per field and row for fp32); a wave owns a run of chunks of one row and marches down RI
// consecutive rows, keeping the 2K+RI rows of u_cur its stencil needs in registers (each row of
// u_cur is loaded once per wave).  A block is 4 waves stacked along the rows, and blocks are
// remapped so that each XCD sweeps one contiguous band of the grid.
//
// Two variants of the column halo:
//   step2d_kernel     a wave owns 64 chunks (256 columns); the left / right neighbour chunks of
//                     the centre row are loaded (cache hits on lines the neighbour lanes fetch);
//   step2d_xl_kernel  waves overlap by one chunk each side (lane 0 and lane 63 only load), so the
//                     neighbour chunks of lanes 1..62 are the centre chunks of lanes l-1 / l+1,
//                     moved with cross-lane shuffles; 248 owned columns per wave.
// Every lane stays active through the row loop (the point passes are wave-cooperative); lanes
// that own no node compute discarded values and store nothing.  The halo layout makes every
// neighbour read an in-bounds element; stores mask the boundary ring and row padding to zero.
// Inside the sponge-free interior the damping terms are skipped (b = 0 makes them exact no-ops).
//
// The arithmetic is the update documented in wavefd.h, in exactly that operation order.
#pragma once
#include "wfd_internal.h"
#include "wfd_points_dev.h"

namespace wfd {

template <typename T>
struct C4 {
    T e[4];
};

template <typename T>
__device__ __forceinline__ C4<T> ld4(const T* p);
template <>
__device__ __forceinline__ C4<float> ld4<float>(const float* p)
{
    const float4 v = *reinterpret_cast<const float4*>(p);
    return {{v.x, v.y, v.z, v.w}};
}
template <>
__device__ __forceinline__ C4<double> ld4<double>(const double* p)
{
    const double2 a = *reinterpret_cast<const double2*>(p);
    const double2 b = *reinterpret_cast<const double2*>(p + 2);
    return {{a.x, a.y, b.x, b.y}};
}

// Streaming access to the fields a step touches once (u_prev, kc, u_next): with NT the loads and
// stores carry the nontemporal hint, which keeps them from displacing u_cur (read by several
// waves) from the caches.  That pays off when the step's fields do not fit the Infinity Cache and
// costs when they do (the hinted accesses then go to HBM), so it is a launch option.
typedef float wfd_v4f __attribute__((ext_vector_type(4)));
typedef double wfd_v2d __attribute__((ext_vector_type(2)));
template <bool NT, typename T>
__device__ __forceinline__ C4<T> sld4(const T* p)
{
    if constexpr (NT && sizeof(T) == 4) {
        const wfd_v4f v = __builtin_nontemporal_load(reinterpret_cast<const wfd_v4f*>(p));
        return {{v.x, v.y, v.z, v.w}};
    } else if constexpr (NT) {
        const wfd_v2d a = __builtin_nontemporal_load(reinterpret_cast<const wfd_v2d*>(p));
        const wfd_v2d b = __builtin_nontemporal_load(reinterpret_cast<const wfd_v2d*>(p + 2));
        return {{a.x, a.y, b.x, b.y}};
    } else {
        return ld4<T>(p);
    }
}

template <typename T>
__device__ __forceinline__ void st4(T* p, const T (&v)[4]);
template <>
__device__ __forceinline__ void st4<float>(float* p, const float (&v)[4])
{
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void st4<double>(double* p, const double (&v)[4])
{
    *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
    *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
}
template <bool NT, typename T>
__device__ __forceinline__ void sst4(T* p, const T (&v)[4])
{
    if constexpr (NT && sizeof(T) == 4) {
        const wfd_v4f w = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(w, reinterpret_cast<wfd_v4f*>(p));
    } else if constexpr (NT) {
        const wfd_v2d a = {v[0], v[1]}, b = {v[2], v[3]};
        __builtin_nontemporal_store(a, reinterpret_cast<wfd_v2d*>(p));
        __builtin_nontemporal_store(b, reinterpret_cast<wfd_v2d*>(p + 2));
    } else {
        st4<T>(p, v);
    }
}

template <typename T>
__device__ __forceinline__ C4<T> zero4()
{
    return {{T(0), T(0), T(0), T(0)}};
}

// value at column offset j in [-4, 8) of a chunk, from the left / centre / right chunks
template <typename T>
__device__ __forceinline__ T pick(const C4<T>& L, const C4<T>& C, const C4<T>& R, int j)
{
    return j < 0 ? L.e[4 + j] : (j < 4 ? C.e[j] : R.e[j - 4]);
}

__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ C4<T> from_lane_left(const C4<T>& v)
{
    C4<T> r;
#pragma unroll
    for (int e = 0; e < 4; e++) r.e[e] = __shfl_up(v.e[e], 1);
    return r;
}
template <typename T>
__device__ __forceinline__ C4<T> from_lane_right(const C4<T>& v)
{
    C4<T> r;
#pragma unroll
    for (int e = 0; e < 4; e++) r.e[e] = __shfl_down(v.e[e], 1);
    return r;
}

// Laplacian of chunk element e: centre row `q` of window `win`, column chunks L / R.
template <typename T, int K, bool VROW>
__device__ __forceinline__ T lap_elem(const T wc, const T* wr, const T* wcol, const C4<T>* win, int q,
                                      const C4<T>& L, const C4<T>& R, int e)
{
    constexpr int KR = VROW ? K : 0;
    const C4<T>& C = win[q + KR];
    T lap = wc * C.e[e];
    if constexpr (VROW) {
#pragma unroll
        for (int m = 1; m <= K; m++) lap = fmaT(wr[m - 1], win[q + KR - m].e[e] + win[q + KR + m].e[e], lap);
    }
#pragma unroll
    for (int m = 1; m <= K; m++) lap = fmaT(wcol[m - 1], pick(L, C, R, e - m) + pick(L, C, R, e + m), lap);
    return lap;
}

// Grouped imaging: the imaging terms (new adjoint value, Laplacian of the forward field) of each
// wave's rows, exchanged through LDS so that one wave sums the group's shots in shot order.
template <typename T, int RI>
__device__ __forceinline__ C4<T> (&img_exchange())[2][IMG_GROUP][RI][64]
{
    __shared__ C4<T> x[2][IMG_GROUP][RI][64];
    return x;
}

// One stencil tile of both variants: column block tx, row block ty of shot `shot`.  XL selects
// the lane-exchange halo; VROW = false is the 1D case (a single row, no vertical stencil).  LA is
// how many rows ahead the per-row operands are loaded (see the row loop); NT selects streaming
// access to u_prev, kc and u_next.  Grouped imaging launches (a.img_group, 2D) pass the shot group
// as `shot`: wave w takes shot IMG_GROUP * group + w on the tile's RI rows, and after the march
// the group's partial row q is updated by wave q % 4 as p = fma(mu_s, L u_s, p) for s in shot
// order (only the order of the float additions differs from one partial per shot); a.uf then holds
// the Laplacian L u of the forward field (stored by the recompute), read pointwise, as it is in
// LUF imaging instances (one partial per shot; the uneven batch).  On a forward
// instance (IMG false) GRP stores that Laplacian of u_cur (the same lap_elem values the update
// uses) into a.ucopy with the nontemporal hint, while u_next rotates in place with u_prev.
template <typename T, int K, bool VROW, bool VARC, bool SPONGE, bool IMG, int RI, bool XL, int LA = 0,
          bool NT = false, bool GRP = false, bool LUF = false>
__device__ __forceinline__ void step2d_tile(const StepArgs<T>& a, uint32_t tx, uint32_t ty, int shot_in, int64_t step)
{
    constexpr int KR = VROW ? K : 0;
    constexpr int NW = RI + 2 * KR;
    constexpr int64_t COLS = XL ? 248 : 256;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);  // wave of the 4-wave tile
    const int64_t base = (int64_t)tx * COLS - (XL ? 4 : 0);  // column of lane 0's chunk
    const int64_t c = base + lane * 4;
    constexpr bool grp = GRP && IMG && VROW;  // grouped imaging instance (a.img_group > 1)
    constexpr bool luf = IMG && (grp || LUF);  // a.uf holds L u of the forward field
    constexpr bool cpy = GRP && !IMG && VROW;  // forward instance storing a second copy (a.ucopy)
    // grouped: a wave of a shot past the last one marches the group's first shot and stores nothing
    const int gshot = grp ? shot_in * IMG_GROUP + wave : shot_in;
    const bool live = !grp || gshot < a.nshots;
    const int shot = live ? gshot : shot_in * IMG_GROUP;
    const int64_t r0 = a.row_lo + (grp ? (int64_t)ty * RI : ((int64_t)ty * 4 + wave) * RI);
    if (r0 >= a.row_hi) return;  // whole wave (grouped: the whole tile, its waves share the rows)
    const int lane_lo = XL ? 1 : 0, lane_hi = XL ? 62 : 63;
    const bool load_ok = c < a.ld;  // no wrap of a lane's chunk into the next row
    const bool own = lane >= lane_lo && lane <= lane_hi && c < a.nc;

    const int64_t sb = (int64_t)shot * a.shot_stride;
    const T* __restrict__ uc = a.uc + sb;
    const T* up = a.up + sb;  // may alias un (in-place leapfrog): same element, same lane
    T* un = a.un + sb;
    const T* __restrict__ coef = VARC ? a.coef + (int64_t)shot * a.coef_shot_stride : nullptr;

    C4<T> win[NW];
#pragma unroll
    for (int t = 0; t < NW; t++) {
        const int64_t r = r0 - KR + t;
        win[t] = (r < a.row_hi + KR && load_ok) ? ld4(uc + r * a.ld + c) : zero4<T>();
    }
    // the forward field's window for its Laplacian (grouped instances read that Laplacian instead)
    C4<T> wf[IMG && !luf ? NW : 1];
    if constexpr (IMG && !luf) {
        const T* __restrict__ uf = a.uf + sb;
#pragma unroll
        for (int t = 0; t < NW; t++) {
            const int64_t r = r0 - KR + t;
            wf[t] = (r < a.row_hi + KR && load_ok) ? ld4(uf + r * a.ld + c) : zero4<T>();
        }
    }

    // does the source list have an entry in this wave's rows? (one range check per wave)
    const int64_t rlast = VROW ? (r0 + RI < a.row_hi ? r0 + RI : a.row_hi) : 1;
    const int64_t rfirst = VROW ? r0 : 0;
    const bool src_here = a.src.count &&
        a.src.row_ptr[(int64_t)shot * (a.src.nrows + 1) + rfirst] != a.src.row_ptr[(int64_t)shot * (a.src.nrows + 1) + rlast];

    bool valid[4];
    T bc[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        valid[e] = own && (c + e >= 1) && (c + e <= a.nc - 2);
        bc[e] = T(0);
        if constexpr (SPONGE) bc[e] = (c + e >= 0 && c + e < a.nc) ? a.beta_c[c + e] : T(0);
    }

    // The per-row operands (u_prev, kc, the sponge row value, the imaging partial and, in the cache
    // variant, the column-halo chunks) are loaded LA rows ahead of the row being updated.  With
    // LA > 0 each row waits for everything outstanding right before its stores, so the next rows'
    // loads are in flight through this row's stencil; a wait placed after a store would also wait
    // for the store (the compiler cannot wait for one load while a store is outstanding without
    // waiting for the store).  u_next may alias u_prev only element for element, and a row's
    // operands are loaded before its own store: the values read are those of a row-by-row update.
    const T* __restrict__ ufp = IMG ? a.uf + sb : nullptr;
    constexpr int NH = XL ? 1 : RI, NHF = (XL || !IMG || luf) ? 1 : RI;
    C4<T> Pv[RI], Kv[VARC ? RI : 1], Gv[IMG ? RI : 1], Lv[NH], Rv[NH], Lfv[NHF], Rfv[NHF], Fv[luf && !grp ? RI : 1];
    T Bv[SPONGE ? RI : 1];
    auto issue = [&](int q) {
        const int64_t r = r0 + q;
        if (q >= RI || r >= a.row_hi) return;  // whole wave
        const bool ld = own && r < a.row_hi;
        const int64_t o = r * a.ld + c;
        Pv[q] = ld ? sld4<NT>(up + o) : zero4<T>();
        if constexpr (VARC) Kv[q] = ld ? sld4<NT>(coef + o) : zero4<T>();
        // Gv: the imaging partial, or for grouped instances the stored Laplacian of the forward
        // field (read once, with the nontemporal hint)
        if constexpr (IMG) Gv[q] = !ld ? zero4<T>() : grp ? sld4<true>(ufp + o) : ld4(a.g + sb + o);
        if constexpr (luf && !grp) Fv[q] = ld ? sld4<true>(ufp + o) : zero4<T>();
        if constexpr (SPONGE) Bv[q] = VROW ? a.beta_r[r] : T(0);
        if constexpr (!XL) {
            Lv[q] = load_ok ? ld4(uc + o - 4) : zero4<T>();
            Rv[q] = load_ok ? ld4(uc + o + 4) : zero4<T>();
            if constexpr (IMG && !luf) {
                Lfv[q] = load_ok ? ld4(ufp + o - 4) : zero4<T>();
                Rfv[q] = load_ok ? ld4(ufp + o + 4) : zero4<T>();
            }
        }
    };
#pragma unroll
    for (int q = 0; q <= LA; q++) issue(q);
#pragma unroll
    for (int q = 0; q < RI; q++) {
        const int64_t r = r0 + q;
        if (r >= a.row_hi) break;  // whole wave
        if (q > 0) issue(q + LA);
        const int64_t o = r * a.ld + c;
        C4<T> L, R, Lf, Rf;
        if constexpr (XL) {
            L = from_lane_left(win[q + KR]);
            R = from_lane_right(win[q + KR]);
            if constexpr (IMG && !luf) {
                Lf = from_lane_left(wf[q + KR]);
                Rf = from_lane_right(wf[q + KR]);
            }
        } else {
            L = Lv[XL ? 0 : q];
            R = Rv[XL ? 0 : q];
            if constexpr (IMG && !luf) {
                Lf = Lfv[NHF > 1 ? q : 0];
                Rf = Rfv[NHF > 1 ? q : 0];
            }
        }
        const C4<T>& P = Pv[q];
        T kc[4];
#pragma unroll
        for (int e = 0; e < 4; e++) kc[e] = VARC ? Kv[VARC ? q : 0].e[e] : a.kconst;
        T b[4] = {T(0), T(0), T(0), T(0)};
        bool damp = false;
        if constexpr (SPONGE) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b[e] = VROW ? Bv[q] + bc[e] : bc[e];
                damp |= b[e] != T(0);
            }
        }
        T val[4], lapc[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T ctr = win[q + KR].e[e];
            const T lap = lap_elem<T, K, VROW>(a.wc, a.wr, a.wcol, win, q, L, R, e);
            lapc[e] = lap;
            const T t = damp ? T(2) * ctr - (T(1) - b[e]) * P.e[e] : T(2) * ctr - P.e[e];
            val[e] = fmaT(kc[e], lap, t);
        }
        const int64_t prow = VROW ? r : 0;  // CSR row of the point lists
        if (src_here) wave_inject(a.src, shot, prow, base, lane_lo, lane_hi, step, lane, kc, val);
        if (damp) {
#pragma unroll
            for (int e = 0; e < 4; e++) val[e] = val[e] / (T(1) + b[e]);
        }
#pragma unroll
        for (int e = 0; e < 4; e++)
            if (!valid[e]) val[e] = T(0);
        T gv[4];
        if constexpr (IMG && grp) {
#pragma unroll
            for (int e = 0; e < 4; e++) gv[e] = Gv[q].e[e];
        } else if constexpr (luf) {
#pragma unroll
            for (int e = 0; e < 4; e++) gv[e] = Fv[q].e[e];
        } else if constexpr (IMG) {
#pragma unroll
            for (int e = 0; e < 4; e++) gv[e] = lap_elem<T, K, VROW>(a.wc, a.wr, a.wcol, wf, q, Lf, Rf, e);
        }
        if constexpr (LA > 0) __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        if (own && live) {
            sst4<NT>(un + o, val);
            if constexpr (cpy) sst4<true>(a.ucopy + sb + o, lapc);
        }
        if constexpr (IMG) {
            if constexpr (grp) {
                auto& xch = img_exchange<T, RI>();
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    xch[0][wave][q][lane].e[e] = val[e];
                    xch[1][wave][q][lane].e[e] = gv[e];
                }
            } else if (own) {
                T* gp = a.g + sb;
                const C4<T>& G = Gv[IMG ? q : 0];
#pragma unroll
                for (int e = 0; e < 4; e++) gv[e] = valid[e] ? fmaT(val[e], gv[e], G.e[e]) : T(0);
                if (a.gsum) {  // fused flush: gsum += partial, partial = 0 (same bits as the flush kernel)
                    double* gs = a.gsum + sb + o;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        if (valid[e]) gs[e] = (a.gsum_first ? 0.0 : gs[e]) + (double)gv[e];
                        gv[e] = T(0);
                    }
                }
                st4(gp + o, gv);
            }
        }
    }
    if constexpr (IMG) {
        if constexpr (grp) {
            // every wave of the tile reaches this point (same rows, same loop); the group's partial
            // row q is updated by wave q % 4 from the four waves' terms, in shot order
            auto& xch = img_exchange<T, RI>();
            const int nlive = min(IMG_GROUP, a.nshots - shot_in * IMG_GROUP);
            T* gp = a.g + (int64_t)shot_in * a.shot_stride;
            C4<T> G[(RI + 3) / 4];
#pragma unroll
            for (int j = 0; j < (RI + 3) / 4; j++) {
                const int q = wave + 4 * j;
                const int64_t r = r0 + q;
                G[j] = q < RI && r < a.row_hi && own ? ld4(gp + r * a.ld + c) : zero4<T>();
            }
            lds_barrier();
#pragma unroll
            for (int j = 0; j < (RI + 3) / 4; j++) {
                const int q = wave + 4 * j;
                const int64_t r = r0 + q;
                if (q >= RI || r >= a.row_hi) break;  // whole wave
                T p[4];
#pragma unroll
                for (int e = 0; e < 4; e++) p[e] = G[j].e[e];
                for (int s = 0; s < nlive; s++) {
                    const C4<T> v = xch[0][s][q][lane], l = xch[1][s][q][lane];
#pragma unroll
                    for (int e = 0; e < 4; e++) p[e] = fmaT(v.e[e], l.e[e], p[e]);
                }
#pragma unroll
                for (int e = 0; e < 4; e++)
                    if (!valid[e]) p[e] = T(0);
                if (own) st4(gp + r * a.ld + c, p);
            }
        }
    }
}

// A step launch: gather blocks copy receiver values, the other blocks update one tile each.
template <typename T, int K, bool VROW, bool VARC, bool SPONGE, bool IMG, int RI, bool XL, bool NT, bool GRP = false,
          bool LUF = false>
__device__ __forceinline__ void step2d_body(const StepArgs<T>& a)
{
    // grouped imaging launches have one tile column per shot group instead of per shot
    const uint32_t nz = GRP && IMG ? (uint32_t)((a.nshots + IMG_GROUP - 1) / IMG_GROUP) : (uint32_t)a.nshots;
    const StepBlock bp = step_block(a.ngb, a.gx, a.gy, nz, a.shot_fast != 0);
    const int64_t step = a.step_dev ? *a.step_dev + a.step : a.step;
    if (bp.gather) {
        gather_points(a, bp.g, step);
        return;
    }
    // one row ahead, except imaging tiles of 4 rows that take the forward field's Laplacian
    // themselves (two windows), where the look-ahead's registers cost a wave per SIMD (2 and 8
    // rows keep their 3 and 2 waves); the instances reading the stored Laplacian stay at 4 waves
    // per SIMD with it (112 / 122 VGPRs; 16 x 1024^2 gradient step 138.3 -> 135.0 us).  Loading a
    // wave's operands up front (LA = RI - 1) measured slower for lack of waves.
    constexpr int LA = RI > 1 && !(IMG && RI == 4 && !GRP && !LUF) ? 1 : 0;
    step2d_tile<T, K, VROW, VARC, SPONGE, IMG, RI, XL, LA, NT, GRP, LUF>(a, bp.x, bp.y, (int)bp.z, step);
}

template <typename T, int K, bool VROW, bool VARC, bool SPONGE, bool IMG, int RI, bool NT, bool GRP = false, bool LUF = false>
__global__ __launch_bounds__(256) void step2d_kernel(const StepArgs<T> a)
{
    step2d_body<T, K, VROW, VARC, SPONGE, IMG, RI, false, NT, GRP, LUF>(a);
}

constexpr int XL_COLS = 248;

template <typename T, int K, bool VARC, bool SPONGE, bool IMG, int RI, bool NT, bool GRP = false, bool LUF = false>
__global__ __launch_bounds__(256) void step2d_xl_kernel(const StepArgs<T> a)
{
    step2d_body<T, K, true, VARC, SPONGE, IMG, RI, true, NT, GRP, LUF>(a);
}

// ------------------------------------------------------------------------------ dispatch
// variant bit 0: lane-exchange halo (XL); bit 1: streaming access (NT)
template <typename T, int K, bool VARC, bool SPONGE, bool IMG, bool NT, bool GRP = false, bool LUF = false>
static hipError_t launch2d_kn(const StepArgs<T>& a0, bool vrow, int variant, int ri, int nshots, hipStream_t s)
{
    const dim3 block(256);
    const int64_t rows = a0.row_hi - a0.row_lo;
    StepArgs<T> a = a0;
    // 1D launch: a.ngb gather blocks, then gx * gy * nshots stencil tiles (grouped imaging: tiles of
    // r rows, one per shot group)
    constexpr bool grp = GRP && IMG;  // grouped imaging instance (2D only)
    auto grid = [&](int64_t cols, int r) {
        a.gx = (uint32_t)((a.nc + cols - 1) / cols);
        a.gy = vrow ? (uint32_t)(grp ? (rows + r - 1) / r : (rows + 4 * r - 1) / (4 * r)) : 1u;
        const int64_t nz = grp ? (nshots + IMG_GROUP - 1) / IMG_GROUP : nshots;
        return dim3(step_grid((uint32_t)((int64_t)a.gx * a.gy * nz), a.ngb));
    };
    if (!vrow) {
        const dim3 g = grid(256, 1);
        hipLaunchKernelGGL((step2d_kernel<T, K, false, VARC, SPONGE, IMG, 1, NT>), g, block, 0, s, a);
        return hipGetLastError();
    }
    if (variant & 1) {
        if constexpr (sizeof(T) == 4) {
            switch (ri) {
                case 1: { const dim3 g = grid(XL_COLS, 1); hipLaunchKernelGGL((step2d_xl_kernel<T, K, VARC, SPONGE, IMG, 1, NT, GRP, LUF>), g, block, 0, s, a); break; }
                case 2: { const dim3 g = grid(XL_COLS, 2); hipLaunchKernelGGL((step2d_xl_kernel<T, K, VARC, SPONGE, IMG, 2, NT, GRP, LUF>), g, block, 0, s, a); break; }
                case 8: { const dim3 g = grid(XL_COLS, 8); hipLaunchKernelGGL((step2d_xl_kernel<T, K, VARC, SPONGE, IMG, 8, NT, GRP, LUF>), g, block, 0, s, a); break; }
                default: { const dim3 g = grid(XL_COLS, 4); hipLaunchKernelGGL((step2d_xl_kernel<T, K, VARC, SPONGE, IMG, 4, NT, GRP, LUF>), g, block, 0, s, a); break; }
            }
        } else {
            const dim3 g = grid(XL_COLS, 4);
            hipLaunchKernelGGL((step2d_xl_kernel<T, K, VARC, SPONGE, IMG, 4, NT, GRP, LUF>), g, block, 0, s, a);
        }
        return hipGetLastError();
    }
    if constexpr (sizeof(T) == 4) {
        switch (ri) {
            case 1: { const dim3 g = grid(256, 1); hipLaunchKernelGGL((step2d_kernel<T, K, true, VARC, SPONGE, IMG, 1, NT, GRP, LUF>), g, block, 0, s, a); break; }
            case 2: { const dim3 g = grid(256, 2); hipLaunchKernelGGL((step2d_kernel<T, K, true, VARC, SPONGE, IMG, 2, NT, GRP, LUF>), g, block, 0, s, a); break; }
            case 8: { const dim3 g = grid(256, 8); hipLaunchKernelGGL((step2d_kernel<T, K, true, VARC, SPONGE, IMG, 8, NT, GRP, LUF>), g, block, 0, s, a); break; }
            default: { const dim3 g = grid(256, 4); hipLaunchKernelGGL((step2d_kernel<T, K, true, VARC, SPONGE, IMG, 4, NT, GRP, LUF>), g, block, 0, s, a); break; }
        }
    } else {
        const dim3 g = grid(256, 4);
        hipLaunchKernelGGL((step2d_kernel<T, K, true, VARC, SPONGE, IMG, 4, NT, GRP, LUF>), g, block, 0, s, a);
    }
    return hipGetLastError();
}

template <typename T, int K, bool NT>
static hipError_t launch2d_flags(const StepArgs<T>& a, bool vrow, bool varc, bool sponge, bool img, int variant,
                                 int ri, int nshots, hipStream_t s)
{
    const int f = (varc ? 4 : 0) | (sponge ? 2 : 0) | (img ? 1 : 0);
    if (img && vrow && a.img_group > 1) {  // grouped imaging instances
        switch (f) {
            case 1: return launch2d_kn<T, K, false, false, true, NT, true>(a, vrow, variant, ri, nshots, s);
            case 3: return launch2d_kn<T, K, false, true, true, NT, true>(a, vrow, variant, ri, nshots, s);
            case 5: return launch2d_kn<T, K, true, false, true, NT, true>(a, vrow, variant, ri, nshots, s);
            default: return launch2d_kn<T, K, true, true, true, NT, true>(a, vrow, variant, ri, nshots, s);
        }
    }
    if (img && vrow && a.img_group == 1) {  // one partial per shot, L u of the forward field read pointwise
        switch (f) {
            case 1: return launch2d_kn<T, K, false, false, true, NT, false, true>(a, vrow, variant, ri, nshots, s);
            case 3: return launch2d_kn<T, K, false, true, true, NT, false, true>(a, vrow, variant, ri, nshots, s);
            case 5: return launch2d_kn<T, K, true, false, true, NT, false, true>(a, vrow, variant, ri, nshots, s);
            default: return launch2d_kn<T, K, true, true, true, NT, false, true>(a, vrow, variant, ri, nshots, s);
        }
    }
    if (!img && vrow && a.ucopy) {  // forward instances that also store L u_cur
        switch (f) {
            case 0: return launch2d_kn<T, K, false, false, false, NT, true>(a, vrow, variant, ri, nshots, s);
            case 2: return launch2d_kn<T, K, false, true, false, NT, true>(a, vrow, variant, ri, nshots, s);
            case 4: return launch2d_kn<T, K, true, false, false, NT, true>(a, vrow, variant, ri, nshots, s);
            default: return launch2d_kn<T, K, true, true, false, NT, true>(a, vrow, variant, ri, nshots, s);
        }
    }
    switch (f) {
        case 0: return launch2d_kn<T, K, false, false, false, NT>(a, vrow, variant, ri, nshots, s);
        case 1: return launch2d_kn<T, K, false, false, true, NT>(a, vrow, variant, ri, nshots, s);
        case 2: return launch2d_kn<T, K, false, true, false, NT>(a, vrow, variant, ri, nshots, s);
        case 3: return launch2d_kn<T, K, false, true, true, NT>(a, vrow, variant, ri, nshots, s);
        case 4: return launch2d_kn<T, K, true, false, false, NT>(a, vrow, variant, ri, nshots, s);
        case 5: return launch2d_kn<T, K, true, false, true, NT>(a, vrow, variant, ri, nshots, s);
        case 6: return launch2d_kn<T, K, true, true, false, NT>(a, vrow, variant, ri, nshots, s);
        default: return launch2d_kn<T, K, true, true, true, NT>(a, vrow, variant, ri, nshots, s);
    }
}

// The kernels of one access mode (NT) for every order and flag set; the float instantiations of
// the two modes live in separate files so that they compile in parallel.
template <typename T, bool NT>
hipError_t launch_step2d_mode(const StepArgs<T>& a, bool vrow, int order, bool varc, bool sponge, bool img,
                              int variant, int ri, int nshots, hipStream_t s)
{
    switch (order) {
        case 2: return launch2d_flags<T, 1, NT>(a, vrow, varc, sponge, img, variant, ri, nshots, s);
        case 4: return launch2d_flags<T, 2, NT>(a, vrow, varc, sponge, img, variant, ri, nshots, s);
        case 6: return launch2d_flags<T, 3, NT>(a, vrow, varc, sponge, img, variant, ri, nshots, s);
        default: return launch2d_flags<T, 4, NT>(a, vrow, varc, sponge, img, variant, ri, nshots, s);
    }
}

extern template hipError_t launch_step2d_mode<float, true>(const StepArgs<float>&, bool, int, bool, bool, bool, int,
                                                           int, int, hipStream_t);

template <typename T>
hipError_t launch_step2d(const StepArgs<T>& a, bool vrow, int order, bool varc, bool sponge, bool img, int variant,
                         int ri, int nshots, hipStream_t s)
{
    if (variant & 2) return launch_step2d_mode<T, true>(a, vrow, order, varc, sponge, img, variant, ri, nshots, s);
    return launch_step2d_mode<T, false>(a, vrow, order, varc, sponge, img, variant, ri, nshots, s);
}

}  // namespace wfd
