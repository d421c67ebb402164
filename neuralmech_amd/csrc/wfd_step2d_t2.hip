// Two leapfrog steps per launch (2D), bit-identical to two single steps.
//
// A wave owns 60 column chunks (240 columns; lanes 2..61) and RI output rows.  It marches down
// the RI + 4K rows of u^n it needs, keeping two rolling register windows of 2K + 1 rows:
//   * rows of u^n: when row b arrives, step 1 produces u^{n+1} at row b - K (its full vertical
//     stencil is in the window, its column neighbours come from lanes l - 1 / l + 1);
//   * rows of u^{n+1}: when row r arrives, step 2 produces u^{n+2} at row r - K from them and
//     the u^n row r - K (still the oldest row of the first window).
// Step-1 values are exact on lanes 1..62 and rows inside the node box (they are zero on the
// boundary ring and outside, exactly as a single step stores them), so step 2 is exact on lanes
// 2..61: the 2K halo rows and two halo chunks per side are computed twice, by both neighbouring
// waves.  Per output node the launch reads u^{n-1}, u^n and kc once and writes u^{n+1} and
// u^{n+2}: 20 bytes per two steps instead of 32.  All four fields are distinct (u^{n+1} cannot
// overwrite u^{n-1} in place: neighbouring waves still read its halo rows).
//
// Sources are injected in both steps (step n on lanes 1..62, step n + 1 on lanes 2..61); the
// gather blocks record u^{n-1} and u^n (the launch's inputs) at the receivers.
#include "wfd_step2d_impl.h"

namespace wfd {

constexpr int T2_COLS = 240;

// Laplacian of element e for the (2K+1)-row window `v` (row r - K .. r + K), centre chunk v[K].
template <typename T, int K>
__device__ __forceinline__ T t2_lap(const StepArgs<T>& a, const C4<T>* v, const C4<T>& L, const C4<T>& R, int e)
{
    return lap_elem<T, K, true>(a.wc, a.wr, a.wcol, v, 0, L, R, e);
}

template <typename T, int K, bool VARC, bool SPONGE, int RI>
__device__ __forceinline__ void t2_tile(const StepArgs<T>& a, uint32_t tx, uint32_t ty, int shot, int64_t step)
{
    constexpr int W = 2 * K + 1;    // stencil window rows
    constexpr int NB = RI + 4 * K;  // rows of u^n read
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
    const int64_t base = (int64_t)tx * T2_COLS - 8;  // column of lane 0's chunk
    const int64_t c = base + lane * 4;
    const int64_t r0 = a.row_lo + ((int64_t)ty * 4 + wave) * RI;
    if (r0 >= a.row_hi) return;  // whole wave
    const int64_t nr = a.nr, nc = a.nc, ld = a.ld;
    const bool col_in = c >= 0 && c < nc;  // a chunk lies inside or outside the node columns
    const bool own = lane >= 2 && lane <= 61 && c < nc;

    const int64_t sb = (int64_t)shot * a.shot_stride;
    const T* A = a.up + sb;
    const T* B = a.uc + sb;
    T* C = a.un + sb;
    T* D = a.un2 + sb;
    const T* coef = VARC ? a.coef + (int64_t)shot * a.coef_shot_stride : nullptr;

    bool valid[4];
    T bc[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        valid[e] = (c + e >= 1) && (c + e <= nc - 2);
        bc[e] = T(0);
        if constexpr (SPONGE) bc[e] = (c + e >= 0 && c + e < nc) ? a.beta_c[c + e] : T(0);
    }
    // source rows of step 1 ([r0 - K, r0 + RI + K)) and of step 2 ([r0, r0 + RI))
    const int32_t* rp = a.src.count ? a.src.row_ptr + (int64_t)shot * (a.src.nrows + 1) : nullptr;
    auto clamp_row = [&](int64_t r) { return r < 0 ? (int64_t)0 : (r > nr ? nr : r); };
    const bool src1 = rp && rp[clamp_row(r0 - K)] != rp[clamp_row(r0 + RI + K)];
    const bool src2 = rp && rp[clamp_row(r0)] != rp[clamp_row(r0 + RI < a.row_hi ? r0 + RI : a.row_hi)];

    auto node_row = [&](int64_t r) { return r >= 0 && r < nr && col_in; };
    auto inner_row = [&](int64_t r) { return r >= 1 && r <= nr - 2 && col_in; };
    // Iteration i brings u^n row rb = r0 - 2K + i; step 1 then produces row r1 = rb - K and, from
    // i = 4K on, step 2 produces row r2 = rb - 2K.  bw holds u^n rows rb - 2K .. rb, sw the step-1
    // rows r1 - 2K .. r1; the operands of iteration i + 1 are loaded during iteration i.
    C4<T> bw[W], sw[W];
#pragma unroll
    for (int t = 0; t < W; t++) sw[t] = zero4<T>();
#pragma unroll
    for (int t = 1; t < W; t++) {
        const int64_t rb = r0 - 2 * K + (t - 1);
        bw[t] = node_row(rb) ? ld4(B + rb * ld + c) : zero4<T>();
    }
    int64_t rb = r0;  // the row iteration 2K brings
    C4<T> nb = node_row(rb) ? ld4(B + rb * ld + c) : zero4<T>();
    C4<T> na = inner_row(rb - K) ? ld4(A + (rb - K) * ld + c) : zero4<T>();
    C4<T> nk = zero4<T>(), nk2 = zero4<T>();
    if constexpr (VARC) nk = inner_row(rb - K) ? ld4(coef + (rb - K) * ld + c) : zero4<T>();
    for (int i = 2 * K; i < NB; i++, rb++) {
        if (i >= 4 * K && rb - 2 * K >= a.row_hi) break;  // no step-2 row left (whole wave)
#pragma unroll
        for (int t = 0; t + 1 < W; t++) bw[t] = bw[t + 1];
        bw[W - 1] = nb;
        const C4<T> P = na;
        const C4<T> kv = nk, kv2 = nk2;
        if (i + 1 < NB) {  // operands of the next iteration
            nb = node_row(rb + 1) ? ld4(B + (rb + 1) * ld + c) : zero4<T>();
            na = inner_row(rb + 1 - K) ? ld4(A + (rb + 1 - K) * ld + c) : zero4<T>();
            if constexpr (VARC) {
                nk = inner_row(rb + 1 - K) ? ld4(coef + (rb + 1 - K) * ld + c) : zero4<T>();
                if (i + 1 >= 4 * K) nk2 = inner_row(rb + 1 - 2 * K) ? ld4(coef + (rb + 1 - 2 * K) * ld + c) : zero4<T>();
            }
        }
        // ---- step 1 at row r1
        const int64_t r1 = rb - K;
        const bool row1 = r1 >= 1 && r1 <= nr - 2;
        const C4<T> L = from_lane_left(bw[K]), R = from_lane_right(bw[K]);
        T kc[4];
#pragma unroll
        for (int e = 0; e < 4; e++) kc[e] = VARC ? kv.e[e] : a.kconst;
        T b[4] = {T(0), T(0), T(0), T(0)};
        bool damp = false;
        if constexpr (SPONGE) {
            const T br = row1 ? a.beta_r[r1] : T(0);
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b[e] = br + bc[e];
                damp |= b[e] != T(0);
            }
        }
        T val[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T lap = t2_lap<T, K>(a, bw, L, R, e);
            const T t = damp ? T(2) * bw[K].e[e] - (T(1) - b[e]) * P.e[e] : T(2) * bw[K].e[e] - P.e[e];
            val[e] = fmaT(kc[e], lap, t);
        }
        if (src1 && row1) wave_inject(a.src, shot, r1, base, 1, 62, step, lane, kc, val);
#pragma unroll
        for (int t = 0; t + 1 < W; t++) sw[t] = sw[t + 1];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T x = damp ? val[e] / (T(1) + b[e]) : val[e];
            sw[W - 1].e[e] = (row1 && valid[e]) ? x : T(0);
        }
        if (i < 4 * K) continue;
        // ---- step 2 at row r2 (a row of this wave's output, r2 < row_hi)
        const int64_t r2 = rb - 2 * K;
        const C4<T> L2 = from_lane_left(sw[K]), R2 = from_lane_right(sw[K]);
        T kc2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) kc2[e] = VARC ? kv2.e[e] : a.kconst;
        T b2[4] = {T(0), T(0), T(0), T(0)};
        bool damp2 = false;
        if constexpr (SPONGE) {
            const T br = a.beta_r[r2];
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b2[e] = br + bc[e];
                damp2 |= b2[e] != T(0);
            }
        }
        T val2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T lap = t2_lap<T, K>(a, sw, L2, R2, e);
            // u_prev of step 2 is u^n at row r2: the oldest row of bw
            const T t = damp2 ? T(2) * sw[K].e[e] - (T(1) - b2[e]) * bw[0].e[e] : T(2) * sw[K].e[e] - bw[0].e[e];
            val2[e] = fmaT(kc2[e], lap, t);
        }
        if (src2) wave_inject(a.src, shot, r2, base, 2, 61, step + 1, lane, kc2, val2);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T x = damp2 ? val2[e] / (T(1) + b2[e]) : val2[e];
            val2[e] = valid[e] ? x : T(0);
        }
        if (own) {
            const int64_t o2 = r2 * ld + c;
            const T s1v[4] = {sw[K].e[0], sw[K].e[1], sw[K].e[2], sw[K].e[3]};
            st4(C + o2, s1v);
            st4(D + o2, val2);
        }
    }
}

// Two steps per launch with the intermediate level in LDS (rows per wave RI = 2, 4 or 8).
//
// A block of 4 waves owns 4 RI output rows x 248 columns (62 chunks; lanes 1..62 of 64, lane 0's
// chunk 4 columns left of the first output column).  Step 1 covers the 4 RI + 2K rows from K above
// to K below the block's rows on all 64 lanes: wave w marches rows w S .. w S + S - 1 of that band
// (S = ceil((4 RI + 2K) / 4)) exactly like a single step (window of 2K + 1 rows of u^n in
// registers, column neighbours loaded from the adjacent chunks), writes u^{n+1} into LDS and, on
// its own rows and lanes, to global memory.  After one barrier, wave w evaluates step 2 on its RI
// rows from the LDS band (its window rolls over LDS rows, column neighbours by lane exchange).
// Step-1 values are zero outside the interior, exactly as a single step leaves them, so step 2 is
// exact on lanes 1..62.  Per output node the launch reads u^{n-1}, kc (1 + 2K / (4 RI) times) and
// u^n (once from the cache-resident level; the halo rows are cache hits) and writes u^{n+1} and
// u^{n+2}.
template <typename T, int K, bool VARC, bool SPONGE, int RI>
__device__ __forceinline__ void t2l_tile(const StepArgs<T>& a, uint32_t tx, uint32_t ty, int shot, int64_t step,
                                         C4<T> (*band)[64])
{
    constexpr int W = 2 * K + 1;
    constexpr int NR1 = 4 * RI + 2 * K;  // step-1 rows of the block
    constexpr int S = (NR1 + 3) / 4;     // step-1 rows per wave
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane((threadIdx.x >> 6) & 3);
    const int64_t base = (int64_t)tx * 248 - 4;  // column of lane 0's chunk
    const int64_t c = base + lane * 4;
    const int64_t R0 = a.row_lo + (int64_t)ty * 4 * RI;  // the block's first output row
    const int64_t nr = a.nr, nc = a.nc, ld = a.ld;
    const bool load_ok = c < ld;  // no wrap of a lane's chunk into the next row
    const bool own = lane >= 1 && lane <= 62 && c < nc;

    const int64_t sb = (int64_t)shot * a.shot_stride;
    const T* A = a.up + sb;
    const T* B = a.uc + sb;
    T* C = a.un + sb;
    T* D = a.un2 + sb;
    const T* coef = VARC ? a.coef + (int64_t)shot * a.coef_shot_stride : nullptr;

    bool valid[4];
    T bc[4];
#pragma unroll
    for (int e = 0; e < 4; e++) {
        valid[e] = (c + e >= 1) && (c + e <= nc - 2);
        bc[e] = T(0);
        if constexpr (SPONGE) bc[e] = (c + e >= 0 && c + e < nc) ? a.beta_c[c + e] : T(0);
    }
    const int32_t* rp = a.src.count ? a.src.row_ptr + (int64_t)shot * (a.src.nrows + 1) : nullptr;
    auto clamp_row = [&](int64_t r) { return r < 0 ? (int64_t)0 : (r > nr ? nr : r); };
    auto node_row = [&](int64_t r) { return r >= 0 && r < nr && load_ok; };

    // Both steps load each row's operands one row ahead (step 1: the window's next row of u^n,
    // the column-halo chunks, u^{n-1} and kc; step 2: u^n and kc) and wait for everything
    // outstanding right before a row's stores: the next row's loads are in flight through this
    // row's stencil, and a wait placed after a store would also wait for the store.  The first
    // step-2 rows' operands are issued before the barrier, which orders LDS only.

    // ---- step 1: rows R0 - K + q, q = q0 .. q1 - 1 of the band
    const int q0 = wave * S, q1 = min(q0 + S, NR1);
    struct Row1 {
        C4<T> nb, L, R, P, kv;
        T br;
    };
    auto issue1 = [&](Row1& o, int q) {
        const int64_t r1 = R0 - K + q;
        o.nb = node_row(r1 + K) ? ld4(B + (r1 + K) * ld + c) : zero4<T>();
        const bool row1 = r1 >= 1 && r1 <= nr - 2;
        const int64_t off = r1 * ld + c;
        o.L = row1 && load_ok ? ld4(B + off - 4) : zero4<T>();
        o.R = row1 && load_ok ? ld4(B + off + 4) : zero4<T>();
        o.P = row1 && load_ok ? ld4(A + off) : zero4<T>();
        o.kv = zero4<T>();
        if constexpr (VARC) o.kv = row1 && load_ok ? ld4(coef + off) : zero4<T>();
        o.br = T(0);
        if constexpr (SPONGE) o.br = row1 ? a.beta_r[r1] : T(0);
    };
    if (q0 < q1) {
        const int64_t rf = R0 - K + q0;  // first step-1 row of this wave
        const bool src1 = rp && rp[clamp_row(rf)] != rp[clamp_row(R0 - K + q1)];
        C4<T> win[W];
#pragma unroll
        for (int t = 1; t < W; t++) {
            const int64_t r = rf - K + (t - 1);
            win[t] = node_row(r) ? ld4(B + r * ld + c) : zero4<T>();
        }
        // rolled (the source pass would be replicated per row otherwise); the look-ahead row is
        // taken over (window shifted, operands moved) after the row's wait and before its store,
        // so the back edge carries no outstanding load
        Row1 cur, nxt;
        issue1(cur, q0);
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
#pragma unroll
        for (int t = 0; t + 1 < W; t++) win[t] = win[t + 1];
        win[W - 1] = cur.nb;
#pragma unroll 1
        for (int q = q0; q < q1; q++) {
            if (q + 1 < q1) issue1(nxt, q + 1);
            const int64_t r1 = R0 - K + q;
            const bool row1 = r1 >= 1 && r1 <= nr - 2;
            C4<T> out = zero4<T>();
            if (row1) {
                T kc[4];
#pragma unroll
                for (int e = 0; e < 4; e++) kc[e] = VARC ? cur.kv.e[e] : a.kconst;
                T b[4] = {T(0), T(0), T(0), T(0)};
                bool damp = false;
                if constexpr (SPONGE) {
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        b[e] = cur.br + bc[e];
                        damp |= b[e] != T(0);
                    }
                }
                T val[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const T lap = lap_elem<T, K, true>(a.wc, a.wr, a.wcol, win, 0, cur.L, cur.R, e);
                    const T t = damp ? T(2) * win[K].e[e] - (T(1) - b[e]) * cur.P.e[e]
                                     : T(2) * win[K].e[e] - cur.P.e[e];
                    val[e] = fmaT(kc[e], lap, t);
                }
                if (src1) wave_inject(a.src, shot, r1, base, 0, 63, step, lane, kc, val);
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const T x = damp ? val[e] / (T(1) + b[e]) : val[e];
                    out.e[e] = valid[e] ? x : T(0);
                }
            }
            band[q][lane] = out;
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
#pragma unroll
            for (int t = 0; t + 1 < W; t++) win[t] = win[t + 1];
            win[W - 1] = nxt.nb;
            cur = nxt;
            if (row1 && own && r1 >= R0 && r1 < R0 + 4 * RI && r1 < a.row_hi) {
                const T v[4] = {out.e[0], out.e[1], out.e[2], out.e[3]};
                st4(C + r1 * ld + c, v);
            }
        }
    }

    // ---- step 2: rows R0 + wave RI .. + RI - 1, band rows q = r2 - (R0 - K)
    const int64_t r2f = R0 + (int64_t)wave * RI;
    struct Row2 {
        C4<T> P, kv;
        T br;
    };
    auto issue2 = [&](Row2& o, int64_t r2) {
        const int64_t off = r2 * ld + c;
        o.P = load_ok ? ld4(B + off) : zero4<T>();  // u_prev of step 2 is u^n
        o.kv = zero4<T>();
        if constexpr (VARC) o.kv = load_ok ? ld4(coef + off) : zero4<T>();
        o.br = T(0);
        if constexpr (SPONGE) o.br = a.beta_r[r2];
    };
    Row2 op2[RI];
    if (r2f < a.row_hi) issue2(op2[0], r2f);
    lds_barrier();
    if (r2f >= a.row_hi) return;  // whole wave; no barrier follows
    const int qf = wave * RI;  // band row of r2f - K
    const int64_t r2l = r2f + RI < a.row_hi ? r2f + RI : a.row_hi;
    const bool src2 = rp && rp[clamp_row(r2f)] != rp[clamp_row(r2l)];
    C4<T> sw[W];
#pragma unroll
    for (int t = 1; t < W; t++) sw[t] = band[qf + t - 1][lane];
#pragma unroll
    for (int i = 0; i < RI; i++) {
        const int64_t r2 = r2f + i;
        if (r2 >= a.row_hi) break;  // whole wave
        if (i + 1 < RI && r2 + 1 < a.row_hi) issue2(op2[i + 1], r2 + 1);
#pragma unroll
        for (int t = 0; t + 1 < W; t++) sw[t] = sw[t + 1];
        sw[W - 1] = band[qf + i + 2 * K][lane];
        const int64_t o = r2 * ld + c;
        const C4<T> L2 = from_lane_left(sw[K]), R2 = from_lane_right(sw[K]);
        T kc2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) kc2[e] = VARC ? op2[i].kv.e[e] : a.kconst;
        T b2[4] = {T(0), T(0), T(0), T(0)};
        bool damp2 = false;
        if constexpr (SPONGE) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                b2[e] = op2[i].br + bc[e];
                damp2 |= b2[e] != T(0);
            }
        }
        T val2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T lap = lap_elem<T, K, true>(a.wc, a.wr, a.wcol, sw, 0, L2, R2, e);
            const T t = damp2 ? T(2) * sw[K].e[e] - (T(1) - b2[e]) * op2[i].P.e[e] : T(2) * sw[K].e[e] - op2[i].P.e[e];
            val2[e] = fmaT(kc2[e], lap, t);
        }
        if (src2) wave_inject(a.src, shot, r2, base, 1, 62, step + 1, lane, kc2, val2);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const T x = damp2 ? val2[e] / (T(1) + b2[e]) : val2[e];
            val2[e] = valid[e] ? x : T(0);
        }
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        if (own) st4(D + o, val2);
    }
}

template <typename T, int K, bool VARC, bool SPONGE, int RI>
__global__ __launch_bounds__(256) void step2d_t2l_kernel(const StepArgs<T> a)
{
    __shared__ C4<T> band[4 * RI + 2 * K][64];
    const StepBlock bp = step_block(a.ngb, a.gx, a.gy);
    const int64_t step = a.step_dev ? *a.step_dev + a.step : a.step;
    if (bp.gather) {  // the gather blocks never reach the tile barrier
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
    t2l_tile<T, K, VARC, SPONGE, RI>(a, bp.x, bp.y, (int)bp.z, step, band);
}

template <typename T, int K, bool VARC, bool SPONGE, int RI>
__global__ __launch_bounds__(256) void step2d_t2_kernel(const StepArgs<T> a)
{
    const StepBlock bp = step_block(a.ngb, a.gx, a.gy);
    const int64_t step = a.step_dev ? *a.step_dev + a.step : a.step;
    if (bp.gather) {
        // first half: u^n into row step + gather_off; second half: u^{n-1} into the row before
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
    t2_tile<T, K, VARC, SPONGE, RI>(a, bp.x, bp.y, (int)bp.z, step);
}

template <typename T, int K, bool VARC, bool SPONGE>
static hipError_t launch_t2_k(const StepArgs<T>& a0, int rows, int nshots, hipStream_t s)
{
    StepArgs<T> a = a0;
    const int64_t nrows = a.row_hi - a.row_lo;
    a.gx = (uint32_t)((a.nc + T2_COLS - 1) / T2_COLS);
    auto grid = [&](int ri) {
        a.gy = (uint32_t)((nrows + 4 * ri - 1) / (4 * ri));
        return dim3(step_grid((uint32_t)((int64_t)a.gx * a.gy * nshots), a.ngb));
    };
    if (rows == 2 || rows == 4 || rows == 8) {  // LDS band: 4 * rows output rows x 248 columns per block
        a.gx = (uint32_t)((a.nc + 247) / 248);
        const int r = rows;
        a.gy = (uint32_t)((nrows + 4 * r - 1) / (4 * r));
        const dim3 g(step_grid((uint32_t)((int64_t)a.gx * a.gy * nshots), a.ngb));
        if (rows == 2)
            hipLaunchKernelGGL((step2d_t2l_kernel<T, K, VARC, SPONGE, 2>), g, dim3(256), 0, s, a);
        else if (rows == 4)
            hipLaunchKernelGGL((step2d_t2l_kernel<T, K, VARC, SPONGE, 4>), g, dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((step2d_t2l_kernel<T, K, VARC, SPONGE, 8>), g, dim3(256), 0, s, a);
        return hipGetLastError();
    }
    switch (rows) {  // register march: 16 or 32 output rows per wave
        case 32: hipLaunchKernelGGL((step2d_t2_kernel<T, K, VARC, SPONGE, 32>), grid(32), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((step2d_t2_kernel<T, K, VARC, SPONGE, 16>), grid(16), dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

template <typename T, int K>
static hipError_t launch_t2_flags(const StepArgs<T>& a, bool varc, bool sponge, int rows, int nshots, hipStream_t s)
{
    if (varc) {
        if (sponge) return launch_t2_k<T, K, true, true>(a, rows, nshots, s);
        return launch_t2_k<T, K, true, false>(a, rows, nshots, s);
    }
    if (sponge) return launch_t2_k<T, K, false, true>(a, rows, nshots, s);
    return launch_t2_k<T, K, false, false>(a, rows, nshots, s);
}

template <typename T>
hipError_t launch_step2d_t2(const StepArgs<T>& a, int order, bool varc, bool sponge, int rows, int nshots,
                            hipStream_t s)
{
    switch (order) {
        case 2: return launch_t2_flags<T, 1>(a, varc, sponge, rows, nshots, s);
        case 4: return launch_t2_flags<T, 2>(a, varc, sponge, rows, nshots, s);
        case 6: return launch_t2_flags<T, 3>(a, varc, sponge, rows, nshots, s);
        default: return launch_t2_flags<T, 4>(a, varc, sponge, rows, nshots, s);
    }
}

template hipError_t launch_step2d_t2<float>(const StepArgs<float>&, int, bool, bool, int, int, hipStream_t);
template hipError_t launch_step2d_t2<double>(const StepArgs<double>&, int, bool, bool, int, int, hipStream_t);

}  // namespace wfd
