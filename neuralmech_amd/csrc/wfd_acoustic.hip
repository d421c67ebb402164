// Variable mass / stiffness wave on the 2D halo layout (orders 2..8) and its adjoint with
// boundary reconstruction of the forward field.
//
// Forward (per interior node x, step n; K = order / 2 bonds per side and axis):
//     q     = sum over rows then columns of sum_m (c_m,near d_m + c_m,far d_m)   (see wavefd.h)
//     t     = 2 p[x] - (1 - b) p_prev[x]               (2 p[x] - p_prev[x] without sponge)
//     num   = t + k[x] q;  for each source s at x (list order): num = num + k[x] sig[n][s]
//     p_next[x] = num / (1 + b)                        (num without sponge)
// The bond operator C is symmetric: sum_x mu_x (C p)_x = -sum_bonds c (dp)(dmu).  Mass,
// volume weights, spacing, closure and the averaging of the stiffness onto bonds are all in k
// and c (set by the caller).
//
// Backward step n (n = N .. 1), from (p^{n+1}, p^n) and (mu^{n+2}, mu^{n+1}):
//     mu^n      = the forward update of mu with the adjoint sources adj[n] (= -dJ/dp^n)
//     p^{n-1}   = (2 p^n + k q(p^n)) + k sig[n] - p^{n+1}       on undamped interior nodes
//     Gm       += mu^{n+1} q(p^n)                                on undamped interior nodes
//     F_bond   += (dp^n)_bond (dmu^{n+1})_bond                   on bonds with an undamped end
// with p^{n-1} on the damped nodes within K of undamped ones ("strip") restored from the values
// the forward recorded.  Then dJ/dk = -(Gm + sum_n mu^{n+1} sig^n) / k and dJ/dc = F.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <mutex>
#include <vector>

#include <cstring>

#include "wfd_internal.h"
#include "wfd_points_dev.h"

namespace wfd {
namespace {


template <typename T>
struct AcArgs {
    const T* up;  // forward: p^{n-1};  backward: p^{n+1} (overwritten in place by p^{n-1})
    const T* uc;  // p^n
    T* un;
    const T* mp;  // backward: mu^{n+2} (overwritten in place by mu^n)
    const T* mc;  // backward: mu^{n+1}
    T* mn;
    const T* k;   // node (0,0) of the shared coefficient fields
    const T* cn;  // K fields, cstride apart: bond to the north neighbour at distance m
    const T* cw;  // K fields: bond to the west neighbour at distance m
    const T* br;  // sponge profile of the row axis (n0 entries) and column axis (n1), or null
    const T* bc;
    T* gm;       // imaging partials (backward): mass term, then K north-bond and K west-bond fields
    double* Gm;  // float64 sums of the partials, same layout
    int32_t flush;  // add the partials to the sums and store zero partials
    int32_t recon;  // backward: reconstruct p^{n-1} and image (0 on the first backward step)
    int64_t shot_stride;
    int64_t cstride;  // elements between the K bond-coefficient fields (one field)
    int64_t pstride;  // elements between the 2K + 1 imaging fields (nshots fields)
    int64_t n0, n1, ld;
    int32_t K;        // bonds per side and axis (order / 2)
    PtsDev<T> src;  // injected into p (forward and reconstruction), row `step`
    PtsDev<T> adj;  // injected into mu (backward), row `step`
    PtsDev<T> ga;   // gather: field ga_f at ga's points -> row step + ga_off
    PtsDev<T> gb;   // gather: field gb_f at gb's points -> row step + gb_off
    PtsDev<T> sc;   // scatter: row step + sc_off of sc's data -> un at sc's points
    const T* ga_f;
    const T* gb_f;
    int64_t ga_off, gb_off, sc_off;
    uint32_t nga, ngb, nsc;  // aux blocks (256 points each)
    uint32_t gx, gy;
    int32_t nshots;
    int64_t step;
};

// rows per wave forced by wfd_ac_set_rows (0: chosen per launch)
std::atomic<int> g_ac_rows{0};

template <typename T>
__device__ __forceinline__ T ld0(const T* f, int64_t o, bool in)
{
    return in ? f[o] : T(0);
}

// Copy between a field and a point list's data row t, 256 points of the (shot, point) list per block.
template <typename T>
__device__ __forceinline__ void ac_points(const AcArgs<T>& a, const PtsDev<T>& p, uint32_t blk, int64_t t,
                                          const T* from, T* to)
{
    if (t < 0 || t >= p.nt) return;
    const int64_t idx = (int64_t)blk * 256 + threadIdx.x;
    const int64_t shot = idx / p.count, q = idx % p.count;
    if (shot >= a.nshots) return;
    const int32_t* rp = p.row_ptr + shot * (p.nrows + 1);
    if (q >= rp[p.nrows]) return;
    const int64_t pq = shot * p.count + q;
    const int64_t off = shot * a.shot_stride + (int64_t)p.row[pq] * a.ld + p.col[pq];
    T* d = p.data + shot * p.shot_stride + t * p.width + p.id[pq];
    if (from)
        *d = from[off];
    else
        to[off] = *d;
}

// num + k * sig for every point of CSR row r in this wave's 64 columns [j0, j0 + 64), in list
// order.  The row's points are read 64 at a time, one per lane; a ballot selects those of this
// wave.  All 64 lanes of the wave take part.
template <typename T>
__device__ __forceinline__ T ac_inject(const PtsDev<T>& p, int64_t shot, int64_t r, int64_t j0, int lane, int64_t t,
                                       T kx, T num)
{
    if (p.count == 0 || t < 0 || t >= p.nt) return num;
    const int32_t* rp = p.row_ptr + shot * (p.nrows + 1);
    const int lo = rp[r], hi = rp[r + 1];
    if (lo >= hi) return num;
    const int32_t* col = p.col + shot * p.count;
    const int32_t* id = p.id + shot * p.count;
    const T* sig = p.data + shot * p.shot_stride + t * p.width;
    for (int b = lo; b < hi; b += 64) {
        const int q = b + lane;
        const int cl = q < hi ? col[q] : -1;
        const int pid = q < hi ? id[q] : -1;
        const T s = pid >= 0 ? sig[pid] : T(0);
        const int64_t off = cl >= 0 ? (int64_t)cl - j0 : -1;
        const bool mine = off >= 0 && off < 64;
        uint64_t m = __ballot(mine);
        const int offi = mine ? (int)off : 0;
        while (m) {
            const int pl = __builtin_ctzll(m);
            m &= m - 1;
            const int o = lane_read(offi, pl);
            const T ts = lane_read(s, pl);
            if (lane == o) num = num + kx * ts;
        }
    }
    return num;
}

// part[o] = v (the loaded partial plus the row's term), or with flush: sum[o] += v, part[o] = 0.
template <typename T>
__device__ __forceinline__ void acc_store(T* part, double* sum, int64_t o, T v, bool flush)
{
    if (flush) {
        sum[o] += (double)v;
        part[o] = T(0);
    } else {
        part[o] = v;
    }
}

// q = sum over the two axes of sum_m (c_m(north/west) d_m + c_m(south/east) d_m), accumulated
// m = 1..K per axis, rows before columns (for K = 1: (cN dN + cS dS) + (cW dW + cE dE)).
template <typename T, int K>
__device__ __forceinline__ T ac_q(const T (&cn)[K], const T (&cs)[K], const T (&cw)[K], const T (&ce)[K],
                                  const T (&dn)[K], const T (&ds)[K], const T (&dw)[K], const T (&de)[K])
{
    T qx = cn[0] * dn[0] + cs[0] * ds[0];
#pragma unroll
    for (int m = 1; m < K; m++) qx = qx + (cn[m] * dn[m] + cs[m] * ds[m]);
    T qy = cw[0] * dw[0] + ce[0] * de[0];
#pragma unroll
    for (int m = 1; m < K; m++) qy = qy + (cw[m] * dw[m] + ce[m] * de[m]);
    return qx + qy;
}

// Column neighbours at distance 1..K of this lane's node (value c): from lanes l -+ m, or the
// loaded values w0 / e0 for the lanes whose neighbour lies in another wave's columns.  Every lane
// of the wave takes part in the shuffles.
template <typename T, int K>
__device__ __forceinline__ void ac_cols(int lane, bool in, T c, const T (&w0)[K], const T (&e0)[K], T (&w)[K],
                                        T (&e)[K])
{
#pragma unroll
    for (int m = 1; m <= K; m++) {
        w[m - 1] = __shfl_up(c, m);
        e[m - 1] = __shfl_down(c, m);
        if (lane < m && in) w[m - 1] = w0[m - 1];
        if (lane >= 64 - m && in) e[m - 1] = e0[m - 1];
    }
}

// Everything a row of the march loads, issued one row ahead.
template <typename T, int K>
struct AcRow {
    T p, m;               // window rows r + K of p and mu
    T cn[K];              // cn_m at row r + m
    T cw[K], ce[K];       // cw_m at (r, j), and at (r, j + m) for the last m lanes
    T pw[K], pe[K];       // p at (r, j - m) for the first m lanes, at (r, j + m) for the last m
    T mw[K], me[K];       // mu likewise (backward)
    T k, prev, upv;       // k; p^{n-1} (forward) or mu^{n+2} (backward); p^{n+1} (reconstruction)
    T g0, gn[K], gw[K];   // imaging partials at the node (mass, north and west bonds)
    T gnf[K], gwf[K];     // imaging partials of the bonds stored at their far end
};

// Launch: 8 * ceil(aux / 8) aux blocks then gx * gy * nshots tiles (see step_block); a tile is 64
// columns x 4 waves x RI rows.  K = order / 2 bonds per side and axis.
//
// A wave marches its RI rows with rolling register windows and loads each row's operands (AcRow)
// before any of the row's stores -- one row ahead where that pays (see LA) -- then waits for
// everything outstanding right before the row's stores: the loads are in flight through the
// row's arithmetic (the next row's, with the look-ahead), and a wait placed after a store would
// also wait for the store.  Rows with sources are found once per wave.  Stores and in-place
// partners (u_next / u_prev, mu_n / mu_prev) touch each element from one thread only, and every
// imaging partial (node or far end of a bond) is read and written by one thread, at one row.
template <typename T, int K, bool SPONGE, bool BWD, int RI>
__global__ __launch_bounds__(256) void ac_kernel(const AcArgs<T> a)
{
    const StepBlock bp = step_block(a.nga + a.ngb + a.nsc, a.gx, a.gy);
    if (bp.gather) {
        uint32_t g = bp.g;
        if (g < a.nga) return ac_points(a, a.ga, g, a.step + a.ga_off, a.ga_f, (T*)nullptr);
        g -= a.nga;
        if (g < a.ngb) return ac_points(a, a.gb, g, a.step + a.gb_off, a.gb_f, (T*)nullptr);
        g -= a.ngb;
        if (g < a.nsc) return ac_points(a, a.sc, g, a.step + a.sc_off, (const T*)nullptr, a.un);
        return;
    }
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t n0 = a.n0, n1 = a.n1, ld = a.ld, cst = a.cstride, pst = a.pstride;
    const int64_t shot = bp.z;
    const int64_t j0 = (int64_t)bp.x * 64, j = j0 + lane;
    const int64_t r0 = 1 + ((int64_t)bp.y * 4 + wave) * RI;
    if (r0 > n0 - 2) return;  // whole wave
    const int64_t rl = min(r0 + RI - 1, n0 - 2);  // last row of the wave
    const bool in = j < n1;
    const bool inter = j >= 1 && j <= n1 - 2;
    const int64_t so = shot * a.shot_stride;
    const T* uc = a.uc + so;
    const T* mc = BWD ? a.mc + so : nullptr;
    const T bcj = SPONGE && in ? a.bc[j] : T(0);
    // sponge rows r0 - K .. rl and columns j - 1 .. j - K (bonds to damped neighbours)
    T brw[RI + K], bcw[K];
#pragma unroll
    for (int t = 0; t < RI + K; t++) brw[t] = SPONGE && r0 - K + t >= 0 && r0 - K + t <= rl ? a.br[r0 - K + t] : T(0);
#pragma unroll
    for (int m = 1; m <= K; m++) bcw[m - 1] = SPONGE && BWD && j - m >= 0 && in ? a.bc[j - m] : T(0);
    // rows of this wave with points of the source / adjoint lists (one check per wave)
    auto has_rows = [&](const PtsDev<T>& p) {
        if (p.count == 0 || a.step < 0 || a.step >= p.nt) return false;
        const int32_t* rp = p.row_ptr + shot * (p.nrows + 1);
        return rp[r0] != rp[rl + 1];
    };
    const bool src_w = has_rows(a.src);
    const bool adj_w = BWD && has_rows(a.adj);
    const bool fl = BWD && a.flush != 0;

    auto load_row = [&](AcRow<T, K>& R, int s) {
        const int64_t r = r0 + s;
        const int64_t o = r * ld + j;
        R.p = ld0(uc, o + K * ld, in);
        R.m = BWD ? ld0(mc, o + K * ld, in) : T(0);
#pragma unroll
        for (int m = 1; m <= K; m++) {
            const T* cwm = a.cw + (m - 1) * cst;
            R.cn[m - 1] = ld0(a.cn + (m - 1) * cst, o + m * ld, in);
            // bonds of the last interior columns reach up to K-1 columns past the nodes (< ld)
            R.cw[m - 1] = ld0(cwm, o, j < n1 + K - 1);
            R.ce[m - 1] = lane >= 64 - m && in ? cwm[o + m] : T(0);
            R.pw[m - 1] = lane < m && in ? uc[o - m] : T(0);
            R.pe[m - 1] = lane >= 64 - m && in ? uc[o + m] : T(0);
            R.mw[m - 1] = BWD && lane < m && in ? mc[o - m] : T(0);
            R.me[m - 1] = BWD && lane >= 64 - m && in ? mc[o + m] : T(0);
        }
        R.k = ld0(a.k, o, in);
        R.prev = BWD ? ld0(a.mp + so, o, in) : ld0(a.up + so, o, in);
        if (BWD && a.recon) {
            const bool und = !SPONGE || brw[s + K] + bcj == T(0);
            const bool ui = inter && und;
            R.upv = ui ? a.up[so + o] : T(0);
            R.g0 = ui ? a.gm[so + o] : T(0);
#pragma unroll
            for (int m = 1; m <= K; m++) {
                const T* fn = a.gm + m * pst + so;
                const T* fw = a.gm + (K + m) * pst + so;
                const bool nund = r - m >= 1 && (!SPONGE || brw[s + K - m] + bcj == T(0));
                const bool wund = j - m >= 1 && (!SPONGE || brw[s + K] + bcw[m - 1] == T(0));
                R.gn[m - 1] = inter && (und || nund) ? fn[o] : T(0);
                R.gw[m - 1] = inter && (und || wund) ? fw[o] : T(0);
                R.gnf[m - 1] = ui && r + m >= n0 - 1 ? fn[o + m * ld] : T(0);
                R.gwf[m - 1] = ui && j + m >= n1 - 1 ? fw[o + m] : T(0);
            }
        }
    };

    // rolling windows: rows r-K .. r+K of p (and mu), bond coefficients cn_m of rows r .. r+m
    T P[2 * K + 1], M[2 * K + 1], CN[K][K + 1];
#pragma unroll
    for (int t = 0; t < 2 * K; t++) {
        P[t] = ld0(uc, (r0 - K + t) * ld + j, in);
        M[t] = BWD ? ld0(mc, (r0 - K + t) * ld + j, in) : T(0);
    }
#pragma unroll
    for (int m = 1; m <= K; m++)
#pragma unroll
        for (int t = 0; t < m; t++) CN[m - 1][t] = ld0(a.cn + (m - 1) * cst, (r0 + t) * ld + j, in);
    // One row ahead: adjoint rows at orders 2 and 4 (higher orders: two rows' worth of operands
    // would cost waves), forward rows of 2-row waves (small grids, latency-bound: 8 x 512^2 order 4
    // plate iteration 122.5 -> 110.5 ms; 4-row waves of the 3300 x 1800 grid ran 3 % slower
    // ahead).  The others load a row's operands at its start.
    constexpr int LA = (BWD ? K <= 2 : RI <= 2) ? 1 : 0;
    AcRow<T, K> rows[RI];
    if (LA) load_row(rows[0], 0);
#pragma unroll
    for (int s = 0; s < RI; s++) {
        const int64_t r = r0 + s;
        if (r > rl) break;  // whole wave
        if (!LA) load_row(rows[s], s);
        else if (s + 1 < RI && r + 1 <= rl) load_row(rows[s + 1], s + 1);
        const AcRow<T, K>& R = rows[s];
        const int64_t o = r * ld + j;
        P[2 * K] = R.p;
        if (BWD) M[2 * K] = R.m;
#pragma unroll
        for (int m = 1; m <= K; m++) CN[m - 1][m] = R.cn[m - 1];
        const T pc = P[K];
        T pw[K], pe[K], cwv[K], cev[K], cnv[K], csv[K], dN[K], dS[K], dW[K], dE[K];
        ac_cols<T, K>(lane, in, pc, R.pw, R.pe, pw, pe);
#pragma unroll
        for (int m = 1; m <= K; m++) {
            cwv[m - 1] = R.cw[m - 1];
            cev[m - 1] = __shfl_down(cwv[m - 1], m);
            if (lane >= 64 - m && in) cev[m - 1] = R.ce[m - 1];
            cnv[m - 1] = CN[m - 1][0];
            csv[m - 1] = CN[m - 1][m];
            dN[m - 1] = P[K - m] - pc;
            dS[m - 1] = P[K + m] - pc;
            dW[m - 1] = pw[m - 1] - pc;
            dE[m - 1] = pe[m - 1] - pc;
        }
        const T kx = R.k;
        const T q = ac_q<T, K>(cnv, csv, cwv, cev, dN, dS, dW, dE);
        const T b = SPONGE ? brw[s + K] + bcj : T(0);
        if (!BWD) {
            const T t = SPONGE ? T(2) * pc - (T(1) - b) * R.prev : T(2) * pc - R.prev;
            T num = t + kx * q;
            if (src_w) num = ac_inject(a.src, shot, r, j0, lane, a.step, kx, num);
            const T outv = SPONGE ? num / (T(1) + b) : num;
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
            if (inter) a.un[so + o] = outv;
        } else {
            const T mcv = M[K];
            T mw[K], me[K], eN[K], eS[K], eW[K], eE[K];
            ac_cols<T, K>(lane, in, mcv, R.mw, R.me, mw, me);
#pragma unroll
            for (int m = 1; m <= K; m++) {
                eN[m - 1] = M[K - m] - mcv;
                eS[m - 1] = M[K + m] - mcv;
                eW[m - 1] = mw[m - 1] - mcv;
                eE[m - 1] = me[m - 1] - mcv;
            }
            const T qm = ac_q<T, K>(cnv, csv, cwv, cev, eN, eS, eW, eE);
            const T tm = SPONGE ? T(2) * mcv - (T(1) - b) * R.prev : T(2) * mcv - R.prev;
            T numm = tm + kx * qm;
            if (adj_w) numm = ac_inject(a.adj, shot, r, j0, lane, a.step, kx, numm);
            const T mnv = SPONGE ? numm / (T(1) + b) : numm;
            const bool und = !SPONGE || b == T(0);
            T unv = T(0), g0 = T(0), gn[K], gw[K], gnf[K], gwf[K];
            if (a.recon) {
                T rv = T(2) * pc + kx * q;
                if (src_w) rv = ac_inject(a.src, shot, r, j0, lane, a.step, kx, rv);
                unv = rv - R.upv;
                g0 = R.g0 + mcv * q;
#pragma unroll
                for (int m = 1; m <= K; m++) {
                    gn[m - 1] = R.gn[m - 1] + dN[m - 1] * eN[m - 1];
                    gw[m - 1] = R.gw[m - 1] + dW[m - 1] * eW[m - 1];
                    gnf[m - 1] = R.gnf[m - 1] + dS[m - 1] * eS[m - 1];
                    gwf[m - 1] = R.gwf[m - 1] + dE[m - 1] * eE[m - 1];
                }
            }
            __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
            if (inter) a.mn[so + o] = mnv;
            if (a.recon) {
                if (inter && und) {
                    a.un[so + o] = unv;
                    acc_store(a.gm + so, a.Gm + so, o, g0, fl);
                }
                if (inter) {
#pragma unroll
                    for (int m = 1; m <= K; m++) {
                        T* fn = a.gm + m * pst + so;
                        double* Fn = a.Gm + m * pst + so;
                        T* fw = a.gm + (K + m) * pst + so;
                        double* Fw = a.Gm + (K + m) * pst + so;
                        // a bond is imaged when either end is an undamped interior node
                        const bool nund = r - m >= 1 && (!SPONGE || brw[s + K - m] + bcj == T(0));
                        if (und || nund) acc_store(fn, Fn, o, gn[m - 1], fl);
                        const bool wund = j - m >= 1 && (!SPONGE || brw[s + K] + bcw[m - 1] == T(0));
                        if (und || wund) acc_store(fw, Fw, o, gw[m - 1], fl);
                        // bonds reaching past the last interior row / column: stored at their far end
                        if (und && r + m >= n0 - 1) acc_store(fn, Fn, o + m * ld, gnf[m - 1], fl);
                        if (und && j + m >= n1 - 1) acc_store(fw, Fw, o + m, gwf[m - 1], fl);
                    }
                }
            }
        }
#pragma unroll
        for (int t = 0; t < 2 * K; t++) {
            P[t] = P[t + 1];
            M[t] = M[t + 1];
        }
#pragma unroll
        for (int m = 1; m <= K; m++)
#pragma unroll
            for (int t = 0; t < m; t++) CN[m - 1][t] = CN[m - 1][t + 1];
    }
}

template <typename T, int K, bool SPONGE, bool BWD>
hipError_t ac_launch_k(AcArgs<T> a, bool tiles_on, hipStream_t s)
{
    // Rows per wave: 8, halved while the launch has fewer waves than a target.  A wave's row march
    // starts with a 2K-row window, so short marches re-read rows (from the caches) but put more
    // loads in flight.  Measured on MI355X: the forward step is fastest with about 16 K waves
    // (3300 x 1800 order 2: 21.9 us at 4 rows vs 23.2 at 8; 8 x 512^2 order 4: 14.9 us at 2 rows
    // vs 15.7 at 4); the adjoint wants 32 K waves while its 8 + 4K fields fit the Infinity Cache
    // (8 x 512^2 order 4: 31.6 us at 1 row vs 43.3 at 4) and 8 K when they stream from HBM
    // (3300 x 1800: 68.5 us at 8 rows vs 71.1 at 4).
    auto waves = [&](int ri) {
        return (uint64_t)a.gx * 4 * (uint64_t)((a.n0 - 2 + 4 * ri - 1) / (4 * ri)) * (uint64_t)a.nshots;
    };
    const uint64_t bwd_bytes = (uint64_t)(8 + 4 * K) * (uint64_t)a.shot_stride * (uint64_t)a.nshots * sizeof(T);
    const uint64_t target = !BWD ? 16384u : (bwd_bytes <= ((uint64_t)192 << 20) ? 32768u : 8192u);
    int ri = 8;
    while (ri > 1 && waves(ri) < target) ri /= 2;
    // orders 6 and 8: an adjoint row's operands take 40-54 registers; 2 rows per wave keep 4 waves
    // per SIMD (113-118 VGPRs), 4 or 8 rows fall to 2-3
    if (BWD && K >= 3 && ri > 2) ri = 2;
    const int forced = g_ac_rows.load(std::memory_order_relaxed);
    if (forced) ri = forced;
    a.gy = (uint32_t)((a.n0 - 2 + 4 * ri - 1) / (4 * ri));
    const uint32_t tiles = tiles_on ? a.gx * a.gy * (uint32_t)a.nshots : 0u;
    const uint32_t blocks = step_grid(tiles, a.nga + a.ngb + a.nsc);
    if (blocks == 0) return hipSuccess;
    switch (ri) {
        case 8: hipLaunchKernelGGL((ac_kernel<T, K, SPONGE, BWD, 8>), dim3(blocks), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((ac_kernel<T, K, SPONGE, BWD, 4>), dim3(blocks), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((ac_kernel<T, K, SPONGE, BWD, 2>), dim3(blocks), dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL((ac_kernel<T, K, SPONGE, BWD, 1>), dim3(blocks), dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

template <typename T, int K>
hipError_t ac_launch_kk(const AcArgs<T>& a, bool tiles, bool sp, bool bwd, hipStream_t s)
{
    if (bwd) return sp ? ac_launch_k<T, K, true, true>(a, tiles, s) : ac_launch_k<T, K, false, true>(a, tiles, s);
    return sp ? ac_launch_k<T, K, true, false>(a, tiles, s) : ac_launch_k<T, K, false, false>(a, tiles, s);
}

template <typename T>
hipError_t ac_launch(AcArgs<T> a, bool tiles_on, bool bwd, hipStream_t s)
{
    auto nb = [](const PtsDev<T>& p, int nshots) {
        return p.count ? (uint32_t)(((int64_t)p.count * nshots + 255) / 256) : 0u;
    };
    a.nga = a.ga_f ? nb(a.ga, a.nshots) : 0;
    a.ngb = a.gb_f ? nb(a.gb, a.nshots) : 0;
    a.nsc = a.sc.count ? nb(a.sc, a.nshots) : 0;
    a.gx = (uint32_t)((a.n1 + 63) / 64);
    const bool sp = a.br != nullptr;
    switch (a.K) {
        case 1: return ac_launch_kk<T, 1>(a, tiles_on, sp, bwd, s);
        case 2: return ac_launch_kk<T, 2>(a, tiles_on, sp, bwd, s);
        case 3: return ac_launch_kk<T, 3>(a, tiles_on, sp, bwd, s);
        default: return ac_launch_kk<T, 4>(a, tiles_on, sp, bwd, s);
    }
}

// ------------------------------------------------------------------------------ host side
int ac_fail(int code, const char* msg) { return set_error(code, msg); }

template <typename T>
int ac_pts(PtsDev<T>& p, const wfd_points* src, int32_t nrows, int nshots, const char* what)
{
    std::memset(&p, 0, sizeof p);
    if (!src || src->count == 0) return WFD_OK;
    if (src->count < 0 || src->nrows != nrows || !src->row_ptr || !src->col || !src->id || !src->row || !src->data)
        return ac_fail(WFD_EINVAL, what);
    if (src->width < src->count || src->nsteps < 0 || src->shot_stride < src->nsteps * src->width)
        return ac_fail(WFD_EINVAL, what);
    (void)nshots;
    p.row_ptr = src->row_ptr;
    p.col = src->col;
    p.id = src->id;
    p.row = src->row;
    p.data = (T*)src->data;
    p.width = src->width;
    p.shot_stride = src->shot_stride;
    p.nt = src->nsteps;
    p.count = src->count;
    p.nrows = src->nrows;
    return WFD_OK;
}

int ac_check(const wfd_grid* g, const wfd_ac_medium* med)
{
    if (!g || !med) return ac_fail(WFD_EINVAL, "wfd_ac: NULL grid or medium");
    if (g->ndim != 2) return ac_fail(WFD_EINVAL, "wfd_ac: 2D grids only");
    if (g->order != 2 && g->order != 4 && g->order != 6 && g->order != 8)
        return ac_fail(WFD_EINVAL, "wfd_ac: order must be 2, 4, 6 or 8");
    if (g->dtype != WFD_F32 && g->dtype != WFD_F64) return ac_fail(WFD_EINVAL, "wfd_ac: bad dtype");
    if (g->nshots < 1 || g->ld <= 0 || g->elems <= 0) return ac_fail(WFD_EINVAL, "wfd_ac: grid not initialised");
    if (!med->k || !med->cn || !med->cw) return ac_fail(WFD_EINVAL, "wfd_ac: NULL coefficient field");
    return WFD_OK;
}

template <typename T>
void ac_base(AcArgs<T>& a, const wfd_grid* g, const wfd_ac_medium* med)
{
    std::memset(&a, 0, sizeof a);
    const int64_t off = g->offset;
    a.k = (const T*)med->k + off;
    a.cn = (const T*)med->cn + off;
    a.cw = (const T*)med->cw + off;
    a.br = med->beta ? (const T*)med->beta : nullptr;
    a.bc = med->beta ? (const T*)med->beta + g->n[0] : nullptr;
    a.shot_stride = g->elems;
    a.cstride = g->elems;
    a.pstride = g->elems * g->nshots;
    a.K = g->order / 2;
    a.n0 = g->n[0];
    a.n1 = g->n[1];
    a.ld = g->ld;
    a.nshots = g->nshots;
}

template <typename T>
int ac_forward_t(const wfd_grid* g, void* ua, void* ub, const wfd_ac_medium* med, const wfd_points* src,
                 const wfd_points* rec, const wfd_points* strip, int64_t step0, int64_t nsteps, hipStream_t s)
{
    AcArgs<T> a;
    ac_base(a, g, med);
    const int32_t nr = (int32_t)g->n[0];
    int rc = ac_pts(a.src, src, nr, g->nshots, "wfd_ac_forward: bad source list");
    if (!rc) rc = ac_pts(a.ga, rec, nr, g->nshots, "wfd_ac_forward: bad receiver list");
    if (!rc) rc = ac_pts(a.gb, strip, nr, g->nshots, "wfd_ac_forward: bad strip list");
    if (rc) return rc;
    T* A = (T*)ua + g->offset;
    T* B = (T*)ub + g->offset;
    a.ga_off = -1;  // receivers: u^n into row n - 1 (trace row n holds u^{n+1})
    a.gb_off = 0;   // strip: u^n into row n
    for (int64_t i = 0; i < nsteps; i++) {
        a.up = (i & 1) ? B : A;
        a.uc = (i & 1) ? A : B;
        a.un = (T*)a.up;
        a.ga_f = a.ga.count ? a.uc : nullptr;
        a.gb_f = a.gb.count ? a.uc : nullptr;
        a.step = step0 + i;
        hipError_t e = ac_launch(a, true, false, s);
        if (e != hipSuccess) return ac_fail(WFD_EHIP, "wfd_ac_forward: launch failed");
    }
    // the newest level into its rows
    const T* last = (nsteps & 1) ? A : B;
    a.ga_f = a.ga.count ? last : nullptr;
    a.gb_f = a.gb.count ? last : nullptr;
    a.step = step0 + nsteps;
    a.sc.count = 0;
    if (ac_launch(a, false, false, s) != hipSuccess) return ac_fail(WFD_EHIP, "wfd_ac_forward: gather failed");
    return WFD_OK;
}

template <typename T>
int ac_backward_t(const wfd_grid* g, void* ua, void* ub, void* ma, void* mb, const wfd_ac_medium* med,
                  const wfd_points* src, const wfd_points* adj, const wfd_points* strip, const wfd_points* mu_src,
                  int64_t nsteps, void* gwork, double* G, hipStream_t s)
{
    AcArgs<T> a;
    ac_base(a, g, med);
    const int32_t nr = (int32_t)g->n[0];
    int rc = ac_pts(a.src, src, nr, g->nshots, "wfd_ac_backward: bad source list");
    if (!rc) rc = ac_pts(a.adj, adj, nr, g->nshots, "wfd_ac_backward: bad adjoint source list");
    if (!rc) rc = ac_pts(a.sc, strip, nr, g->nshots, "wfd_ac_backward: bad strip list");
    if (!rc) rc = ac_pts(a.ga, mu_src, nr, g->nshots, "wfd_ac_backward: bad mu_src list");
    if (rc) return rc;
    const int64_t fbytes = g->elems * g->nshots * (int64_t)sizeof(T);
    const int64_t field = g->elems * g->nshots;
    hipError_t e = launch_zero(ma, fbytes, s);
    if (e == hipSuccess) e = launch_zero(mb, fbytes, s);
    const int64_t nf = 2 * (g->order / 2) + 1;
    if (e == hipSuccess) e = launch_zero(gwork, nf * fbytes, s);
    if (e == hipSuccess) e = launch_zero(G, nf * field * (int64_t)sizeof(double), s);
    if (e != hipSuccess) return ac_fail(WFD_EHIP, "wfd_ac_backward: clearing work fields failed");
    const int64_t off = g->offset;
    T* PA = (T*)ua + off;  // p^{N-1}
    T* PB = (T*)ub + off;  // p^N
    T* MA = (T*)ma + off;
    T* MB = (T*)mb + off;
    a.gm = (T*)gwork + off;
    a.Gm = G + off;
    a.ga_off = 0;   // mu^{n+1} at the sources into row n
    a.sc_off = -1;  // strip row n - 1 into p^{n-1}
    int64_t imaged = 0;
    for (int64_t n = nsteps; n >= 1; n--) {
        const int64_t i = nsteps - n;
        a.mp = (i & 1) ? MB : MA;
        a.mc = (i & 1) ? MA : MB;
        a.mn = (T*)a.mp;
        a.step = n;
        a.recon = n < nsteps;
        if (a.recon) {
            const int64_t e2 = nsteps - 1 - n;
            a.up = (e2 & 1) ? PA : PB;
            a.uc = (e2 & 1) ? PB : PA;
            imaged++;
            a.flush = (imaged % 32 == 0) || n == 1;
        } else {
            a.up = PB;
            a.uc = PA;
            a.flush = 0;
        }
        a.un = (T*)a.up;
        a.ga_f = a.ga.count ? a.mc : nullptr;
        PtsDev<T> sc = a.sc;
        if (!a.recon) a.sc.count = 0;
        e = ac_launch(a, true, true, s);
        a.sc = sc;
        if (e != hipSuccess) return ac_fail(WFD_EHIP, "wfd_ac_backward: launch failed");
    }
    if (nsteps >= 1 && a.ga.count) {  // mu^1 at the sources into row 0
        AcArgs<T> f = a;
        f.ga_f = ((nsteps - 1) & 1) ? MB : MA;
        f.step = 0;
        f.sc.count = 0;
        f.gb.count = 0;
        if (ac_launch(f, false, true, s) != hipSuccess) return ac_fail(WFD_EHIP, "wfd_ac_backward: gather failed");
    }
    return WFD_OK;
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" int wfd_ac_forward(const wfd_grid* g, void* u_a, void* u_b, const wfd_ac_medium* med,
                              const wfd_points* src, const wfd_points* rec, const wfd_points* strip, int64_t step0,
                              int64_t nsteps, wfd_stream_t stream)
{
    int rc = ac_check(g, med);
    if (rc) return rc;
    if (!u_a || !u_b || u_a == u_b) return ac_fail(WFD_EINVAL, "wfd_ac_forward: NULL or aliased fields");
    if (nsteps < 0 || step0 < 0) return ac_fail(WFD_EINVAL, "wfd_ac_forward: negative step count");
    hipStream_t s = (hipStream_t)stream;
    return g->dtype == WFD_F64 ? ac_forward_t<double>(g, u_a, u_b, med, src, rec, strip, step0, nsteps, s)
                               : ac_forward_t<float>(g, u_a, u_b, med, src, rec, strip, step0, nsteps, s);
}

extern "C" int wfd_ac_backward(const wfd_grid* g, void* u_a, void* u_b, void* mu_a, void* mu_b,
                               const wfd_ac_medium* med, const wfd_points* src, const wfd_points* adj,
                               const wfd_points* strip, const wfd_points* mu_src, int64_t nsteps, void* g_work,
                               double* G, wfd_stream_t stream)
{
    int rc = ac_check(g, med);
    if (rc) return rc;
    if (!u_a || !u_b || !mu_a || !mu_b || !g_work || !G)
        return ac_fail(WFD_EINVAL, "wfd_ac_backward: NULL field");
    const void* f[4] = {u_a, u_b, mu_a, mu_b};
    for (int x = 0; x < 4; x++)
        for (int y = x + 1; y < 4; y++)
            if (f[x] == f[y]) return ac_fail(WFD_EINVAL, "wfd_ac_backward: aliased fields");
    if (nsteps < 0) return ac_fail(WFD_EINVAL, "wfd_ac_backward: negative step count");
    hipStream_t s = (hipStream_t)stream;
    return g->dtype == WFD_F64
               ? ac_backward_t<double>(g, u_a, u_b, mu_a, mu_b, med, src, adj, strip, mu_src, nsteps, g_work, G, s)
               : ac_backward_t<float>(g, u_a, u_b, mu_a, mu_b, med, src, adj, strip, mu_src, nsteps, g_work, G, s);
}

// ------------------------------------------------ allocation-free adjoint (file-scope scratch)
// The work fields of wfd_ac_backward (mu_a, mu_b and the 2K+1 partial-sum fields) in one
// file-scope device array: each device holds its own copy until the process ends.  Host code
// never uses the array's name as an address; it uses the device address hipGetSymbolAddress
// returns for the current device, looked up once per device.
static __device__ __align__(256) unsigned char scratch[WFD_AC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;  // per device, nullptr until looked up

int scratch_address(void** out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0) return ac_fail(WFD_EHIP, "wfd_ac_backward_noalloc: hipGetDevice failed");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            char msg[160];
            snprintf(msg, sizeof msg, "wfd_ac_backward_noalloc: hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
            return ac_fail(WFD_EHIP, msg);
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return WFD_OK;
}

}  // namespace

extern "C" int wfd_ac_backward_noalloc(const wfd_grid* g, void* u_a, void* u_b, const wfd_ac_medium* med,
                                       const wfd_points* src, const wfd_points* adj, const wfd_points* strip,
                                       const wfd_points* mu_src, int64_t nsteps, double* G, wfd_stream_t stream)
{
    int rc = ac_check(g, med);
    if (rc) return rc;
    if (!u_a || !u_b || u_a == u_b || !G) return ac_fail(WFD_EINVAL, "wfd_ac_backward_noalloc: NULL or aliased field");
    if (nsteps < 0) return ac_fail(WFD_EINVAL, "wfd_ac_backward_noalloc: negative step count");
    // mu_a, mu_b, then the 2K+1 partial-sum fields, each 256-byte aligned
    const uint64_t es = g->dtype == WFD_F64 ? 8 : 4;
    const uint64_t fb = ((uint64_t)g->elems * (uint64_t)g->nshots * es + 255) / 256 * 256;
    const uint64_t nf = 2 * (uint64_t)(g->order / 2) + 1;
    const uint64_t need = (2 + nf) * fb;
    if (need > (uint64_t)WFD_AC_SCRATCH_BYTES) {
        char msg[240];
        snprintf(msg, sizeof msg,
                 "wfd_ac_backward_noalloc: the work fields need %llu bytes; the allocation-free variant holds "
                 "at most %llu (use wfd_ac_backward with a workspace)",
                 (unsigned long long)need, (unsigned long long)WFD_AC_SCRATCH_BYTES);
        return ac_fail(WFD_EINVAL, msg);
    }
    void* base = nullptr;
    rc = scratch_address(&base);
    if (rc) return rc;
    char* p = static_cast<char*>(base);
    hipStream_t s = (hipStream_t)stream;
    // the kernels clear mu_a, mu_b and the partial sums before use (fill kernels, no memset)
    return g->dtype == WFD_F64
               ? ac_backward_t<double>(g, u_a, u_b, p, p + fb, med, src, adj, strip, mu_src, nsteps, p + 2 * fb, G, s)
               : ac_backward_t<float>(g, u_a, u_b, p, p + fb, med, src, adj, strip, mu_src, nsteps, p + 2 * fb, G, s);
}

extern "C" int wfd_ac_set_rows(int32_t rows)
{
    if (rows != 0 && rows != 1 && rows != 2 && rows != 4 && rows != 8)
        return set_error(WFD_EINVAL, "wfd_ac_set_rows: rows must be 0, 1, 2, 4 or 8");
    g_ac_rows.store(rows, std::memory_order_relaxed);
    return WFD_OK;
}
