// Variable mass / stiffness wave on the 2D halo layout (order 2) and its adjoint with boundary
// reconstruction of the forward field.
//
// Forward (per interior node x, step n):
//     q     = (cn[x] (p[N] - p[x]) + cn[S] (p[S] - p[x])) + (cw[x] (p[W] - p[x]) + cw[E] (p[E] - p[x]))
//     t     = 2 p[x] - (1 - b) p_prev[x]               (2 p[x] - p_prev[x] without sponge)
//     num   = t + k[x] q;  for each source s at x (list order): num = num + k[x] sig[n][s]
//     p_next[x] = num / (1 + b)                        (num without sponge)
// N/S/W/E are the four neighbours; cn[y] is the coefficient of the face between y and its north
// neighbour, cw[y] of the face between y and its west neighbour, so the face operator C is
// symmetric: sum_x mu_x (C p)_x = -sum_faces c_f (dp)_f (dmu)_f.  Mass, volume weights, spacing,
// closure and the averaging of the stiffness onto faces are all in k and c (set by the caller).
//
// Backward step n (n = N .. 1), from (p^{n+1}, p^n) and (mu^{n+2}, mu^{n+1}):
//     mu^n      = the forward update of mu with the adjoint sources adj[n] (= -dJ/dp^n)
//     p^{n-1}   = (2 p^n + k q(p^n)) + k sig[n] - p^{n+1}       on undamped interior nodes
//     Gm       += mu^{n+1} q(p^n)                                on undamped interior nodes
//     F_face   += (dp^n)_f (dmu^{n+1})_f                         on faces with an undamped end
// with p^{n-1} on the damped nodes next to undamped ones ("strip") restored from the values the
// forward recorded.  Then dJ/dk = -(Gm + sum_n mu^{n+1} sig^n) / k and dJ/dc_f = F_f.
#include <hip/hip_runtime.h>

#include <cstring>

#include "wfd_internal.h"
#include "wfd_points_dev.h"

namespace wfd {
namespace {

constexpr int AC_RI = 8;  // rows marched per wave

template <typename T>
struct AcArgs {
    const T* up;  // forward: p^{n-1};  backward: p^{n+1} (overwritten in place by p^{n-1})
    const T* uc;  // p^n
    T* un;
    const T* mp;  // backward: mu^{n+2} (overwritten in place by mu^n)
    const T* mc;  // backward: mu^{n+1}
    T* mn;
    const T* k;  // node (0,0) of the shared coefficient fields
    const T* cn;
    const T* cw;
    const T* br;  // sponge profile of the row axis (n0 entries) and column axis (n1), or null
    const T* bc;
    T* gm;  // imaging partials (backward): mass term, north faces, west faces
    T* fn;
    T* fw;
    double* Gm;  // float64 sums of the partials
    double* Fn;
    double* Fw;
    int32_t flush;  // add the partials to the sums and store zero partials
    int32_t recon;  // backward: reconstruct p^{n-1} and image (0 on the first backward step)
    int64_t shot_stride;
    int64_t n0, n1, ld;
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

template <typename T>
__device__ __forceinline__ void acc(T* part, double* sum, int64_t o, T add, bool flush)
{
    const T v = part[o] + add;
    if (flush) {
        sum[o] += (double)v;
        part[o] = T(0);
    } else {
        part[o] = v;
    }
}

// Launch: 8 * ceil(aux / 8) aux blocks then gx * gy * nshots tiles (see step_block); a tile is 64
// columns x 4 waves x AC_RI rows.
template <typename T, bool SPONGE, bool BWD>
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
    const int64_t n0 = a.n0, n1 = a.n1, ld = a.ld;
    const int64_t shot = bp.z;
    const int64_t j0 = (int64_t)bp.x * 64, j = j0 + lane;
    const int64_t r0 = 1 + ((int64_t)bp.y * 4 + wave) * AC_RI;
    if (r0 > n0 - 2) return;  // whole wave
    const bool in = j < n1;
    const bool inter = j >= 1 && j <= n1 - 2;
    const int64_t so = shot * a.shot_stride;
    const T* uc = a.uc + so;
    const T* mc = BWD ? a.mc + so : nullptr;

    // column profile of this node and of its west neighbour
    T bcj = T(0), bcw = T(0);
    if (SPONGE) {
        bcj = in ? a.bc[j] : T(0);
        bcw = __shfl_up(bcj, 1);
        if (lane == 0) bcw = j0 > 0 ? a.bc[j0 - 1] : T(0);
    }
    // rolling rows: north neighbour, centre, north-face coefficient of the centre row
    T pN = ld0(uc, (r0 - 1) * ld + j, in), pC = ld0(uc, r0 * ld + j, in);
    T mN = T(0), mC = T(0);
    if (BWD) {
        mN = ld0(mc, (r0 - 1) * ld + j, in);
        mC = ld0(mc, r0 * ld + j, in);
    }
    T cN = ld0(a.cn, r0 * ld + j, in);
#pragma unroll
    for (int s = 0; s < AC_RI; s++) {
        const int64_t r = r0 + s;
        if (r > n0 - 2) break;  // whole wave
        const int64_t o = r * ld + j;
        const T pS = ld0(uc, o + ld, in);
        const T cS = ld0(a.cn, o + ld, in);
        const T cW = ld0(a.cw, o, in);
        T cE = __shfl_down(cW, 1);
        if (lane == 63) cE = j0 + 64 < n1 ? a.cw[r * ld + j0 + 64] : T(0);
        T pW = __shfl_up(pC, 1), pE = __shfl_down(pC, 1);
        if (lane == 0) pW = j0 > 0 ? uc[r * ld + j0 - 1] : T(0);
        if (lane == 63) pE = j0 + 64 < n1 ? uc[r * ld + j0 + 64] : T(0);
        const T kx = ld0(a.k, o, in);
        const T dN = pN - pC, dS = pS - pC, dW = pW - pC, dE = pE - pC;
        const T q = (cN * dN + cS * dS) + (cW * dW + cE * dE);
        const T b = SPONGE ? a.br[r] + bcj : T(0);
        if (!BWD) {
            const T pp = ld0(a.up + so, o, in);
            const T t = SPONGE ? T(2) * pC - (T(1) - b) * pp : T(2) * pC - pp;
            T num = t + kx * q;
            num = ac_inject(a.src, shot, r, j0, lane, a.step, kx, num);
            if (inter) a.un[so + o] = SPONGE ? num / (T(1) + b) : num;
        } else {
            const T mS = ld0(mc, o + ld, in);
            T mW = __shfl_up(mC, 1), mE = __shfl_down(mC, 1);
            if (lane == 0) mW = j0 > 0 ? mc[r * ld + j0 - 1] : T(0);
            if (lane == 63) mE = j0 + 64 < n1 ? mc[r * ld + j0 + 64] : T(0);
            const T eN = mN - mC, eS = mS - mC, eW = mW - mC, eE = mE - mC;
            const T qm = (cN * eN + cS * eS) + (cW * eW + cE * eE);
            const T mpv = ld0(a.mp + so, o, in);
            const T tm = SPONGE ? T(2) * mC - (T(1) - b) * mpv : T(2) * mC - mpv;
            T numm = tm + kx * qm;
            numm = ac_inject(a.adj, shot, r, j0, lane, a.step, kx, numm);
            if (inter) a.mn[so + o] = SPONGE ? numm / (T(1) + b) : numm;
            if (a.recon) {
                const bool und = !SPONGE || b == T(0);
                T rv = T(2) * pC + kx * q;
                rv = ac_inject(a.src, shot, r, j0, lane, a.step, kx, rv);
                const bool fl = a.flush != 0;
                if (inter && und) {
                    a.un[so + o] = rv - a.up[so + o];
                    acc(a.gm + so, a.Gm + so, o, mC * q, fl);
                }
                if (inter) {
                    // north face: imaged when either end is an undamped interior node
                    const bool nund = r - 1 >= 1 && (!SPONGE || a.br[r - 1] + bcj == T(0));
                    if (und || nund) acc(a.fn + so, a.Fn + so, o, dN * eN, fl);
                    const bool wund = j - 1 >= 1 && (!SPONGE || a.br[r] + bcw == T(0));
                    if (und || wund) acc(a.fw + so, a.Fw + so, o, dW * eW, fl);
                    // faces to the boundary ring on the far sides are stored at the ring node
                    if (r == n0 - 2 && und) acc(a.fn + so, a.Fn + so, o + ld, dS * eS, fl);
                    if (j == n1 - 2 && und) acc(a.fw + so, a.Fw + so, o + 1, dE * eE, fl);
                }
            }
            mN = mC;
            mC = mS;
        }
        pN = pC;
        pC = pS;
        cN = cS;
    }
}

template <typename T, bool SPONGE, bool BWD>
hipError_t ac_launch_k(const AcArgs<T>& a, uint32_t tiles, hipStream_t s)
{
    const uint32_t blocks = step_grid(tiles, a.nga + a.ngb + a.nsc);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((ac_kernel<T, SPONGE, BWD>), dim3(blocks), dim3(256), 0, s, a);
    return hipGetLastError();
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
    a.gy = (uint32_t)((a.n0 - 2 + 4 * AC_RI - 1) / (4 * AC_RI));
    const uint32_t tiles = tiles_on ? a.gx * a.gy * (uint32_t)a.nshots : 0u;
    const bool sp = a.br != nullptr;
    if (bwd) return sp ? ac_launch_k<T, true, true>(a, tiles, s) : ac_launch_k<T, false, true>(a, tiles, s);
    return sp ? ac_launch_k<T, true, false>(a, tiles, s) : ac_launch_k<T, false, false>(a, tiles, s);
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
    if (g->ndim != 2 || g->order != 2) return ac_fail(WFD_EINVAL, "wfd_ac: 2D grids of order 2 only");
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
    if (e == hipSuccess) e = launch_zero(gwork, 3 * fbytes, s);
    if (e == hipSuccess) e = launch_zero(G, 3 * field * (int64_t)sizeof(double), s);
    if (e != hipSuccess) return ac_fail(WFD_EHIP, "wfd_ac_backward: clearing work fields failed");
    const int64_t off = g->offset;
    T* PA = (T*)ua + off;  // p^{N-1}
    T* PB = (T*)ub + off;  // p^N
    T* MA = (T*)ma + off;
    T* MB = (T*)mb + off;
    a.gm = (T*)gwork + off;
    a.fn = a.gm + field;
    a.fw = a.fn + field;
    a.Gm = G + off;
    a.Fn = a.Gm + field;
    a.Fw = a.Fn + field;
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
