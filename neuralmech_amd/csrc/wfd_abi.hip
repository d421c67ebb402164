// Host side of the wavefd C ABI: validation, layout, point lists, launch dispatch, hipGraph
// replay and the checkpointed gradient plan.  No exceptions cross the ABI; errors are returned
// as status codes with a thread-local message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "wfd_internal.h"

using namespace wfd;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    return fail(WFD_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                          \
    do {                                             \
        hipError_t e_ = (expr);                      \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

int64_t gcd64(int64_t a, int64_t b)
{
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    while (b) {
        int64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

// a_0..a_K, each rounded once from the exact rational (see wavefd.h).
void fd_coeffs(int order, double* a)
{
    const int K = order / 2;
    int64_t f[17];
    f[0] = 1;
    for (int i = 1; i <= 16; i++) f[i] = f[i - 1] * i;
    int64_t sn = 0, sd = 1;
    for (int k = 1; k <= K; k++) {
        int64_t num = 2 * f[K] * f[K];
        int64_t den = (int64_t)k * k * f[K - k] * f[K + k];
        int64_t q = gcd64(num, den);
        num /= q;
        den /= q;
        if (k % 2 == 0) num = -num;
        a[k] = (double)num / (double)den;
        const int64_t nn = sn * den + num * sd, nd = sd * den;
        q = gcd64(nn, nd);
        sn = nn / q;
        sd = nd / q;
    }
    a[0] = (double)(-2 * sn) / (double)sd;
}

bool valid_order(int order) { return order == 2 || order == 4 || order == 6 || order == 8; }

int check_grid(const wfd_grid* g)
{
    if (!g) return fail(WFD_EINVAL, "grid is NULL");
    if (g->ndim < 1 || g->ndim > 3) return fail(WFD_EINVAL, "ndim must be 1, 2 or 3 (got %d)", g->ndim);
    if (!valid_order(g->order)) return fail(WFD_EINVAL, "order must be 2, 4, 6 or 8 (got %d)", g->order);
    if (g->dtype != WFD_F32 && g->dtype != WFD_F64) return fail(WFD_EINVAL, "dtype must be WFD_F32 or WFD_F64");
    if (g->nshots < 1) return fail(WFD_EINVAL, "nshots must be >= 1");
    if (g->ld <= 0 || g->elems <= 0) return fail(WFD_EINVAL, "grid layout not initialised (wfd_grid_init)");
    return WFD_OK;
}

// axis extents as (planes, rows, columns)
void extents(const wfd_grid* g, int64_t& np, int64_t& nr, int64_t& nc)
{
    if (g->ndim == 3) {
        np = g->n[0];
        nr = g->n[1];
        nc = g->n[2];
    } else if (g->ndim == 2) {
        np = 1;
        nr = g->n[0];
        nc = g->n[1];
    } else {
        np = 1;
        nr = 1;
        nc = g->n[0];
    }
}

size_t esize(const wfd_grid* g) { return g->dtype == WFD_F64 ? 8 : 4; }

template <typename T>
void fill_weights(const wfd_grid* g, T& wc, T (*w)[4])
{
    double a[5];
    fd_coeffs(g->order, a);
    const int K = g->order / 2;
    double s = 0.0;
    for (int d = 0; d < g->ndim; d++) s += 1.0 / (g->h[d] * g->h[d]);
    wc = (T)(a[0] * s);
    for (int d = 0; d < g->ndim; d++)
        for (int m = 1; m <= 4; m++) w[d][m - 1] = m <= K ? (T)(a[m] / (g->h[d] * g->h[d])) : T(0);
}

template <typename T>
int fill_points(PtsDev<T>& p, const wfd_points* src, int32_t nrows, const char* what, bool gathered)
{
    std::memset(&p, 0, sizeof p);
    if (!src || src->count == 0) return WFD_OK;
    if (src->count < 0) return fail(WFD_EINVAL, "%s: negative point count", what);
    if (src->nrows != nrows) return fail(WFD_EINVAL, "%s: CSR has %d rows, grid needs %d", what, src->nrows, nrows);
    if (!src->row_ptr || !src->col || !src->id || !src->data || (gathered && !src->row))
        return fail(WFD_EINVAL, "%s: NULL CSR or data pointer", what);
    if (src->width < src->count || src->nsteps < 0 || src->shot_stride < src->nsteps * src->width)
        return fail(WFD_EINVAL, "%s: data narrower than the point count or shorter than nsteps rows", what);
    p.row_ptr = src->row_ptr;
    p.col = src->col;
    p.id = src->id;
    p.row = src->row;
    p.data = static_cast<T*>(src->data);
    p.width = src->width;
    p.shot_stride = src->shot_stride;
    p.nt = src->nsteps;
    p.count = src->count;
    p.nrows = src->nrows;
    return WFD_OK;
}

int nrows_of(const wfd_grid* g)
{
    if (g->ndim == 3) return (int)(g->n[0] * g->n[1]);
    if (g->ndim == 2) return (int)g->n[0];
    return 1;
}

struct Tuned {
    int ndim, order, dtype, nshots;
    int64_t n[3];
    LaunchCfg cfg;
};
std::mutex g_tune_mu;
std::vector<Tuned> g_tuned;

bool same_shape(const Tuned& t, const wfd_grid* g)
{
    return t.ndim == g->ndim && t.order == g->order && t.dtype == g->dtype && t.nshots == g->nshots &&
           t.n[0] == g->n[0] && t.n[1] == g->n[1] && t.n[2] == g->n[2];
}

void remember_cfg(const wfd_grid* g, const LaunchCfg& c)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (Tuned& t : g_tuned)
        if (same_shape(t, g)) {
            t.cfg = c;
            return;
        }
    Tuned t{g->ndim, g->order, g->dtype, g->nshots, {g->n[0], g->n[1], g->n[2]}, c};
    g_tuned.push_back(t);
}

void forget_cfg(const wfd_grid* g)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned.erase(std::remove_if(g_tuned.begin(), g_tuned.end(), [&](const Tuned& t) { return same_shape(t, g); }),
                  g_tuned.end());
}

// Launch shape used when nothing was tuned or set for this grid: enough waves to give every
// SIMD (1024 on MI355X) several, then as many rows per wave as that allows; streaming access when
// the four fields of a step (u_prev, u_cur, u_next, kc) exceed twice the 256 MB Infinity Cache (a
// working set that still fits in good part gains more from the cache hits: 16 shots of 1024^2,
// 279 MB, run the gradient plan 7 % slower streaming); 2D grids of several shots take their tiles
// shot-fastest (16 shots of 1024^2, order 8, lane-exchange halo, 2 rows: 39.1 -> 37.1 us per
// forward step; cache halo 38.6 -> 34.4 us).
LaunchCfg default_cfg(const wfd_grid* g)
{
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (const Tuned& t : g_tuned)
            if (same_shape(t, g)) return t.cfg;
    }
    LaunchCfg c{};
    const int64_t step_bytes = 4 * g->elems * (int64_t)g->nshots * (g->dtype == WFD_F64 ? 8 : 4);
    const int stream = step_bytes > (int64_t)512 << 20 ? WFD_VARIANT_STREAM : 0;
    if (g->ndim == 3) {
        c.ri = 32;
        c.tile = stream;
    } else if (g->ndim == 2) {
        c.tile = 1 | stream | (g->nshots > 1 ? WFD_VARIANT_SHOTS : 0);
        const int64_t chunks = ((g->n[1] + 247) / 248) * (int64_t)g->nshots;
        const int64_t rows = g->n[0] - 2;
        int ri = 8;
        while (ri > 1 && chunks * ((rows + ri - 1) / ri) < 4096) ri /= 2;
        c.ri = g->dtype == WFD_F64 ? 4 : ri;
    } else {
        c.ri = 1;
        c.tile = 0;
    }
    return c;
}

// Planes per block of a 3D imaging launch: the fewest blocks along the marching axis whose tiles
// fill whole rounds of the resident slots (float: 2 blocks of the imaging kernel per CU, float64:
// 1), or else give at least four rounds.  512^3: 2 blocks of 255 planes, 512 tiles = one round on
// 256 CUs (832 us per adjoint step, against 892 us at 64 planes and 1094 us at 170).
int img_depth(const wfd_grid* g)
{
    constexpr int kDevs = 64;
    static std::atomic<int> cus[kDevs];  // CUs of each device, 0 until looked up
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) dev = 0;
    if (dev < kDevs) ncu = cus[dev].load(std::memory_order_relaxed);
    if (ncu <= 0) {
        if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
        if (dev < kDevs) cus[dev].store(ncu, std::memory_order_relaxed);
    }
    const int64_t per = ((g->n[2] + 63) / 64) * ((g->n[1] - 2 + 15) / 16) * (int64_t)g->nshots;  // tiles per z-block
    const int64_t slots = (int64_t)ncu * (g->dtype == WFD_F64 ? 1 : 2);
    const int64_t np = g->n[0] - 2;
    for (int64_t m = 1; m <= np; m++) {
        const int64_t rk = (np + m - 1) / m, t = per * ((np + rk - 1) / rk);
        if (t >= slots && (t % slots == 0 || t >= 4 * slots)) return (int)rk;
    }
    return 1;
}

struct Launch {
    const wfd_grid* g;
    const void* up;
    const void* uc;
    void* un;
    const wfd_medium* med;
    const wfd_points* src;
    const wfd_points* rec;  // gathered: u_cur at these nodes -> trace row step + gather_off
    int64_t step;
    const int64_t* step_dev;
    const void* uf;
    void* gpart;
    bool img;
    LaunchCfg cfg;
    int64_t gather_off;
    void* un2 = nullptr;  // two-step launch: u^{n+2} (un receives u^{n+1})
    int t2rows = 16;
    bool img_group = false;  // 2D plan imaging: uf holds L u; several shots share a partial per IMG_GROUP
    void* ucopy = nullptr;   // 2D forward: receives L u_cur (see StepArgs::ucopy)
};

template <typename T>
int launch_t(const Launch& L, hipStream_t s)
{
    const wfd_grid* g = L.g;
    StepArgs<T> a;
    std::memset(&a, 0, sizeof a);
    const int64_t off = g->offset;
    a.up = L.up ? static_cast<const T*>(L.up) + off : nullptr;
    a.uc = static_cast<const T*>(L.uc) + off;
    a.un = L.un ? static_cast<T*>(L.un) + off : nullptr;
    a.ucopy = L.ucopy && !L.img && g->ndim == 2 ? static_cast<T*>(L.ucopy) + off : nullptr;
    const wfd_medium* m = L.med;
    const bool varc = m && m->coef;
    const bool sponge = m && m->beta;
    if (varc) {
        a.coef = static_cast<const T*>(m->coef) + off;
        a.coef_shot_stride = m->coef_shot_stride;
    }
    a.kconst = m ? (T)m->kconst : T(0);
    extents(g, a.np, a.nr, a.nc);
    if (sponge) {
        const T* b = static_cast<const T*>(m->beta);
        if (g->ndim == 3) {
            a.beta_p = b;
            a.beta_r = b + a.np;
            a.beta_c = b + a.np + a.nr;
        } else if (g->ndim == 2) {
            a.beta_r = b;
            a.beta_c = b + a.nr;
        } else {
            a.beta_c = b;
        }
    }
    if (L.img) {
        a.uf = static_cast<const T*>(L.uf) + off;
        a.g = static_cast<T*>(L.gpart) + off;
        a.img_group = L.img_group && g->ndim == 2 ? (g->nshots > 1 ? IMG_GROUP : 1) : 0;
    }
    a.shot_stride = g->elems;
    a.ld = g->ld;
    a.ps = g->plane * g->ld;
    if (g->ndim == 1) {
        a.row_lo = 0;
        a.row_hi = 1;
    } else {
        a.row_lo = 1;
        a.row_hi = a.nr - 1;
    }
    T w[3][4];
    fill_weights<T>(g, a.wc, w);
    // weights by role: the marching/plane axis, the row axis, the column (fastest) axis
    const int d_c = g->ndim - 1;
    for (int mm = 0; mm < 4; mm++) {
        a.wcol[mm] = w[d_c][mm];
        a.wr[mm] = g->ndim >= 2 ? w[d_c - 1][mm] : T(0);
        a.wp[mm] = g->ndim == 3 ? w[0][mm] : T(0);
    }
    const int32_t nrows = nrows_of(g);
    int rc = fill_points<T>(a.src, L.src, nrows, L.img ? "adjoint sources" : "sources", false);
    if (rc) return rc;
    rc = fill_points<T>(a.rec, L.rec, nrows, "receivers", true);
    if (rc) return rc;
    a.gather_off = L.gather_off;
    a.nshots = g->nshots;
    a.shot_fast = g->ndim == 2 && g->nshots > 1 && (L.cfg.tile & WFD_VARIANT_SHOTS) && !L.un2 ? 1 : 0;
    a.ngb = a.rec.count ? (uint32_t)(((int64_t)g->nshots * a.rec.count + 255) / 256) : 0u;
    a.step = L.step;
    a.step_dev = L.step_dev;
    hipError_t e;
    if (L.un2) {  // two steps: the gather blocks record both input levels
        a.un2 = static_cast<T*>(L.un2) + off;
        a.ngb *= 2;
        e = g->ndim == 3 ? launch_step3d_t2<T>(a, g->order, varc, sponge, L.t2rows, g->nshots, s)
                         : launch_step2d_t2<T>(a, g->order, varc, sponge, L.t2rows, g->nshots, s);
        if (e != hipSuccess) return hip_fail(e, "two-step kernel launch");
        return WFD_OK;
    }
    if (!L.un) {  // gather only: L.uc is the field, row = L.step + L.gather_off
        e = launch_gather<T>(a, s);
        if (e != hipSuccess) return hip_fail(e, "gather kernel launch");
        return WFD_OK;
    }
    if (g->ndim == 3) {
        // The tuner times forward launches; imaging launches (2 blocks per CU instead of 3) march
        // the depth that fills whole rounds of their own resident slots (any depth gives the same
        // bits), or the forced one.
        const int rk = L.img ? (L.cfg.img_ri > 0 ? L.cfg.img_ri : img_depth(g)) : std::max(1, L.cfg.ri);
        e = launch_step3d<T>(a, g->order, varc, sponge, L.img, rk, (L.cfg.tile & WFD_VARIANT_STREAM) != 0, g->nshots,
                             s);
    } else {
        e = launch_step2d<T>(a, g->ndim == 2, g->order, varc, sponge, L.img, L.cfg.tile, L.cfg.ri, g->nshots, s);
    }
    if (e != hipSuccess) return hip_fail(e, "step kernel launch");
    return WFD_OK;
}

int do_launch(const Launch& L, hipStream_t s)
{
    return L.g->dtype == WFD_F64 ? launch_t<double>(L, s) : launch_t<float>(L, s);
}

int check_kconst(const wfd_grid* g, const wfd_medium* med);

int check_step_args(const wfd_grid* g, const void* up, const void* uc, void* un, const wfd_medium* med)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!up || !uc || !un) return fail(WFD_EINVAL, "NULL field pointer");
    if (uc == un) return fail(WFD_EINVAL, "u_next may alias u_prev only, not u_cur");
    if (up == uc) return fail(WFD_EINVAL, "u_prev and u_cur must be distinct fields");
    if (!med) return fail(WFD_EINVAL, "medium is NULL");
    if (med->coef && med->coef_shot_stride != 0 && med->coef_shot_stride != g->elems)
        return fail(WFD_EINVAL, "coef_shot_stride must be 0 or elems");
    rc = check_kconst(g, med);
    if (rc) return rc;
    const uintptr_t align = 4 * esize(g);
    if (((uintptr_t)up | (uintptr_t)uc | (uintptr_t)un) % align ||
        (med->coef && (uintptr_t)med->coef % align))
        return fail(WFD_EINVAL, "fields must be %zu-byte aligned", (size_t)align);
    return WFD_OK;
}

// Largest stable kc = v^2 dt^2 of the grid's order and spacing: the leapfrog (damped or not) is
// stable for kc * sum_d S / h_d^2 <= 4, S = |a_0| + 2 sum |a_k| (the symbol's largest magnitude);
// a relative slack of 1e-6 admits a dt computed at the limit and rounded to the grid dtype.
double kc_limit(const wfd_grid* g)
{
    double a[5];
    fd_coeffs(g->order, a);
    double S = std::fabs(a[0]);
    for (int k = 1; k <= g->order / 2; k++) S += 2.0 * std::fabs(a[k]);
    double q = 0.0;
    for (int d = 0; d < g->ndim; d++) q += S / (g->h[d] * g->h[d]);
    return 4.0 / q * (1.0 + 1e-6);
}

int check_kconst(const wfd_grid* g, const wfd_medium* med)
{
    if (med->coef) return WFD_OK;
    if (!(med->kconst > 0.0 && std::isfinite(med->kconst)))
        return fail(WFD_ENONFINITE, "constant coefficient must be finite and > 0");
    if (med->kconst > kc_limit(g))
        return fail(WFD_EINVAL, "CFL violation: kc = %.6g exceeds the order-%d limit %.6g for this spacing",
                    med->kconst, g->order, kc_limit(g));
    return WFD_OK;
}

}  // namespace

// A forward run of `chunk` steps captured once into a hipGraph, with the step index of every
// launch read from its own device counter (set before each run, advanced by the graph's last
// node), so that one graph serves every chunk of a run and graphs on different streams never
// share a counter.
struct wfd_fgraph {
    wfd_grid g;
    wfd_medium med;
    wfd_points src, rec;
    bool has_src, has_rec;
    void* u_a;
    void* u_b;
    int32_t chunk;
    int64_t* ctr;
    hipGraph_t graph;
    hipGraphExec_t exec;
    hipEvent_t done;  // recorded after each run: destruction waits for it
    bool ran;
};

namespace {

void fgraph_free(wfd_fgraph* f)
{
    if (!f) return;
    if (f->ran && f->done) (void)hipEventSynchronize(f->done);
    if (f->exec) (void)hipGraphExecDestroy(f->exec);
    if (f->graph) (void)hipGraphDestroy(f->graph);
    if (f->done) (void)hipEventDestroy(f->done);
    if (f->ctr) (void)hipFree(f->ctr);
    delete f;
}

int fgraph_capture(wfd_fgraph* f)
{
    HIP_TRY(hipMalloc(&f->ctr, 64), "hipMalloc(step counter)");
    HIP_TRY(hipEventCreateWithFlags(&f->done, hipEventDisableTiming), "hipEventCreate");
    hipStream_t cap;
    HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking), "hipStreamCreate");
    const LaunchCfg cfg = default_cfg(&f->g);
    int rc = WFD_OK;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        void* p = f->u_a;
        void* c = f->u_b;
        for (int k = 0; k < f->chunk && rc == WFD_OK; k++) {
            Launch L{&f->g, p, c, p, &f->med, f->has_src ? &f->src : nullptr, f->has_rec ? &f->rec : nullptr,
                     k, f->ctr, nullptr, nullptr, false, cfg, -1};
            rc = do_launch(L, cap);
            std::swap(p, c);
        }
        if (rc == WFD_OK) {
            e = launch_add_i64(f->ctr, f->chunk, cap);
            if (e != hipSuccess) rc = hip_fail(e, "counter kernel");
        }
        hipError_t e2 = hipStreamEndCapture(cap, &f->graph);
        if (rc == WFD_OK && e2 != hipSuccess) rc = hip_fail(e2, "hipStreamEndCapture");
    } else {
        rc = hip_fail(e, "hipStreamBeginCapture");
    }
    if (rc == WFD_OK) {
        e = hipGraphInstantiate(&f->exec, f->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
    (void)hipStreamDestroy(cap);
    return rc;
}

int fgraph_make(wfd_fgraph** out, const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med,
                const wfd_points* src, const wfd_points* rec, int32_t chunk)
{
    *out = nullptr;
    int rc = check_step_args(g, u_a, u_b, u_a, med);
    if (rc) return rc;
    if (chunk < 2 || chunk % 2) return fail(WFD_EINVAL, "chunk must be even and >= 2 (got %d)", chunk);
    wfd_fgraph* f = new wfd_fgraph{};
    f->g = *g;
    f->med = *med;
    f->has_src = src && src->count > 0;
    f->has_rec = rec && rec->count > 0;
    if (f->has_src) f->src = *src;
    if (f->has_rec) f->rec = *rec;
    f->u_a = u_a;
    f->u_b = u_b;
    f->chunk = chunk;
    rc = fgraph_capture(f);
    if (rc) {
        fgraph_free(f);
        return rc;
    }
    *out = f;
    return WFD_OK;
}

// True when `st` is being captured into a graph by the caller.  A failed query (which HIP also
// leaves as the last error, read by the launch checks) counts as "not capturing".
bool capturing(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return cs != hipStreamCaptureStatusNone;
}

int fgraph_run(wfd_fgraph* f, int64_t step0, int64_t nsteps, hipStream_t st)
{
    if (nsteps < 0) return fail(WFD_EINVAL, "nsteps must be >= 0");
    const wfd_points* src = f->has_src ? &f->src : nullptr;
    const wfd_points* rec = f->has_rec ? &f->rec : nullptr;
    // Under the caller's own capture the steps are recorded as plain launches (same kernels,
    // same results), so the caller's graph references nothing this object owns.
    if (capturing(st)) return wfd_forward(&f->g, f->u_a, f->u_b, &f->med, src, rec, step0, nsteps, st);
    const int64_t reps = nsteps / f->chunk;
    if (reps > 0) {
        // counter = step0 (a stream-ordered kernel), then the replays
        hipError_t e = launch_set_i64(f->ctr, step0, st);
        for (int64_t r = 0; e == hipSuccess && r < reps; r++) e = hipGraphLaunch(f->exec, st);
        if (e == hipSuccess) e = hipEventRecord(f->done, st);
        if (e != hipSuccess) return hip_fail(e, "graph replay");
        f->ran = true;
    }
    const int64_t rem = nsteps - reps * f->chunk;
    // after an even number of replayed steps the roles of u_a / u_b are unchanged; the
    // remainder's first launch records the last replayed level, otherwise record it here
    if (rem == 0 && reps > 0) return wfd::gather_launch(&f->g, f->u_b, rec, step0 + nsteps - 1, st);
    int rc = wfd_forward(&f->g, f->u_a, f->u_b, &f->med, src, rec, step0 + reps * f->chunk, rem, st);
    if (rc == WFD_OK && rem > 0) {
        hipError_t e = hipEventRecord(f->done, st);
        if (e != hipSuccess) return hip_fail(e, "hipEventRecord");
        f->ran = true;
    }
    return rc;
}

// wfd_forward_graph keeps its graphs between calls, keyed by everything a capture bakes in
// (stream, grid, buffers, medium, point descriptors, chunk, launch shape); each entry has its own
// counter, so calls on different streams never share one.  At most kGraphCache entries; the
// least recently used one is destroyed (after its last run has finished) to make room.
struct CachedGraph {
    hipStream_t stream;
    std::string key;
    wfd_fgraph* f;
    uint64_t used;
};
constexpr size_t kGraphCache = 16;
std::mutex g_graph_mu;
std::vector<CachedGraph> g_graphs;
uint64_t g_graph_tick = 0;

template <typename T>
void key_add(std::string& k, const T& v)
{
    k.append(reinterpret_cast<const char*>(&v), sizeof v);
}

std::string graph_key(const wfd_grid* g, const void* u_a, const void* u_b, const wfd_medium* med,
                      const wfd_points* src, const wfd_points* rec, int32_t chunk)
{
    std::string k;
    key_add(k, *g);
    key_add(k, u_a);
    key_add(k, u_b);
    key_add(k, *med);
    const wfd_points none{};
    key_add(k, src && src->count ? *src : none);
    key_add(k, rec && rec->count ? *rec : none);
    key_add(k, chunk);
    const LaunchCfg c = default_cfg(g);
    key_add(k, c);
    int dev = -1;
    (void)hipGetDevice(&dev);
    key_add(k, dev);
    return k;
}

}  // namespace

int wfd::set_error(int code, const char* msg) { return fail(code, "%s", msg); }

int wfd::check_medium_args(const wfd_grid* g, const wfd_medium* med)
{
    if (!med) return fail(WFD_EINVAL, "medium is NULL");
    if (med->coef && med->coef_shot_stride != 0 && med->coef_shot_stride != g->elems)
        return fail(WFD_EINVAL, "coef_shot_stride must be 0 or elems");
    return check_kconst(g, med);
}

int wfd::step_launch(const wfd_grid* g, const void* up, const void* uc, void* un, const wfd_medium* med,
                     const wfd_points* src, const wfd_points* rec, int64_t step, int64_t gather_off, const void* uf,
                     void* gpart, bool img, hipStream_t s, bool img_group, void* ucopy)
{
    Launch L{g, up, uc, un, med, src, rec, step, nullptr, uf, gpart, img, default_cfg(g), gather_off};
    L.img_group = img_group;
    L.ucopy = ucopy;
    return do_launch(L, s);
}

int wfd::gather_launch(const wfd_grid* g, const void* field, const wfd_points* rec, int64_t row, hipStream_t s)
{
    if (!rec || rec->count == 0) return WFD_OK;
    if (!field) return fail(WFD_EINVAL, "gather: NULL field");
    Launch L{g, nullptr, field, nullptr, nullptr, nullptr, rec, row, nullptr, nullptr, nullptr, false, default_cfg(g), 0};
    return do_launch(L, s);
}

// ======================================================================== exported C ABI
extern "C" {

const char* wfd_last_error(void) { return g_err.c_str(); }
const char* wfd_version(void) { return "wavefd 0.1.0 gfx950"; }

int wfd_grid_init(wfd_grid* g, int ndim, const int64_t* n, const double* h, double dt, int order, int dtype,
                  int nshots)
{
    if (!g || !n || !h) return fail(WFD_EINVAL, "NULL argument");
    std::memset(g, 0, sizeof *g);
    if (ndim < 1 || ndim > 3) return fail(WFD_EINVAL, "ndim must be 1, 2 or 3 (got %d)", ndim);
    if (!valid_order(order)) return fail(WFD_EINVAL, "order must be 2, 4, 6 or 8 (got %d)", order);
    if (dtype != WFD_F32 && dtype != WFD_F64) return fail(WFD_EINVAL, "dtype must be WFD_F32 or WFD_F64");
    if (nshots < 1) return fail(WFD_EINVAL, "nshots must be >= 1");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(WFD_EINVAL, "dt must be finite and > 0");
    for (int d = 0; d < ndim; d++) {
        if (n[d] < 3) return fail(WFD_EINVAL, "axis %d has %lld nodes; at least 3 needed", d, (long long)n[d]);
        if (n[d] > (int64_t)1 << 30) return fail(WFD_EINVAL, "axis %d too long", d);
        if (!(h[d] > 0.0) || !std::isfinite(h[d])) return fail(WFD_EINVAL, "spacing of axis %d must be > 0", d);
    }
    g->ndim = ndim;
    g->order = order;
    g->dtype = dtype;
    g->nshots = nshots;
    for (int d = 0; d < 3; d++) {
        g->n[d] = d < ndim ? n[d] : 1;
        g->h[d] = d < ndim ? h[d] : 1.0;
    }
    g->dt = dt;
    const int64_t nfast = n[ndim - 1];
    g->ld = (nfast + HALO + 31) / 32 * 32;
    if (ndim == 3) {
        g->plane = n[1] + HALO;
        g->offset = (int64_t)HALO * g->plane * g->ld;
        g->elems = (n[0] + 2 * HALO) * g->plane * g->ld;
    } else {
        g->plane = 1;
        const int64_t rows = ndim == 2 ? n[0] : 1;
        g->offset = (int64_t)HALO * g->ld;
        g->elems = (rows + 2 * HALO) * g->ld;
    }
    if (nrows_of(g) > (1 << 30)) return fail(WFD_EINVAL, "grid too large for 32-bit point rows");
    return WFD_OK;
}

double wfd_stable_dt(int ndim, const double* h, double vmax, int order)
{
    if (ndim < 1 || ndim > 3 || !h || !valid_order(order) || !(vmax > 0.0)) {
        fail(WFD_EINVAL, "wfd_stable_dt: bad arguments");
        return 0.0;
    }
    double a[5];
    fd_coeffs(order, a);
    double S = std::fabs(a[0]);
    for (int k = 1; k <= order / 2; k++) S += 2.0 * std::fabs(a[k]);
    double q = 0.0;
    for (int d = 0; d < ndim; d++) q += S / (h[d] * h[d]);
    return 2.0 / std::sqrt(q) / vmax;
}

int wfd_stencil_weights(const wfd_grid* g, double* wc, double* w)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!wc || !w) return fail(WFD_EINVAL, "NULL output");
    const int K = g->order / 2;
    if (g->dtype == WFD_F64) {
        double c, ww[3][4];
        fill_weights<double>(g, c, ww);
        *wc = c;
        for (int d = 0; d < g->ndim; d++)
            for (int m = 0; m < K; m++) w[d * K + m] = ww[d][m];
    } else {
        float c, ww[3][4];
        fill_weights<float>(g, c, ww);
        *wc = c;
        for (int d = 0; d < g->ndim; d++)
            for (int m = 0; m < K; m++) w[d * K + m] = ww[d][m];
    }
    return WFD_OK;
}

int32_t wfd_points_nrows(const wfd_grid* g) { return check_grid(g) ? -1 : nrows_of(g); }

int wfd_points_csr(const wfd_grid* g, int32_t npts, const int64_t* nodes, int32_t* row_ptr, int32_t* col,
                   int32_t* id, int32_t* row)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (npts < 0 || (npts > 0 && (!nodes || !col || !id || !row)) || !row_ptr) return fail(WFD_EINVAL, "bad point arrays");
    int64_t np, nr, nc;
    extents(g, np, nr, nc);
    const int32_t nrows = nrows_of(g);
    std::vector<int64_t> key(npts);
    for (int32_t p = 0; p < npts; p++) {
        const int64_t x = nodes[p];
        if (x < 0 || x >= np * nr * nc) return fail(WFD_EINVAL, "point %d: node %lld outside the grid", p, (long long)x);
        const int64_t j = x % nc, r = x / nc, i = r % nr, k = r / nr;
        const bool inner = j >= 1 && j <= nc - 2 && (g->ndim < 2 || (i >= 1 && i <= nr - 2)) &&
                           (g->ndim < 3 || (k >= 1 && k <= np - 2));
        if (!inner) return fail(WFD_EINVAL, "point %d: node %lld is not an interior node", p, (long long)x);
        key[p] = x;  // flat node order = (row, column) order
    }
    std::vector<int32_t> perm(npts);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
    std::fill(row_ptr, row_ptr + nrows + 1, 0);
    for (int32_t q = 0; q < npts; q++) {
        const int32_t p = perm[q];
        const int64_t r = key[p] / nc;
        col[q] = (int32_t)(key[p] % nc);
        id[q] = p;
        row[q] = (int32_t)r;
        row_ptr[r + 1]++;
    }
    for (int32_t r = 0; r < nrows; r++) row_ptr[r + 1] += row_ptr[r];
    return WFD_OK;
}

int wfd_step(const wfd_grid* g, const void* u_prev, const void* u_cur, void* u_next, const wfd_medium* med,
             const wfd_points* src, const wfd_points* rec, int64_t step, wfd_stream_t stream)
{
    int rc = check_step_args(g, u_prev, u_cur, u_next, med);
    if (rc) return rc;
    rc = wfd::step_launch(g, u_prev, u_cur, u_next, med, src, nullptr, step, 0, nullptr, nullptr, false,
                          (hipStream_t)stream);
    if (rc) return rc;
    return wfd::gather_launch(g, u_next, rec, step, (hipStream_t)stream);
}

int wfd_forward(const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med, const wfd_points* src,
                const wfd_points* rec, int64_t step0, int64_t nsteps, wfd_stream_t stream)
{
    int rc = check_step_args(g, u_a, u_b, u_a, med);
    if (rc) return rc;
    if (nsteps < 0) return fail(WFD_EINVAL, "nsteps must be >= 0");
    void* p = u_a;
    void* c = u_b;
    for (int64_t k = 0; k < nsteps; k++) {
        // each launch also records u^n (its u_cur) into trace row n-1
        rc = wfd::step_launch(g, p, c, p, med, src, rec, step0 + k, -1, nullptr, nullptr, false, (hipStream_t)stream);
        if (rc) return rc;
        std::swap(p, c);
    }
    if (nsteps > 0) return wfd::gather_launch(g, c, rec, step0 + nsteps - 1, (hipStream_t)stream);
    return WFD_OK;
}

int wfd_forward2(const wfd_grid* g, void* u_a, void* u_b, void* u_c, void* u_d, const wfd_medium* med,
                 const wfd_points* src, const wfd_points* rec, int64_t step0, int64_t nsteps, int32_t rows,
                 wfd_stream_t stream)
{
    int rc = check_step_args(g, u_a, u_b, u_a, med);
    if (rc) return rc;
    if (g->ndim == 1) return fail(WFD_EINVAL, "wfd_forward2: 2D and 3D grids only");
    if (nsteps < 0) return fail(WFD_EINVAL, "nsteps must be >= 0");
    if (g->ndim == 2 && rows != 0 && rows != 2 && rows != 4 && rows != 8 && rows != 16 && rows != 32)
        return fail(WFD_EINVAL, "wfd_forward2: rows must be 0, 2, 4, 8, 16 or 32");
    if (g->ndim == 3 && (rows < 0 || rows > 4096))
        return fail(WFD_EINVAL, "wfd_forward2: planes per block must be 0 (default) or 1 .. 4096");
    if (g->ndim == 3 && g->dtype != WFD_F32) return fail(WFD_EINVAL, "wfd_forward2: 3D grids in float32 only");
    if (!u_c || !u_d) return fail(WFD_EINVAL, "wfd_forward2: NULL work field");
    const void* f[4] = {u_a, u_b, u_c, u_d};
    for (int i = 0; i < 4; i++)
        for (int j = i + 1; j < 4; j++)
            if (f[i] == f[j]) return fail(WFD_EINVAL, "wfd_forward2: the four fields must be distinct");
    const uintptr_t align = 4 * esize(g);
    if (((uintptr_t)u_c | (uintptr_t)u_d) % align) return fail(WFD_EINVAL, "fields must be %zu-byte aligned", (size_t)align);
    hipStream_t st = (hipStream_t)stream;
    void* P = u_a;  // u^{n-1}
    void* C = u_b;  // u^n
    void* X = u_c;
    void* Y = u_d;
    int64_t k = 0;
    if (nsteps % 2) {  // one single step first; it records u^n into row n-1
        rc = wfd::step_launch(g, P, C, P, med, src, rec, step0, -1, nullptr, nullptr, false, st);
        if (rc) return rc;
        std::swap(P, C);
        k = 1;
    }
    for (; k + 2 <= nsteps; k += 2) {
        // u^{n+1} -> X, u^{n+2} -> Y; the launch records u^{n-1}, u^n into rows n-2, n-1
        Launch L{g, P, C, X, med, src, rec, step0 + k, nullptr, nullptr, nullptr, false, default_cfg(g), -1};
        L.un2 = Y;
        L.t2rows = rows ? rows : (g->ndim == 3 ? 32 : 8);
        rc = do_launch(L, st);
        if (rc) return rc;
        std::swap(P, X);
        std::swap(C, Y);
    }
    if (nsteps == 0) return WFD_OK;
    if (nsteps >= 2) {
        rc = wfd::gather_launch(g, P, rec, step0 + nsteps - 2, st);
        if (rc) return rc;
    }
    rc = wfd::gather_launch(g, C, rec, step0 + nsteps - 1, st);
    if (rc) return rc;
    // leave (u^{N-1}, u^N) in (u_a, u_b)
    const size_t bytes = (size_t)g->elems * (size_t)g->nshots * esize(g);
    auto copy = [&](void* dst, const void* from) {
        return hipMemcpyAsync(dst, from, bytes, hipMemcpyDeviceToDevice, st);
    };
    if (P == u_b && C == u_a) {  // swapped: go through a work field
        HIP_TRY(copy(u_c, u_a), "wfd_forward2: copy of u^N");
        HIP_TRY(copy(u_a, u_b), "wfd_forward2: copy of u^{N-1}");
        HIP_TRY(copy(u_b, u_c), "wfd_forward2: copy of u^N");
    } else {
        if (P != u_a) HIP_TRY(copy(u_a, P), "wfd_forward2: copy of u^{N-1}");
        if (C != u_b) HIP_TRY(copy(u_b, C), "wfd_forward2: copy of u^N");
    }
    return WFD_OK;
}

int wfd_forward_graph(const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med, const wfd_points* src,
                      const wfd_points* rec, int64_t step0, int64_t nsteps, int32_t chunk, wfd_stream_t stream)
{
    int rc = check_step_args(g, u_a, u_b, u_a, med);
    if (rc) return rc;
    if (nsteps < 0) return fail(WFD_EINVAL, "nsteps must be >= 0");
    if (chunk < 2 || chunk % 2) return fail(WFD_EINVAL, "chunk must be even and >= 2 (got %d)", chunk);
    hipStream_t st = (hipStream_t)stream;
    if (nsteps < chunk || capturing(st)) return wfd_forward(g, u_a, u_b, med, src, rec, step0, nsteps, stream);
    const std::string key = graph_key(g, u_a, u_b, med, src, rec, chunk);
    std::lock_guard<std::mutex> lk(g_graph_mu);
    wfd_fgraph* f = nullptr;
    for (CachedGraph& c : g_graphs)
        if (c.stream == st && c.key == key) {
            c.used = ++g_graph_tick;
            f = c.f;
            break;
        }
    if (!f) {
        if (g_graphs.size() >= kGraphCache) {
            auto lru = std::min_element(g_graphs.begin(), g_graphs.end(),
                                        [](const CachedGraph& a, const CachedGraph& b) { return a.used < b.used; });
            fgraph_free(lru->f);
            g_graphs.erase(lru);
        }
        rc = fgraph_make(&f, g, u_a, u_b, med, src, rec, chunk);
        if (rc) return rc;
        g_graphs.push_back({st, key, f, ++g_graph_tick});
    }
    return fgraph_run(f, step0, nsteps, st);
}

int wfd_fgraph_create(wfd_fgraph** fg, const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med,
                      const wfd_points* src, const wfd_points* rec, int32_t chunk)
{
    if (!fg) return fail(WFD_EINVAL, "NULL graph handle");
    return fgraph_make(fg, g, u_a, u_b, med, src, rec, chunk);
}

int wfd_fgraph_run(wfd_fgraph* fg, int64_t step0, int64_t nsteps, wfd_stream_t stream)
{
    if (!fg) return fail(WFD_EINVAL, "NULL graph");
    return fgraph_run(fg, step0, nsteps, (hipStream_t)stream);
}

int wfd_fgraph_destroy(wfd_fgraph* fg)
{
    fgraph_free(fg);
    return WFD_OK;
}

int wfd_adjoint_step(const wfd_grid* g, const void* mu_prev, const void* mu_cur, void* mu_next,
                     const wfd_medium* med, const wfd_points* adj, const wfd_points* rec, int64_t step,
                     const void* u_fwd, void* g_partial, wfd_stream_t stream)
{
    int rc = check_step_args(g, mu_prev, mu_cur, mu_next, med);
    if (rc) return rc;
    if (!u_fwd || !g_partial) return fail(WFD_EINVAL, "NULL forward field or gradient partial");
    if (u_fwd == mu_next || g_partial == mu_next || g_partial == mu_cur || g_partial == u_fwd)
        return fail(WFD_EINVAL, "adjoint buffers must not alias");
    rc = wfd::step_launch(g, mu_prev, mu_cur, mu_next, med, adj, nullptr, step, 0, u_fwd, g_partial, true,
                          (hipStream_t)stream);
    if (rc) return rc;
    return wfd::gather_launch(g, mu_next, rec, step, (hipStream_t)stream);
}

int wfd_grad_flush(const wfd_grid* g, void* g_partial, double* G, wfd_stream_t stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!g_partial || !G) return fail(WFD_EINVAL, "NULL pointer");
    hipError_t e = g->dtype == WFD_F64
                       ? launch_flush<double>((double*)g_partial, G, g->elems, g->nshots, (hipStream_t)stream)
                       : launch_flush<float>((float*)g_partial, G, g->elems, g->nshots, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "flush kernel");
    return WFD_OK;
}

int wfd_check_medium_async(const wfd_grid* g, const wfd_medium* med, int32_t* flag, wfd_stream_t stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!med || !flag) return fail(WFD_EINVAL, "NULL argument");
    if (med->coef && med->coef_shot_stride != 0 && med->coef_shot_stride != g->elems)
        return fail(WFD_EINVAL, "coef_shot_stride must be 0 or elems");
    rc = check_kconst(g, med);
    if (rc) return rc;
    int64_t np, nr, nc;
    extents(g, np, nr, nc);
    const int64_t nbeta = (g->ndim == 3 ? np + nr : g->ndim == 2 ? nr : 0) + nc;
    const int64_t ps = g->plane * g->ld;
    const double kmax = kc_limit(g);
    hipError_t e = hipSuccess;
    const int nsets = med->coef && med->coef_shot_stride ? g->nshots : 1;
    for (int s = 0; s < nsets && e == hipSuccess; s++) {
        const int64_t so = (int64_t)s * med->coef_shot_stride + g->offset;
        if (g->dtype == WFD_F64)
            e = launch_check<double>(med->coef ? (const double*)med->coef + so : nullptr, np, nr, nc, g->ld, ps,
                                     kmax, s == 0 ? (const double*)med->beta : nullptr, nbeta, flag,
                                     (hipStream_t)stream);
        else
            e = launch_check<float>(med->coef ? (const float*)med->coef + so : nullptr, np, nr, nc, g->ld, ps,
                                    (float)kmax, s == 0 ? (const float*)med->beta : nullptr, nbeta, flag,
                                    (hipStream_t)stream);
    }
    if (e != hipSuccess) return hip_fail(e, "check kernel");
    return WFD_OK;
}

}  // extern "C"

namespace {

// The host-synchronising medium check on a caller-provided device flag (cleared here by a kernel).
int check_medium_on(const wfd_grid* g, const wfd_medium* med, int32_t* flag, hipStream_t st)
{
    hipError_t e = launch_zero(flag, sizeof(int32_t), st);
    if (e != hipSuccess) return hip_fail(e, "flag clear");
    int rc = wfd_check_medium_async(g, med, flag, st);
    if (rc) return rc;
    int32_t host = 0;
    HIP_TRY(hipMemcpyAsync(&host, flag, sizeof host, hipMemcpyDeviceToHost, st), "flag copy");
    HIP_TRY(hipStreamSynchronize(st), "medium check");
    if (host & 1) return fail(WFD_ENONFINITE, "medium holds a NaN/Inf, a non-positive coefficient or negative damping");
    if (host & 2)
        return fail(WFD_EINVAL, "CFL violation: a coefficient kc = v^2 dt^2 exceeds the order-%d limit %.6g for this spacing",
                    g->order, kc_limit(g));
    return WFD_OK;
}

}  // namespace

extern "C" {

int wfd_check_medium(const wfd_grid* g, const wfd_medium* med, wfd_stream_t stream)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (!med) return fail(WFD_EINVAL, "NULL medium");
    rc = check_kconst(g, med);
    if (rc) return rc;
    int32_t* flag = nullptr;
    HIP_TRY(hipMalloc(&flag, sizeof(int32_t)), "hipMalloc(flag)");
    rc = check_medium_on(g, med, flag, (hipStream_t)stream);
    (void)hipFree(flag);
    return rc;
}

// ------------------------------------------------------------------------- launch tuning
int wfd_set_config(const wfd_grid* g, int32_t variant, int32_t rows)
{
    int rc = check_grid(g);
    if (rc) return rc;
    if (rows == 0) {
        forget_cfg(g);
        return WFD_OK;
    }
    const bool ok2d = g->ndim == 2 && variant >= 0 && variant <= (1 | WFD_VARIANT_STREAM | WFD_VARIANT_SHOTS) &&
                      (g->dtype == WFD_F64 ? rows == 4 : (rows == 1 || rows == 2 || rows == 4 || rows == 8));
    const bool ok3d = g->ndim == 3 && (variant == 0 || variant == WFD_VARIANT_STREAM) && rows >= 1 && rows <= 1024;
    const bool ok1d = g->ndim == 1 && variant == 0 && rows == 1;
    if (!(ok1d || ok2d || ok3d)) return fail(WFD_EINVAL, "unsupported launch config (variant %d, rows %d)", variant, rows);
    LaunchCfg c{};
    c.tile = variant;
    c.ri = rows;
    c.img_ri = g->ndim == 3 ? rows : 0;
    remember_cfg(g, c);
    return WFD_OK;
}

int wfd_get_config(const wfd_grid* g, int32_t* variant, int32_t* rows)
{
    int rc = check_grid(g);
    if (rc) return rc;
    const LaunchCfg c = default_cfg(g);
    if (variant) *variant = c.tile;
    if (rows) *rows = c.ri;
    return WFD_OK;
}

int wfd_tune(const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med, int32_t steps, int32_t* variant,
             int32_t* rows, double* us_per_step, wfd_stream_t stream)
{
    int rc = check_step_args(g, u_a, u_b, u_a, med);
    if (rc) return rc;
    if (steps < 2) return fail(WFD_EINVAL, "wfd_tune: steps must be >= 2");
    std::vector<LaunchCfg> cands;
    if (g->ndim == 2) {
        const int f32rows[] = {1, 2, 4, 8};
        const int nv = g->nshots > 1 ? 8 : 4;  // the shot-fastest order only matters for several shots
        for (int v = 0; v < nv; v++) {
            if (g->dtype == WFD_F64) {
                cands.push_back({4, v});
            } else {
                for (int r : f32rows) cands.push_back({r, v});
            }
        }
    } else if (g->ndim == 3) {
        // plus the depths that split the interior planes into 2, 3 or 6 equal marches: with 3 or 4
        // resident blocks per CU one of them fills whole rounds of the launch
        std::vector<int> depths = {4, 8, 16, 32, 64, 128};
        for (int m : {2, 3, 6}) {
            const int d = (int)((g->n[0] - 2 + m - 1) / m);
            if (d > 0 && std::find(depths.begin(), depths.end(), d) == depths.end()) depths.push_back(d);
        }
        for (int v : {0, (int)WFD_VARIANT_STREAM})
            for (int r : depths) cands.push_back({r, v});
    } else {
        cands.push_back({1, 0});
    }
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&e1), "hipEventCreate");
    float best_ms = 1e30f;
    LaunchCfg best = cands[0];
    for (const LaunchCfg& c : cands) {
        for (int k = 0; k < steps + 2 && rc == WFD_OK; k++) {
            if (k == 2) (void)hipEventRecord(e0, st);
            void* p = (k % 2) ? u_b : u_a;
            void* q = (k % 2) ? u_a : u_b;
            Launch L{g, p, q, p, med, nullptr, nullptr, k, nullptr, nullptr, nullptr, false, c, 0};
            rc = do_launch(L, st);
        }
        if (rc) break;
        (void)hipEventRecord(e1, st);
        hipError_t e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = hip_fail(e, "wfd_tune timing");
            break;
        }
        if (ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    remember_cfg(g, best);
    if (variant) *variant = best.tile;
    if (rows) *rows = best.ri;
    if (us_per_step) *us_per_step = 1e3 * best_ms / steps;
    return WFD_OK;
}

}  // extern "C"
