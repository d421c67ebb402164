// Checkpointed gradient plan: forward run that keeps sqrt(n)-spaced checkpoints, then a
// backward pass that recomputes one segment at a time into a segment buffer and runs the
// adjoint with the imaging condition over it (newest step first).
//
// Schedule (seg = C, N steps, checkpoints k = 0..ceil(N/C)-1):
//   checkpoint k holds (u^{kC-1}, u^{kC}); checkpoint 0 is the rest state (zero).
//   forward:  step n writes u^{n+1} straight into checkpoint storage when n+1 is kC-1 or kC,
//             otherwise into a work field; no copies.
//   backward: for k = last..0: u^{kC} is checkpoint k, u^{kC+1..s1-1} are recomputed into the
//             segment slots (s1 = min(kC+C, N)); then adjoint steps m = s1..kC+1 read u^{m-1}.
// The imaging partial (grid dtype) is flushed into the float64 gradient whenever
// (m-1) % flush == 0, independently of C, so the gradient bits do not depend on the
// checkpoint spacing.  2D grids of several shots keep one partial per group of IMG_GROUP shots
// (the imaging launch adds the group's terms in shot order; the gradient is their sum anyway),
// flushed in group order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "wfd_internal.h"

extern "C" int wfd_adjoint_step(const wfd_grid* g, const void* mu_prev, const void* mu_cur, void* mu_next,
                                const wfd_medium* med, const wfd_points* adj, const wfd_points* rec,
                                int64_t step, const void* u_fwd, void* g_partial, wfd_stream_t stream);

struct wfd_plan {
    wfd_grid g;
    int64_t nsteps = 0, seg = 0, nckpt = 0, flush = 32;
    // 2D grids: the adjoint's imaging reads L u^{m-1} from a stored field instead
    // of taking the Laplacian of u^{m-1}.  The recompute (every level kept: the forward) steps in
    // place in the two work fields, as a plain run does, and stores the Laplacian of its current
    // level as a second, nontemporal output of the launch: C slots per segment (L u^{s0} ..
    // L u^{s1-1}) instead of C - 1, one more recompute step per segment.  1D and 3D grids step into
    // the kept fields directly and the adjoint reads u.
    bool lu = false;
    size_t bytes = 0;
    size_t batch_bytes = 0;  // one field batch (nshots fields)
    char* mem = nullptr;
    std::vector<void*> ck_prev, ck_cur, slot;
    void* work[2] = {nullptr, nullptr};
    void* mu[2] = {nullptr, nullptr};
    void* gpart = nullptr;
    bool forward_done = false;
};

namespace {

int plan_fail(int code, const char* msg) { return wfd::set_error(code, msg); }

// G += the imaging partials (one per shot, or one per shot group), which are cleared
int flush_partials(const wfd_plan* p, double* G, hipStream_t st)
{
    const int groups = wfd::img_groups(&p->g);
    const int parts = groups ? groups : p->g.nshots;
    const hipError_t e = p->g.dtype == WFD_F64
                             ? wfd::launch_flush<double>((double*)p->gpart, G, p->g.elems, parts, st)
                             : wfd::launch_flush<float>((float*)p->gpart, G, p->g.elems, parts, st);
    return e == hipSuccess ? WFD_OK : plan_fail(WFD_EHIP, "wfd_plan_backward: flush kernel failed");
}

size_t fields_needed(int64_t N, int64_t C, bool lu)
{
    const int64_t nck = (N + C - 1) / C;
    // C == 1: checkpoint k's older level is checkpoint k-1's newer one; only u^{-1} is extra
    // (with lu, checkpoint field k holds L u^k)
    const int64_t ck_fields = C == 1 ? nck + 1 : 2 * nck;
    const int64_t slots = C == 1 ? 0 : (lu ? C : C - 1);
    return (size_t)(ck_fields + slots + 2 + 2 + 1);
}

void* field_at(const wfd_plan* p, int64_t n, int64_t s0, int64_t k)
{
    return n == s0 ? p->ck_cur[k] : p->slot[n - s0 - 1];
}

}  // namespace

extern "C" {

int wfd_plan_create(wfd_plan** out, const wfd_grid* g, int64_t nsteps, int64_t seg, size_t max_bytes)
{
    if (!out || !g) return plan_fail(WFD_EINVAL, "wfd_plan_create: NULL argument");
    *out = nullptr;
    if (g->ld <= 0 || g->elems <= 0 || g->nshots < 1) return plan_fail(WFD_EINVAL, "wfd_plan_create: grid not initialised");
    if (nsteps < 1) return plan_fail(WFD_EINVAL, "wfd_plan_create: nsteps must be >= 1");
    if ((seg < 0 && seg != WFD_SEG_FASTEST) || seg > nsteps)
        return plan_fail(WFD_EINVAL, "wfd_plan_create: seg must be in [0, nsteps] or WFD_SEG_FASTEST");
    const size_t batch = (size_t)g->elems * (size_t)g->nshots * (g->dtype == WFD_F64 ? 8 : 4);
    const bool lu = g->ndim == 2;
    if (seg == WFD_SEG_FASTEST) {
        // every level kept (no recompute: the forward and the adjoint pass only)
        // when that store fits the budget
        size_t budget = max_bytes;
        if (!budget) {
            size_t free_b = 0, total_b = 0;
            if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return plan_fail(WFD_EHIP, "wfd_plan_create: hipMemGetInfo failed");
            budget = (size_t)(0.8 * (double)free_b);
        }
        seg = fields_needed(nsteps, 1, lu) * batch <= budget ? 1 : 0;
    }
    if (seg == 0) {
        int64_t best = 1;
        size_t best_f = fields_needed(nsteps, 1, lu);
        for (int64_t C = 2; C <= nsteps; C++) {
            const size_t f = fields_needed(nsteps, C, lu);
            if (f < best_f) {
                best_f = f;
                best = C;
            }
        }
        seg = best;
    }
    if (max_bytes) {
        // shrink toward the memory-optimal spacing until the store fits
        if (fields_needed(nsteps, seg, lu) * batch > max_bytes) {
            int64_t best = -1;
            for (int64_t C = 1; C <= nsteps; C++)
                if (fields_needed(nsteps, C, lu) * batch <= max_bytes) {
                    best = C;
                    break;
                }
            if (best < 0) return plan_fail(WFD_ENOMEM, "wfd_plan_create: no checkpoint spacing fits max_bytes");
            seg = best;
        }
    }
    wfd_plan* p = new (std::nothrow) wfd_plan;
    if (!p) return plan_fail(WFD_ENOMEM, "wfd_plan_create: host allocation failed");
    p->g = *g;
    p->nsteps = nsteps;
    p->seg = seg;
    p->lu = lu;
    p->nckpt = (nsteps + seg - 1) / seg;
    p->batch_bytes = batch;
    const size_t nf = fields_needed(nsteps, seg, lu);
    p->bytes = nf * batch;
    hipError_t e = hipMalloc(&p->mem, p->bytes);
    if (e != hipSuccess) {
        delete p;
        return plan_fail(WFD_ENOMEM, "wfd_plan_create: hipMalloc of the checkpoint store failed");
    }
    // The kernels write interior nodes only: halo, padding and boundary nodes of every field in
    // the store must start (and so stay) zero.
    e = hipMemset(p->mem, 0, p->bytes);
    if (e != hipSuccess) {
        (void)hipFree(p->mem);
        delete p;
        return plan_fail(WFD_EHIP, "wfd_plan_create: clearing the checkpoint store failed");
    }
    char* q = p->mem;
    auto take = [&]() {
        void* r = q;
        q += batch;
        return r;
    };
    p->ck_cur.resize(p->nckpt);
    p->ck_prev.resize(p->nckpt);
    for (int64_t k = 0; k < p->nckpt; k++) p->ck_cur[k] = take();
    for (int64_t k = 0; k < p->nckpt; k++) p->ck_prev[k] = seg == 1 ? (k ? p->ck_cur[k - 1] : nullptr) : take();
    if (seg == 1) p->ck_prev[0] = take();  // u^{-1}
    p->slot.resize(seg > 1 ? (lu ? seg : seg - 1) : 0);
    for (auto& s : p->slot) s = take();
    p->work[0] = take();
    p->work[1] = take();
    p->mu[0] = take();
    p->mu[1] = take();
    p->gpart = take();
    if ((size_t)(q - p->mem) > p->bytes) {
        (void)hipFree(p->mem);
        delete p;
        return plan_fail(WFD_EINVAL, "wfd_plan_create: internal layout error");
    }
    *out = p;
    return WFD_OK;
}

int wfd_plan_destroy(wfd_plan* p)
{
    if (!p) return WFD_OK;
    hipError_t e = hipFree(p->mem);
    delete p;
    return e == hipSuccess ? WFD_OK : plan_fail(WFD_EHIP, "wfd_plan_destroy: hipFree failed");
}

int wfd_plan_info(const wfd_plan* p, int64_t* seg, int64_t* nckpt, size_t* bytes)
{
    if (!p) return plan_fail(WFD_EINVAL, "wfd_plan_info: NULL plan");
    if (seg) *seg = p->seg;
    if (nckpt) *nckpt = p->nckpt;
    if (bytes) *bytes = p->bytes;
    return WFD_OK;
}

int wfd_plan_forward(wfd_plan* p, const wfd_medium* med, const wfd_points* src, const wfd_points* rec,
                     void* u_last, wfd_stream_t stream)
{
    if (!p) return plan_fail(WFD_EINVAL, "wfd_plan_forward: NULL plan");
    int rc = wfd::check_medium_args(&p->g, med);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const int64_t N = p->nsteps, C = p->seg;
    // every level kept with stored Laplacians: the levels rotate in the work fields and step n
    // stores L u^n into checkpoint field n; otherwise checkpoints receive u directly
    const bool lu_all = p->lu && C == 1;
    void* P = lu_all ? p->work[0] : p->ck_prev[0];
    void* Cu = lu_all ? p->work[1] : p->ck_cur[0];
    hipError_t e = wfd::launch_zero(P, (int64_t)p->batch_bytes, st);
    if (e == hipSuccess) e = wfd::launch_zero(Cu, (int64_t)p->batch_bytes, st);
    if (e != hipSuccess) return plan_fail(WFD_EHIP, "wfd_plan_forward: zero fill failed");
    for (int64_t n = 0; n < N; n++) {
        const int64_t m = n + 1;
        void* keep = nullptr;  // the checkpoint field that keeps u^m, if any
        if (!lu_all) {
            if (C > 1 && m % C == C - 1 && (m + 1) / C < p->nckpt)
                keep = p->ck_prev[(m + 1) / C];
            else if (m % C == 0 && m / C < p->nckpt)
                keep = p->ck_cur[m / C];
        }
        void* D = nullptr;
        if (keep)
            D = keep;
        else if (P == p->work[0] || P == p->work[1])
            D = P;  // in place over the oldest level
        else
            D = Cu == p->work[0] ? p->work[1] : p->work[0];
        // the launch also records u^n (its current level) into trace row n-1
        rc = wfd::step_launch(&p->g, P, Cu, D, med, src, rec, n, -1, nullptr, nullptr, false, st, false,
                              lu_all ? p->ck_cur[n] : nullptr);
        if (rc) return rc;
        P = Cu;
        Cu = D;
    }
    rc = wfd::gather_launch(&p->g, Cu, rec, N - 1, st);
    if (rc) return rc;
    if (u_last) {
        e = hipMemcpyAsync(u_last, P, p->batch_bytes, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync((char*)u_last + p->batch_bytes, Cu, p->batch_bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return plan_fail(WFD_EHIP, "wfd_plan_forward: copy of the final state failed");
    }
    p->forward_done = true;
    return WFD_OK;
}

int wfd_plan_backward(wfd_plan* p, const wfd_medium* med, const wfd_points* src, const wfd_points* adj,
                      const wfd_points* src_rec, double* G, wfd_stream_t stream)
{
    if (!p || !G) return plan_fail(WFD_EINVAL, "wfd_plan_backward: NULL argument");
    if (!p->forward_done) return plan_fail(WFD_EINVAL, "wfd_plan_backward: run wfd_plan_forward first");
    int rc0 = wfd::check_medium_args(&p->g, med);
    if (rc0) return rc0;
    hipStream_t st = (hipStream_t)stream;
    const int64_t N = p->nsteps, C = p->seg;
    const size_t gbytes = (size_t)p->g.elems * sizeof(double);
    hipError_t e = wfd::launch_zero(p->mu[0], (int64_t)p->batch_bytes, st);
    if (e == hipSuccess) e = wfd::launch_zero(p->mu[1], (int64_t)p->batch_bytes, st);
    if (e == hipSuccess) e = wfd::launch_zero(p->gpart, (int64_t)p->batch_bytes, st);
    if (e == hipSuccess) e = wfd::launch_zero(G, (int64_t)gbytes, st);
    if (e != hipSuccess) return plan_fail(WFD_EHIP, "wfd_plan_backward: zero fill failed");
    void* muP = p->mu[0];  // mu^{m+2}
    void* muC = p->mu[1];  // mu^{m+1}
    for (int64_t k = p->nckpt - 1; k >= 0; k--) {
        const int64_t s0 = k * C, s1 = std::min(s0 + C, N);
        if (p->lu) {
            // recompute in the two work fields (in place, like the forward); step n stores L u^n
            // into slot n - s0 (C == 1: the forward stored them)
            const void* P = p->ck_prev[k];
            const void* Cu = p->ck_cur[k];
            for (int64_t n = s0; C > 1 && n < s1; n++) {
                void* D = P == p->work[0] || P == p->work[1] ? const_cast<void*>(P)
                                                              : (Cu == p->work[0] ? p->work[1] : p->work[0]);
                int rc = wfd::step_launch(&p->g, P, Cu, D, med, src, nullptr, n, 0, nullptr, nullptr, false, st,
                                          false, p->slot[n - s0]);
                if (rc) return rc;
                P = Cu;
                Cu = D;
            }
        } else {
            for (int64_t n = s0; n + 1 < s1; n++) {
                const void* prev = n == s0 ? p->ck_prev[k] : field_at(p, n - 1, s0, k);
                int rc = wfd::step_launch(&p->g, prev, field_at(p, n, s0, k), field_at(p, n + 1, s0, k), med, src,
                                          nullptr, n, 0, nullptr, nullptr, false, st);
                if (rc) return rc;
            }
        }
        for (int64_t m = s1; m > s0; m--) {
            // adjoint step m (signal row m-1) makes mu^m; its launch records mu^{m+1} (its current
            // level) into row m of src_rec
            const void* F = !p->lu ? field_at(p, m - 1, s0, k) : C == 1 ? p->ck_cur[m - 1] : p->slot[m - 1 - s0];
            int rc = wfd::step_launch(&p->g, muP, muC, muP, med, adj, src_rec, m - 1, +1, F, p->gpart, true, st, true);
            if (rc) return rc;
            std::swap(muP, muC);  // muC = mu^m, muP = mu^{m+1}
            if ((m - 1) % p->flush == 0) {
                rc = flush_partials(p, G, st);
                if (rc) return rc;
            }
        }
    }
    return wfd::gather_launch(&p->g, muC, src_rec, 0, st);  // mu^1 into row 0
}

}  // extern "C"
