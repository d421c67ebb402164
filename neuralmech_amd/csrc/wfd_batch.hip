// Uneven shot batch: the checkpointed forward run, the trace residual and the adjoint run with
// the imaging condition of many independent 2D shots (each with its own grid size, order, step
// count and medium) in ONE persistent kernel launch.
//
// Host side.  Every shot's work is compiled into a list of ops over the fields of its own store
// (the schedule of the gradient plan: forward steps writing straight into checkpoint storage,
// then per segment a recompute into the segment slots and the adjoint steps over it), one op per
// time step.  A list scheduler gives each shot a start phase so that the interior cells in
// flight stay near a target (a working set that stays in the last-level cache while filling the
// GPU); phase p then holds one item per shot active at p.  Items of a phase are laid end to end
// as tiles (stencil tiles, then the 256-point gather tiles of the item).
//
// Device side.  A grid of resident blocks (one per CU) is split into 8 groups by the XCD each
// block runs on (the XCC_ID hardware register; blocks register at kernel start to learn their rank
// and their group's size).  Shots are assigned to groups longest-processing-time first and each
// group walks its own phases: its blocks stride through the tiles of the phase, then meet at a
// group barrier.  All of a shot's fields are touched by one XCD only, so they stay coherent in
// that XCD's L2: the barrier needs no L2 write-back, only the store-completion wait and one cache
// invalidation per block, and no group ever waits for another.  Every wait is bounded by a
// wall-clock limit that sets an error word and ends the kernel, so the launch always terminates
// even if the grid were not co-resident.  The per-phase mode (one launch per phase of a single
// global schedule) shares the tile code.
//
// Fusions (the same bits as the separate launches of the gradient plan): the imaging partial is
// flushed to the float64 sum inside the adjoint step that would be followed by a flush; the
// receiver gather of the forward run also writes the adjoint source -(trace - observed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "wfd_internal.h"
#include "wfd_step2d_impl.h"

namespace wfd {

// ------------------------------------------------------------------------------ device side
constexpr int BATCH_MAX_ITEMS = 256;  // items per phase (a shot has one or two per phase)
constexpr uint32_t F_INJECT = 1, F_FLUSH = 2, F_FIRST = 4, F_GREC = 8, F_GSRC = 16;
constexpr int K_GATHER = 0, K_STEP = 1, K_ADJ = 2;

// Tile shapes of the batch kernel (2D), codes 0..3: variant (1 = lane-exchange halo) and rows per wave.
__host__ __device__ constexpr int cfg_variant(int c) { return c < 2 ? 1 : 0; }
__host__ __device__ constexpr int cfg_rows(int c) { return (c & 1) ? 2 : 8; }

template <typename T>
struct BShot {
    StepArgs<T> a;  // geometry, medium, sponge, weights (field pointers set per item)
    T* fields;      // field 0 of the shot's store, node (0, 0)
    int64_t fstride;
    PtsDev<T> src, adj, rec, msrc;  // sources, receivers with the adjoint signal, receivers
                                    // recording traces, sources recording mu
    const T* obs;                   // observed traces [nsteps][rec.width]
    double* G;                      // node (0, 0) of the imaging sum
    uint32_t gx, gy, stiles;        // tile grid of forward / recompute steps
    uint32_t gy_adj, stiles_adj;    // adjoint steps (ADJ_ROWS rows per wave, float64: 2)
    int32_t K, cfg;
};

struct BItem {
    uint32_t tile0;   // first tile of the item in its phase
    uint32_t ntiles;  // stencil tiles + gather tiles
    uint16_t shot;
    uint8_t kind;   // K_GATHER, K_STEP (forward / recompute), K_ADJ (adjoint + imaging)
    uint8_t flags;  // F_*
    int16_t up, uc, un, uf, gp, gf;  // field indices (uf / gp: imaging; gf: gathered field)
    int32_t step;  // signal row of the step
    int32_t grow;  // trace row of the gather
};
static_assert(sizeof(BItem) == 32, "BItem layout");

struct BatchSync {
    uint32_t arrive[XCDS * 32];   // per-group arrival counters, 128 bytes apart
    uint32_t release[XCDS * 32];  // per-group released generation
    uint32_t reg[XCDS * 32];      // blocks registered per XCD
    uint32_t reg_all[32];         // blocks registered in total
    uint32_t error[32];           // bit 0: a wait timed out; bit 1: a group with work has no block
};

template <typename T>
struct BatchArgs {
    const BShot<T>* shots;
    const BItem* items;
    const uint32_t* phase_off;  // [nphases + 1]
    int64_t p0, p1;             // phases of this launch
    BatchSync* sync;
    uint64_t timeout;  // wall-clock ticks per barrier wait
    int32_t persistent;
};

// Phase barrier of one XCD group.  Every block of the group runs on the same XCD (checked against
// the XCC_ID hardware register when the kernel starts), so the group's fields live in one L2: a
// wave's stores are visible to the group once they have completed (s_waitcnt vmcnt(0)) and no L2
// write-back is needed.  What a CU may still hold are stale L1 lines of tiles other CUs rewrote;
// on gfx950 only the device-scope invalidate clears them (a workgroup-scope one leaves the L1 of a
// non-split dispatch alone), so one thread per block issues it and waits before the block goes on.
// The last block to arrive publishes the generation in a release word the others poll (polling
// the arrival counter itself slows the arrivals); every wait is bounded.
__device__ __forceinline__ bool group_barrier(BatchSync* s, uint32_t g, uint32_t gen, uint32_t bpx, uint64_t timeout)
{
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are in the L2
    __syncthreads();
    __shared__ uint32_t s_ok;
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        const uint32_t old = __hip_atomic_fetch_add(&s->arrive[g * 32], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (old + 1 == gen * bpx) {
            __hip_atomic_store(&s->release[g * 32], gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (uint32_t it = 0;
                 __hip_atomic_load(&s->release[g * 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen; it++) {
                if ((it & 31) == 31 && (__hip_atomic_load(&s->error[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                                        __builtin_amdgcn_s_memrealtime() - t0 > timeout)) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!ok) __hip_atomic_fetch_or(&s->error[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        asm volatile("buffer_inv sc1" ::: "memory");  // this CU's L1 (and the XCD's L2)
        __builtin_amdgcn_s_waitcnt(0x0F70);
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// XCC (XCD) that runs the calling wave, from the XCC_ID hardware register.
__device__ __forceinline__ uint32_t hw_xcc_id() { return __builtin_amdgcn_s_getreg((15 << 11) | 20) & 15u; }

// Forward / recompute tiles march cfg_rows(cfg) rows per wave (8: fewer re-reads of the window
// rows and fewer tiles per phase; 3.44 s against 3.53 s with 4 rows for the 64-shot batch);
// float adjoint tiles march ADJ_ROWS = 4 (8 once they read the stored Laplacian instead of a
// second window: 3.165 s against 3.148 s) and float64 tiles 2, which keeps
// the kernel at 3 waves per SIMD.  At that occupancy the 8-row tiles load their rows' operands 2
// rows ahead and the float adjoint tiles 1 row ahead (157 VGPRs, no spills, since they read the
// stored Laplacian: 3.149 -> 3.119 s; with the forward field's window it had spilled 41 VGPRs),
// the others row by row.  4-row adjoint tiles re-read 12 window rows per 4 output rows instead of 10 per
// 2 and halve the adjoint tiles per phase: 3.350 -> 3.258 s for the 64-shot batch, though the
// kernel then spills 3 VGPRs (2 slots per CU with 8-row adjoint tiles and no spills: 3.368 s).
constexpr int ADJ_ROWS = 4;
// Adjoint tiles read the Laplacian L u^{m-1} that the recompute stored (LUF); recompute tiles
// (a.ucopy set) store it as a second output while their levels rotate in the work fields.
template <typename T, int K, bool IMG, bool CPY>
__device__ __forceinline__ void batch_tile_kc(const StepArgs<T>& a, int cfg, uint32_t tx, uint32_t ty, int64_t step)
{
    if constexpr (sizeof(T) == 8) {
        if (cfg_variant(cfg) == 1)
            step2d_tile<T, K, true, true, true, IMG, 2, true, 0, false, CPY, IMG>(a, tx, ty, 0, step);
        else
            step2d_tile<T, K, true, true, true, IMG, 2, false, 0, false, CPY, IMG>(a, tx, ty, 0, step);
    } else if constexpr (IMG) {
        if (cfg_variant(cfg) == 1)
            step2d_tile<T, K, true, true, true, IMG, ADJ_ROWS, true, 1, false, false, true>(a, tx, ty, 0, step);
        else
            step2d_tile<T, K, true, true, true, IMG, ADJ_ROWS, false, 1, false, false, true>(a, tx, ty, 0, step);
    } else {
        switch (cfg) {
            case 0: step2d_tile<T, K, true, true, true, IMG, 8, true, 2, false, CPY>(a, tx, ty, 0, step); break;
            case 1: step2d_tile<T, K, true, true, true, IMG, 2, true, 0, false, CPY>(a, tx, ty, 0, step); break;
            case 2: step2d_tile<T, K, true, true, true, IMG, 8, false, 2, false, CPY>(a, tx, ty, 0, step); break;
            default: step2d_tile<T, K, true, true, true, IMG, 2, false, 0, false, CPY>(a, tx, ty, 0, step); break;
        }
    }
}

template <typename T, int K, bool IMG>
__device__ __forceinline__ void batch_tile_k(const StepArgs<T>& a, int cfg, uint32_t tx, uint32_t ty, int64_t step)
{
    if (!IMG && a.ucopy)
        batch_tile_kc<T, K, IMG, !IMG>(a, cfg, tx, ty, step);
    else
        batch_tile_kc<T, K, IMG, false>(a, cfg, tx, ty, step);
}

template <typename T, bool IMG>
__device__ __forceinline__ void batch_tile(const StepArgs<T>& a, int K, int cfg, uint32_t tx, uint32_t ty,
                                           int64_t step)
{
    switch (K) {
        case 1: batch_tile_k<T, 1, IMG>(a, cfg, tx, ty, step); break;
        case 2: batch_tile_k<T, 2, IMG>(a, cfg, tx, ty, step); break;
        case 3: batch_tile_k<T, 3, IMG>(a, cfg, tx, ty, step); break;
        default: batch_tile_k<T, 4, IMG>(a, cfg, tx, ty, step); break;
    }
}

template <typename T>
__device__ __forceinline__ void batch_item(const BShot<T>& sh, const BItem& it, uint32_t local)
{
    T* const F = sh.fields;
    const int64_t fs = sh.fstride;
    const uint32_t stiles = it.kind == K_ADJ ? sh.stiles_adj : (it.kind == K_STEP ? sh.stiles : 0u);
    if (local < stiles) {
        StepArgs<T> a = sh.a;
        a.up = F + it.up * fs;
        a.uc = F + it.uc * fs;
        a.un = F + it.un * fs;
        const bool img = it.kind == K_ADJ;
        a.src = img ? sh.adj : sh.src;
        if (!(it.flags & F_INJECT)) a.src.count = 0;
        const uint32_t tx = local % sh.gx, ty = local / sh.gx;
        a.ucopy = !img && it.uf >= 0 ? F + it.uf * fs : nullptr;  // recompute: L u^n into field uf
        if (img) {
            a.uf = F + it.uf * fs;
            a.g = F + it.gp * fs;
            a.gsum = (it.flags & F_FLUSH) ? sh.G : nullptr;
            a.gsum_first = (it.flags & F_FIRST) ? 1 : 0;
            batch_tile<T, true>(a, sh.K, sh.cfg, tx, ty, it.step);
        } else {
            batch_tile<T, false>(a, sh.K, sh.cfg, tx, ty, it.step);
        }
        return;
    }
    // gather tile: 256 points of the item's receivers (+ adjoint source) or sources (mu)
    const uint32_t gt = local - stiles;
    const bool rec = (it.flags & F_GREC) != 0;
    const PtsDev<T>& p = rec ? sh.rec : sh.msrc;
    const int64_t q = (int64_t)gt * 256 + (threadIdx.x & 255);
    if (q >= p.count || it.grow < 0 || it.grow >= p.nt) return;
    const T* f = F + it.gf * fs;
    const T u = f[(int64_t)p.row[q] * sh.a.ld + p.col[q]];
    const int64_t d = (int64_t)it.grow * p.width + p.id[q];
    p.data[d] = u;
    if (rec) sh.adj.data[d] = -(u - sh.obs[d]);
}

struct GroupPhases {
    int64_t b[XCDS + 1];  // group g runs phases [b[g], b[g+1])
};

// Threads per block of the persistent launch: one block per CU, made of 256-thread tile slots
// (4 waves each).  float32: 3 slots (12 waves, 3 per SIMD at <= 168 VGPRs); float64: 1 slot.
template <typename T>
constexpr int batch_block() { return sizeof(T) == 4 ? 768 : 256; }

// Read-only tables as separate restrict parameters: their loads are provably unclobbered by
// the field stores, so uniform ones become scalar loads.
//
// persistent != 0: group g = the XCD a block runs on (XCC_ID register) runs its own phases; the
// tile slots of the group's blocks stride through each phase's tiles and meet at group_barrier.
// persistent == 0: one launch per phase (p0 = p1 - 1), tiles split into 8 contiguous ranges, one
// per XCD under the usual round-robin placement of blocks.
template <typename T>
__global__ __launch_bounds__(batch_block<T>(), sizeof(T) == 4 ? 3 : 1) void batch_kernel(
    const BShot<T>* __restrict__ shots, const BItem* __restrict__ items, const uint32_t* __restrict__ phase_off,
    int64_t p0, int64_t p1, GroupPhases gp, BatchSync* sync, uint64_t timeout, int32_t persistent)
{
    __shared__ uint32_t s_t0[BATCH_MAX_ITEMS + 1];
    const uint32_t x = blockIdx.x % XCDS;
    const uint32_t nsub = blockDim.x >> 8, sub = threadIdx.x >> 8;  // tile slots of this block
    uint32_t g = x, bi = blockIdx.x / XCDS, bpx = gridDim.x / XCDS;
    if (persistent) {
        // group = the XCD this block runs on; rank and count of the group's blocks by registration
        // (the XCD the hardware starts its round-robin placement on varies between launches)
        __shared__ uint32_t s_reg[3];
        if (threadIdx.x == 0) {
            const uint32_t xcc = hw_xcc_id();
            uint32_t ok = xcc < XCDS ? 1u : 0u;
            const uint32_t gg = ok ? xcc : 0u;
            const uint32_t rank = __hip_atomic_fetch_add(&sync->reg[gg * 32], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&sync->reg_all[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
            for (uint32_t it = 0;
                 __hip_atomic_load(&sync->reg_all[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gridDim.x; it++) {
                if ((it & 31) == 31 && (__hip_atomic_load(&sync->error[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                                        __builtin_amdgcn_s_memrealtime() - t0 > timeout)) {
                    ok = 0;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            for (uint32_t k = 0; k < XCDS && ok; k++)  // every group with phases needs a block
                if (gp.b[k + 1] > gp.b[k] &&
                    __hip_atomic_load(&sync->reg[k * 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
                    ok = 0;
            if (!ok) __hip_atomic_fetch_or(&sync->error[0], 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            s_reg[0] = ok;
            s_reg[1] = gg;
            s_reg[2] = rank;
        }
        __syncthreads();
        if (!s_reg[0]) return;
        g = s_reg[1];
        bi = s_reg[2];
        bpx = __hip_atomic_load(&sync->reg[g * 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        p0 = gp.b[g];
        p1 = gp.b[g + 1];
    }
    for (int64_t p = p0; p < p1; p++) {
        const uint32_t i0 = phase_off[p], n = phase_off[p + 1] - i0;
        for (uint32_t t = threadIdx.x; t < n; t += blockDim.x) s_t0[t] = items[i0 + t].tile0;
        if (threadIdx.x == 0) s_t0[n] = n ? items[i0 + n - 1].tile0 + items[i0 + n - 1].ntiles : 0u;
        __syncthreads();
        const uint32_t total = s_t0[n];
        uint32_t lo = 0, hi = total;
        if (!persistent) {
            const uint32_t per = total / XCDS, rem = total % XCDS;
            lo = x * per + (x < rem ? x : rem);
            hi = lo + per + (x < rem ? 1u : 0u);
        }
        for (uint32_t j = lo + bi * nsub + sub; j < hi; j += bpx * nsub) {
            uint32_t a = 0, b = n;  // largest e with s_t0[e] <= j
            while (b - a > 1) {
                const uint32_t m = (a + b) >> 1;
                if (s_t0[m] <= j) a = m;
                else b = m;
            }
            const uint32_t e = __builtin_amdgcn_readfirstlane(a);
            const uint32_t local = __builtin_amdgcn_readfirstlane(j - s_t0[e]);
            const BItem& it = items[i0 + e];
            batch_item<T>(shots[it.shot], it, local);
        }
        if (p + 1 < p1) {
            if (persistent) {
                if (!group_barrier(sync, g, (uint32_t)(p - p0 + 1), bpx, timeout)) return;
            } else {
                __syncthreads();  // s_t0 reuse
            }
        }
    }
}

template <typename T>
hipError_t launch_batch(const BatchArgs<T>& A, const GroupPhases& gp, uint32_t blocks, int threads, hipStream_t s)
{
    hipLaunchKernelGGL((batch_kernel<T>), dim3(blocks), dim3(threads), 0, s, A.shots, A.items, A.phase_off, A.p0,
                       A.p1, gp, A.sync, A.timeout, A.persistent);
    return hipGetLastError();
}

template <typename T>
int batch_occupancy(int* blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, batch_kernel<T>, batch_block<T>(), 0) == hipSuccess
               ? 0
               : 1;
}

}  // namespace wfd

// ------------------------------------------------------------------------------ host side
using namespace wfd;

namespace {

int bfail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return set_error(code, buf);
}

struct HostOp {
    int32_t kind, flags, up, uc, un, uf, gp, gf, step, grow;
    int32_t phase;  // phase of the shot's schedule the op runs in (ops of one phase run concurrently)
};

// Fields of a shot's store: checkpoints (as in the gradient plan), two sets of C segment slots
// holding L u^n for the imaging (the backward recomputes segment k-1 into one set while the
// adjoint of segment k reads the other), two work fields, two adjoint fields and the imaging
// partial.  With C = 1 checkpoint field n holds L u^n.
size_t batch_fields(int64_t N, int64_t C)
{
    const int64_t nck = (N + C - 1) / C;
    const int64_t ck_fields = C == 1 ? nck + 1 : 2 * nck;
    return (size_t)(ck_fields + (C == 1 ? 0 : (nck >= 2 ? 2 : 1) * C) + 5);
}

int64_t optimal_seg(int64_t N)
{
    int64_t best = 1;
    size_t best_f = batch_fields(N, 1);
    for (int64_t C = 2; C <= N; C++) {
        const size_t f = batch_fields(N, C);
        if (f < best_f) {
            best_f = f;
            best = C;
        }
    }
    return best;
}

// The ops of one shot over field indices: the forward run of wfd_plan_forward (steps writing
// straight into checkpoint storage; with C = 1 the levels rotate in the work fields and each step
// stores L u^n into checkpoint field n), then the backward of wfd_plan_backward with the recompute
// of segment k-1 running in the same phases as the adjoint of segment k (they touch disjoint
// fields: segments alternate between the two slot sets).  The recompute steps in place in the
// work fields and stores L u^n into the segment's slot n - s0; the adjoint reads it pointwise.  A
// checkpoint field that is never written (checkpoint 0's newer level, or with C = 1 u^{-1})
// doubles as the zero field the adjoint starts from.  Every value, and the order of every
// floating-point operation, is the plan's.  Returns the number of fields.
int shot_ops(int64_t N, int64_t C, bool has_src, bool has_rec, std::vector<HostOp>& ops)
{
    const int64_t nck = (N + C - 1) / C;
    int nf = 0;
    std::vector<int> ck_cur(nck), ck_prev(nck), slot_a(C > 1 ? C : 0), slot_b(C > 1 && nck >= 2 ? C : 0);
    for (int64_t k = 0; k < nck; k++) ck_cur[k] = nf++;
    if (C == 1) {
        ck_prev[0] = nf++;
        for (int64_t k = 1; k < nck; k++) ck_prev[k] = ck_cur[k - 1];
    } else {
        for (int64_t k = 0; k < nck; k++) ck_prev[k] = nf++;
    }
    for (auto& f : slot_a) f = nf++;
    for (auto& f : slot_b) f = nf++;
    const int w0 = nf++, w1 = nf++, mu0 = nf++, mu1 = nf++, gpart = nf++;
    const int Z = C == 1 ? ck_prev[0] : ck_cur[0];
    ops.clear();
    int32_t ph = 0;
    auto mk = [&](int kind, int flags, int up, int uc, int un, int uf, int gp, int gf, int64_t step, int64_t grow) {
        return HostOp{kind, flags, up, uc, un, uf, gp, gf, (int32_t)step, (int32_t)grow, 0};
    };
    auto emit = [&](HostOp o) {
        o.phase = ph++;
        ops.push_back(o);
    };
    // forward
    int P = ck_prev[0], Cu = C == 1 ? Z : ck_cur[0];  // u^{-1} = u^0 = 0
    for (int64_t n = 0; n < N; n++) {
        const int64_t m = n + 1;
        int D;
        if (C > 1 && m % C == C - 1 && (m + 1) / C < nck)
            D = ck_prev[(m + 1) / C];
        else if (C > 1 && m % C == 0 && m / C < nck)
            D = ck_cur[m / C];
        else if (P == w0 || P == w1)
            D = P;
        else
            D = Cu == w0 ? w1 : w0;
        const int fl = (has_src ? (int)F_INJECT : 0) | (has_rec && n >= 1 ? (int)F_GREC : 0);
        emit(mk(K_STEP, fl, P, Cu, D, C == 1 ? ck_cur[n] : -1, -1, Cu, n, n - 1));
        P = Cu;
        Cu = D;
    }
    const int last = Cu;  // u^N: receiver row N-1 is gathered from it
    // backward
    auto slots = [&](int64_t k) -> const std::vector<int>& { return ((nck - 1 - k) & 1) ? slot_b : slot_a; };
    // L u^n of a segment's level n (C = 1: checkpoint field n)
    auto lu_at = [&](int64_t n, int64_t s0, int64_t k) { return C == 1 ? ck_cur[n] : slots(k)[n - s0]; };
    auto recompute = [&](int64_t k) {
        std::vector<HostOp> r;
        if (C == 1) return r;
        const int64_t s0 = k * C, s1 = std::min(s0 + C, N);
        int P = ck_prev[k], Cu = ck_cur[k];
        for (int64_t n = s0; n < s1; n++) {
            // in place over the oldest work level; never into u^N (gathered in the first phase)
            const int D = (P == w0 || P == w1) ? P : (Cu == w0 || Cu == w1) ? (Cu == w0 ? w1 : w0) : (last == w0 ? w1 : w0);
            r.push_back(mk(K_STEP, has_src ? F_INJECT : 0, P, Cu, D, lu_at(n, s0, k), -1, -1, n, -1));
            P = Cu;
            Cu = D;
        }
        return r;
    };
    int muP = Z, muC = Z;
    bool first = true;
    auto adjoint = [&](int64_t k) {
        std::vector<HostOp> a;
        const int64_t s0 = k * C, s1 = std::min(s0 + C, N);
        for (int64_t m = s1; m > s0; m--) {
            const int dest = muP != Z ? muP : (muC == mu0 ? mu1 : mu0);
            int fl = (has_rec ? (int)F_INJECT : 0) | (has_src && m <= N - 1 ? (int)F_GSRC : 0);
            if ((m - 1) % 32 == 0) {
                fl |= F_FLUSH | (first ? (int)F_FIRST : 0);
                first = false;
            }
            a.push_back(mk(K_ADJ, fl, muP, muC, dest, lu_at(m - 1, s0, k), gpart, muC, m - 1, m));
            muP = muC;
            muC = dest;
        }
        return a;
    };
    // the last segment's recompute runs alone; the final receiver gather shares its first phase
    std::vector<HostOp> r = recompute(nck - 1);
    bool gathered = !has_rec;
    for (HostOp& o : r) {
        if (!gathered) {
            o.flags |= F_GREC;
            o.gf = last;
            o.grow = (int32_t)(N - 1);
            gathered = true;
        }
        emit(o);
    }
    if (!gathered) emit(mk(K_GATHER, F_GREC, -1, -1, -1, -1, -1, last, -1, N - 1));
    for (int64_t k = nck - 1; k >= 0; k--) {
        const std::vector<HostOp> a = adjoint(k);
        const std::vector<HostOp> rr = k >= 1 ? recompute(k - 1) : std::vector<HostOp>();
        for (size_t i = 0; i < std::max(a.size(), rr.size()); i++) {
            if (i < a.size()) {
                ops.push_back(a[i]);
                ops.back().phase = ph;
            }
            if (i < rr.size()) {
                ops.push_back(rr[i]);
                ops.back().phase = ph;
            }
            ph++;
        }
    }
    if (has_src) emit(mk(K_GATHER, F_GSRC, -1, -1, -1, -1, -1, muC, -1, 0));
    return nf;
}

int size_class(const wfd_grid& g)
{
    const int64_t n = std::max(g.n[0], g.n[1]);
    if (n <= 768) return 0;
    if (n <= 1024) return 1;
    if (n <= 1280) return 2;
    if (n <= 1536) return 3;
    return 4;
}
constexpr int NCLASS = 5;

}  // namespace

// A schedule: the phase tables on the device and, per group, its range of phases.
struct Sched {
    BItem* items = nullptr;
    uint32_t* off = nullptr;
    int64_t nphases = 0;
    GroupPhases gp{};
    std::vector<uint32_t> tiles;          // tiles per phase
    std::vector<int64_t> start, group;    // per shot: global first phase, group
    void release()
    {
        (void)hipFree(items);
        (void)hipFree(off);
        items = nullptr;
        off = nullptr;
    }
};

struct wfd_batch {
    int dtype = WFD_F32;
    int nshots = 0;
    wfd_batch_opts opts{};
    std::vector<wfd_batch_shot> in;
    std::vector<std::vector<HostOp>> ops;
    std::vector<int64_t> seg, nfields, fbytes;
    std::vector<int> cfg, cls;
    size_t bytes = 0;
    char* store = nullptr;  // all shots' fields
    void* zprof = nullptr;  // zero damping profile for shots without a sponge
    void* shots_dev = nullptr;
    BatchSync* sync = nullptr;
    Sched sched;
    uint32_t blocks = 0;
    uint64_t timeout_ticks = 0;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = false;  // a run failed: the partial sums may be non-zero
};

namespace {

template <typename T>
void fill_pts(PtsDev<T>& p, const wfd_points& w, void* data, int64_t width, int64_t nt)
{
    std::memset(&p, 0, sizeof p);
    if (w.count <= 0) return;
    p.row_ptr = w.row_ptr;
    p.col = w.col;
    p.id = w.id;
    p.row = w.row;
    p.data = static_cast<T*>(data);
    p.width = width;
    p.shot_stride = nt * width;
    p.nt = nt;
    p.count = w.count;
    p.nrows = w.nrows;
}

void stiles_of(const wfd_grid& g, int cfg, bool adj, uint32_t& gx, uint32_t& gy)
{
    const int64_t cols = cfg_variant(cfg) == 1 ? 248 : 256;
    const int64_t ri = g.dtype == WFD_F64 ? 2 : (adj ? ADJ_ROWS : cfg_rows(cfg));  // as batch_tile_k
    gx = (uint32_t)((g.n[1] + cols - 1) / cols);
    gy = (uint32_t)((g.n[0] - 2 + 4 * ri - 1) / (4 * ri));
}

template <typename T>
int upload_shots(wfd_batch* b)
{
    std::vector<BShot<T>> hs(b->nshots);
    char* base = b->store;
    for (int s = 0; s < b->nshots; s++) {
        const wfd_batch_shot& in = b->in[s];
        const wfd_grid& g = in.grid;
        BShot<T>& h = hs[s];
        std::memset(&h, 0, sizeof h);
        StepArgs<T>& a = h.a;
        const int64_t off = g.offset;
        a.coef = static_cast<const T*>(in.med.coef) + off;
        a.coef_shot_stride = 0;
        a.kconst = T(0);
        a.np = 1;
        a.nr = g.n[0];
        a.nc = g.n[1];
        const T* beta = in.med.beta ? static_cast<const T*>(in.med.beta) : static_cast<const T*>(b->zprof);
        a.beta_r = beta;
        a.beta_c = in.med.beta ? beta + a.nr : beta;
        a.shot_stride = g.elems;
        a.ld = g.ld;
        a.ps = g.plane * g.ld;
        a.row_lo = 1;
        a.row_hi = a.nr - 1;
        double wc, w[8];
        wfd_stencil_weights(&g, &wc, w);
        const int K = g.order / 2;
        a.wc = (T)wc;
        for (int m = 0; m < 4; m++) {
            a.wr[m] = m < K ? (T)w[m] : T(0);
            a.wcol[m] = m < K ? (T)w[K + m] : T(0);
        }
        a.nshots = 1;
        h.fields = reinterpret_cast<T*>(base) + off;
        h.fstride = b->fbytes[s] / (int64_t)sizeof(T);
        base += b->fbytes[s] * b->nfields[s];
        const int64_t N = in.nsteps;
        fill_pts<T>(h.src, in.src, in.src.data, in.src.width, in.src.nsteps);
        fill_pts<T>(h.rec, in.rec, in.rec.data, in.rec.width, N);
        fill_pts<T>(h.adj, in.rec, in.adj, in.rec.width, N);
        fill_pts<T>(h.msrc, in.src, in.mu_src, in.src.width, N);
        h.obs = static_cast<const T*>(in.observed);
        h.G = in.G + off;
        h.K = K;
        h.cfg = b->cfg[s];
        stiles_of(g, h.cfg, false, h.gx, h.gy);
        h.stiles = h.gx * h.gy;
        uint32_t gxa = 0;
        stiles_of(g, h.cfg, true, gxa, h.gy_adj);
        h.stiles_adj = gxa * h.gy_adj;
    }
    if (!b->shots_dev) {
        if (hipMalloc(&b->shots_dev, sizeof(BShot<T>) * b->nshots) != hipSuccess)
            return bfail(WFD_ENOMEM, "wfd_batch: hipMalloc of the shot table failed");
    }
    if (hipMemcpy(b->shots_dev, hs.data(), sizeof(BShot<T>) * b->nshots, hipMemcpyHostToDevice) != hipSuccess)
        return bfail(WFD_EHIP, "wfd_batch: upload of the shot table failed");
    return WFD_OK;
}

int upload_shot_table(wfd_batch* b) { return b->dtype == WFD_F64 ? upload_shots<double>(b) : upload_shots<float>(b); }

uint32_t op_tiles(const wfd_batch* b, int s, const HostOp& o)
{
    uint32_t t = 0;
    if (o.kind != K_GATHER) {
        uint32_t gx, gy;
        stiles_of(b->in[s].grid, b->cfg[s], o.kind == K_ADJ, gx, gy);
        t = gx * gy;
    }
    if (o.flags & F_GREC) t += (uint32_t)((b->in[s].rec.count + 255) / 256);
    if (o.flags & F_GSRC) t += (uint32_t)((b->in[s].src.count + 255) / 256);
    return t;
}

int64_t shot_cells(const wfd_batch* b, int s) { return (b->in[s].grid.n[0] - 2) * (b->in[s].grid.n[1] - 2); }

// Shots -> groups (persistent: the 8 XCDs, longest processing time first on cells x ops), then
// per group a list schedule: longest shots first, each started at the first phase where the
// group's cells in flight stay within its share of max_active_cells.
int make_schedule(wfd_batch* b, const std::vector<std::vector<HostOp>>& ops, Sched& out)
{
    const int S = b->nshots;
    const bool grouped = b->opts.mode == 0;
    const int G = grouped ? (int)XCDS : 1;
    std::vector<int> order;
    for (int s = 0; s < S; s++)
        if (!ops[s].empty()) order.push_back(s);
    auto nph = [&](int s) { return ops[s].empty() ? (int64_t)0 : (int64_t)ops[s].back().phase + 1; };
    auto work = [&](int s) { return (double)shot_cells(b, s) * (double)ops[s].size(); };
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int64_t a = nph(x), c = nph(y);
        return a != c ? a > c : shot_cells(b, x) > shot_cells(b, y);
    });
    out.group.assign(S, 0);
    if (grouped) {
        std::vector<int> by_work(order);
        std::stable_sort(by_work.begin(), by_work.end(), [&](int x, int y) { return work(x) > work(y); });
        std::vector<double> load(G, 0.0);
        for (int s : by_work) {
            const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            out.group[s] = g;
            load[g] += work(s);
        }
    }
    // default: 16 Mi cells (64-shot batch of 512^2-1536^2 grids on one box: 12 Mi 3.260 s, 16 Mi
    // 3.218 s, 20 Mi 3.213 s, 24 Mi 3.337 s -- beyond that the shots in flight leave the cache)
    const int64_t total_target = b->opts.max_active_cells > 0 ? b->opts.max_active_cells : (int64_t)16 * 1024 * 1024;
    const int64_t target = std::max<int64_t>(1, total_target / G);
    std::vector<int64_t> local(S, 0), gphases(G, 0);
    for (int g = 0; g < G; g++) {
        using Ev = std::pair<int64_t, int64_t>;  // (end phase, cells)
        std::priority_queue<Ev, std::vector<Ev>, std::greater<Ev>> act;
        int64_t t = 0, cur = 0;
        for (int s : order) {
            if (out.group[s] != g) continue;
            while (!act.empty() && (cur + shot_cells(b, s) > target || 2 * (int)act.size() >= BATCH_MAX_ITEMS)) {
                t = std::max(t, act.top().first);
                while (!act.empty() && act.top().first <= t) {
                    cur -= act.top().second;
                    act.pop();
                }
            }
            local[s] = t;
            act.push({t + nph(s), shot_cells(b, s)});
            cur += shot_cells(b, s);
            gphases[g] = std::max<int64_t>(gphases[g], t + nph(s));
        }
    }
    std::memset(&out.gp, 0, sizeof out.gp);
    for (int g = 0; g < (int)XCDS; g++) out.gp.b[g + 1] = out.gp.b[g] + (g < G ? gphases[g] : 0);
    const int64_t P = out.gp.b[XCDS];
    out.start.assign(S, 0);
    for (int s = 0; s < S; s++) out.start[s] = out.gp.b[out.group[s]] + local[s];
    std::vector<uint32_t> count(P + 1, 0);
    for (int s = 0; s < S; s++)
        for (const HostOp& o : ops[s]) count[out.start[s] + o.phase]++;
    std::vector<uint32_t> off(P + 1, 0);
    for (int64_t p = 0; p < P; p++) {
        if (count[p] > (uint32_t)BATCH_MAX_ITEMS) return bfail(WFD_EINVAL, "wfd_batch: more than %d items in one phase", BATCH_MAX_ITEMS);
        off[p + 1] = off[p] + count[p];
    }
    std::vector<BItem> items(off[P]);
    std::vector<uint32_t> fill(off.begin(), off.end() - 1);
    out.tiles.assign(P, 0);
    for (int s = 0; s < S; s++) {  // shot order inside a phase
        for (const HostOp& o : ops[s]) {
            const int64_t p = out.start[s] + o.phase;
            BItem& it = items[fill[p]++];
            it.ntiles = op_tiles(b, s, o);
            it.tile0 = out.tiles[p];
            out.tiles[p] += it.ntiles;
            it.shot = (uint16_t)s;
            it.kind = (uint8_t)o.kind;
            it.flags = (uint8_t)o.flags;
            it.up = (int16_t)o.up;
            it.uc = (int16_t)o.uc;
            it.un = (int16_t)o.un;
            it.uf = (int16_t)o.uf;
            it.gp = (int16_t)o.gp;
            it.gf = (int16_t)o.gf;
            it.step = o.step;
            it.grow = o.grow;
        }
    }
    if (hipMalloc(&out.items, std::max<size_t>(1, items.size()) * sizeof(BItem)) != hipSuccess ||
        hipMalloc(&out.off, off.size() * sizeof(uint32_t)) != hipSuccess) {
        out.release();
        return bfail(WFD_ENOMEM, "wfd_batch: hipMalloc of the phase tables failed");
    }
    if ((!items.empty() && hipMemcpy(out.items, items.data(), items.size() * sizeof(BItem), hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(out.off, off.data(), off.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        out.release();
        return bfail(WFD_EHIP, "wfd_batch: upload of the phase tables failed");
    }
    out.nphases = P;
    return WFD_OK;
}

int rebuild(wfd_batch* b)
{
    b->sched.release();
    if (b->exec) {
        (void)hipGraphExecDestroy(b->exec);
        b->exec = nullptr;
    }
    int rc = upload_shot_table(b);
    if (rc) return rc;
    return make_schedule(b, b->ops, b->sched);
}

template <typename T>
int launch_t(wfd_batch* b, const Sched& sc, hipStream_t s)
{
    hipError_t e = launch_zero(b->sync, sizeof(BatchSync), s);
    if (e != hipSuccess) return bfail(WFD_EHIP, "wfd_batch: reset of the barrier state failed");
    BatchArgs<T> A;
    std::memset(&A, 0, sizeof A);
    A.shots = static_cast<const BShot<T>*>(b->shots_dev);
    A.items = sc.items;
    A.phase_off = sc.off;
    A.sync = b->sync;
    A.timeout = b->timeout_ticks;
    if (b->opts.mode == 0) {
        A.persistent = 1;
        e = launch_batch<T>(A, sc.gp, b->blocks, batch_block<T>(), s);
        if (e != hipSuccess) return bfail(WFD_EHIP, "wfd_batch: persistent launch failed (%s)", hipGetErrorString(e));
        return WFD_OK;
    }
    for (int64_t p = 0; p < sc.nphases; p++) {
        A.p0 = p;
        A.p1 = p + 1;
        const uint32_t blocks = std::max<uint32_t>(XCDS, (sc.tiles[p] + XCDS - 1) / XCDS * XCDS);
        e = launch_batch<T>(A, sc.gp, blocks, 256, s);
        if (e != hipSuccess) return bfail(WFD_EHIP, "wfd_batch: phase launch failed (%s)", hipGetErrorString(e));
    }
    return WFD_OK;
}

int launch_sched(wfd_batch* b, const Sched& sc, hipStream_t s)
{
    return b->dtype == WFD_F64 ? launch_t<double>(b, sc, s) : launch_t<float>(b, sc, s);
}

int check_shot(const wfd_batch_shot& sh, int i, int dtype)
{
    const wfd_grid& g = sh.grid;
    if (g.ndim != 2 || g.nshots != 1) return bfail(WFD_EINVAL, "wfd_batch: shot %d: 2D single-shot grids only", i);
    if (g.order != 2 && g.order != 4 && g.order != 6 && g.order != 8) return bfail(WFD_EINVAL, "wfd_batch: shot %d: bad order", i);
    if (g.dtype != dtype) return bfail(WFD_EINVAL, "wfd_batch: shot %d: every shot needs the same dtype", i);
    if (g.ld <= 0 || g.elems <= 0 || g.n[0] < 3 || g.n[1] < 3) return bfail(WFD_EINVAL, "wfd_batch: shot %d: grid not initialised", i);
    if (sh.nsteps < 1 || sh.nsteps > INT32_MAX / 2) return bfail(WFD_EINVAL, "wfd_batch: shot %d: nsteps out of range", i);
    if (sh.seg < 0 || sh.seg > sh.nsteps) return bfail(WFD_EINVAL, "wfd_batch: shot %d: seg must be in [0, nsteps]", i);
    if (!sh.med.coef || (sh.med.coef_shot_stride != 0 && sh.med.coef_shot_stride != g.elems))
        return bfail(WFD_EINVAL, "wfd_batch: shot %d: a coefficient field (coef) is required", i);
    if (!sh.G) return bfail(WFD_EINVAL, "wfd_batch: shot %d: G is NULL", i);
    const size_t al = 4 * (dtype == WFD_F64 ? 8 : 4);
    if ((uintptr_t)sh.med.coef % al || (uintptr_t)sh.G % 32) return bfail(WFD_EINVAL, "wfd_batch: shot %d: misaligned field", i);
    auto pts = [&](const wfd_points& p, const char* what, bool need_data) {
        if (p.count == 0) return WFD_OK;
        if (p.count < 0 || p.nrows != g.n[0]) return bfail(WFD_EINVAL, "wfd_batch: shot %d: %s CSR needs %lld rows", i, what, (long long)g.n[0]);
        if (!p.row_ptr || !p.col || !p.id || !p.row || (need_data && !p.data))
            return bfail(WFD_EINVAL, "wfd_batch: shot %d: %s: NULL CSR or data pointer", i, what);
        if (p.width < p.count) return bfail(WFD_EINVAL, "wfd_batch: shot %d: %s: width < count", i, what);
        return WFD_OK;
    };
    int rc = pts(sh.src, "sources", true);
    if (rc) return rc;
    rc = pts(sh.rec, "receivers", true);
    if (rc) return rc;
    if (sh.src.count && sh.src.nsteps < sh.nsteps) return bfail(WFD_EINVAL, "wfd_batch: shot %d: source signal shorter than nsteps", i);
    if (sh.rec.count && (!sh.observed || !sh.adj)) return bfail(WFD_EINVAL, "wfd_batch: shot %d: observed / adj NULL", i);
    if (sh.src.count && !sh.mu_src) return bfail(WFD_EINVAL, "wfd_batch: shot %d: mu_src NULL", i);
    return WFD_OK;
}

}  // namespace

extern "C" {

int wfd_batch_create(wfd_batch** out, int32_t nshots, const wfd_batch_shot* shots, const wfd_batch_opts* opts)
{
    if (!out || !shots) return bfail(WFD_EINVAL, "wfd_batch_create: NULL argument");
    *out = nullptr;
    if (nshots < 1 || nshots > 65535) return bfail(WFD_EINVAL, "wfd_batch_create: nshots must be in [1, 65535]");
    const int dtype = shots[0].grid.dtype;
    if (dtype != WFD_F32 && dtype != WFD_F64) return bfail(WFD_EINVAL, "wfd_batch_create: bad dtype");
    for (int s = 0; s < nshots; s++) {
        int rc = check_shot(shots[s], s, dtype);
        if (rc) return rc;
        // values of the velocity field and the damping (NaN/Inf, kc <= 0, CFL), host-synchronising
        rc = wfd_check_medium(&shots[s].grid, &shots[s].med, nullptr);
        if (rc) {
            const std::string msg = wfd_last_error();
            return bfail(rc, "wfd_batch_create: shot %d: %s", s, msg.c_str());
        }
    }
    wfd_batch* b = new (std::nothrow) wfd_batch;
    if (!b) return bfail(WFD_ENOMEM, "wfd_batch_create: host allocation failed");
    b->dtype = dtype;
    b->nshots = nshots;
    if (opts) b->opts = *opts;
    if (b->opts.mode != 0 && b->opts.mode != 1) {
        delete b;
        return bfail(WFD_EINVAL, "wfd_batch_create: mode must be 0 (persistent) or 1 (per phase)");
    }
    b->in.assign(shots, shots + nshots);
    b->ops.resize(nshots);
    b->seg.resize(nshots);
    b->nfields.resize(nshots);
    b->fbytes.resize(nshots);
    b->cfg.resize(nshots);
    b->cls.resize(nshots);
    const size_t es = dtype == WFD_F64 ? 8 : 4;
    size_t total = 0;
    int64_t maxn = 0;
    for (int s = 0; s < nshots; s++) {
        const wfd_batch_shot& sh = shots[s];
        const int64_t C = sh.seg ? sh.seg : optimal_seg(sh.nsteps);
        b->seg[s] = C;
        const int nf = shot_ops(sh.nsteps, C, sh.src.count > 0, sh.rec.count > 0, b->ops[s]);
        if (nf > 32767) {
            delete b;
            return bfail(WFD_EINVAL, "wfd_batch_create: shot %d needs more than 32767 fields (seg too small)", s);
        }
        b->nfields[s] = nf;
        b->fbytes[s] = (int64_t)((sh.grid.elems * es + 255) / 256 * 256);
        total += (size_t)b->fbytes[s] * nf;
        b->cls[s] = size_class(sh.grid);
        b->cfg[s] = dtype == WFD_F64 ? 1 : 0;
        maxn = std::max(maxn, sh.grid.n[0] + sh.grid.n[1]);
    }
    b->bytes = total;
    auto fail_free = [&](int code, const char* msg) {
        wfd_batch_destroy(b);
        return bfail(code, "%s", msg);
    };
    if (hipMalloc(&b->store, total) != hipSuccess) return fail_free(WFD_ENOMEM, "wfd_batch_create: hipMalloc of the field store failed");
    // halos, padding and boundary nodes of every field stay zero; checkpoint 0 is the rest state
    if (hipMemset(b->store, 0, total) != hipSuccess) return fail_free(WFD_EHIP, "wfd_batch_create: clearing the store failed");
    if (hipMalloc(&b->zprof, (size_t)maxn * es) != hipSuccess || hipMemset(b->zprof, 0, (size_t)maxn * es) != hipSuccess)
        return fail_free(WFD_ENOMEM, "wfd_batch_create: zero profile allocation failed");
    if (hipMalloc(&b->sync, sizeof(BatchSync)) != hipSuccess) return fail_free(WFD_ENOMEM, "wfd_batch_create: hipMalloc of the barrier state failed");
    int dev = 0, ncu = 0, rate_khz = 0, bpc = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
        return fail_free(WFD_EHIP, "wfd_batch_create: device query failed");
    if ((dtype == WFD_F64 ? batch_occupancy<double>(&bpc) : batch_occupancy<float>(&bpc)) || bpc < 1)
        return fail_free(WFD_EHIP, "wfd_batch_create: occupancy query failed");
    if (b->opts.blocks_per_cu > 0) bpc = std::min(bpc, b->opts.blocks_per_cu);
    b->blocks = (uint32_t)std::max<int64_t>(XCDS, (int64_t)bpc * ncu / XCDS * XCDS);
    const int64_t ms = b->opts.timeout_ms > 0 ? b->opts.timeout_ms : 2000;
    b->timeout_ticks = (uint64_t)ms * (uint64_t)std::max(rate_khz, 1);
    int rc = rebuild(b);
    if (rc) {
        wfd_batch_destroy(b);
        return rc;
    }
    *out = b;
    return WFD_OK;
}

int wfd_batch_destroy(wfd_batch* b)
{
    if (!b) return WFD_OK;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    b->sched.release();
    (void)hipFree(b->shots_dev);
    (void)hipFree(b->sync);
    (void)hipFree(b->zprof);
    (void)hipFree(b->store);
    delete b;
    return WFD_OK;
}

int wfd_batch_info(const wfd_batch* b, int64_t* phases, int64_t* blocks, size_t* bytes, int64_t* cell_steps)
{
    if (!b) return bfail(WFD_EINVAL, "wfd_batch_info: NULL batch");
    if (phases) *phases = b->sched.nphases;
    if (blocks) *blocks = b->opts.mode == 0 ? (int64_t)b->blocks : 0;
    if (bytes) *bytes = b->bytes;
    if (cell_steps) {
        int64_t c = 0;
        for (const auto& s : b->in) c += (s.grid.n[0] - 2) * (s.grid.n[1] - 2) * s.nsteps;
        *cell_steps = c;
    }
    return WFD_OK;
}

int wfd_batch_shot_info(const wfd_batch* b, int32_t shot, int32_t* size_cls, int32_t* variant, int32_t* rows,
                        int64_t* seg, int64_t* start_phase, int64_t* nphases)
{
    if (!b || shot < 0 || shot >= b->nshots) return bfail(WFD_EINVAL, "wfd_batch_shot_info: bad batch or shot");
    if (size_cls) *size_cls = b->cls[shot];
    if (variant) *variant = cfg_variant(b->cfg[shot]);
    if (rows) *rows = b->dtype == WFD_F64 ? 2 : cfg_rows(b->cfg[shot]);
    if (seg) *seg = b->seg[shot];
    if (start_phase) *start_phase = b->sched.start[shot];
    if (nphases) *nphases = b->ops[shot].empty() ? 0 : (int64_t)b->ops[shot].back().phase + 1;
    return WFD_OK;
}

int wfd_batch_run(wfd_batch* b, int32_t use_graph, wfd_stream_t stream)
{
    if (!b) return bfail(WFD_EINVAL, "wfd_batch_run: NULL batch");
    hipStream_t s = (hipStream_t)stream;
    if (b->dirty) {  // a failed run may have left partial sums: clear the imaging partials
        for (int i = 0; i < b->nshots; i++) {
            // gpart is the last field of the shot's store
            char* base = b->store;
            for (int j = 0; j < i; j++) base += b->fbytes[j] * b->nfields[j];
            if (launch_zero(base + b->fbytes[i] * (b->nfields[i] - 1), b->fbytes[i], s) != hipSuccess)
                return bfail(WFD_EHIP, "wfd_batch_run: clearing the imaging partial failed");
        }
        b->dirty = false;
    }
    if (!use_graph) return launch_sched(b, b->sched, s);
    if (!b->exec) {
        if (!b->cap_stream && hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess)
            return bfail(WFD_EHIP, "wfd_batch_run: stream creation failed");
        if (hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return bfail(WFD_EHIP, "wfd_batch_run: hipStreamBeginCapture failed");
        int rc = launch_sched(b, b->sched, b->cap_stream);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->cap_stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return bfail(WFD_EHIP, "wfd_batch_run: hipStreamEndCapture failed");
        const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            b->exec = nullptr;
            return bfail(WFD_EHIP, "wfd_batch_run: hipGraphInstantiate failed");
        }
    }
    if (hipGraphLaunch(b->exec, s) != hipSuccess) return bfail(WFD_EHIP, "wfd_batch_run: hipGraphLaunch failed");
    return WFD_OK;
}

int wfd_batch_status(wfd_batch* b, wfd_stream_t stream)
{
    if (!b) return bfail(WFD_EINVAL, "wfd_batch_status: NULL batch");
    uint32_t err = 0;
    if (hipStreamSynchronize((hipStream_t)stream) != hipSuccess ||
        hipMemcpy(&err, b->sync->error, sizeof err, hipMemcpyDeviceToHost) != hipSuccess)
        return bfail(WFD_EHIP, "wfd_batch_status: reading the barrier state failed");
    if (err) {
        b->dirty = true;
        if (err & 2u)
            return bfail(WFD_EHIP, "wfd_batch: the persistent grid was not resident on every XCD with work "
                                   "(use mode 1)");
        return bfail(WFD_EHIP, "wfd_batch: a group barrier wait timed out (blocks not co-resident?)");
    }
    return WFD_OK;
}

int wfd_batch_tune(wfd_batch* b, int32_t steps, wfd_stream_t stream)
{
    if (!b) return bfail(WFD_EINVAL, "wfd_batch_tune: NULL batch");
    if (steps < 2) steps = 2;
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        return bfail(WFD_EHIP, "wfd_batch_tune: event creation failed");
    int rc = WFD_OK;
    std::vector<int> cands = b->dtype == WFD_F64 ? std::vector<int>{1, 3} : std::vector<int>{0, 1, 2, 3};
    for (int c = 0; c < NCLASS && !rc; c++) {
        std::vector<std::vector<HostOp>> probe(b->nshots);
        bool any = false;
        for (int sh = 0; sh < b->nshots; sh++) {
            if (b->cls[sh] != c) continue;
            any = true;
            // forward steps on the two work fields, no points: the bandwidth-bound shape of a step
            const int w0 = (int)b->nfields[sh] - 5, w1 = w0 + 1;
            for (int i = 0; i < steps; i++) {
                const int A = i % 2 ? w1 : w0, B = i % 2 ? w0 : w1;
                probe[sh].push_back({K_STEP, 0, A, B, A, -1, -1, -1, i, -1, i});
            }
        }
        if (!any) continue;
        double best_t = 1e30;
        int best = cands[0];
        for (int cand : cands) {
            for (int sh = 0; sh < b->nshots; sh++)
                if (b->cls[sh] == c) b->cfg[sh] = cand;
            rc = upload_shot_table(b);
            if (rc) break;
            Sched probe_sched;
            rc = make_schedule(b, probe, probe_sched);
            if (rc) break;
            double t = 1e30;
            for (int r = 0; r < 3 && !rc; r++) {
                (void)hipEventRecord(e0, s);
                rc = launch_sched(b, probe_sched, s);
                (void)hipEventRecord(e1, s);
                if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = bfail(WFD_EHIP, "wfd_batch_tune: launch failed");
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (r) t = std::min(t, (double)ms);  // first run warms up
            }
            if (!rc) rc = wfd_batch_status(b, stream);
            probe_sched.release();
            if (rc) break;
            if (t < best_t) {
                best_t = t;
                best = cand;
            }
        }
        for (int sh = 0; sh < b->nshots; sh++)
            if (b->cls[sh] == c) b->cfg[sh] = best;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return rebuild(b);
}

int wfd_batch_set_config(wfd_batch* b, int32_t size_cls, int32_t variant, int32_t rows)
{
    if (!b) return bfail(WFD_EINVAL, "wfd_batch_set_config: NULL batch");
    if (variant != 0 && variant != 1) return bfail(WFD_EINVAL, "wfd_batch_set_config: variant must be 0 or 1");
    if (rows != 2 && rows != 8) return bfail(WFD_EINVAL, "wfd_batch_set_config: rows must be 2 or 8");
    if (b->dtype == WFD_F64 && rows != 2) return bfail(WFD_EINVAL, "wfd_batch_set_config: float64 batches march 2 rows");
    const int cfg = (variant == 1 ? 0 : 2) + (rows == 2 ? 1 : 0);
    for (int s = 0; s < b->nshots; s++)
        if (size_cls < 0 || b->cls[s] == size_cls) b->cfg[s] = cfg;
    return rebuild(b);
}

int wfd_batch_shot_ops(int64_t nsteps, int64_t seg, int32_t has_src, int32_t has_rec, int32_t* ops, int64_t cap,
                       int64_t* count, int32_t* nfields)
{
    if (nsteps < 1 || seg < 0 || seg > nsteps) return bfail(WFD_EINVAL, "wfd_batch_shot_ops: bad nsteps / seg");
    std::vector<HostOp> v;
    const int nf = shot_ops(nsteps, seg ? seg : optimal_seg(nsteps), has_src != 0, has_rec != 0, v);
    if (count) *count = (int64_t)v.size();
    if (nfields) *nfields = nf;
    if (ops) {
        if (cap < (int64_t)v.size()) return bfail(WFD_EINVAL, "wfd_batch_shot_ops: output holds %lld ops, %zu needed", (long long)cap, v.size());
        for (size_t i = 0; i < v.size(); i++) std::memcpy(ops + 11 * i, &v[i], sizeof(HostOp));
    }
    return WFD_OK;
}

}  // extern "C"
