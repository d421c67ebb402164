// Internal declarations shared by the HIP kernels and the C-ABI host code of wavefd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "wavefd.h"

namespace wfd {

constexpr int HALO = 4;  // zero halo of the layout = half-width of the widest stencil

// Device view of a CSR point list (see wfd_points in wavefd.h).
template <typename T>
struct PtsDev {
    const int32_t* row_ptr;
    const int32_t* col;
    const int32_t* id;
    const int32_t* row;  // CSR row of each point (receivers: gathered by dedicated blocks)
    T* data;
    int64_t width;
    int64_t shot_stride;
    int64_t nt;  // time rows of `data` per shot
    int32_t count;
    int32_t nrows;
};

// Arguments of one step launch.  Field pointers address node (0,..,0) of shot 0.
template <typename T>
struct StepArgs {
    const T* up;
    const T* uc;
    T* un;
    const T* coef;
    int64_t coef_shot_stride;
    const T* beta_p;  // sponge profile of the plane axis (3D), row axis, column axis
    const T* beta_r;
    const T* beta_c;
    const T* uf;  // forward field for the imaging condition (adjoint only)
    T* g;         // imaging partial (adjoint only)
    // Fused flush of the imaging partial (2D adjoint of a single-shot grid only): when set, the
    // step adds its new partial to gsum (float64, field layout) and stores a zero partial;
    // gsum_first makes the sum start from 0.0 instead of reading gsum.
    double* gsum;
    int32_t gsum_first;
    int64_t shot_stride;
    int64_t np, nr, nc;      // nodes along plane / row / column axis (1 if absent)
    int64_t row_lo, row_hi;  // output rows [lo, hi) of the 1D/2D kernels
    int64_t ld, ps;          // row stride, plane stride (elements)
    T kconst;
    T wc;
    T wp[4], wr[4], wcol[4];  // weights of the plane / row / column axis
    PtsDev<T> src, rec;
    int64_t step;
    const int64_t* step_dev;  // graph replay: the step index is *step_dev + step
    // Receivers are not written by the stencil blocks: `rec` lists nodes whose value in u_cur is
    // copied to trace row (step + gather_off) by the first `ngb` blocks of the launch (the
    // forward records u^n, written by the previous launch, into row n-1: gather_off = -1).
    int64_t gather_off;
    uint32_t ngb;        // gather blocks (256 points each)
    uint32_t gx, gy;     // stencil block grid (x fastest); the launch is 1D: ngb + gx*gy*gz blocks
    int32_t nshots;
    int32_t shot_fast;  // 2D step launches: tile index -> (shot fastest, then column, then row block)
    // 2D imaging launches of the gradient plan (a.uf holds L u of the forward field): IMG_GROUP =
    // the 4 waves of a tile take the same rows of 4 consecutive shots, and shot group j's imaging
    // partial (field j of a.g) accumulates their terms in shot order; 1 = one shot, one partial;
    // 0 = a.uf holds u (its Laplacian is taken in the launch), one partial per shot
    int32_t img_group;
    // Two-step launches: un receives u^{n+1}, un2 u^{n+2}; their gather blocks record up (u^{n-1})
    // and uc (u^n) into trace rows step + gather_off - 1 and step + gather_off.
    T* un2;
    // 2D forward launches of the gradient plan: receives L u^n (the Laplacian of uc that the update
    // uses), written with the nontemporal hint, for the adjoint's imaging; un rotates in place
    T* ucopy;
};

// Block barrier that orders LDS only: each wave waits for its own LDS accesses (lgkmcnt), then
// meets the others at s_barrier.  __syncthreads() also waits for every outstanding global load
// and store of the wave (vmcnt(0)), which drains loads issued for later planes or rows; with this
// barrier they stay in flight.  Use it only where no thread of the block reads global memory
// another thread of the block wrote before the barrier.
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// s_waitcnt immediate (gfx9 encoding) that waits for every outstanding vector memory access of
// the wave and for nothing else: vmcnt 0, expcnt 7, lgkmcnt 15.
constexpr int WAIT_VMCNT0 = 0x0F70;

constexpr int IMG_GROUP = 4;  // shots per imaging partial in grouped imaging launches

// Position of this block in a step launch: a gather block or a stencil tile.  The hardware hands
// block b to XCD b % 8 (each XCD has its own L2) and starts an XCD's blocks in order b / 8.  The
// first `gper` blocks of every XCD are gather blocks (so they start first); the remaining blocks
// of XCD x take one contiguous range of stencil tiles, start(x) + i, the same ranges whether or
// not gather blocks are present.  Neighbouring tiles -- and the halo rows / planes they share --
// then stay in one L2.  The launch has 8 * gper + tiles blocks (see step_grid).  With shot_fast
// (gz shots) the tile index runs over the shots first: the tiles of one region in every shot are
// neighbours, so a coefficient field shared by the shots is read from the L2 after the first.
struct StepBlock {
    bool gather;
    uint32_t g;        // gather block index
    uint32_t x, y, z;  // stencil tile
};
constexpr uint32_t XCDS = 8;
__device__ __forceinline__ StepBlock step_block(uint32_t ngb, uint32_t gx, uint32_t gy, uint32_t gz = 1,
                                                bool shot_fast = false)
{
    const uint32_t gper = (ngb + XCDS - 1) / XCDS;
    const uint32_t b = blockIdx.x, x = b % XCDS, i = b / XCDS;
    if (i < gper) return {true, i * XCDS + x, 0, 0, 0};  // index >= ngb: an idle gather block
    const uint32_t nt = gridDim.x - gper * XCDS;         // stencil tiles
    const uint32_t per = nt / XCDS, rem = nt % XCDS;
    const uint32_t j = x * per + (x < rem ? x : rem) + (i - gper);
    if (shot_fast) {
        const uint32_t t = j / gz;
        return {false, 0, t % gx, t / gx, j % gz};
    }
    return {false, 0, j % gx, (j / gx) % gy, j / (gx * gy)};
}
// blocks of a step launch with `tiles` stencil tiles and ngb gather blocks
__host__ __device__ inline uint32_t step_grid(uint32_t tiles, uint32_t ngb)
{
    return tiles + (ngb + XCDS - 1) / XCDS * XCDS;
}

// Copy u_cur at gather block `blk`'s points (256 of the flattened (shot, point) list) to trace
// row step + gather_off, when that row exists.
template <typename T>
__device__ __forceinline__ void gather_points(const StepArgs<T>& a, uint32_t blk, int64_t step)
{
    const PtsDev<T>& p = a.rec;
    const int64_t t = step + a.gather_off;
    if (t < 0 || t >= p.nt) return;
    const int64_t idx = (int64_t)blk * 256 + threadIdx.x;
    const int64_t shot = idx / p.count, q = idx % p.count;
    if (shot >= a.nshots) return;
    const int32_t* rp = p.row_ptr + shot * (p.nrows + 1);
    if (q >= rp[p.nrows]) return;  // this shot has fewer points
    const int64_t pq = shot * p.count + q;
    const int64_t row = p.row[pq], col = p.col[pq];
    const int64_t off = (row / a.nr) * a.ps + (row % a.nr) * a.ld + col;  // 1D/2D: row / nr == 0 or nr == 1
    p.data[shot * p.shot_stride + t * p.width + p.id[pq]] = a.uc[shot * a.shot_stride + off];
}

struct LaunchCfg {
    int ri;    // rows (2D) or planes (3D) marched per wave / block
    int tile;  // variant: bit 0 = lane-exchange halo (2D), bit 1 = streaming access (WFD_VARIANT_STREAM),
               // bit 2 = shot-fastest tile order (WFD_VARIANT_SHOTS, 2D)
    int img_ri = 0;  // 3D imaging launches: planes per block when forced (wfd_set_config), 0 = by grid size
};

template <typename T>
hipError_t launch_flush(T* g, double* G, int64_t elems, int nshots, hipStream_t s);

template <typename T>
hipError_t launch_check(const T* origin, int64_t np, int64_t nr, int64_t nc, int64_t ld, int64_t ps, T kmax,
                        const T* beta, int64_t nbeta, int32_t* flag, hipStream_t s);

template <typename T>
hipError_t launch_step2d(const StepArgs<T>& a, bool vrow, int order, bool varc, bool sponge, bool img,
                         int variant, int ri, int nshots, hipStream_t s);
template <typename T>
hipError_t launch_step3d(const StepArgs<T>& a, int order, bool varc, bool sponge, bool img, int rk, bool nt,
                         int nshots, hipStream_t s);

template <typename T>
hipError_t launch_gather(const StepArgs<T>& a, hipStream_t s);

// Two steps per launch (2D): u^{n+1} -> a.un, u^{n+2} -> a.un2 from (a.up = u^{n-1}, a.uc = u^n);
// all four fields distinct.  rows = output rows per wave (8, 16 or 32).
template <typename T>
hipError_t launch_step2d_t2(const StepArgs<T>& a, int order, bool varc, bool sponge, int rows, int nshots,
                            hipStream_t s);

// Two steps per launch (3D), same contract; rk = planes marched per block.
template <typename T>
hipError_t launch_step3d_t2(const StepArgs<T>& a, int order, bool varc, bool sponge, int rk, int nshots,
                            hipStream_t s);

hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t s);
hipError_t launch_set_i64(int64_t* p, int64_t v, hipStream_t s);
hipError_t launch_zero(void* p, int64_t bytes, hipStream_t s);

// record an error message for wfd_last_error and return `code`
int set_error(int code, const char* msg);

// Host checks of a medium descriptor for a stream-ordered call: non-NULL, coefficient stride 0 or
// elems, a constant coefficient finite, > 0 and within the CFL limit of the grid.
int check_medium_args(const wfd_grid* g, const wfd_medium* med);

// One step launch (stencil tiles + gather blocks copying u_cur at `rec` into trace row
// step + gather_off), and a gather-only launch of `field` into trace row `row`.  Used by the
// public entry points and the gradient plan.
int step_launch(const wfd_grid* g, const void* up, const void* uc, void* un, const wfd_medium* med,
                const wfd_points* src, const wfd_points* rec, int64_t step, int64_t gather_off, const void* uf,
                void* gpart, bool img, hipStream_t s, bool img_group = false, void* ucopy = nullptr);
// Shot groups of a grouped imaging launch on g (0: one partial per shot).
inline int img_groups(const wfd_grid* g) { return g->ndim == 2 && g->nshots > 1 ? (g->nshots + IMG_GROUP - 1) / IMG_GROUP : 0; }
int gather_launch(const wfd_grid* g, const void* field, const wfd_points* rec, int64_t row, hipStream_t s);

}  // namespace wfd
