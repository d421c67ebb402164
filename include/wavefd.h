/* wavefd — C ABI of the MI355X (gfx950) finite-difference scalar-wave solver.
 *
 * Plain pointers and sizes only.  Every device pointer is a HIP device allocation owned by the
 * caller (typically a torch tensor's data_ptr()); the library allocates device memory only
 * inside wfd_plan, wfd_batch and wfd_fgraph objects (wfd_forward_graph keeps its wfd_fgraph
 * objects in a cache) and for the flag of the host-synchronising wfd_check_medium.  Every launch
 * is stream-ordered on the caller's hipStream_t (passed as void*, NULL = the default stream) and
 * the library never synchronises the host except where a function says so.  Functions return
 * WFD_OK (0) or a positive status; wfd_last_error() gives a message (thread-local).  Calls are
 * re-entrant across threads and streams; the process-wide state is the launch-shape table
 * (wfd_tune / wfd_set_config, guarded by a mutex), the rows override of wfd_ac_set_rows (atomic)
 * -- neither of which changes results -- and the graph cache of wfd_forward_graph (guarded by a
 * mutex; one step counter per cached graph and stream).  Two calls that write the same fields
 * race, as any two writers do.
 *
 * Field layout ("halo layout").  A field of a grid with nodes n[0..ndim-1] (slowest axis first,
 * boundary ring included) is stored as
 *     rows of `ld` elements (ld >= n_fast + 4, a multiple of 32),
 *     `plane` rows per plane in 3D (plane >= n[1] + 4),
 *     4 zero halo rows (2D) / planes (3D) before and after the nodes,
 * so that every neighbour of an interior node within 4 nodes along any axis is an in-bounds
 * element that is either a node or zero.  `offset` is the element index of node (0,..,0) and
 * `elems` the element count of one field.  Batches of `nshots` fields are contiguous with stride
 * `elems`.  The outside of the node box (halo, padding) must be zero and stays zero; boundary
 * nodes stay zero (homogeneous Dirichlet, as step2D_wave.cu:70 never writes them).
 *
 * Update (per interior node x, step n; K = order/2, kc = v^2 dt^2, b = sponge damping):
 *     lap   = wc*u[x];  for each axis d (slowest first), m = 1..K:
 *             lap = fma(w[d][m-1], u[x - m e_d] + u[x + m e_d], lap)
 *     t     = 2*u[x] - (1 - b)*u_prev[x]            (t = 2*u[x] - u_prev[x] without sponge)
 *     num   = fma(kc, lap, t);   for each source s at x: num = fma(kc, sig[n][s], num)
 *     u_next[x] = num / (1 + b)                     (num without sponge)
 *     trace[n][r] = u_next[rec_r]
 * with w[d][m-1] = (T)(a_m / h_d^2), wc = (T)(a_0 * sum_d 1/h_d^2), a_k the central
 * second-derivative weights of the order.  This extends the reference update
 * u2 = -u0 + 2u1 + dt^2 c^2 (Lx + Ly) of projects/18_simple_wave_fd/step2D_wave.cu:72-79.
 */
#ifndef WAVEFD_H
#define WAVEFD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFD_OK 0
#define WFD_EINVAL 1      /* bad shape, order, CFL violation, aliasing, bad point list */
#define WFD_EHIP 2        /* a HIP runtime call or launch failed */
#define WFD_ENOMEM 3      /* plan allocation failed or exceeds the caller's byte budget */
#define WFD_ENONFINITE 4  /* a medium field holds NaN/Inf or a non-positive coefficient */

#define WFD_F32 0
#define WFD_F64 1

typedef void* wfd_stream_t; /* hipStream_t */

/* Grid descriptor.  Fill with wfd_grid_init; treat the layout fields as read-only. */
typedef struct wfd_grid {
    int32_t ndim;   /* 1, 2 or 3 */
    int32_t order;  /* 2, 4, 6 or 8 */
    int32_t dtype;  /* WFD_F32 or WFD_F64 */
    int32_t nshots; /* independent fields per launch (batch), >= 1 */
    int64_t n[3];   /* nodes per axis incl. the boundary ring, slowest first (unused = 1) */
    double h[3];    /* node spacing per axis */
    double dt;      /* time step */
    int64_t ld;     /* elements per row */
    int64_t plane;  /* rows per plane (3D), 1 otherwise */
    int64_t offset; /* element index of node (0,...,0) */
    int64_t elems;  /* elements per field */
} wfd_grid;

/* Medium: kc = v^2 dt^2 per node (a field in the halo layout, shared by all shots when
 * coef_shot_stride == 0) or the constant kconst when coef == NULL; optional sponge profiles
 * beta = concatenation of one per-axis profile (n[d] values, slowest axis first), the damping
 * of node x being the sum of its axes' entries. */
typedef struct wfd_medium {
    const void* coef;
    double kconst;
    int64_t coef_shot_stride; /* 0 (shared) or elems */
    const void* beta;         /* NULL = no sponge */
} wfd_medium;

/* Point list (sources or receivers) in device memory, CSR over rows: a row is every axis but the
 * fastest (n[0] rows in 2D, n[0]*n[1] in 3D, 1 in 1D).  Build it with wfd_points_csr.
 * `data` is the signal (sources: read) or the traces (receivers: written), per shot
 * [nsteps][width], shots `shot_stride` elements apart, dtype of the grid.  Receivers are written
 * by gather blocks that read the field a launch receives as its current level, i.e. one launch
 * after the step that produced it, plus a final gather at the end of each call: a call's traces
 * are complete when the call's work has finished on the stream. */
typedef struct wfd_points {
    int32_t count;          /* points per shot (0 = none) */
    int32_t nrows;          /* CSR rows per shot */
    const int32_t* row_ptr; /* [nshots][nrows + 1], offsets into this shot's col/id/row */
    const int32_t* col;     /* [nshots][count], fastest-axis index, ascending within a row */
    const int32_t* id;      /* [nshots][count], column of the point in `data` */
    const int32_t* row;     /* [nshots][count], CSR row of the point (needed for receivers) */
    void* data;
    int64_t width;          /* entries per time step in data (>= count) */
    int64_t nsteps;         /* time rows of data per shot */
    int64_t shot_stride;    /* elements between shots in data (>= nsteps * width) */
} wfd_points;

/* --------------------------------------------------------------------------- host-only */
const char* wfd_last_error(void);
const char* wfd_version(void);

/* Fill the descriptor and its layout.  Replaces the reference's manual padding
 * (NY_PADDED = ceil32(NY), projects/18_simple_wave_fd/2D_wave_cuda_padding.py:43-47) and
 * cuwave's ScalarWave(Nx, dx, N, dt, ..., space_order) grid (projects/18_fwi/helper.py:63-72). */
int wfd_grid_init(wfd_grid* g, int ndim, const int64_t* n, const double* h, double dt, int order,
                  int dtype, int nshots);

/* Largest stable dt for max speed vmax: vmax dt <= 2 / sqrt(sum_d S / h_d^2),
 * S = |a_0| + 2 sum |a_k|.  Order 2 in 2D gives the reference's min(dx,dy)/c/sqrt(2)
 * (2D_wave_cupy.py:38); replaces cuwave.wave.stable_dt (projects/18_fwi/helper.py:52). */
double wfd_stable_dt(int ndim, const double* h, double vmax, int order);

/* The stencil weights the kernels use, as doubles (rounded to the grid dtype first):
 * wc and w[d*K + m-1]. */
int wfd_stencil_weights(const wfd_grid* g, double* wc, double* w);

/* Sort `npts` logical flat node indices (C order over n) into a CSR point list of this grid.
 * Host arrays: row_ptr[nrows+1], col[npts], id[npts] (id = the point's position in `nodes`),
 * row[npts].  Points must be interior nodes (WFD_EINVAL otherwise); points sharing a node keep
 * their order.  Replaces cuwave's point_source / Sensors node lookup (projects/18_fwi/plate.py:51). */
int wfd_points_csr(const wfd_grid* g, int32_t npts, const int64_t* nodes, int32_t* row_ptr,
                   int32_t* col, int32_t* id, int32_t* row);
int32_t wfd_points_nrows(const wfd_grid* g);

/* ------------------------------------------------------------------------------ device */
/* One step of the update above: u_next = step(u_prev, u_cur).  u_next may alias u_prev (the
 * in-place leapfrog of 2D_wave_cuda_padding.py:84), no other aliasing.  `src`/`rec` may be NULL.
 * `step` indexes the signal and trace rows.  Replaces padded_step (step2D_wave.cu:59-81) and the
 * per-step work inside cuwave.wave.simulate. */
int wfd_step(const wfd_grid* g, const void* u_prev, const void* u_cur, void* u_next,
             const wfd_medium* med, const wfd_points* src, const wfd_points* rec, int64_t step,
             wfd_stream_t stream);

/* nsteps steps from (u_a = u^{n-1}, u_b = u^n), n = step0, rotating in place.  On return the
 * newest level is in u_b when nsteps is even and in u_a when it is odd.  Replaces the host loop
 * of 2D_wave_cuda_padding.py:83-85 and cuwave.wave.simulate/measure. */
int wfd_forward(const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med,
                const wfd_points* src, const wfd_points* rec, int64_t step0, int64_t nsteps,
                wfd_stream_t stream);

/* nsteps steps like wfd_forward, two per launch (2D, 3D): each launch reads u^{n-1}, u^n and kc
 * once and writes u^{n+1}, u^{n+2} (20 bytes per node per two steps instead of 32), recomputing a
 * halo of step 1; bit-identical to single steps.  u_c, u_d are work fields of the same size (all
 * four distinct).  On return u_a = u^{N-1}, u_b = u^N.  2D: rows = output rows per wave: 2, 4, 8
 * (0 = 8) keep the intermediate level of a 4-wave block's rows in LDS; 16, 32 march both levels
 * in registers.  3D (float32 only): rows = planes marched per block, 1 .. 4096 (0 = 32); step 1
 * on a 48 x 64 in-plane tile, step 2 on its inner 40 x 56.  Replaces temporal_step
 * (step2D_wave.cu:244-304) and its 4-array rotation (2D_wave_cuda_temporal.py:110-112) for the
 * general update. */
int wfd_forward2(const wfd_grid* g, void* u_a, void* u_b, void* u_c, void* u_d,
                 const wfd_medium* med, const wfd_points* src, const wfd_points* rec,
                 int64_t step0, int64_t nsteps, int32_t rows, wfd_stream_t stream);

/* A forward run of `chunk` steps (chunk even) on the fields (u_a, u_b) with this medium and these
 * point lists, captured once into a hipGraph and replayed per chunk: the reference's captured
 * 20-step graph, uploaded once and launched repeatedly (2D_wave_cuda_graph.py:73-79,91,94-95).
 * The step index of every launch is read from a device counter the object owns, set before each
 * run (by a kernel) and advanced by the graph's last node, so one graph serves every chunk of a
 * run and objects never share a counter.  wfd_fgraph_run(step0, nsteps) equals
 * wfd_forward(u_a, u_b, ..., step0, nsteps) bit for bit: nsteps / chunk replays, the remainder as
 * plain launches; u_a, u_b rotate as there.  The buffers, the medium's fields and the point
 * lists' arrays must stay allocated while the object exists (their descriptors are copied at
 * creation).  Runs of one object must not overlap in time (they share its counter); objects on
 * different streams run concurrently.  wfd_fgraph_destroy waits for the object's last run.  On a
 * stream that the caller is capturing into a graph, a run records plain launches instead, so the
 * caller's graph holds nothing the object owns. */
typedef struct wfd_fgraph wfd_fgraph;
int wfd_fgraph_create(wfd_fgraph** fg, const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med,
                      const wfd_points* src, const wfd_points* rec, int32_t chunk);
int wfd_fgraph_run(wfd_fgraph* fg, int64_t step0, int64_t nsteps, wfd_stream_t stream);
int wfd_fgraph_destroy(wfd_fgraph* fg);

/* Same as wfd_forward, replayed from a captured hipGraph of `chunk` steps (chunk even).  Graphs
 * are kept between calls in a process-wide cache (16 entries, least recently used dropped) keyed
 * by the stream and everything a capture depends on, so repeated calls capture once and calls on
 * different streams never share a step counter; runs shorter than one chunk, and calls on a
 * stream being captured by the caller, are plain launches.
 * Same contract as wfd_fgraph_run (replaces the loop of 2D_wave_cuda_graph.py:91-96). */
int wfd_forward_graph(const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med,
                      const wfd_points* src, const wfd_points* rec, int64_t step0,
                      int64_t nsteps, int32_t chunk, wfd_stream_t stream);

/* One adjoint step with imaging: mu_next = step(mu_prev, mu_cur) with the adjoint sources
 * `adj` (signal = -residual, injected like sources) and recording `rec` (e.g. mu at the forward
 * sources), then g[x] = fma(mu_next[x], (L u_fwd)[x], g[x]) on interior nodes. */
int wfd_adjoint_step(const wfd_grid* g, const void* mu_prev, const void* mu_cur, void* mu_next,
                     const wfd_medium* med, const wfd_points* adj, const wfd_points* rec,
                     int64_t step, const void* u_fwd, void* g_partial, wfd_stream_t stream);

/* G[x] += sum over shots of g_partial[shot][x] (G float64, one field), then g_partial = 0. */
int wfd_grad_flush(const wfd_grid* g, void* g_partial, double* G, wfd_stream_t stream);

/* Medium check.  wfd_check_medium_async ORs into *flag (device int32, cleared by the caller):
 * 1 if a coefficient or damping entry is NaN/Inf, a coefficient is <= 0 or a damping entry < 0;
 * 2 if a coefficient exceeds the stability (CFL) limit of the grid's order and spacing,
 * kc * sum_d S / h_d^2 <= 4 with S = |a_0| + 2 sum |a_k| (wfd_stable_dt's bound; relative slack
 * 1e-6).  A constant coefficient (coef == NULL) is checked on the host: WFD_ENONFINITE or, above
 * the limit, WFD_EINVAL, returned at once.  wfd_check_medium is the host-synchronising variant:
 * it returns WFD_ENONFINITE for flag 1 and WFD_EINVAL for flag 2.  The stream-ordered entry
 * points check a constant coefficient the same way (WFD_EINVAL above the limit); a coefficient
 * field is checked only by these two functions, which the Python layer runs once per medium.
 * The reference's driver fixes dt below this limit by construction (2D_wave_cupy.py:37-39). */
int wfd_check_medium_async(const wfd_grid* g, const wfd_medium* med, int32_t* flag,
                           wfd_stream_t stream);
int wfd_check_medium(const wfd_grid* g, const wfd_medium* med, wfd_stream_t stream);

/* ---------------------------------------------------------------------- launch shapes */
/* The step kernels come in variants (2D: 0 = halo chunks from cache, 1 = halo chunks exchanged
 * between lanes; either, and the 3D kernel's 0, | WFD_VARIANT_STREAM = the fields a step touches
 * once (u_prev, kc, u_next) read and written with the nontemporal hint, which pays off when the
 * step's fields exceed the 256 MB Infinity Cache and costs when they fit; 2D grids of several
 * shots also | WFD_VARIANT_SHOTS = tiles taken shot-fastest, so that the tiles of one region in
 * every shot run back to back on one XCD and read a coefficient field the shots share from its
 * L2) and march `rows` rows (2D) or planes (3D) per wave / block.  The choice is kept per grid
 * shape (ndim, order, dtype, n, nshots) in a process-wide table; untuned grids stream when their
 * four step fields exceed 512 MiB and take the tiles of several shots shot-fastest.  Results do not depend on it (every variant evaluates the same expression in the
 * same order).
 * wfd_tune times every candidate for `steps` steps on the caller's scratch fields u_a, u_b
 * (overwritten; host-synchronising) and records the fastest; the reference's fixed launch
 * shapes (2D_wave_cuda_padding.py:22,62-63) are what it replaces.
 * wfd_set_config forces a shape (rows = 0 forgets it); wfd_get_config reports the one in use. */
#define WFD_VARIANT_STREAM 2
#define WFD_VARIANT_SHOTS 4
int wfd_tune(const wfd_grid* g, void* u_a, void* u_b, const wfd_medium* med, int32_t steps,
             int32_t* variant, int32_t* rows, double* us_per_step, wfd_stream_t stream);
int wfd_set_config(const wfd_grid* g, int32_t variant, int32_t rows);
int wfd_get_config(const wfd_grid* g, int32_t* variant, int32_t* rows);

/* ------------------------------------------------- gradient plan (sqrt(n) checkpointing) */
/* A plan owns the checkpoint store, the segment buffer and the working fields for a forward
 * run of `nsteps` steps from rest and its adjoint.  seg = steps per segment (0: chosen to
 * minimise 2*ceil(nsteps/seg) + seg fields); max_bytes caps the allocation (0: no cap) and seg
 * shrinks the store to fit.  seg = WFD_SEG_FASTEST keeps every level (seg 1: nsteps + 6 fields,
 * no recompute in the backward pass) when that store fits max_bytes -- or, with max_bytes 0,
 * 0.8 of the device memory free at creation -- and is memory-optimal otherwise.  The gradient
 * does not depend on seg (same bits).  Replaces cuwave.utils.misfit_gradient's internal memory
 * strategy (projects/18_fwi/2D_fwi.py:40) with checkpoint + recompute. */
#define WFD_SEG_FASTEST (-1)
typedef struct wfd_plan wfd_plan;
int wfd_plan_create(wfd_plan** plan, const wfd_grid* g, int64_t nsteps, int64_t seg,
                    size_t max_bytes);
int wfd_plan_destroy(wfd_plan* plan);
int wfd_plan_info(const wfd_plan* plan, int64_t* seg, int64_t* nckpt, size_t* bytes);
/* Forward from rest over all nsteps, recording `rec` and saving checkpoints.  The final two
 * levels are copied to u_last (optional, 2 fields: u^{N-1} then u^N). */
int wfd_plan_forward(wfd_plan* plan, const wfd_medium* med, const wfd_points* src,
                     const wfd_points* rec, void* u_last, wfd_stream_t stream);
/* Backward: recompute each segment from its checkpoint, run the adjoint with `adj` injected,
 * `src_rec` recording mu at the source nodes, and accumulate G = sum_n mu^{n+1} (L u^n)
 * (float64, one field summed over shots, overwritten).  The terms are summed in grid-dtype
 * partials flushed into G every 32 adjoint steps (at (m-1) % 32 == 0); a 2D grid of several
 * shots keeps one partial per group of 4 consecutive shots (terms added in shot order, groups
 * flushed in order), other grids one per shot, so G is deterministic whatever the launch shape
 * and checkpoint spacing.  On 2D grids the recompute also stores L u^n for the imaging (one more
 * field per checkpoint segment in the plan's store). */
int wfd_plan_backward(wfd_plan* plan, const wfd_medium* med, const wfd_points* src,
                      const wfd_points* adj, const wfd_points* src_rec, double* G,
                      wfd_stream_t stream);

/* ---------------------------------------- uneven shot batch (one persistent launch) */
/* Many independent 2D shots -- each with its own grid size, order, step count, checkpoint
 * spacing, medium, source and receivers -- run their whole misfit gradient in one launch: for
 * every shot the checkpointed forward run of wfd_plan_forward, the residual
 * r = traces - observed, and the adjoint run of wfd_plan_backward injecting -r.  Per shot the
 * results are bit-identical to those two calls on that shot alone with adj = -(traces - observed):
 *   rec.data  traces [nsteps][rec.width]
 *   adj       -(traces - observed), same shape (written; the adjoint source)
 *   mu_src    mu^{n+1} at the sources in row n [nsteps][src.width]
 *   G         sum_n mu^{n+1} (L u^n) on the interior nodes, float64, one field (entries outside
 *             the interior are not written and must be zero).
 * The launch is a persistent kernel: a grid of resident blocks walks a schedule of phases (one
 * time step of every shot in flight per phase) separated by grid barriers whose waits are bounded
 * by timeout_ms (a timed-out run ends early and wfd_batch_status reports it).  Shots are started
 * longest first, as soon as the interior cells in flight stay within max_active_cells.
 * Replaces the per-shot loop over sources inside cuwave.utils.misfit_gradient
 * (projects/18_fwi/2D_fwi.py:40; projects/9_structural_optimization/fwi_penalization.py:143) for
 * shots whose grids differ.  The caller keeps every buffer a shot names alive while the batch is
 * used; runs on one device must not overlap. */
typedef struct wfd_batch_shot {
    wfd_grid grid;        /* ndim 2, nshots 1; one dtype for the whole batch */
    int64_t nsteps;
    int64_t seg;          /* checkpoint spacing, 0 = memory-optimal as in wfd_plan_create */
    wfd_medium med;       /* coef (kc field) required, kconst unused; beta optional */
    wfd_points src;       /* data = source signal, >= nsteps rows */
    wfd_points rec;       /* data = traces (written), nsteps rows */
    const void* observed; /* [nsteps][rec.width] */
    void* adj;            /* [nsteps][rec.width] */
    void* mu_src;         /* [nsteps][src.width] */
    double* G;            /* elems */
} wfd_batch_shot;

typedef struct wfd_batch_opts {
    int64_t max_active_cells; /* 0 = 16 Mi cells (about 16 shots of 1024^2) */
    int32_t blocks_per_cu;    /* persistent grid: resident blocks per CU, 0 = occupancy limit */
    int32_t mode;             /* 0 = one persistent launch; 1 = one launch per phase */
    int32_t timeout_ms;       /* bound of each grid-barrier wait, 0 = 2000 */
    int32_t reserved;
} wfd_batch_opts;

typedef struct wfd_batch wfd_batch;
/* Validate the shots, allocate and clear the field store (host-synchronising), build the
 * schedule.  opts may be NULL. */
int wfd_batch_create(wfd_batch** batch, int32_t nshots, const wfd_batch_shot* shots,
                     const wfd_batch_opts* opts);
int wfd_batch_destroy(wfd_batch* batch);
/* Run every shot's gradient (use_graph != 0: replay a hipGraph captured at the first call). */
int wfd_batch_run(wfd_batch* batch, int32_t use_graph, wfd_stream_t stream);
/* Synchronise `stream` and report whether the last run completed (WFD_EHIP if a barrier wait
 * timed out; the next run then clears the partial sums first). */
int wfd_batch_status(wfd_batch* batch, wfd_stream_t stream);
/* Tile shapes per size class (largest grid side <= 768, 1024, 1280, 1536, larger): time `steps`
 * forward steps of each class's shots under every candidate (variant 0/1, 2/8 rows per wave)
 * inside the persistent kernel and keep the fastest; host-synchronising.  Replaces the fixed
 * launch shapes of the reference drivers (2D_wave_cuda_padding.py:22,62-63). */
int wfd_batch_tune(wfd_batch* batch, int32_t steps, wfd_stream_t stream);
/* Force a tile shape for one size class (-1 = all). */
int wfd_batch_set_config(wfd_batch* batch, int32_t size_class, int32_t variant, int32_t rows);
int wfd_batch_info(const wfd_batch* batch, int64_t* phases, int64_t* blocks, size_t* bytes,
                   int64_t* cell_steps);
int wfd_batch_shot_info(const wfd_batch* batch, int32_t shot, int32_t* size_class,
                        int32_t* variant, int32_t* rows, int64_t* seg, int64_t* start_phase,
                        int64_t* nphases);
/* Host-only: the op list of one shot (11 int32 per op: kind 0 gather / 1 step / 2 adjoint step,
 * flags 1 inject / 2 flush / 4 first flush / 8 gather receivers / 16 gather sources, field
 * indices up, uc, un, uf, gp, gf (-1 = none), signal row, trace row, phase; ops of one phase run
 * concurrently).  ops may be NULL to query the count. */
int wfd_batch_shot_ops(int64_t nsteps, int64_t seg, int32_t has_src, int32_t has_rec,
                       int32_t* ops, int64_t cap, int64_t* count, int32_t* nfields);


/* ------------------------------------- reference-layout steps (projects/18_simple_wave_fd) */
/* Fields are N rows of M_padded elements with the boundary ring inside the array (node (i, j)
 * at i*M_padded + j); only interior nodes are written.  The update is the slice form of
 * 2D_wave_cupy.py:54-62 in fp32 with IEEE division and no contraction, coef = f32(dt^2 c^2),
 * divisors f32(dx^2), f32(dy^2): bit-identical to NumPy evaluating fd_step on float32 arrays.
 * u2 may alias u0 (the in-place loop of 2D_wave_cuda_padding.py:83-85). */

/* Replaces padded_step (step2D_wave.cu:59-81; M_padded = ceil32(M)), basic_step / coalesced_step
 * (:8-55; M_padded = M) and, with variant = 1 (LDS-staged tile), shared_step (:187-240). */
int wfd_ref_step2d(const float* u0, const float* u1, float* u2, double wavespeed, double dt, double dx,
                   double dy, int32_t N, int32_t M, int32_t M_padded, int32_t variant, wfd_stream_t stream);
/* Replaces step1D (step1D_wave.cu:4-17). */
int wfd_ref_step1d(const float* u0, const float* u1, float* u2, double wavespeed, double dt, double dx,
                   int32_t N, wfd_stream_t stream);
/* `steps` (2, 4 or 8) steps per launch on LDS tiles, writing the older / newer level to (v0, v1);
 * bit-identical to `steps` single steps.  Replaces temporal_step (step2D_wave.cu:244-304). */
int wfd_ref_temporal(const float* u0, const float* u1, float* v0, float* v1, double wavespeed, double dt,
                     double dx, double dy, int32_t N, int32_t M, int32_t M_padded, int32_t steps,
                     wfd_stream_t stream);
/* fp16 storage (IEEE binary16 bit patterns), fp32 arithmetic, round-to-nearest-even stores;
 * M_padded even and >= M + 2.  Replaces mixed_step (step2D_wave.cu:85-124). */
int wfd_ref_mixed_step(const void* u0, const void* u1, void* u2, double wavespeed, double dt, double dx,
                       double dy, int32_t N, int32_t M, int32_t M_padded, wfd_stream_t stream);
/* int16 values with one fp32 scale per 32 nodes of a row (s has M_padded/32 entries per row,
 * M_padded % 128 == 0); re-encoded every step by the block maximum.  Replaces scaled_step
 * (step2D_wave.cu:128-183). */
int wfd_ref_scaled_step(const int16_t* q0, const float* s0, const int16_t* q1, const float* s1, int16_t* q2,
                        float* s2, double wavespeed, double dt, double dx, double dy, int32_t N, int32_t M,
                        int32_t M_padded, wfd_stream_t stream);

/* ---------------------------------- variable mass / stiffness model (2D, orders 2..8) */
/* m p_tt = div(s grad p) + f with a mass m and stiffness s per node, cuwave's AcousticWave
 * (projects/8_physics_drivers/topopt_acoustic2D.py:119-131; 1/rho and 1/kappa interpolated by
 * the indicator, projects/1_benchmark_2D/README.md:141-143) and ScalarWave's density indicator
 * (projects/18_fwi/helper.py:31-33).  The library sees it as a symmetric bond operator with
 * K = order/2 bonds per side and axis, per interior node x:
 *     qx = cN_1 dN_1 + cS_1 dS_1;  for m = 2..K: qx = qx + (cN_m dN_m + cS_m dS_m)
 *     qy = cW_1 dW_1 + cE_1 dE_1;  for m = 2..K: qy = qy + (cW_m dW_m + cE_m dE_m)
 *     q = qx + qy,   dN_m = p[x - m rows] - p[x] (S: + m rows, W / E: -+ m columns)
 *     t = 2 p[x] - (1 - b) p_prev[x]     (2 p[x] - p_prev[x] without sponge)
 *     num = t + k[x] q;  per source at x, in list order: num = num + k[x] sig
 *     p_next[x] = num / (1 + b)          (num without sponge)
 * with cN_m = cn_m[x], cS_m = cn_m[x + m rows], cW_m = cw_m[x], cE_m = cw_m[x + m columns]:
 * cn_m[y] couples y to the node m rows above it, cw_m[y] to the node m columns left of it.  The
 * field is zero outside the node box.  k, cn and cw are fields in the halo layout of the grid's
 * dtype shared by every shot: k one field, cn and cw K fields each (`elems` apart, field m-1 for
 * distance m), entries up to K-1 rows / columns past the last node included.  The caller folds
 * mass, control-volume weights, spacing, closure and the averaging of s onto bonds into them (a
 * zero bond carries nothing).  beta is as in wfd_medium. */
typedef struct wfd_ac_medium {
    const void* k;
    const void* cn;
    const void* cw;
    const void* beta;
} wfd_ac_medium;

/* nsteps steps from (u_a = p^{n-1}, u_b = p^n), n = step0, rotating in place as wfd_forward.
 * Records trace row n = p^{n+1} at `rec` (as wfd_forward) and row n = p^n at `strip` for
 * n = step0 .. step0 + nsteps (the values the reconstruction of wfd_ac_backward restores).
 * Replaces the forward of cuwave.wave.simulate / cuwave.utils.response / measure for
 * AcousticWave and a density indicator (topopt_acoustic2D.py:179-181;
 * projects/1_benchmark_2D/forward/wave.py:126; projects/18_fwi/plate.py:62-64). */
int wfd_ac_forward(const wfd_grid* g, void* u_a, void* u_b, const wfd_ac_medium* med,
                   const wfd_points* src, const wfd_points* rec, const wfd_points* strip,
                   int64_t step0, int64_t nsteps, wfd_stream_t stream);

/* Adjoint of an nsteps forward from rest with boundary reconstruction: u_a = p^{N-1},
 * u_b = p^N as the forward left them (consumed).  The forward field is rebuilt backwards in
 * time on nodes without sponge damping; `strip` (its data as wfd_ac_forward wrote it) restores
 * the damped nodes within K nodes of them.  `adj` injects row n (= -dJ/dp^n at the receivers,
 * rows 0..N) into the adjoint field mu = k lambda; `mu_src` (data rows 0..N-1) receives mu^{n+1}
 * at the sources.  mu_a, mu_b are work fields, g_work 2K+1 fields per shot of the grid dtype,
 * and on return G (float64, [2K+1][nshots][elems]) holds per shot
 *     G[0]   = sum_n mu^{n+1} q(p^n)                  on undamped interior nodes,
 *     G[m]   = sum_n (dp^n)(dmu^{n+1}) over each node's bond to the node m rows above,
 *     G[K+m] = the same over bonds to the node m columns left,
 * bonds summed when either end is an undamped interior node (bonds from the last interior rows /
 * columns to nodes past them are stored at that far node).  With J(traces):
 *     dJ/dk = -(G[0] + sum_n mu^{n+1} sig^n) / k,   dJ/dcn_m = G[m],   dJ/dcw_m = G[K+m].
 * Replaces cuwave.sensitivity.reconstruction_sensitivity (analog_rnn_train.py:142-146;
 * strip footprint (N+2) x strip x itemsize, analog_rnn_fixture.py:84-87) and the adjoint of
 * cuwave.utils.misfit_gradient for a density indicator (projects/18_fwi/2D_fwi.py:40). */
int wfd_ac_backward(const wfd_grid* g, void* u_a, void* u_b, void* mu_a, void* mu_b,
                    const wfd_ac_medium* med, const wfd_points* src, const wfd_points* adj,
                    const wfd_points* strip, const wfd_points* mu_src, int64_t nsteps,
                    void* g_work, double* G, wfd_stream_t stream);

/* Allocation-free variant of wfd_ac_backward: the work fields (mu_a, mu_b and g_work) live in
 * a file-scope device array of WFD_AC_SCRATCH_BYTES that each device holds from the library's
 * first use on it until the process ends, so the call makes no hipMalloc / hipFree and no
 * allocating library call; the caller passes only the inputs (u_a, u_b as the forward left them,
 * consumed; medium; point lists) and the outputs (G, mu_src data), with the same results, bit for
 * bit, as wfd_ac_backward.  Its work fields need (2 + 2K+1) * nshots * elems * itemsize bytes
 * (each field rounded up to 256); larger problems return WFD_EINVAL.  Every call on one device
 * shares that device's array, whose contents persist between calls (the kernels clear what must
 * start at zero): calls that overlap in time on one device race, so run them on one stream or
 * serialise them.  No counterpart in the reference: it exists for callers that cannot allocate
 * device memory -- code under hipGraph capture, tight serving loops, harnesses that forbid it. */
#define WFD_AC_SCRATCH_BYTES (64ull << 20)
int wfd_ac_backward_noalloc(const wfd_grid* g, void* u_a, void* u_b, const wfd_ac_medium* med,
                            const wfd_points* src, const wfd_points* adj, const wfd_points* strip,
                            const wfd_points* mu_src, int64_t nsteps, double* G, wfd_stream_t stream);

/* Rows per wave of the wfd_ac_* kernels: 0 (default) chooses per launch (the forward halves 8
 * rows until about 16 K waves are in flight; the adjoint aims at 32 K waves while its 8 + 4K
 * fields fit in 192 MiB and at 8 K otherwise); 1, 2, 4 or 8 forces that count for every later
 * launch in the process.  Results do not depend on it (the same expression at every node); it
 * exists so that every launch shape can be tested and timed.  No counterpart in the reference. */
int wfd_ac_set_rows(int32_t rows);

/* ------------------------------------------------------------- total focusing method */
/* Delay and sum of a full matrix capture onto P pixels: h = analytic traces, complex64 as
 * (re, im) float pairs, [E][E][Nt] (emitter, receiver, sample); element e at (xe[e], ye[e]),
 * pixel p at (xp[p], yp[p]); image[p] = |sum_pairs h_er((t0 + (|p-x_e| + |p-x_r|)/c) / dt)|
 * with linear interpolation in time.  reciprocal = 1: pairs r >= e, r > e weighted 2, samples
 * with k >= Nt-1 skipped (the tfm RawKernel, projects/18_fwi/tfm.py:12-36); reciprocal = 0: all
 * E^2 pairs, k clamped to Nt-2 and skipped there (the array form, tfm.py:39-53).  fp32
 * arithmetic; 1 <= E <= 64; t0 >= 0.  All arrays on the device.  Images of fewer than 2^20
 * pixels split each pixel's pairs over 2-16 lanes, which changes the fp32 summation order (not
 * the terms); larger images sum the pairs in emitter-major order. */
int wfd_tfm(const void* h, const float* xe, const float* ye, int32_t E, int32_t Nt, double dt,
            double t0, double c, const float* xp, const float* yp, int64_t P, float* image,
            int32_t reciprocal, wfd_stream_t stream);

/* ------------------------------------------------------------------- measurement probe */
/* Peak fp32 FMA rate: `blocks` x 256 threads each run `n` iterations of 4 dependent fused
 * multiply-adds (8 flop) in registers and write one float to out[blocks * 256].  Time it with
 * events on `stream`.  Replaces the fma_chain RawKernel of the reference's roofline harness
 * (projects/18_simple_wave_fd/2D_wave_roofline.py:55-68); built without packed fp32
 * instructions, like every kernel here. */
int wfd_fma_peak(float* out, int32_t blocks, int32_t n, wfd_stream_t stream);

/* HBM streaming ceilings for the access mixes of the step kernels, the hand-written counterpart
 * of the reference harness's library copy (2D_wave_roofline.py:46-57): n floats per array (a
 * multiple of 4, 16-byte aligned), one 16-byte chunk per lane.  mode 0: o = a (copy); 1: o = a +
 * b + c (three reads, one write: the stencil's mix); 2: mode 1 with the nontemporal hint on every
 * access; 3: read a only, block sums of 4096 floats into o[0 .. ceil(n / 4096)).  Time it with
 * events on `stream`. */
int wfd_stream_bench(const float* a, const float* b, const float* c, float* o, int64_t n, int32_t mode,
                     wfd_stream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* WAVEFD_H */
