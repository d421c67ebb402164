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
    T* data;
    int64_t width;
    int64_t shot_stride;
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
};

// Block index remapped so that each of the 8 XCDs (whose L2 caches are separate) processes one
// contiguous range of the grid's blocks.  The hardware hands block b (linear id, x fastest) to
// XCD b % 8; giving that block the position start(b % 8) + b / 8 keeps neighbouring tiles -- and
// the halo rows they share -- inside one L2.  A bijection for any grid size.
struct BlockPos {
    uint32_t x, y, z;
};
__device__ __forceinline__ BlockPos xcd_block()
{
    constexpr uint32_t XCDS = 8;
    const uint32_t gx = gridDim.x, gy = gridDim.y;
    const uint32_t nb = gx * gy * gridDim.z;
    const uint32_t b = (blockIdx.z * gy + blockIdx.y) * gx + blockIdx.x;
    const uint32_t per = nb / XCDS, rem = nb % XCDS, x = b % XCDS, i = b / XCDS;
    const uint32_t n = x * per + (x < rem ? x : rem) + i;
    return {n % gx, (n / gx) % gy, n / (gx * gy)};
}

struct LaunchCfg {
    int ri;    // rows (2D) or planes (3D) marched per wave / block
    int tile;  // 3D: tile variant
};

template <typename T>
hipError_t launch_flush(T* g, double* G, int64_t elems, int nshots, hipStream_t s);

template <typename T>
hipError_t launch_check(const T* origin, int64_t np, int64_t nr, int64_t nc, int64_t ld, int64_t ps,
                        const T* beta, int64_t nbeta, int32_t* flag, hipStream_t s);

template <typename T>
hipError_t launch_step2d(const StepArgs<T>& a, bool vrow, int order, bool varc, bool sponge, bool img,
                         int variant, int ri, int nshots, hipStream_t s);
template <typename T>
hipError_t launch_step3d(const StepArgs<T>& a, int order, bool varc, bool sponge, bool img, int rk,
                         int nshots, hipStream_t s);

hipError_t launch_add_i64(int64_t* p, int64_t v, hipStream_t s);
hipError_t launch_set_i64(int64_t* p, int64_t v, hipStream_t s);
hipError_t launch_zero(void* p, int64_t bytes, hipStream_t s);

// record an error message for wfd_last_error and return `code`
int set_error(int code, const char* msg);

}  // namespace wfd
