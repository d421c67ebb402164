// Total focusing method: delay and sum of a full matrix capture onto image pixels.
//
// image[p] = | sum over element pairs (e, r) of  h_er(t_er(p)) |,  t_er(p) = t0 + (|p - x_e| + |p - x_r|) / c,
// with h the analytic traces (complex, [E][E][Nt]) linearly interpolated at t / dt.  Two forms, as
// the reference has both (projects/18_fwi/tfm.py:12-36 and :39-53):
//   reciprocal: pairs r >= e, the pair r > e weighted 2 (h_er = h_re), a sample with k >= Nt-1
//               skipped;   w = (t - k) f, v = f - w (f = 1 or 2),  s += v h[k] + w h[k+1]
//   all pairs:  k = min(k, Nt-2), w = t - k,  s += ((1-w) h[k] + w h[k+1]) if k < Nt-2
// One lane per pixel: the element delays sit in LDS (a column per lane), the pair loop reads h
// through the cache (neighbouring pixels read neighbouring samples of the same pair).
#include <hip/hip_runtime.h>

#include "wfd_internal.h"

namespace wfd {
namespace {

template <int ECAP, bool RECIP>
__global__ __launch_bounds__(256) void tfm_kernel(const float2* __restrict__ h, const float* __restrict__ xe,
                                                  const float* __restrict__ ye, int E, int Nt, float dt, float t0,
                                                  float c, const float* __restrict__ xp,
                                                  const float* __restrict__ yp, int64_t P, float* image)
{
    __shared__ float tau[ECAP][256];  // this lane's delays, one bank per lane
    const int tid = threadIdx.x;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + tid;
    const bool in = p < P;
    const float px = in ? xp[p] : 0.f, py = in ? yp[p] : 0.f;
    for (int e = 0; e < E; e++) tau[e][tid] = hypotf(px - xe[e], py - ye[e]) / c;
    if (!in) return;  // no barrier below: each lane reads only its own column of tau
    float sx = 0.f, sy = 0.f;
    for (int e = 0; e < E; e++) {
        const float te = t0 + tau[e][tid];
        const float2* he = h + (int64_t)e * E * Nt;
        for (int r = RECIP ? e : 0; r < E; r++) {
            const float t = (te + tau[r][tid]) / dt;
            int k = (int)t;
            const float2* hr = he + (int64_t)r * Nt;
            if (RECIP) {
                if (k >= Nt - 1) continue;
                const float f = r > e ? 2.f : 1.f;
                const float w = (t - (float)k) * f, v = f - w;
                const float2 a = hr[k], b = hr[k + 1];
                sx += v * a.x + w * b.x;
                sy += v * a.y + w * b.y;
            } else {
                k = k < Nt - 2 ? k : Nt - 2;
                if (k >= Nt - 2) continue;
                const float w = t - (float)k, v = 1.f - w;
                const float2 a = hr[k], b = hr[k + 1];
                sx += v * a.x + w * b.x;
                sy += v * a.y + w * b.y;
            }
        }
    }
    image[p] = hypotf(sx, sy);
}

template <int ECAP>
hipError_t tfm_launch(const float2* h, const float* xe, const float* ye, int E, int Nt, float dt, float t0, float c,
                      const float* xp, const float* yp, int64_t P, float* image, int recip, hipStream_t s)
{
    const dim3 grid((unsigned)((P + 255) / 256));
    if (recip)
        hipLaunchKernelGGL((tfm_kernel<ECAP, true>), grid, dim3(256), 0, s, h, xe, ye, E, Nt, dt, t0, c, xp, yp, P,
                           image);
    else
        hipLaunchKernelGGL((tfm_kernel<ECAP, false>), grid, dim3(256), 0, s, h, xe, ye, E, Nt, dt, t0, c, xp, yp, P,
                           image);
    return hipGetLastError();
}

}  // namespace
}  // namespace wfd

using namespace wfd;

extern "C" int wfd_tfm(const void* h, const float* xe, const float* ye, int32_t E, int32_t Nt, double dt, double t0,
                       double c, const float* xp, const float* yp, int64_t P, float* image, int32_t reciprocal,
                       wfd_stream_t stream)
{
    if (!h || !xe || !ye || !xp || !yp || !image) return set_error(WFD_EINVAL, "wfd_tfm: NULL pointer");
    if (E < 1 || E > 64) return set_error(WFD_EINVAL, "wfd_tfm: 1 <= E <= 64 elements");
    if (Nt < 3) return set_error(WFD_EINVAL, "wfd_tfm: need at least 3 samples per trace");
    if (!(dt > 0) || !(c > 0) || !(t0 >= 0)) return set_error(WFD_EINVAL, "wfd_tfm: dt, c > 0 and t0 >= 0");
    if (P < 0) return set_error(WFD_EINVAL, "wfd_tfm: negative pixel count");
    if (P == 0) return WFD_OK;
    const float2* hh = (const float2*)h;
    hipStream_t s = (hipStream_t)stream;
    hipError_t e;
    if (E <= 16)
        e = tfm_launch<16>(hh, xe, ye, E, Nt, (float)dt, (float)t0, (float)c, xp, yp, P, image, reciprocal, s);
    else if (E <= 32)
        e = tfm_launch<32>(hh, xe, ye, E, Nt, (float)dt, (float)t0, (float)c, xp, yp, P, image, reciprocal, s);
    else
        e = tfm_launch<64>(hh, xe, ye, E, Nt, (float)dt, (float)t0, (float)c, xp, yp, P, image, reciprocal, s);
    if (e != hipSuccess) return set_error(WFD_EHIP, "wfd_tfm: launch failed");
    return WFD_OK;
}
