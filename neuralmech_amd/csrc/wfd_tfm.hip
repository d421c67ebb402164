// Total focusing method: delay and sum of a full matrix capture onto image pixels.
//
// image[p] = | sum over element pairs (e, r) of  h_er(t_er(p)) |,  t_er(p) = t0 + (|p - x_e| + |p - x_r|) / c,
// with h the analytic traces (complex, [E][E][Nt]) linearly interpolated at t / dt.  Two forms, as
// the reference has both (projects/18_fwi/tfm.py:12-36 and :39-53):
//   reciprocal: pairs r >= e, the pair r > e weighted 2 (h_er = h_re), a sample with k >= Nt-1
//               skipped;   w = (t - k) f, v = f - w (f = 1 or 2),  s += v h[k] + w h[k+1]
//   all pairs:  k = min(k, Nt-2), w = t - k,  s += ((1-w) h[k] + w h[k+1]) if k < Nt-2
// A group of LPP lanes per pixel (LPP = 1 for large images): the pixel's element delays sit in LDS
// (one column per pixel), lane s of the group takes the receivers r = s (mod LPP) of every
// emitter, and the group's partial sums are added by lane exchange (a fixed order).  Small images
// use several lanes per pixel so that enough waves are in flight to hide the latency of the
// per-pair sample loads; neighbouring pixels read neighbouring samples of the same pair through
// the cache.
#include <hip/hip_runtime.h>

#include "wfd_internal.h"

namespace wfd {
namespace {

template <int ECAP, bool RECIP, int LPP>
__global__ __launch_bounds__(256) void tfm_kernel(const float2* __restrict__ h, const float* __restrict__ xe,
                                                  const float* __restrict__ ye, int E, int Nt, float dt, float t0,
                                                  float c, const float* __restrict__ xp,
                                                  const float* __restrict__ yp, int64_t P, float* image)
{
    constexpr int PPB = 256 / LPP;  // pixels per block
    __shared__ float tau[ECAP][PPB];
    const int tid = threadIdx.x;
    const int slot = tid / LPP, sub = tid % LPP;
    const int64_t p = (int64_t)blockIdx.x * PPB + slot;
    const bool in = p < P;
    const float px = in ? xp[p] : 0.f, py = in ? yp[p] : 0.f;
    for (int e = sub; e < E; e += LPP) tau[e][slot] = hypotf(px - xe[e], py - ye[e]) / c;
    if (LPP > 1) __syncthreads();  // the group's lanes wrote the pixel's column together
    float sx = 0.f, sy = 0.f;
    if (in) {
        for (int e = 0; e < E; e++) {
            const float te = t0 + tau[e][slot];
            const float2* he = h + (int64_t)e * E * Nt;
            const int rlo = RECIP ? e : 0;
            for (int r = rlo + ((sub - rlo % LPP) + LPP) % LPP; r < E; r += LPP) {
                const float t = (te + tau[r][slot]) / dt;
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
    }
#pragma unroll
    for (int off = LPP / 2; off > 0; off >>= 1) {  // groups are aligned runs of LPP lanes in a wave
        sx += __shfl_xor(sx, off);
        sy += __shfl_xor(sy, off);
    }
    if (in && sub == 0) image[p] = hypotf(sx, sy);
}

template <int ECAP, int LPP>
hipError_t tfm_launch_l(const float2* h, const float* xe, const float* ye, int E, int Nt, float dt, float t0, float c,
                        const float* xp, const float* yp, int64_t P, float* image, int recip, hipStream_t s)
{
    const dim3 grid((unsigned)((P + 256 / LPP - 1) / (256 / LPP)));
    if (recip)
        hipLaunchKernelGGL((tfm_kernel<ECAP, true, LPP>), grid, dim3(256), 0, s, h, xe, ye, E, Nt, dt, t0, c, xp, yp,
                           P, image);
    else
        hipLaunchKernelGGL((tfm_kernel<ECAP, false, LPP>), grid, dim3(256), 0, s, h, xe, ye, E, Nt, dt, t0, c, xp, yp,
                           P, image);
    return hipGetLastError();
}

// Lanes per pixel: about a million lanes in flight (16 waves on each of the 1024 SIMDs), at most 16.
inline int tfm_lanes(int64_t P)
{
    int l = 1;
    while (l < 16 && P * l < (int64_t)1 << 20) l *= 2;
    return l;
}

template <int ECAP>
hipError_t tfm_launch(const float2* h, const float* xe, const float* ye, int E, int Nt, float dt, float t0, float c,
                      const float* xp, const float* yp, int64_t P, float* image, int recip, int lanes, hipStream_t s)
{
    switch (lanes) {
        case 1: return tfm_launch_l<ECAP, 1>(h, xe, ye, E, Nt, dt, t0, c, xp, yp, P, image, recip, s);
        case 2: return tfm_launch_l<ECAP, 2>(h, xe, ye, E, Nt, dt, t0, c, xp, yp, P, image, recip, s);
        case 4: return tfm_launch_l<ECAP, 4>(h, xe, ye, E, Nt, dt, t0, c, xp, yp, P, image, recip, s);
        case 8: return tfm_launch_l<ECAP, 8>(h, xe, ye, E, Nt, dt, t0, c, xp, yp, P, image, recip, s);
        default: return tfm_launch_l<ECAP, 16>(h, xe, ye, E, Nt, dt, t0, c, xp, yp, P, image, recip, s);
    }
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
    const int lanes = tfm_lanes(P);
    hipError_t e;
    if (E <= 16)
        e = tfm_launch<16>(hh, xe, ye, E, Nt, (float)dt, (float)t0, (float)c, xp, yp, P, image, reciprocal, lanes, s);
    else if (E <= 32)
        e = tfm_launch<32>(hh, xe, ye, E, Nt, (float)dt, (float)t0, (float)c, xp, yp, P, image, reciprocal, lanes, s);
    else
        e = tfm_launch<64>(hh, xe, ye, E, Nt, (float)dt, (float)t0, (float)c, xp, yp, P, image, reciprocal, lanes, s);
    if (e != hipSuccess) return set_error(WFD_EHIP, "wfd_tfm: launch failed");
    return WFD_OK;
}
