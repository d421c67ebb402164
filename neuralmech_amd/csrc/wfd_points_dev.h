// Wave-cooperative source injection for the step kernels (receivers are recorded by dedicated
// gather blocks, see gather_points).
//
// A wave updates one row of 4-node chunks at a time: lane l owns nodes [base + 4l, base + 4l + 4)
// of a CSR row, for lanes lane_lo..lane_hi.  The row's points are read 64 at a time with
// coalesced loads (one point per lane, all loads independent); a ballot picks the points this
// wave owns, and for each (in list order) the owning lane applies num = fma(kc, sig, num) with
// the value broadcast by readlane.  No dependent chains of memory loads (a per-lane binary
// search would serialise one load latency per probe).  All 64 lanes of the wave must be active.
#pragma once

#include "wfd_internal.h"

namespace wfd {

__device__ __forceinline__ int lane_read(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ float lane_read(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double lane_read(double v, int l)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(x & 0xffffffff), l);
    const int hi = __builtin_amdgcn_readlane((int)(x >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ float pt_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double pt_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Points of a row are handled PB batches of 64 at a time: all loads of a pass are issued before
// any is used, so a pass costs one load latency (plus one for the signal gather) however many
// batches it holds.
constexpr int PB = 4;

template <typename T>
__device__ __forceinline__ void wave_inject(const PtsDev<T>& p, int shot, int64_t row, int64_t base, int lane_lo,
                                            int lane_hi, int64_t step, int lane, const T (&kc)[4], T (&num)[4])
{
    const int32_t* rp = p.row_ptr + (int64_t)shot * (p.nrows + 1);
    const int lo = rp[row], hi = rp[row + 1];
    if (lo >= hi) return;
    const int32_t* col = p.col + (int64_t)shot * p.count;
    const int32_t* id = p.id + (int64_t)shot * p.count;
    const T* sig = p.data + (int64_t)shot * p.shot_stride + step * p.width;
    for (int b = lo; b < hi; b += 64 * PB) {
        int cl[PB], pid[PB];
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int q = b + 64 * j + lane;
            cl[j] = q < hi ? col[q] : -1;
            pid[j] = q < hi ? id[q] : -1;
        }
        T s[PB];
#pragma unroll
        for (int j = 0; j < PB; j++) s[j] = pid[j] >= 0 ? sig[pid[j]] : T(0);
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int64_t off = cl[j] >= 0 ? (int64_t)cl[j] - base : -1;
            const bool mine = off >= 4 * lane_lo && off < 4 * (lane_hi + 1);
            uint64_t m = __ballot(mine);
            const int offi = mine ? (int)off : 0;
            while (m) {
                const int pl = __builtin_ctzll(m);
                m &= m - 1;
                const int o = lane_read(offi, pl);
                const T ts = lane_read(s[j], pl);
                if (lane == (o >> 2)) {
                    const int e = o & 3;
#pragma unroll
                    for (int k = 0; k < 4; k++)
                        if (e == k) num[k] = pt_fma(kc[k], ts, num[k]);
                }
            }
        }
    }
}

}  // namespace wfd
