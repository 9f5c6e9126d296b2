// Cross-lane moves within a 64-lane wavefront without the LDS pipeline
// (ds_bpermute): DPP lane permutations inside rows of 16 lanes and
// broadcasts of one lane through a scalar register.
#pragma once
#include <hip/hip_runtime.h>

namespace capsd {

template <int CTRL>
__device__ __forceinline__ int dpp_i(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x)
{
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double, make_int2(dpp_i<CTRL>(v.x), dpp_i<CTRL>(v.y)));
}
// value of lane l (uniform l)
__device__ __forceinline__ double lane_d(double x, int l)
{
    const int2 v = __builtin_bit_cast(int2, x);
    return __builtin_bit_cast(double,
                              make_int2(__builtin_amdgcn_readlane(v.x, l), __builtin_amdgcn_readlane(v.y, l)));
}

// DPP controls: exchanges within quads (lane i <-> i^1, i <-> i^2) and the
// mirrors of 8- and 16-lane groups (i <-> 7-i, i <-> 15-i)
constexpr int DPP_QUAD_X1 = 0xB1, DPP_QUAD_X2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

// Sum over the wavefront, the same value in every lane, with the additions
// of the xor butterfly (pairs 1, 2, 4, 8, 16, 32): after the quad stages all
// lanes of a quad hold the same partial sum, so the mirrors pair the same
// operands as xor 4 and xor 8, and the last two stages are
// (row0 + row1) + (row2 + row3) of the 16-lane rows.  Bit-identical to the
// butterfly (each addition has the same two operands).
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_d<DPP_QUAD_X1>(v);
    v += dpp_d<DPP_QUAD_X2>(v);
    v += dpp_d<DPP_HALF_MIRROR>(v);
    v += dpp_d<DPP_MIRROR>(v);
    return (lane_d(v, 0) + lane_d(v, 16)) + (lane_d(v, 32) + lane_d(v, 48));
}

__device__ __forceinline__ int wave_max_int(int v)
{
    v = max(v, dpp_i<DPP_QUAD_X1>(v));
    v = max(v, dpp_i<DPP_QUAD_X2>(v));
    v = max(v, dpp_i<DPP_HALF_MIRROR>(v));
    v = max(v, dpp_i<DPP_MIRROR>(v));
    return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
               max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// Kahan sum of x[0..n) in index order (the recurrence of a single thread
// running through the array), by a whole wavefront: 64 consecutive entries
// per load (the next chunk loaded while this one is summed), each lane
// running the same recurrence on the lane broadcasts.  Same value in every
// lane, bit-identical to the sequential loop.
__device__ __forceinline__ double wave_kahan_sum(const double* __restrict__ x, int n)
{
    const int lane = threadIdx.x & 63;
    double s = 0, c = 0;
    double v = lane < n ? x[lane] : 0.0;
    for (int i0 = 0; i0 < n; i0 += 64) {
        const double nv = i0 + 64 + lane < n ? x[i0 + 64 + lane] : 0.0;
        const int m = n - i0 < 64 ? n - i0 : 64;
        for (int t = 0; t < m; t++) {
            const double y = lane_d(v, t) - c;
            const double u = s + y;
            c = (u - s) - y;
            s = u;
        }
        v = nv;
    }
    return s;
}

// Sum of x[0..n) by a whole wavefront in double-double arithmetic: lane l
// adds x[l], x[l+64], ... with error-free transformations (TwoSum) into a
// head and a tail, then the 64 head/tail pairs are combined in a butterfly;
// the rounded head + tail is returned in every lane.  As accurate as a
// compensated (Kahan) sum in index order, with a dependent chain of n/64
// steps instead of n.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

__device__ __forceinline__ double wave_sum_dd(const double* __restrict__ x, int n)
{
    const int lane = threadIdx.x & 63;
    double s = 0, c = 0;
    for (int i = lane; i < n; i += 64) {
        double t, e;
        two_sum(s, x[i], t, e);
        s = t;
        c += e;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const double so = __shfl_xor(s, off), co = __shfl_xor(c, off);
        double t, e;
        two_sum(s, so, t, e);
        s = t;
        c = (c + co) + e;
    }
    return s + c;
}

}  // namespace capsd
