/* grid64_common.h -- the 8 x 8 lane grid of the one-problem-per-wavefront
 * kernels (qr_grid64.hip, qr_apply64.hip): lane (r, c) = lane bits [2:0] and
 * [5:3], rows r + 8a, columns c + 8b; the cross-lane moves they share. */
#ifndef CBK_GRID64_COMMON_H
#define CBK_GRID64_COMMON_H

#include "cbk_internal.h"
#include "quad_common.h"

namespace cbk {
namespace qrgrid64 {

using team::dpp;
using team::fmaT;
using team::sfor;
using team::sfor_down;
using quad::qx;

constexpr int RS = 8; /* row slots: rows r + 8a */
constexpr int CS = 8; /* column slots: columns c + 8b */
constexpr int KMAX = 64;

/* the sum over the 8 row lanes of this lane's column lane, in each of them */
template <class T> __device__ __forceinline__ T row_sum(T x)
{
    x = x + qx<1>(x);
    x = x + qx<2>(x);
    return x + dpp<0x141>(x); /* row_half_mirror: the other quad of the 8 */
}

/* value of lane (r, C) (lane r + 8C), for every lane (r, c) */
template <int C> __device__ __forceinline__ double colbc(double x, int raddr)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_ds_bpermute(raddr + 32 * C, w.x);
    w.y = __builtin_amdgcn_ds_bpermute(raddr + 32 * C, w.y);
    return __builtin_bit_cast(double, w);
}
template <int C> __device__ __forceinline__ float colbc(float x, int raddr)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(raddr + 32 * C, __builtin_bit_cast(int, x)));
}

/* value of lane (R, c) of this lane's column lane (ds_swizzle within 32
 * lanes: and 0x18, or R) */
template <int R> __device__ __forceinline__ double rowbc(double x)
{
    constexpr int pat = 0x18 | (R << 5);
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_ds_swizzle(w.x, pat);
    w.y = __builtin_amdgcn_ds_swizzle(w.y, pat);
    return __builtin_bit_cast(double, w);
}
template <int R> __device__ __forceinline__ float rowbc(float x)
{
    constexpr int pat = 0x18 | (R << 5);
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, x), pat));
}

/* x of lane L, in every lane (wave-uniform) */
template <int L> __device__ __forceinline__ double lane_value(double x)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_readlane(w.x, L);
    w.y = __builtin_amdgcn_readlane(w.y, L);
    return __builtin_bit_cast(double, w);
}
template <int L> __device__ __forceinline__ float lane_value(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), L));
}

template <class T> __device__ __forceinline__ T opaque(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

} /* namespace qrgrid64 */
} /* namespace cbk */

#endif /* CBK_GRID64_COMMON_H */
