/* quad_common.h -- helpers of the quad kernels: one matrix per 4 lanes.
 *
 * A quad is 4 consecutive lanes (a DPP quad); a wavefront carries 16 matrices,
 * the 16 consecutive slots v0 .. v0+15 of a group. Lane t of a quad owns rows
 * i = t, t+4, t+8, ... of its matrix (slot k holds row t + 4k), so a column
 * element of the 16 matrices is one 128-byte line per row at V >= 16 in FP64
 * and a load instruction moves four full lines. Each lane keeps the lower
 * triangle of its rows in registers: slot k holds columns 0 .. 4k+3 (the
 * entries right of the diagonal ride along as scratch), N(N+4)/8 values.
 *
 * A value of lane s reaches the other lanes of its quad by one DPP move
 * (quad_perm [s,s,s,s]; two 32-bit moves for a double -- the 64-bit DPP forms
 * accept row broadcasts only). A lane then applies it to all its rows with
 * plain fma's, so a broadcast serves up to N/4 fma's and the FP32 fma's run at
 * the full FP32 rate (the DPP-fused FP32 fma runs at the FP64 rate). The
 * builtins let the compiler insert the DPP wait states.
 */
#ifndef CBK_QUAD_COMMON_H
#define CBK_QUAD_COMMON_H

#include "team_common.h"

namespace cbk {
namespace quad {

using team::sfor;
using team::sfor_down;
using team::cmin;
using team::kOOB;

constexpr int QL = 4;          /* lanes per matrix */
constexpr int kMatsPerWave = 16;

/* value of lane S of this lane's quad, in every lane of the quad */
template <int S> __device__ __forceinline__ float qbc(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x),
                                                              S * 0x55, 0xF, 0xF, true));
}
template <int S> __device__ __forceinline__ double qbc(double x)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_mov_dpp(w.x, S * 0x55, 0xF, 0xF, true);
    w.y = __builtin_amdgcn_mov_dpp(w.y, S * 0x55, 0xF, 0xF, true);
    return __builtin_bit_cast(double, w);
}

/* value of lane t ^ M of this lane's quad (M = 1: quad_perm [1,0,3,2]; M = 2:
 * [2,3,0,1]) */
template <int M> constexpr int qx_ctrl() { return M == 1 ? 0xB1 : 0x4E; }
template <int M> __device__ __forceinline__ float qx(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x),
                                                              qx_ctrl<M>(), 0xF, 0xF, true));
}
template <int M> __device__ __forceinline__ double qx(double x)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_mov_dpp(w.x, qx_ctrl<M>(), 0xF, 0xF, true);
    w.y = __builtin_amdgcn_mov_dpp(w.y, qx_ctrl<M>(), 0xF, 0xF, true);
    return __builtin_bit_cast(double, w);
}

/* max over the quad, in every lane, NaN-propagating */
template <class T> __device__ __forceinline__ T quad_max(T x)
{
    x = team::fmax_nan(x, qx<1>(x));
    x = team::fmax_nan(x, qx<2>(x));
    return x;
}

/* whether c holds in any lane of the calling lane's quad */
__device__ __forceinline__ bool quad_any(bool c)
{
    const unsigned long long m = __builtin_amdgcn_ballot_w64(c);
    return ((m >> (threadIdx.x & 60u)) & 0xFull) != 0;
}

/* LDS writes of one lane made visible to the other lanes of its wavefront (the
 * wavefront's LDS operations complete in order; the fences keep the compiler
 * from moving LDS accesses across) */
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Where a lane's matrix sits: the wavefront's first matrix mat0 (16 per
 * wavefront), its first group w0 (the buffer resources are based there, so
 * the bases stay in SGPRs), and the lane's byte offset of its slot in a
 * compact operand of group stride gs -- ((g - w0) * gs + v) * sz, kOOB for a
 * lane past the last slot (it loads zeros, its stores are dropped). */
struct QuadPlace {
    long mat0, mat, w0, groups_left;
    int t;
    bool live;
    __device__ __forceinline__ unsigned slot(size_t gs, int V, unsigned sz) const
    {
        if (!live) return kOOB;
        const long g = mat / V, v = mat % V;
        return (unsigned)(((g - w0) * (long)gs + v) * (long)sz);
    }
};

/* wavefront w of the grid (4 per 256-thread block) takes matrices 16w ..
 * 16w+15; returns false for a wavefront past the last slot */
__device__ __forceinline__ bool quad_place(long nmat, int V, QuadPlace &q)
{
    q.t = threadIdx.x & 3;
    q.mat0 = ((long)blockIdx.x * 4 + team::wave_in_block()) * kMatsPerWave;
    if (q.mat0 >= nmat) return false;
    q.mat = q.mat0 + (long)((threadIdx.x & 63) >> 2);
    q.live = q.mat < nmat;
    q.w0 = __builtin_amdgcn_readfirstlane((int)(q.mat0 / V));
    q.groups_left = nmat / V - q.w0;
    return true;
}

} /* namespace quad */
} /* namespace cbk */

#endif /* CBK_QUAD_COMMON_H */
