/* team_common.h -- helpers of the team kernels: one matrix per 16-lane team.
 *
 * A team is one DPP row of a wavefront (lanes 16k .. 16k+15), so a value held
 * by lane s of the team reaches every lane of the team with a single
 * v_mov_b32_dpp row_newbcast:s per 32-bit half -- no LDS round trip, no
 * barrier. Indices that select a source lane or a register must be
 * compile-time constants, hence the sfor<> loops below: every loop over rows,
 * columns and pivots of a team kernel is unrolled at compile time.
 */
#ifndef CBK_TEAM_COMMON_H
#define CBK_TEAM_COMMON_H

#include <hip/hip_runtime.h>

#include <type_traits>

namespace cbk {
namespace team {

constexpr int TS = 16; /* lanes per team */

template <int B, int E, class F> __device__ __forceinline__ void sfor(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(f);
    }
}

/* descending: E-1, E-2, ..., B */
template <int B, int E, class F> __device__ __forceinline__ void sfor_down(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, E - 1>{});
        sfor_down<B, E - 1>(f);
    }
}

constexpr int cmin(int a, int b) { return a < b ? a : b; }

/* value of lane S of this lane's 16-lane row, in every lane of the row */
template <int S> __device__ __forceinline__ float bcast(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x),
                                                              0x150 + S, 0xF, 0xF, true));
}
template <int S> __device__ __forceinline__ double bcast(double x)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_mov_dpp(w.x, 0x150 + S, 0xF, 0xF, true);
    w.y = __builtin_amdgcn_mov_dpp(w.y, 0x150 + S, 0xF, 0xF, true);
    return __builtin_bit_cast(double, w);
}

/* acc := fma(-x_S, a, acc), with x_S the value of x in lane S of this lane's
 * row: the broadcast rides on the fma as its DPP source operand
 * (v_fmac_*_dpp row_newbcast), one VALU op instead of a move and an fma. The
 * s_nop covers the two wait states a DPP read needs after the VALU write of
 * its source. Same rounding as fma(-a, x_S, acc). */
template <int S> __device__ __forceinline__ void fnma_bc(double &acc, double x, double a)
{
    asm("s_nop 1\n\tv_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}
template <int S> __device__ __forceinline__ void fnma_bc(float &acc, float x, float a)
{
    asm("s_nop 1\n\tv_fmac_f32_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}

/* acc := fma(x_S, a, acc), the broadcast riding on the fma as above */
template <int S> __device__ __forceinline__ void fma_bc(double &acc, double x, double a)
{
    asm("s_nop 1\n\tv_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}
template <int S> __device__ __forceinline__ void fma_bc(float &acc, float x, float a)
{
    asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}

/* fnma_bc / fma_bc without their s_nop, for broadcast sources made ready by
 * dpp_ready: a sweep or an update whose source registers do not change
 * between the reads pays the two wait states once instead of per fma. */
template <int S> __device__ __forceinline__ void fnma_bcr(double &acc, double x, double a)
{
    asm("v_fmac_f64_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}
template <int S> __device__ __forceinline__ void fnma_bcr(float &acc, float x, float a)
{
    asm("v_fmac_f32_dpp %0, -%1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}

template <int S> __device__ __forceinline__ void fma_bcr(double &acc, double x, double a)
{
    asm("v_fmac_f64_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}
template <int S> __device__ __forceinline__ void fma_bcr(float &acc, float x, float a)
{
    asm("v_fmac_f32_dpp %0, %1, %2 row_newbcast:%3 row_mask:0xf bank_mask:0xf"
        : "+v"(acc)
        : "v"(x), "v"(a), "i"(S));
}

/* Two wait states after the last VALU write of the registers named, before
 * any DPP read of them that follows: the s_nop sits in an asm that names them
 * as outputs, so no write of them can be scheduled after it. (Each call
 * covers only its own operands.) tests/test_dpp_hazards.py checks the built
 * code for every DPP read. */
template <class T> __device__ __forceinline__ void dpp_ready(T &a)
{
    asm volatile("s_nop 1" : "+v"(a));
}
template <class T> __device__ __forceinline__ void dpp_ready(T &a, T &b)
{
    asm volatile("s_nop 1" : "+v"(a), "+v"(b));
}
template <class T> __device__ __forceinline__ void dpp_ready(T &a, T &b, T &c, T &d)
{
    asm volatile("s_nop 1" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
}
/* entries B..R-1 of x */
template <int B, class T, int R> __device__ __forceinline__ void dpp_ready_from(T (&x)[R])
{
    constexpr int n = R - B;
    if constexpr (n == 1) dpp_ready(x[B]);
    else if constexpr (n == 2) dpp_ready(x[B], x[B + 1]);
    else if constexpr (n == 3) {
        dpp_ready(x[B], x[B + 1]);
        dpp_ready(x[B + 2]);
    }
    else if constexpr (n >= 4) {
        dpp_ready(x[B], x[B + 1], x[B + 2], x[B + 3]);
        if constexpr (n > 4) dpp_ready_from<B + 4>(x);
    }
}

/* column J of a, rows B..R-1 */
template <int B, int J, class T, int R, int C> __device__ __forceinline__ void dpp_ready_col(T (&a)[R][C])
{
    constexpr int n = R - B;
    if constexpr (n == 1) dpp_ready(a[B][J]);
    else if constexpr (n >= 2) {
        dpp_ready(a[B][J], a[B + 1][J]);
        if constexpr (n > 2) dpp_ready_col<B + 2, J>(a);
    }
}

/* the value of x in lane l of the wavefront (l wave-uniform) */
__device__ __forceinline__ double rdlane(double x, int l)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_readlane(w.x, l);
    w.y = __builtin_amdgcn_readlane(w.y, l);
    return __builtin_bit_cast(double, w);
}
__device__ __forceinline__ float rdlane(float x, int l)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l));
}

/* Index of the calling wavefront in its block, as a value the compiler knows
 * to be wave-uniform. Bases derived from it (group pointers, buffer
 * resources) stay in SGPRs; derived from threadIdx.x directly they would be
 * treated as divergent and every buffer access would become a waterfall loop. */
__device__ __forceinline__ int wave_in_block()
{
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

/* XCD-aware block order. MI355X deals consecutive workgroups round-robin to
 * its 8 XCDs, each with its own L2. Logical blocks 4k..4k+3 (whose matrices
 * share 128-byte lines at V = 64 when a block holds 4 matrices) are mapped to
 * one XCD's consecutive dispatch slots, so a line is fetched from HBM once.
 * The grid is padded to a multiple of 32 blocks; padding blocks get a logical
 * index >= nblocks and exit. */
__device__ __forceinline__ unsigned xcd_block(unsigned phys)
{
    const unsigned xcd = phys & 7u, slot = phys >> 3;
    return (slot >> 2) * 32u + xcd * 4u + (slot & 3u);
}
constexpr unsigned xcd_grid(unsigned nblocks) { return (nblocks + 31u) & ~31u; }

/* value of another lane of the row under a DPP pattern (quad_perm, mirrors) */
template <int CTRL> __device__ __forceinline__ float dpp(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), CTRL,
                                                              0xF, 0xF, true));
}
template <int CTRL> __device__ __forceinline__ double dpp(double x)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_mov_dpp(w.x, CTRL, 0xF, 0xF, true);
    w.y = __builtin_amdgcn_mov_dpp(w.y, CTRL, 0xF, 0xF, true);
    return __builtin_bit_cast(double, w);
}

/* max that propagates NaN (a NaN residual must never look converged) */
template <class T> __device__ __forceinline__ T fmax_nan(T a, T b)
{
    return (a > b || a != a) ? a : b;
}

/* ---- buffer addressing --------------------------------------------------
 * A team kernel addresses one compact group through a buffer resource whose
 * base is the group (wave-uniform: the 4 teams of a wave are 4 consecutive
 * slots of one group, V being a multiple of 4), a 32-bit per-lane byte offset
 * (row and slot) and a scalar byte offset (column): one VGPR per row block
 * instead of a 64-bit address per element. */
constexpr unsigned kOOB = 0x80000000u; /* a byte offset past every group's range */

/* the byte range a group's resource covers: the rest of the buffer, capped so
 * that kOOB stays out of range (a load there reads 0, a store is dropped) */
__device__ __forceinline__ unsigned group_range(long groups_left, size_t gstride, unsigned sz)
{
    const unsigned long long left = (unsigned long long)groups_left * gstride * sz;
    return left < 0x7fffffffull ? (unsigned)left : 0x7fffffffu;
}

__device__ __forceinline__ __amdgpu_buffer_rsrc_t group_rsrc(const void *base, unsigned range)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, range, 0x00020000);
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double bload(double, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                        unsigned soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
__device__ __forceinline__ float bload(float, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                       unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ void bstore(double x, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                       unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), r, voff, soff, 0);
}
__device__ __forceinline__ void bstore(float x, __amdgpu_buffer_rsrc_t r, unsigned voff,
                                       unsigned soff)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, x), r, voff, soff, 0);
}

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtT(float x) { return __builtin_sqrtf(x); }

} /* namespace team */
} /* namespace cbk */

#endif /* CBK_TEAM_COMMON_H */
