/* ormqr_reg.hip -- ormqr_compact (C := op(Q) C or C op(Q)) with one lane per
 * matrix and a block of C's columns in registers, for reflector lengths
 * nq <= 24 (and nq <= 32 in FP32): register buckets NQ in {8, 16, 24, 32},
 * guard-free at nq == NQ, with wave-uniform row guards on the run-time nq
 * below it (rows past nq are neither loaded, computed nor stored; the guarded
 * instances run at one wavefront per SIMD fewer, which keeps them free of
 * scratch).
 *
 * The lane kernel (lane_kernels.hip, ormqr_lane) applies each reflector in
 * place in HBM, so every element of C is read and written once per
 * reflector. Here a lane holds NCB columns of C' (C, or C^T for side R; nq x
 * ncols) in registers, streams each reflector's body V(kk+1:nq, kk) and
 * tau(kk) from HBM once per column block (at V = 64 a wavefront's 64 lanes read
 * one 512-byte run per element), and stores C' once. The operations are the
 * reference's larf (src/cbk_ormqr_compact.hpp:68-105), per element in the same
 * order as ormqr_lane:
 *     w = C'(kk,q) + sum_{i>kk} V(i,kk) C'(i,q)  (i ascending),  w *= tau(kk),
 *     C'(kk,q) -= w,  C'(i,q) -= V(i,kk) w,
 * reflectors forward when side L and trans T (or side R and N), backward
 * otherwise (src/cbk_ormqr_compact.hpp:131-156), so C is bit-identical to the
 * lane kernel's, hence to the reference's. Every register index is a
 * compile-time constant; the reflector count k is a wave-uniform guard.
 */

#include "cbk_internal.h"
#include "team_common.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace cbk {
namespace {

template <int B, int E, class F> __device__ __forceinline__ void rfor(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        rfor<B + 1, E>(f);
    }
}
template <int B, int E, class F> __device__ __forceinline__ void rfor_down(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, E - 1>{});
        rfor_down<B, E - 1>(f);
    }
}

__device__ __forceinline__ double fmaR(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaR(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

constexpr int kBlock = 256;

template <class T, int NQ, int NCB, bool EXACT>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu((NQ * NCB * sizeof(T) <= 256 ? 3 : 2) - (EXACT ? 0 : 1)))) void ormqr_reg_kernel(bool forward, bool ctr, int nq_arg, int ncols,
                                                           int k, Operand<T> A, const T *tau,
                                                           Operand<T> C, int V, long nlanes)
{
    constexpr unsigned SZ = sizeof(T);
    const int nq = EXACT ? NQ : nq_arg;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    /* the wavefront's first group (wave-uniform) addresses a buffer resource;
     * a lane's slot is a byte offset in it, an element's place a scalar one */
    const long w0 = __builtin_amdgcn_readfirstlane((int)(((long)blockIdx.x * blockDim.x +
                                                          (threadIdx.x & ~63u)) / V));
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const long groups_left = nlanes / V - w0;
    const int asi = A.row ? A.ld : 1, asj = A.row ? 1 : A.ld;
    int csi = C.row ? C.ld : 1, csj = C.row ? 1 : C.ld;
    if (ctr) {
        const int tmp = csi;
        csi = csj;
        csj = tmp;
    }
    const __amdgpu_buffer_rsrc_t ra = team::group_rsrc(A.p + w0 * A.gstride, team::group_range(groups_left, A.gstride, SZ));
    const __amdgpu_buffer_rsrc_t rc = team::group_rsrc(C.p + w0 * C.gstride, team::group_range(groups_left, C.gstride, SZ));
    const __amdgpu_buffer_rsrc_t rt = team::group_rsrc(tau + w0 * (size_t)k * V, team::group_range(groups_left, (size_t)k * V, SZ));
    const unsigned av = (unsigned)(((g - w0) * A.gstride + v) * SZ);
    const unsigned cv = (unsigned)(((g - w0) * C.gstride + v) * SZ);
    const unsigned tv = (unsigned)(((g - w0) * (long)k * V + v) * SZ);
    const unsigned VS = (unsigned)V * SZ;
    /* scalar element offsets are formed where they are used from strides the
     * compiler cannot see through (an empty asm per use site), so it does not
     * keep one scalar register per element for the whole kernel */
    auto opaque = [](unsigned x) {
        asm volatile("" : "+s"(x));
        return x;
    };
    unsigned asiV = (unsigned)asi * VS, asjV = (unsigned)asj * VS;
    unsigned csiV = (unsigned)csi * VS, csjV = (unsigned)csj * VS;

    /* one column block per grid row: no loop around the reflectors, whose
     * addresses would otherwise be kept live across its iterations */
    const int j0 = (int)blockIdx.y * NCB;
    T c[NQ][NCB];
    {
        const unsigned si = opaque(csiV), sj = opaque(csjV);
        rfor<0, NQ>([&](auto i) {
            rfor<0, NCB>([&](auto q) {
                /* a column past ncols reads zeros (out of the resource's range) */
                c[i][q] = team::bload(T(), rc, (j0 + q < ncols && i < nq) ? cv : team::kOOB, i * si + (j0 + q) * sj);
            });
        });
    }
    auto apply = [&](auto kc) {
        constexpr int kk = kc;
        if (kk < k) {
            const T t = team::bload(T(), rt, tv, (unsigned)kk * VS);
            const unsigned si = opaque(asiV), col = kk * opaque(asjV);
            T vv[NQ];
            rfor<kk + 1, NQ>([&](auto i) { vv[i] = team::bload(T(), ra, i < nq ? av : team::kOOB, col + i * si); });
            __builtin_amdgcn_sched_barrier(0); /* one reflector's loads in flight at a time */
            rfor<0, NCB>([&](auto q) {
                T w = c[kk][q];
                rfor<kk + 1, NQ>([&](auto i) {
                    if (i < nq) w = fmaR(vv[i], c[i][q], w);
                });
                w = w * t;
                c[kk][q] = c[kk][q] - w;
                rfor<kk + 1, NQ>([&](auto i) {
                    if (i < nq) c[i][q] = fmaR(-vv[i], w, c[i][q]);
                });
            });
        }
        __builtin_amdgcn_sched_barrier(0);
    };
    if (forward)
        rfor<0, NQ>(apply);
    else
        rfor_down<0, NQ>(apply);
    {
        const unsigned si = opaque(csiV), sj = opaque(csjV);
        rfor<0, NQ>([&](auto i) {
            rfor<0, NCB>([&](auto q) {
                team::bstore(c[i][q], rc, (j0 + q < ncols && i < nq) ? cv : team::kOOB, i * si + (j0 + q) * sj);
            });
        });
    }
}

template <class T, int NQ, int NCB>
hipError_t launch(bool forward, bool ctr, int nq, int ncols, int k, Operand<T> A, const T *tau,
                  Operand<T> C, int V, long nl)
{
    const dim3 grid((unsigned)((nl + kBlock - 1) / kBlock), (unsigned)((ncols + NCB - 1) / NCB));
    if (nq == NQ)
        hipLaunchKernelGGL((ormqr_reg_kernel<T, NQ, NCB, true>), grid, dim3(kBlock), 0,
                           current_stream(), forward, ctr, nq, ncols, k, A, tau, C, V, nl);
    else
        hipLaunchKernelGGL((ormqr_reg_kernel<T, NQ, NCB, false>), grid, dim3(kBlock), 0,
                           current_stream(), forward, ctr, nq, ncols, k, A, tau, C, V, nl);
    return hipGetLastError();
}

} /* namespace */

/* FP64 past nq = 24 would spill (the 16-lane team kernel serves nq = 32) */
bool ormqr_reg_supports(int nq, bool fp64) { return nq >= 1 && nq <= (fp64 ? 24 : 32); }

template <class T>
hipError_t launch_ormqr_reg(bool left, bool tran, int m, int n, int k, Operand<T> A,
                            const T *tau, Operand<T> C, int V, int nm)
{
    const long nl = (long)((nm + V - 1) / V) * V;
    const int nq = left ? m : n, ncols = left ? n : m;
    const bool forward = (left == tran);
    if (nq <= 8) return launch<T, 8, 4>(forward, !left, nq, ncols, k, A, tau, C, V, nl);
    if (nq <= 16) return launch<T, 16, 2>(forward, !left, nq, ncols, k, A, tau, C, V, nl);
    if (nq <= 24) return launch<T, 24, 2>(forward, !left, nq, ncols, k, A, tau, C, V, nl);
    if constexpr (sizeof(T) == 4) {
        if (nq <= 32) return launch<T, 32, 2>(forward, !left, nq, ncols, k, A, tau, C, V, nl);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_ormqr_reg<double>(bool, bool, int, int, int, Operand<double>,
                                             const double *, Operand<double>, int, int);
template hipError_t launch_ormqr_reg<float>(bool, bool, int, int, int, Operand<float>,
                                            const float *, Operand<float>, int, int);

} /* namespace cbk */
