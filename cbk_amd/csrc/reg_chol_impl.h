/* reg_chol_impl.h -- potrf/posv/potrs and sytrfnp/sysvnp/sytrsnp for small
 * orders (n <= 16 in FP64, n <= 24 in FP32) with one lane per matrix and the
 * whole lower triangle in registers (at the largest orders a lane holds up
 * to 300 values: the compiler parks part of them in accumulation registers,
 * one wavefront per SIMD, no scratch).
 *
 * The general lane kernel (lane_kernels.hip) works in place in HBM; at small n
 * the arithmetic per byte is low and every element access is a round trip. Here
 * a lane loads its lower triangle and right-hand side once (at V = 64 each
 * element of a wavefront's 64 matrices is one 512-byte transaction), runs the
 * same operations as the lane kernel -- chol_lane's right-looking per-pivot
 * order (panel form within the reference's 8-wide panels, trailing form
 * beyond), then potrs / sytrsnp as the reference's trsm routing computes them
 * (the row-dot form, and the axpy form where the router picks it for a
 * single leftover column of the backward sweep; in the forward sweep the two
 * forms perform the same operations per element) -- on registers, and stores
 * the results once. Elements are addressed through buffer resources based at
 * the wavefront's first group (a lane's slot a 32-bit offset, an element's
 * place a scalar one); rows past n read zeros and drop their stores. Loops run to the
 * compile-time bound NMAX with wave-uniform guards on the run-time n, so
 * every register index is a constant; when n is NMAX exactly (EXACT) the
 * guards fold away (a third fewer instructions at n = 24). Results are bit-identical to the lane
 * kernel, hence to the reference.
 */

#pragma once

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

__device__ __forceinline__ double sqrtR(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtR(float x) { return __builtin_sqrtf(x); }

constexpr int kBlock = 256;
constexpr int kNB = 8; /* the reference's panel width, src/cbk_potrf_compact.hpp:71-74 */

template <class T, int NMAX, bool LDL, bool EXACT>
__global__ __launch_bounds__(kBlock) void reg_chol_kernel(bool factor, bool up, int n_arg, int nrhs,
                                                          Operand<T> A, Operand<T> B, int V,
                                                          long nlanes, int nm, int *info)
{
    constexpr unsigned SZ = sizeof(T);
    /* n == NMAX: every row guard folds away */
    const int n = EXACT ? NMAX : n_arg;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long w0 = __builtin_amdgcn_readfirstlane(
        (int)(((long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) / V));
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const long groups_left = nlanes / V - w0;
    const unsigned VS = (unsigned)V * SZ;
    auto opaque = [](unsigned x) {
        asm volatile("" : "+s"(x));
        return x;
    };
    /* the lower view: uplo 'U' reads A transposed (src/cbk_common.hpp:201-206) */
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (up) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(A.p + w0 * A.gstride, team::group_range(groups_left, A.gstride, SZ));
    const unsigned av = (unsigned)(((g - w0) * A.gstride + v) * SZ);
    const unsigned siV = (unsigned)si * VS, sjV = (unsigned)sj * VS;

    T a[NMAX][NMAX];
    {
        const unsigned s_i = opaque(siV), s_j = opaque(sjV);
        rfor<0, NMAX>([&](auto i) {
            rfor<0, i + 1>([&](auto j) { a[i][j] = team::bload(T(), ra, i < n ? av : team::kOOB, i * s_i + j * s_j); });
        });
    }

    const bool real = idx < nm;
    int linfo = 0;
    if (factor) {
        /* chol_lane (lane_kernels.hip): per pivot j, panel form on the columns
         * of j's panel, scaling of the column, trailing form beyond */
        rfor<0, NMAX>([&](auto jc) {
            constexpr int j = jc;
            if (j < n) {
                const int pe = min((j / kNB) * kNB + kNB, n);
                const T piv = a[j][j];
                if (linfo == 0 && (LDL ? (piv == T(0)) : !(piv > T(0)))) linfo = j + 1;
                const T inv = T(1) / piv;
                rfor<j + 1, NMAX>([&](auto c) {
                    if (c < pe) {
                        const T w = a[c][j] * inv;
                        rfor<c, NMAX>([&](auto i) {
                            if (i < n) a[i][c] = team::fmaT(-a[i][j], w, a[i][c]);
                        });
                    }
                });
                T s;
                if (LDL) {
                    s = inv;
                }
                else {
                    const T d = sqrtR(piv);
                    a[j][j] = d;
                    s = d * inv;
                }
                rfor<j + 1, NMAX>([&](auto i) {
                    if (i < n) a[i][j] = a[i][j] * s;
                });
                rfor<j + 1, NMAX>([&](auto c) {
                    if (c >= pe && c < n) {
                        const T w = LDL ? a[c][j] * a[j][j] : a[c][j];
                        rfor<c, NMAX>([&](auto i) {
                            if (i < n) a[i][c] = team::fmaT(-a[i][j], w, a[i][c]);
                        });
                    }
                });
            }
        });
        if (info && real) info[idx] = linfo;
        const unsigned s_i = opaque(siV), s_j = opaque(sjV);
        rfor<0, NMAX>([&](auto i) {
            rfor<0, i + 1>([&](auto j) { team::bstore(a[i][j], ra, i < n ? av : team::kOOB, i * s_i + j * s_j); });
        });
    }
    if (nrhs <= 0 || linfo != 0 && info) return;

    /* potrs / sytrsnp on the plain operands: forward with L, then (sytrsnp)
     * the reciprocal-then-multiply by D, then backward with L^T (the dot form
     * of the reference's trsm, src/cbk_trsm_compact.hpp:73-104) */
    const int bsi = B.row ? B.ld : 1, bsj = B.row ? 1 : B.ld;
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(B.p + w0 * B.gstride, team::group_range(groups_left, B.gstride, SZ));
    const unsigned bv = (unsigned)(((g - w0) * B.gstride + v) * SZ);
    const unsigned bsiV = (unsigned)bsi * VS, bsjV = (unsigned)bsj * VS;
    T invd[NMAX];
    if (LDL) {
        rfor<0, NMAX>([&](auto i) { invd[i] = (i < n) ? T(1) / a[i][i] : T(0); });
    }
    /* the reference's trsm router (src/cbk_trsm_compact.hpp:110-145, 239-254)
     * decides on strides: with B's row stride 1 it takes the tuned path
     * (operands as given when A's row stride is 1, flipped when A's column
     * stride is 1), whose single leftover column of the 4/2/1 blocking gets
     * the axpy form when the operand it sees is not transposed. For the
     * backward sweep (trsm with tran = !up) that is: */
    const int asi = A.row ? A.ld : 1, asj = A.row ? 1 : A.ld;
    const bool axpy_back = bsi == 1 && (asi == 1 ? up : (asj == 1 ? !up : false));
    const bool single_tail = (nrhs % 4 == 1) || (nrhs % 4 == 3);
    for (int q = 0; q < nrhs; ++q) {
        T b[NMAX];
        const unsigned b_i = opaque(bsiV), col = (unsigned)q * opaque(bsjV);
        rfor<0, NMAX>([&](auto i) { b[i] = team::bload(T(), rb, i < n ? bv : team::kOOB, col + i * b_i); });
        rfor<0, NMAX>([&](auto i) {
            if (i < n) {
                T w = b[i] * T(1);
                rfor<0, i>([&](auto l) { w = team::fmaT(-a[i][l], b[l], w); });
                b[i] = LDL ? w : w / a[i][i];
            }
        });
        if (LDL) {
            rfor<0, NMAX>([&](auto i) {
                if (i < n) b[i] = b[i] * invd[i];
            });
        }
        if (axpy_back && single_tail && q == nrhs - 1) {
            /* the axpy form of the backward sweep: pivots descending, each
             * column of L^T applied at once (a different order per element) */
            rfor<0, NMAX>([&](auto kc) {
                constexpr int k = NMAX - 1 - kc;
                if (k < n) {
                    const T x = LDL ? b[k] : b[k] / a[k][k];
                    b[k] = x;
                    rfor<0, k>([&](auto i) { b[i] = team::fmaT(-a[k][i], x, b[i]); });
                }
            });
        }
        else {
            rfor<0, NMAX>([&](auto ic) {
                constexpr int i = NMAX - 1 - ic;
                if (i < n) {
                    T w = b[i] * T(1);
                    rfor<i + 1, NMAX>([&](auto l) {
                        if (l < n) w = team::fmaT(-a[l][i], b[l], w);
                    });
                    b[i] = LDL ? w : w / a[i][i];
                }
            });
        }
        rfor<0, NMAX>([&](auto i) { team::bstore(b[i], rb, i < n ? bv : team::kOOB, col + i * b_i); });
    }
}

} /* namespace */

/* one (order bound, form) of the register kernel: the guard-free instance at
 * n == NMAX, the guarded one below it (instantiated per translation unit, so
 * the large unrolled instances build in parallel) */
template <class T, int NMAX, bool LDL>
hipError_t launch_reg_chol_n(bool factor, bool up, int n, int nrhs, Operand<T> A, Operand<T> B,
                             int V, int nm, int *info)
{
    const long nl = (long)((nm + V - 1) / V) * V;
    const dim3 grid((unsigned)((nl + kBlock - 1) / kBlock)), block(kBlock);
    const hipStream_t s = current_stream();
    /* FP64 LDL^T at 16 keeps the guarded instance at n == 16 too: it measured
     * 6-7 % faster than the guard-free one there (a different register
     * allocation); both are bit-identical */
    constexpr bool exact_ok = !(sizeof(T) == 8 && NMAX == 16 && LDL);
    bool exact = false;
    if constexpr (exact_ok) {
        if (n == NMAX) {
            exact = true;
            hipLaunchKernelGGL((reg_chol_kernel<T, NMAX, LDL, true>), grid, block, 0, s, factor, up,
                               n, nrhs, A, B, V, nl, nm, info);
        }
    }
    if (!exact)
        hipLaunchKernelGGL((reg_chol_kernel<T, NMAX, LDL, false>), grid, block, 0, s, factor, up, n,
                           nrhs, A, B, V, nl, nm, info);
    return hipGetLastError();
}

} /* namespace cbk */
