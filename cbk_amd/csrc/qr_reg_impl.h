/* qr_reg_impl.h -- geqrf_compact, the overdetermined gels_compact and
 * orgqr_compact with one lane per matrix and the whole tall view F (p x q,
 * F = A or A^T; Q for orgqr) in registers, for p <= 16 with q <= 8 and
 * p <= 32 with q <= 4 (q <= 8 in FP32). Register buckets P x Q hold the view;
 * at p x q = P x Q the loops are guard-free, otherwise wave-uniform guards on
 * the run-time p and q skip the rows and columns past the view (they are
 * neither loaded, computed nor stored, so no padding enters a sum).
 *
 * The lane kernels (lane_kernels.hip) keep F in HBM and read and write it once
 * per reflector; the team kernels spread a matrix over 16 lanes and are bound
 * by the latency of their cross-lane sums. Here a lane loads F once, factors
 * it in registers, stores it once, and then takes B one column at a time
 * (load, apply the q reflectors, back substitution, store), so F, tau and B
 * cross HBM once each; at V = 64 a wavefront's 64 lanes read one 512-byte run
 * per element. Every register index is a compile-time constant; the
 * right-hand side count is a loop bound.
 *
 * The operations and their order are the lane kernel's, hence the
 * reference's, so every output is bit-identical to it:
 *   geqrf  larfg + larf per column (src/cbk_geqrf_compact.hpp:74-131):
 *          tail = sum_{i>kk} F(i,kk)^2 (i ascending), norm = sqrt(x0^2 + tail),
 *          beta = -sign(x0) norm, tau = (beta - x0)/beta, the body scaled by
 *          1/(x0 - beta); each later column p: w = F(kk,p) + sum V(i) F(i,p),
 *          w *= tau, F(kk,p) -= w, F(i,p) -= V(i) w.
 *   gels   (src/cbk_gels_compact.hpp:84-146): Q^T B by the same larf per
 *          column, reflectors forward (src/cbk_ormqr_compact.hpp:109-124),
 *          then R X = B's top q rows by trsm_compact_group's form
 *          (src/cbk_trsm_compact.hpp:73-145, 239-254): the dot form (x_i = (b_i
 *          - sum_{l>i} R(i,l) x_l) / R(i,i), l ascending) for every column,
 *          except the last one of an odd count when F and B are both unit
 *          row stride, which takes the axpy form (x_k = b_k / R(k,k), then
 *          b_i -= R(i,k) x_k for i < k). With info, an exact zero on R's
 *          diagonal leaves B as Q^T B and reports its index + 1 (LAPACK
 *          dgels's dtrtrs test), as the lane kernel does.
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
template <int B, int E, class F> __device__ __forceinline__ void rfor_down(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, E - 1>{});
        rfor_down<B, E - 1>(f);
    }
}

__device__ __forceinline__ double sqrtR(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtR(float x) { return __builtin_sqrtf(x); }

constexpr int kBlock = 256;

template <class T> struct RegQrArgs {
    T *a;
    int a_si, a_sj; /* element strides of the tall view F */
    size_t a_gs;
    T *tau;
    T *b; /* null: geqrf */
    int b_si, b_sj;
    size_t b_gs;
    int rows, cols; /* p x q of the view (rows >= cols) */
    int nrhs;
    bool axpy_last; /* the last column takes trsm's axpy form */
    int V;
    long nlanes;
    int nm;
    int *info;
};

template <class T, int P, int Q, bool EXACT>
__global__ __launch_bounds__(kBlock) void qr_reg_kernel(RegQrArgs<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    /* the view is rp x rq (rp >= rq); rows and columns past it are neither
     * loaded, computed nor stored */
    const int rp = EXACT ? P : p.rows, rq = EXACT ? Q : p.cols;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    /* the wavefront's first group (wave-uniform) bases the buffer resources;
     * a lane's slot is a byte offset in them, an element's place a scalar one */
    const long w0 = __builtin_amdgcn_readfirstlane(
        (int)(((long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) / p.V));
    if (idx >= p.nlanes) return;
    const long g = idx / p.V;
    const int v = (int)(idx % p.V);
    const long groups_left = p.nlanes / p.V - w0;
    const unsigned VS = (unsigned)p.V * SZ;
    auto opaque = [](unsigned x) {
        asm volatile("" : "+s"(x));
        return x;
    };
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + w0 * p.a_gs, team::group_range(groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rt = team::group_rsrc(
        p.tau + w0 * (size_t)rq * p.V, team::group_range(groups_left, (size_t)rq * p.V, SZ));
    const unsigned av = (unsigned)(((g - w0) * p.a_gs + v) * SZ);
    const unsigned tv = (unsigned)(((g - w0) * (long)rq * p.V + v) * SZ);
    const unsigned asiV = (unsigned)p.a_si * VS, asjV = (unsigned)p.a_sj * VS;

    T f[P][Q];
    {
        const unsigned si = opaque(asiV), sj = opaque(asjV);
        rfor<0, P>([&](auto i) {
            rfor<0, Q>([&](auto j) {
                f[i][j] = team::bload(T(), ra, (i < rp && j < rq) ? av : team::kOOB, i * si + j * sj);
            });
        });
    }
    /* geqrf (src/cbk_geqrf_compact.hpp:74-131) */
    T tau[Q];
    rfor<0, Q>([&](auto kc) {
        constexpr int kk = kc;
        if (kk < rq) {
            T tail = T(0);
            rfor<kk + 1, P>([&](auto i) {
                if (i < rp) tail = team::fmaT(f[i][kk], f[i][kk], tail);
            });
            const T x0 = f[kk][kk];
            const T norm = sqrtR(team::fmaT(x0, x0, tail));
            const T beta = (x0 >= T(0)) ? -norm : norm;
            const bool has = tail > T(0);
            const T t = has ? (beta - x0) / beta : T(0);
            const T inv = has ? T(1) / (x0 - beta) : T(0);
            f[kk][kk] = has ? beta : x0;
            tau[kk] = t;
            rfor<kk + 1, P>([&](auto i) {
                if (i < rp) f[i][kk] = f[i][kk] * inv;
            });
            rfor<kk + 1, Q>([&](auto c) {
                if (c < rq) {
                    T w = f[kk][c];
                    rfor<kk + 1, P>([&](auto i) {
                        if (i < rp) w = team::fmaT(f[i][kk], f[i][c], w);
                    });
                    w = w * t;
                    f[kk][c] = f[kk][c] - w;
                    rfor<kk + 1, P>([&](auto i) {
                        if (i < rp) f[i][c] = team::fmaT(-f[i][kk], w, f[i][c]);
                    });
                }
            });
        }
    });
    {
        const unsigned si = opaque(asiV), sj = opaque(asjV), ts = opaque(VS);
        rfor<0, P>([&](auto i) {
            rfor<0, Q>([&](auto j) {
                team::bstore(f[i][j], ra, (i < rp && j < rq) ? av : team::kOOB, i * si + j * sj);
            });
        });
        rfor<0, Q>([&](auto j) { team::bstore(tau[j], rt, j < rq ? tv : team::kOOB, j * ts); });
    }
    if (!p.b) return;

    /* info (real matrices only): the first exact zero on R's diagonal */
    int bad = 0;
    if (p.info && idx < p.nm) {
        rfor_down<0, Q>([&](auto i) {
            if (i < rq && f[i][i] == T(0)) bad = i + 1;
        });
        p.info[idx] = bad;
    }
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + w0 * p.b_gs, team::group_range(groups_left, p.b_gs, SZ));
    const unsigned bv = (unsigned)(((g - w0) * p.b_gs + v) * SZ);
    const unsigned bsiV = (unsigned)p.b_si * VS, bsjV = (unsigned)p.b_sj * VS;
    for (int j = 0; j < p.nrhs; ++j) {
        const unsigned si = opaque(bsiV), col = (unsigned)j * opaque(bsjV);
        T b[P];
        rfor<0, P>([&](auto i) { b[i] = team::bload(T(), rb, i < rp ? bv : team::kOOB, col + i * si); });
        /* Q^T b, reflectors forward (src/cbk_ormqr_compact.hpp:68-124) */
        rfor<0, Q>([&](auto kc) {
            constexpr int kk = kc;
            if (kk < rq) {
                T w = b[kk];
                rfor<kk + 1, P>([&](auto i) {
                    if (i < rp) w = team::fmaT(f[i][kk], b[i], w);
                });
                w = w * tau[kk];
                b[kk] = b[kk] - w;
                rfor<kk + 1, P>([&](auto i) {
                    if (i < rp) b[i] = team::fmaT(-f[i][kk], w, b[i]);
                });
            }
        });
        if (!bad) {
            if (p.axpy_last && j == p.nrhs - 1) {
                /* trsm_axpy_col (src/cbk_trsm_compact.hpp:49-67) */
                rfor_down<0, Q>([&](auto kc) {
                    constexpr int kk = kc;
                    if (kk < rq) {
                        const T xk = b[kk] / f[kk][kk];
                        b[kk] = xk;
                        rfor<0, kk>([&](auto i) { b[i] = team::fmaT(-f[i][kk], xk, b[i]); });
                    }
                });
            }
            else {
                /* trsm_dot_block, one column (src/cbk_trsm_compact.hpp:73-104) */
                rfor_down<0, Q>([&](auto ic) {
                    constexpr int i = ic;
                    if (i < rq) {
                        T w = b[i];
                        rfor<i + 1, Q>([&](auto l) {
                            if (l < rq) w = team::fmaT(-f[i][l], b[l], w);
                        });
                        b[i] = w / f[i][i];
                    }
                });
            }
        }
        rfor<0, P>([&](auto i) { team::bstore(b[i], rb, i < rp ? bv : team::kOOB, col + i * si); });
    }
}

template <class T, int P, int Q> hipError_t launch(const RegQrArgs<T> &a)
{
    const unsigned nb = (unsigned)((a.nlanes + kBlock - 1) / kBlock);
    if (a.rows == P && a.cols == Q)
        hipLaunchKernelGGL((qr_reg_kernel<T, P, Q, true>), dim3(nb), dim3(kBlock), 0, current_stream(), a);
    else
        hipLaunchKernelGGL((qr_reg_kernel<T, P, Q, false>), dim3(nb), dim3(kBlock), 0, current_stream(), a);
    return hipGetLastError();
}


/* orgqr_compact_group (src/cbk_orgqr_compact.hpp:51-85) with Q (p x q) in
 * registers: the columns j >= k start as e_j, then for kk = k-1 down to 0 the
 * reflector kk is applied to the columns right of it by the same larf and its
 * own column becomes -tau v below the diagonal, 1 - tau on it and 0 above */
template <class T, int P, int Q, bool EXACT>
__global__ __launch_bounds__(kBlock) void orgqr_reg_kernel(int m_arg, int n_arg, int k, Operand<T> A,
                                                           const T *tau_p, int V, long nlanes)
{
    constexpr unsigned SZ = sizeof(T);
    const int rp = EXACT ? P : m_arg, rq = EXACT ? Q : n_arg;
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
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(A.p + w0 * A.gstride, team::group_range(groups_left, A.gstride, SZ));
    const __amdgpu_buffer_rsrc_t rt = team::group_rsrc(
        tau_p + w0 * (size_t)k * V, team::group_range(groups_left, (size_t)k * V, SZ));
    const unsigned av = (unsigned)(((g - w0) * A.gstride + v) * SZ);
    const unsigned tv = (unsigned)(((g - w0) * (long)k * V + v) * SZ);
    const unsigned asiV = (unsigned)(A.row ? A.ld : 1) * VS, asjV = (unsigned)(A.row ? 1 : A.ld) * VS;

    T f[P][Q];
    {
        const unsigned si = opaque(asiV), sj = opaque(asjV);
        rfor<0, P>([&](auto i) {
            rfor<0, Q>([&](auto j) {
                /* only the reflector bodies (columns < k) are read */
                f[i][j] = team::bload(T(), ra, (j < k && i < rp) ? av : team::kOOB, i * si + j * sj);
                if (j >= k) f[i][j] = (i == j) ? T(1) : T(0);
            });
        });
    }
    rfor_down<0, Q>([&](auto kc) {
        constexpr int kk = kc;
        if (kk < k) {
            const T t = team::bload(T(), rt, tv, (unsigned)kk * opaque(VS));
            rfor<kk + 1, Q>([&](auto c) {
                if (c < rq) {
                    T w = f[kk][c];
                    rfor<kk + 1, P>([&](auto i) {
                        if (i < rp) w = team::fmaT(f[i][kk], f[i][c], w);
                    });
                    w = w * t;
                    f[kk][c] = f[kk][c] - w;
                    rfor<kk + 1, P>([&](auto i) {
                        if (i < rp) f[i][c] = team::fmaT(-f[i][kk], w, f[i][c]);
                    });
                }
            });
            const T mt = -t;
            rfor<kk + 1, P>([&](auto i) {
                if (i < rp) f[i][kk] = f[i][kk] * mt;
            });
            f[kk][kk] = T(1) - t;
            rfor<0, kk>([&](auto i) { f[i][kk] = T(0); });
        }
    });
    {
        const unsigned si = opaque(asiV), sj = opaque(asjV);
        rfor<0, P>([&](auto i) {
            rfor<0, Q>([&](auto j) {
                team::bstore(f[i][j], ra, (i < rp && j < rq) ? av : team::kOOB, i * si + j * sj);
            });
        });
    }
}

template <class T, int P, int Q>
hipError_t launch_orgqr(int m, int n, int k, Operand<T> A, const T *tau, int V, long nl)
{
    const unsigned nb = (unsigned)((nl + kBlock - 1) / kBlock);
    if (m == P && n == Q)
        hipLaunchKernelGGL((orgqr_reg_kernel<T, P, Q, true>), dim3(nb), dim3(kBlock), 0,
                           current_stream(), m, n, k, A, tau, V, nl);
    else
        hipLaunchKernelGGL((orgqr_reg_kernel<T, P, Q, false>), dim3(nb), dim3(kBlock), 0,
                           current_stream(), m, n, k, A, tau, V, nl);
    return hipGetLastError();
}

template <int X> using ic = std::integral_constant<int, X>;

/* the register bucket P x Q (8x4, 8x8, 16x4, 16x8, 32x4, FP32 32x8) that
 * holds a p x q view: its kernel runs guard-free when p x q is the bucket
 * itself, with wave-uniform row and column guards otherwise */
template <class T, class F> hipError_t by_bucket(int p, int q, F &&go)
{
    if (p <= 8) return q <= 4 ? go(ic<8>{}, ic<4>{}) : go(ic<8>{}, ic<8>{});
    if (p <= 16) return q <= 4 ? go(ic<16>{}, ic<4>{}) : go(ic<16>{}, ic<8>{});
    if constexpr (sizeof(T) == 4) {
        if (q > 4) return go(ic<32>{}, ic<8>{});
    }
    return go(ic<32>{}, ic<4>{});
}

} /* namespace */

template <class T>
hipError_t launch_qr_reg(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                         Operand<T> B, int V, int nm, int *info)
{
    RegQrArgs<T> a;
    a.a = A.p;
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (transposed) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    a.a_si = si;
    a.a_sj = sj;
    a.a_gs = A.gstride;
    a.tau = tau;
    a.b = B.p;
    a.b_si = B.row ? B.ld : 1;
    a.b_sj = B.row ? 1 : B.ld;
    a.b_gs = B.gstride;
    a.nrhs = B.p ? nrhs : 0;
    /* trsm_compact_group's first branch (unit row strides of both views)
     * takes the axpy form for the last column of an odd count */
    a.axpy_last = (a.b_si == 1 && a.a_si == 1 && (nrhs % 4 == 1 || nrhs % 4 == 3));
    a.V = V;
    a.nlanes = (long)((nm + V - 1) / V) * V;
    a.nm = nm;
    a.info = B.p ? info : nullptr;
    a.rows = p;
    a.cols = q;
    if (!qr_reg_supports(p, q, sizeof(T) == 8)) return hipErrorInvalidValue;
    return by_bucket<T>(p, q, [&](auto P, auto Q) { return launch<T, P, Q>(a); });
}

template <class T>
hipError_t launch_orgqr_reg(int m, int n, int k, Operand<T> A, const T *tau, int V, int nm)
{
    const long nl = (long)((nm + V - 1) / V) * V;
    if (!qr_reg_supports(m, n, sizeof(T) == 8)) return hipErrorInvalidValue;
    return by_bucket<T>(m, n, [&](auto P, auto Q) { return launch_orgqr<T, P, Q>(m, n, k, A, tau, V, nl); });
}

} /* namespace cbk */
