/* mixed_batch.hip -- the kernels of dsgels_batch, the mixed-precision form
 * of an uneven least-squares batch (batch_plan.cpp records them into the
 * plan's graph; include/cbk_ext.h states the contract):
 *
 *   mixed_prep_kernel     A32 := fl32(A), B32 := fl32(B) for the problems
 *                         the FP32 path serves; X := B, iter := -1 for the
 *                         others
 *   (the FP32 batch gels of batch_gels_impl.h on A32 / B32: R of each
 *    problem in A32's upper triangle, its FP32 solution in B32)
 *   mixed_refine_kernel   FP64 refinement of that solution, one problem per
 *                         wavefront; a problem that does not converge gets
 *                         X := B and joins the FP64 fallback list
 *   (the FP64 batch gels over the fallback lists, in place on A and X)
 *
 * Refinement (the per-column scheme of dsgels_compact, started from the FP32
 * solution x_0), with R^-1 formed once per problem in FP32 (column j by back
 * substitution, in lane j) so that no step waits on a pivot chain: per step
 *     r = b - A x          FP64, per row, fma over columns ascending
 *     g = A^T r            FP64, per column, fma over rows ascending
 *     y = R^-T fl32(g)     FP32, lane j sums k <= j ascending
 *     h = R^-1 y           FP32, lane i sums k >= i ascending
 *     x += h
 * until ||h||_inf <= ||x||_inf eps sqrt(n) (eps = 2^-53), or the correction
 * stopped halving once below ||x||_inf 2^-40; a correction that does not
 * contract, a NaN, a zero or non-finite R(k,k) or 30 steps send the problem
 * to FP64.
 *
 * Layout: lane i holds row i of A in registers for the residual; A is also
 * staged once in LDS (row-major, padded stride) for A^T r, lane = column;
 * lane j holds column j and row j of R^-1 for the two triangular products;
 * x, r, g and y pass between the phases through LDS as broadcasts, so there
 * are no cross-lane reductions. Column q's convergence is decided by lane q
 * over the |h| and |x| the lanes leave in LDS. HBM traffic per problem: A, B
 * and the FP32 R and x_0 read once, X written once.
 */

#include "cbk_internal.h"
#include "team_common.h"

#include <hip/hip_runtime.h>

namespace cbk {
namespace mixedbatch {

using team::fmax_nan;

__device__ __forceinline__ float shfl_f(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v)));
}

__device__ __forceinline__ double shfl_xor_d(double v, int mask)
{
    const int src = ((int)threadIdx.x & 63) ^ mask;
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b & 0xffffffffll));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

/* ---- prep ------------------------------------------------------------------ */

/* dst(i, j) := src(i, j) for the rows x cols block of a column-major problem
 * (leading dimension ld): the block's elements dealt over the wavefront's
 * lanes in column order, four loads in flight a lane, so that a problem costs
 * a few memory round trips rather than one per column */
template <class S, class D>
__device__ __forceinline__ void copy_block(D *dst, const S *src, int rows, int cols, int ld, int lane)
{
    constexpr int U = 4;
    const int total = rows * cols;
    for (int e0 = 0; e0 < total; e0 += 64 * U) {
        long long off[U];
        S v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + lane + 64 * u;
            const int c = e < total ? e / rows : 0;
            off[u] = e < total ? (long long)c * ld + (e - c * rows) : -1;
            v[u] = off[u] >= 0 ? src[off[u]] : S(0);
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (off[u] >= 0) dst[off[u]] = (D)v[u];
    }
}

/* one wavefront per problem: convert (mode 0, MixedProblem list) or copy B's
 * max(m, n) x nrhs block into X (mode 1, BatchProblem list) */
__global__ __launch_bounds__(256) void mixed_prep_kernel(MixedPrepArgs p)
{
    const long k = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= p.count) return;
    if (p.conv) {
        const MixedProblem d = p.conv[k];
        copy_block(p.a32 + d.a_off, p.a + d.a_off, d.m, d.n, d.lda, lane);
        copy_block(p.b32 + d.b_off, p.b + d.b_off, d.m, d.nrhs, d.ldb, lane);
    }
    else {
        const BatchProblem d = p.copy[k];
        const int rows = d.m > d.n ? d.m : d.n;
        copy_block(p.x + d.b_off, p.b + d.b_off, rows, d.nrhs, d.ldb, lane);
        if (lane == 0) p.iter[d.idx] = -1;
    }
}

/* ---- refinement -------------------------------------------------------------- */

template <int MB, int NB, int Q>
__global__ __launch_bounds__(64) void mixed_refine_kernel(MixedRefineArgs p)
{
    constexpr int SA = NB + 1; /* LDS row stride of A, R and R^-1 */
    __shared__ double sA[MB * SA];  /* A(i, c) for the A^T r pass */
    __shared__ double sR[MB * Q];   /* r(i, q) */
    __shared__ double sX[NB * Q];   /* x(c, q) */
    __shared__ float sF[NB * SA];   /* R(k, j), j >= k; then R^-1(k, j) */
    __shared__ float sV[NB * Q];    /* fl32(g)(k, q); then y(k, q) */
    __shared__ float sD[NB * Q];    /* |h(k, q)| */
    __shared__ double sN[NB * Q];   /* |x(k, q)| */

    const long kq = blockIdx.x;
    if (kq >= p.count) return;
    const int lane = threadIdx.x;
    const MixedProblem d = p.desc[kq];
    const int m = d.m, n = d.n, nrhs = d.nrhs;
    const double *A = p.a + d.a_off;
    const double *B = p.b + d.b_off;
    const float *A32 = p.a32 + d.a_off;
    const float *B32 = p.b32 + d.b_off;

    /* row i of A in registers (the residual), A and R in LDS */
    double arow[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) {
        arow[c] = (c < n && lane < m) ? A[(long long)c * d.lda + lane] : 0.0;
        if (c < n && lane < m) sA[lane * SA + c] = arow[c];
        if (c < n && lane <= c) sF[lane * SA + c] = A32[(long long)c * d.lda + lane];
    }
    double bq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        bq[q] = (lane < m && q < nrhs) ? B[(long long)q * d.ldb + lane] : 0.0;
    /* x_0 = the FP32 solution, lane = row j < n */
    double xq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        xq[q] = (lane < n && q < nrhs) ? (double)B32[(long long)q * d.ldb + lane] : 0.0;
        if (lane < NB) sX[lane * Q + q] = xq[q];
    }
    __syncthreads();
    bool ok = true;
    for (int k = 0; k < n; ++k) {
        const float r = sF[k * SA + k];
        ok = ok && r != 0.0f && __builtin_isfinite(r);
    }
    if (!ok) goto fallback;
    {
        /* column j = lane of R^-1 by back substitution: z_j = 1 / R(j,j), then
         * for i < j descending z_i = -(sum_{k=i+1..j} R(i,k) z_k) / R(i,i) */
        float z[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k)
            z[k] = (k == lane && k < n) ? 1.0f / sF[k * SA + k] : 0.0f;
#pragma unroll
        for (int i = NB - 2; i >= 0; --i) {
            if (i + 1 < n) {
                float s = 0.0f;
#pragma unroll
                for (int k = i + 1; k < NB; ++k)
                    if (k < n) s = __builtin_fmaf(sF[i * SA + k], z[k], s);
                if (i < lane) z[i] = -s / sF[i * SA + i];
            }
        }
        __syncthreads(); /* every lane has read R */
        float rcol[NB], rrow[NB];
#pragma unroll
        for (int k = 0; k < NB; ++k) {
            rcol[k] = z[k];
            if (lane < NB) sF[k * SA + lane] = z[k]; /* R^-1(k, lane) */
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NB; ++k)
            rrow[k] = (lane < n && k >= lane && k < n) ? sF[lane * SA + k] : 0.0f;

        const double cte = 0x1p-53 * __builtin_sqrt((double)n), floor_ = 0x1p-40;
        /* lanes q < nrhs keep column q's state */
        double prev = __builtin_inf();
        bool done = lane >= nrhs, fail = false;
        int steps = 0;
        unsigned long long done_mask = 0; /* columns settled, bit q */
        for (int it = 0; it <= p.itermax; ++it) {
            /* r = b - A x, lane = row, columns ascending; guarded by chunks of 8
             * columns (A and x are zero past n), so that a chunk's LDS reads
             * issue together */
            if (lane < m) {
                double r[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) r[q] = bq[q];
#pragma unroll
                for (int c8 = 0; c8 < NB; c8 += 8) {
                    if (c8 < n) {
#pragma unroll
                        for (int c = c8; c < c8 + 8; ++c) {
#pragma unroll
                            for (int q = 0; q < Q; ++q) r[q] = __builtin_fma(-sX[c * Q + q], arow[c], r[q]);
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) sR[lane * Q + q] = r[q];
            }
            __syncthreads();
            /* g = A^T r, lane = column, rows ascending */
            if (lane < n) {
                double g[Q];
#pragma unroll
                for (int q = 0; q < Q; ++q) g[q] = 0.0;
#pragma unroll 8
                for (int i = 0; i < m; ++i) {
                    const double a = sA[i * SA + lane];
#pragma unroll
                    for (int q = 0; q < Q; ++q) g[q] = __builtin_fma(a, sR[i * Q + q], g[q]);
                }
#pragma unroll
                for (int q = 0; q < Q; ++q) sV[lane * Q + q] = (float)g[q];
            }
            else if (lane < NB) {
#pragma unroll
                for (int q = 0; q < Q; ++q) sV[lane * Q + q] = 0.0f; /* the chunks below read past n */
            }
            __syncthreads();
            /* y = R^-T fl32(g): y_j = sum_{k <= j} R^-1(k, j) g_k, k ascending */
            float y[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) y[q] = 0.0f;
#pragma unroll
            for (int k8 = 0; k8 < NB; k8 += 8) {
                if (k8 < n) { /* R^-1 and g are zero past n */
#pragma unroll
                    for (int k = k8; k < k8 + 8; ++k) {
#pragma unroll
                        for (int q = 0; q < Q; ++q) y[q] = __builtin_fmaf(rcol[k], sV[k * Q + q], y[q]);
                    }
                }
            }
            __syncthreads(); /* every lane has read g */
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (lane < NB) sV[lane * Q + q] = y[q];
            __syncthreads();
            /* h = R^-1 y: h_i = sum_{k >= i} R^-1(i, k) y_k, k ascending */
            float h[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) h[q] = 0.0f;
#pragma unroll
            for (int k8 = 0; k8 < NB; k8 += 8) {
                if (k8 < n) { /* R^-1 and y are zero past n */
#pragma unroll
                    for (int k = k8; k < k8 + 8; ++k) {
#pragma unroll
                        for (int q = 0; q < Q; ++q) h[q] = __builtin_fmaf(rrow[k], sV[k * Q + q], h[q]);
                    }
                }
            }
            /* x += h on the open columns; |h|, |x| to LDS */
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const bool open = ((done_mask >> q) & 1ull) == 0;
                if (lane < n && open) xq[q] = xq[q] + (double)h[q];
                if (lane < NB) {
                    sD[lane * Q + q] = lane < n ? __builtin_fabsf(h[q]) : 0.0f;
                    sN[lane * Q + q] = lane < n ? __builtin_fabs(xq[q]) : 0.0;
                    sX[lane * Q + q] = xq[q];
                }
            }
            __syncthreads();
            /* lane q decides column q (NaN-propagating maxima) */
            if (lane < Q && !done) {
                double dn = 0.0, xn = 0.0;
                for (int k = 0; k < n; ++k) {
                    dn = fmax_nan(dn, (double)sD[k * Q + lane]);
                    xn = fmax_nan(xn, sN[k * Q + lane]);
                }
                steps = it > steps ? it : steps;
                if (dn <= xn * cte || (it > 0 && dn > 0.5 * prev && dn <= xn * floor_))
                    done = true;
                else if ((!(dn < prev) && it > 0) || it == p.itermax)
                    fail = true; /* not contracting, NaN, or out of steps */
                prev = dn;
            }
            done_mask = __builtin_amdgcn_ballot_w64(lane < Q && done);
            if (__builtin_amdgcn_ballot_w64(lane < Q && fail) != 0) break;
            if ((done_mask & ((1ull << Q) - 1ull)) == ((1ull << Q) - 1ull)) break;
        }
        ok = (done_mask & ((1ull << Q) - 1ull)) == ((1ull << Q) - 1ull);
        /* steps: the most any column took */
#pragma unroll
        for (int mask = 1; mask < Q; mask <<= 1) {
            const int o = __builtin_amdgcn_ds_bpermute(((lane ^ mask) & 63) << 2, steps);
            steps = o > steps ? o : steps;
        }
        if (ok) {
            double *X = p.x + d.b_off;
#pragma unroll
            for (int q = 0; q < Q; ++q)
                if (lane < n && q < nrhs) X[(long long)q * d.ldb + lane] = xq[q];
            if (lane == 0) {
                p.iter[d.idx] = steps;
                p.info[d.idx] = 0;
            }
            return;
        }
    }
fallback:
    /* the FP64 path: X := B (m rows), then the problem joins its kernel's list */
    {
        double *X = p.x + d.b_off;
        for (int q = 0; q < nrhs; ++q)
            if (lane < m) X[(long long)q * d.ldb + lane] = B[(long long)q * d.ldb + lane];
        if (lane == 0) {
            p.iter[d.idx] = -p.itermax - 1;
            const unsigned slot = atomicAdd(p.fb_fill + d.k64, 1u);
            BatchProblem e;
            e.m = m;
            e.n = n;
            e.nrhs = nrhs;
            e.lda = d.lda;
            e.ldb = d.ldb;
            e.idx = d.idx;
            e.a_off = d.a_off;
            e.b_off = d.b_off;
            e.tau_off = 0;
            p.fb[p.fb_begin[d.k64] + slot] = e;
        }
    }
}

template <int MB, int NB, int Q>
hipError_t launch_refine(const MixedRefineArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL((mixed_refine_kernel<MB, NB, Q>), dim3((unsigned)a.count), dim3(64), 0, s, a);
    return hipGetLastError();
}

} /* namespace mixedbatch */

int mixed_refine_bucket(int m, int n, int nrhs)
{
    const int mb = m <= 32 ? 0 : 1;
    const int nb = n <= 16 ? 0 : 1;
    const int qb = nrhs <= 4 ? 0 : 1;
    /* (32,16), (32,32), (64,16), (64,32), times Q in {4, 8} */
    return (mb * 2 + nb) * 2 + qb;
}

hipError_t launch_mixed_refine(int bucket, const MixedRefineArgs &a, hipStream_t s)
{
    using namespace mixedbatch;
    if (a.count <= 0) return hipSuccess;
    switch (bucket) {
    case 0: return launch_refine<32, 16, 4>(a, s);
    case 1: return launch_refine<32, 16, 8>(a, s);
    case 2: return launch_refine<32, 32, 4>(a, s);
    case 3: return launch_refine<32, 32, 8>(a, s);
    case 4: return launch_refine<64, 16, 4>(a, s);
    case 5: return launch_refine<64, 16, 8>(a, s);
    case 6: return launch_refine<64, 32, 4>(a, s);
    case 7: return launch_refine<64, 32, 8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_mixed_prep(const MixedPrepArgs &a, hipStream_t s)
{
    if (a.count <= 0) return hipSuccess;
    hipLaunchKernelGGL(mixedbatch::mixed_prep_kernel, dim3((unsigned)((a.count + 3) / 4)), dim3(256), 0,
                       s, a);
    return hipGetLastError();
}

} /* namespace cbk */
