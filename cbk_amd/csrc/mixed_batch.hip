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
 * solution x_0): per step
 *     r = b - A x          FP64, per row, fma over columns ascending
 *     g = A^T r            FP64, per column, fma over rows ascending
 *     R^T R h = fl32(g)    FP32: forward sweep right-looking (per element in
 *                          the order of the oracle's), backward sweep
 *                          right-looking from the last row
 *     x += h
 * until ||h||_inf <= ||x||_inf eps sqrt(n) (eps = 2^-53), or the correction
 * stopped halving once below ||x||_inf 2^-40; a correction that does not
 * contract, a NaN, a zero or non-finite R(k,k) or 30 steps send the problem
 * to FP64.
 *
 * Layout: A (m <= MB rows, n <= NB columns) is staged once in LDS, row-major
 * with a padded stride, so the residual (lane = row) and A^T r (lane =
 * column) both read it conflict-free with no cross-lane reductions; R sits in
 * LDS beside it; the solves run with lane = (row group, right-hand side),
 * the pivot broadcast by ds_bpermute. HBM traffic per problem: A, B and the
 * FP32 R and x_0 read once, X written once.
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

/* one wavefront per problem: convert (mode 0, MixedProblem list) or copy B's
 * max(m, n) x nrhs block into X (mode 1, BatchProblem list) */
__global__ __launch_bounds__(256) void mixed_prep_kernel(MixedPrepArgs p)
{
    const long k = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (k >= p.count) return;
    if (p.conv) {
        const MixedProblem d = p.conv[k];
        for (int c = 0; c < d.n; ++c)
            for (int i = lane; i < d.m; i += 64)
                p.a32[d.a_off + (long long)c * d.lda + i] = (float)p.a[d.a_off + (long long)c * d.lda + i];
        for (int q = 0; q < d.nrhs; ++q)
            for (int i = lane; i < d.m; i += 64)
                p.b32[d.b_off + (long long)q * d.ldb + i] = (float)p.b[d.b_off + (long long)q * d.ldb + i];
    }
    else {
        const BatchProblem d = p.copy[k];
        const int rows = d.m > d.n ? d.m : d.n;
        for (int q = 0; q < d.nrhs; ++q)
            for (int i = lane; i < rows; i += 64)
                p.x[d.b_off + (long long)q * d.ldb + i] = p.b[d.b_off + (long long)q * d.ldb + i];
        if (lane == 0) p.iter[d.idx] = -1;
    }
}

/* counters := their initial values (no memset node in the graph) */
__global__ void init_counters_kernel(unsigned *dst, const unsigned *init, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = init[i];
}

/* ---- refinement -------------------------------------------------------------- */

template <int MB, int NB, int Q>
__global__ __launch_bounds__(64) void mixed_refine_kernel(MixedRefineArgs p)
{
    constexpr int SA = NB + 1;      /* LDS row stride of A and R */
    constexpr int G = 64 / Q;       /* row groups of the solve layout */
    constexpr int T = (NB + G - 1) / G;
    __shared__ double sA[MB * SA];
    __shared__ double sR[MB * Q];   /* r(i, q) */
    __shared__ double sX[NB * Q];   /* x(c, q) */
    __shared__ float sF[NB * SA];   /* R(k, j), j >= k */
    __shared__ float sH[NB * Q];

    const long kq = blockIdx.x;
    if (kq >= p.count) return;
    const int lane = threadIdx.x;
    const MixedProblem d = p.desc[kq];
    const int m = d.m, n = d.n, nrhs = d.nrhs;
    const double *A = p.a + d.a_off;
    const double *B = p.b + d.b_off;
    const float *A32 = p.a32 + d.a_off;
    const float *B32 = p.b32 + d.b_off;

    /* stage A (FP64) and R (FP32) */
#pragma unroll 4
    for (int c = 0; c < n; ++c) {
        if (lane < m) sA[lane * SA + c] = A[(long long)c * d.lda + lane];
        if (lane <= c) sF[lane * SA + c] = A32[(long long)c * d.lda + lane];
    }
    /* b in the residual layout (lane = row) */
    double bq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q)
        bq[q] = (lane < m && q < nrhs) ? B[(long long)q * d.ldb + lane] : 0.0;
    /* x_0 = the FP32 solution, in the solve layout: lane (rg, q), rows rg + G t */
    const int qs = lane % Q, rg = lane / Q;
    const bool qlive = qs < nrhs;
    double xv[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int i = rg + G * t;
        xv[t] = (i < n && qlive) ? (double)B32[(long long)qs * d.ldb + i] : 0.0;
        if (i < NB) sX[i * Q + qs] = xv[t];
    }
    __syncthreads();
    bool ok = true;
    for (int k = 0; k < n; ++k) {
        const float r = sF[k * SA + k];
        ok = ok && r != 0.0f && __builtin_isfinite(r);
    }

    const double cte = 0x1p-53 * __builtin_sqrt((double)n), floor_ = 0x1p-40;
    double prev = __builtin_inf();
    bool done = !qlive; /* columns beyond nrhs take no part */
    bool fail = false;
    int steps = 0;
    for (int it = 0; ok && it <= p.itermax; ++it) {
        /* r = b - A x, lane = row */
        if (lane < m) {
            double r[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) r[q] = bq[q];
            for (int c = 0; c < n; ++c) {
                const double a = sA[lane * SA + c];
#pragma unroll
                for (int q = 0; q < Q; ++q) r[q] = __builtin_fma(-sX[c * Q + q], a, r[q]);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) sR[lane * Q + q] = r[q];
        }
        __syncthreads();
        /* g = A^T r, lane = column */
        if (lane < n) {
            double g[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) g[q] = 0.0;
            for (int i = 0; i < m; ++i) {
                const double a = sA[i * SA + lane];
#pragma unroll
                for (int q = 0; q < Q; ++q) g[q] = __builtin_fma(a, sR[i * Q + q], g[q]);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) sH[lane * Q + q] = (float)g[q];
        }
        __syncthreads();
        /* R^T R h = g in FP32, lane (rg, qs) holding rows rg + G t */
        float hv[T];
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = rg + G * t;
            hv[t] = i < n ? sH[i * Q + qs] : 0.0f;
        }
        /* forward, R^T y = g: y_k = y_k / R(k,k), then y_i -= R(k,i) y_k */
#pragma unroll
        for (int kt = 0; kt < T; ++kt) {
            for (int ko = 0; ko < G; ++ko) {
                const int k = ko + G * kt;
                if (k >= n) break;
                if (rg == ko) hv[kt] = hv[kt] / sF[k * SA + k];
                const float hk = shfl_f(hv[kt], ko * Q + qs);
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int i = rg + G * t;
                    if (i > k && i < n) hv[t] = __builtin_fmaf(-sF[k * SA + i], hk, hv[t]);
                }
            }
        }
        /* backward, R h = y: h_k = y_k / R(k,k), then y_i -= R(i,k) h_k, k descending */
#pragma unroll
        for (int kt = T - 1; kt >= 0; --kt) {
            for (int ko = G - 1; ko >= 0; --ko) {
                const int k = ko + G * kt;
                if (k >= n) continue;
                if (rg == ko) hv[kt] = hv[kt] / sF[k * SA + k];
                const float hk = shfl_f(hv[kt], ko * Q + qs);
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    const int i = rg + G * t;
                    if (i < k) hv[t] = __builtin_fmaf(-sF[i * SA + k], hk, hv[t]);
                }
            }
        }
        /* x += h for the live columns; ||h||_inf, ||x||_inf per column */
        double dn = 0.0, xn = 0.0;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = rg + G * t;
            if (i < n) {
                dn = fmax_nan(dn, (double)__builtin_fabsf(hv[t]));
                if (!done) xv[t] = xv[t] + (double)hv[t];
                xn = fmax_nan(xn, __builtin_fabs(xv[t]));
            }
        }
#pragma unroll
        for (int mask = Q; mask < 64; mask <<= 1) {
            dn = fmax_nan(dn, shfl_xor_d(dn, mask));
            xn = fmax_nan(xn, shfl_xor_d(xn, mask));
        }
        if (!done) {
            steps = it > steps ? it : steps;
            if (dn <= xn * cte || (it > 0 && dn > 0.5 * prev && dn <= xn * floor_))
                done = true;
            else if ((!(dn < prev) && it > 0) || it == p.itermax)
                fail = true; /* not contracting, NaN, or out of steps */
            prev = dn;
        }
        /* every column settled, or one failed: the problem is decided */
        if (__builtin_amdgcn_ballot_w64(fail) != 0) break;
        if (__builtin_amdgcn_ballot_w64(!done) == 0) break;
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = rg + G * t;
            if (i < NB) sX[i * Q + qs] = xv[t];
        }
        __syncthreads();
    }
    ok = ok && __builtin_amdgcn_ballot_w64(fail || !done) == 0;
    /* steps: the most any column took */
#pragma unroll
    for (int mask = 1; mask < 64; mask <<= 1) {
        const int o = __builtin_amdgcn_ds_bpermute(((lane ^ mask) & 63) << 2, steps);
        steps = o > steps ? o : steps;
    }
    double *X = p.x + d.b_off;
    if (ok) {
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const int i = rg + G * t;
            if (i < n && qlive) X[(long long)qs * d.ldb + i] = xv[t];
        }
        if (lane == 0) {
            p.iter[d.idx] = steps;
            p.info[d.idx] = 0;
        }
        return;
    }
    /* the FP64 path: X := B (m rows), then the problem joins its kernel's list */
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
    const int nb = n <= 16 ? 0 : n <= 32 ? 1 : 2;
    const int qb = nrhs <= 4 ? 0 : 1;
    /* (32,16), (32,32), (64,16), (64,32), (64,64), times Q in {4, 8} */
    const int shape = mb == 0 ? nb : 2 + nb;
    return shape * 2 + qb;
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
    case 8: return launch_refine<64, 64, 4>(a, s);
    case 9: return launch_refine<64, 64, 8>(a, s);
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

hipError_t launch_init_counters(unsigned *dst, const unsigned *init, int n, hipStream_t s)
{
    hipLaunchKernelGGL(mixedbatch::init_counters_kernel, dim3((n + 255) / 256), dim3(256), 0, s, dst,
                       init, n);
    return hipGetLastError();
}

} /* namespace cbk */
