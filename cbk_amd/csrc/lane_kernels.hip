/* lane_kernels.hip -- the general path: one GPU lane per matrix.
 *
 * Every routine of the reference, for every layout, uplo, side, trans, diag,
 * size and V, runs here one matrix per lane, working in place in HBM through
 * a strided lane view (the reference's BatchView, src/cbk_common.hpp:156-177,
 * narrowed to one slot). With V = 64 a wavefront covers one group and every
 * element access is one contiguous 512-byte transaction; with smaller V a
 * wavefront spans 64/V groups.
 *
 * The arithmetic is the reference's, operation for operation: the same
 * divides, square roots and multiplies, and an fma() wherever the reference's
 * compiler fuses a multiply-add (the library is built without contraction, so
 * nothing else fuses). Results are therefore bit-identical to the reference
 * kernels. The team kernels (sym_team_impl.h, qr_team_impl.h, mixed_team_impl.h) take the hot
 * shapes; this file is the fallback that covers everything.
 */

#include "cbk_internal.h"

#include <hip/hip_runtime.h>

namespace cbk {
namespace {

constexpr int kBlock = 256; /* 4 wavefronts */
constexpr int kNB = 8;      /* the reference's panel width, src/cbk_potrf_compact.hpp:71-74 */

/* element (i,j) of one lane: p[(i*si + j*sj)*V] */
template <class T> struct LaneView {
    T *p;
    int si, sj, V;
    __device__ __forceinline__ T &operator()(int i, int j) const
    {
        return p[(size_t)(i * si + j * sj) * V];
    }
    __device__ __forceinline__ LaneView tr() const { return {p, sj, si, V}; }
};

template <class T>
__device__ __forceinline__ LaneView<T> lane_view(const Operand<T> &o, long g, int v, int V)
{
    return {o.p + (size_t)g * o.gstride + v, o.row ? o.ld : 1, o.row ? 1 : o.ld, V};
}

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double sqrtT(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float sqrtT(float x) { return __builtin_sqrtf(x); }

/* ------------------------------------------------------------------ */
/* Cholesky / LDL^T, right-looking per pivot                          */
/* ------------------------------------------------------------------ */

/* The reference factors recursively (src/cbk_potrf_compact.hpp:112-124,
 * src/cbk_sytrfnp_compact.hpp:242-265) with leaf panels [8p, min(8p+8, n)).
 * Every element (i,c) then receives its updates in ascending pivot order k:
 * pivots of earlier panels in the trailing-update form
 *     A(i,c) = fma(-L(i,k), w, A(i,c)),  w = L(c,k) (potrf) or L(c,k)*D(k) (LDL^T)
 * (src/cbk_sytrfnp_compact.hpp:132-173) and pivots of its own panel in the
 * panel form on the unscaled column with the parked reciprocal
 *     A(i,c) = fma(-A(i,k), A(c,k)*(1/a_k), A(i,c))
 * (src/cbk_sytrfnp_compact.hpp:96-112). Applying each pivot right after it is
 * formed, to every later column, in the form its panel dictates, performs the
 * same operations on every element in the same order. */
template <class T, bool LDL>
__device__ void chol_lane(int n, const LaneView<T> &A, int *info)
{
    for (int j = 0; j < n; ++j) {
        const int pe = min((j / kNB) * kNB + kNB, n);
        const T a = A(j, j);
        if (info && *info == 0 && (LDL ? (a == T(0)) : !(a > T(0)))) *info = j + 1;
        const T inv = T(1) / a;
        /* panel form on the columns of this panel */
        for (int c = j + 1; c < pe; ++c) {
            const T w = A(c, j) * inv;
            for (int i = c; i < n; ++i)
                A(i, c) = fmaT(-A(i, j), w, A(i, c));
        }
        /* scale the pivot column (src/cbk_potrf_compact.hpp:98-102;
         * src/cbk_sytrfnp_compact.hpp:230-233) */
        T s;
        if (LDL) {
            s = inv;
        }
        else {
            const T d = sqrtT(a);
            A(j, j) = d;
            s = d * inv;
        }
        for (int i = j + 1; i < n; ++i)
            A(i, j) = A(i, j) * s;
        /* trailing form on the columns of later panels */
        for (int c = pe; c < n; ++c) {
            const T w = LDL ? A(c, j) * A(j, j) : A(c, j);
            for (int i = c; i < n; ++i)
                A(i, c) = fmaT(-A(i, j), w, A(i, c));
        }
    }
}

/* ------------------------------------------------------------------ */
/* triangular solves: src/cbk_trsm_compact.hpp                        */
/* ------------------------------------------------------------------ */

/* trsm_axpy_col (src/cbk_trsm_compact.hpp:49-67); a column-major A, b one
 * column, stride V between rows */
template <class T>
__device__ void trsm_axpy_col(bool upper, bool unit, int m, const T *a, int lda, T *b, int V,
                              T alpha)
{
    for (int i = 0; i < m; ++i)
        b[(size_t)i * V] = b[(size_t)i * V] * alpha;
    for (int t = 0; t < m; ++t) {
        const int kk = upper ? m - 1 - t : t;
        const T *ak = a + (size_t)kk * lda * V;
        const T xk = unit ? b[(size_t)kk * V] : b[(size_t)kk * V] / ak[(size_t)kk * V];
        if (!unit) b[(size_t)kk * V] = xk;
        const int lo = upper ? 0 : kk + 1, hi = upper ? kk : m;
        for (int i = lo; i < hi; ++i)
            b[(size_t)i * V] = fmaT(-ak[(size_t)i * V], xk, b[(size_t)i * V]);
    }
}

/* trsm_dot_block, one column (src/cbk_trsm_compact.hpp:73-104) */
template <class T>
__device__ void trsm_dot_col(bool upper, bool tran, bool unit, int m, const T *a, int lda,
                             T *b, int V, T alpha)
{
    const bool back = (upper != tran);
    for (int t = 0; t < m; ++t) {
        const int i = back ? m - 1 - t : t;
        T w = b[(size_t)i * V] * alpha;
        const int lo = back ? i + 1 : 0, hi = back ? m : i;
        for (int l = lo; l < hi; ++l) {
            const T av = tran ? a[(size_t)(i * lda + l) * V] : a[(size_t)(l * lda + i) * V];
            w = fmaT(-av, b[(size_t)l * V], w);
        }
        if (!unit) w = w / a[(size_t)(i * lda + i) * V];
        b[(size_t)i * V] = w;
    }
}

/* trsm_left_dot_tb (src/cbk_trsm_compact.hpp:110-145): the single leftover
 * column of the 4/2/1 blocking takes the axpy form when !tran */
template <class T>
__device__ void trsm_left_dot(bool upper, bool tran, bool unit, int m, int n, T alpha,
                              const T *a, int lda, T *b, int ldb, int V)
{
    const bool single_tail = (n % 4 == 1) || (n % 4 == 3);
    for (int j = 0; j < n; ++j) {
        T *bj = b + (size_t)j * ldb * V;
        if (single_tail && j == n - 1 && !tran)
            trsm_axpy_col(upper, unit, m, a, lda, bj, V, alpha);
        else
            trsm_dot_col(upper, tran, unit, m, a, lda, bj, V, alpha);
    }
}

/* trsm_compact_group_strided (src/cbk_trsm_compact.hpp:179-228) */
template <class T>
__device__ void trsm_strided(bool left, bool upper, bool tran, bool unit, int m, int n,
                             T alpha, const LaneView<T> &A, const LaneView<T> &B)
{
    if (left) {
        const bool back = (upper != tran);
        for (int j = 0; j < n; ++j)
            for (int t = 0; t < m; ++t) {
                const int i = back ? m - 1 - t : t;
                T w = B(i, j) * alpha;
                const int lo = back ? i + 1 : 0, hi = back ? m : i;
                for (int l = lo; l < hi; ++l)
                    w = fmaT(-(tran ? A(l, i) : A(i, l)), B(l, j), w);
                B(i, j) = unit ? w : w / A(i, i);
            }
    }
    else {
        const bool fwd = (upper != tran);
        for (int t = 0; t < n; ++t) {
            const int j = fwd ? t : n - 1 - t;
            for (int i = 0; i < m; ++i)
                B(i, j) = B(i, j) * alpha;
            const int lo = fwd ? 0 : j + 1, hi = fwd ? j : n;
            for (int l = lo; l < hi; ++l) {
                const T c = tran ? A(j, l) : A(l, j);
                for (int i = 0; i < m; ++i)
                    B(i, j) = fmaT(-c, B(i, l), B(i, j));
            }
            if (!unit) {
                const T d = A(j, j);
                for (int i = 0; i < m; ++i)
                    B(i, j) = B(i, j) / d;
            }
        }
    }
}

/* the view router of trsm_compact_group (src/cbk_trsm_compact.hpp:239-254) */
template <class T>
__device__ void trsm_group(bool left, bool upper, bool tran, bool unit, int m, int n, T alpha,
                           const LaneView<T> &A, const LaneView<T> &B)
{
    if (left && B.si == 1 && A.si == 1)
        trsm_left_dot(upper, tran, unit, m, n, alpha, A.p, A.sj, B.p, B.sj, A.V);
    else if (left && B.si == 1 && A.sj == 1)
        trsm_left_dot(!upper, !tran, unit, m, n, alpha, A.p, A.si, B.p, B.sj, A.V);
    else
        trsm_strided(left, upper, tran, unit, m, n, alpha, A, B);
}

/* potrs_compact_group / sytrsnp_compact_group (src/cbk_potrs_compact.hpp:46-58,
 * src/cbk_sytrsnp_compact.hpp:52-78) on the plain (not lower-) views */
template <class T, bool LDL>
__device__ void chol_solve_lane(bool upper, int n, int nrhs, const LaneView<T> &A,
                                const LaneView<T> &B)
{
    trsm_group<T>(true, upper, upper, LDL, n, nrhs, T(1), A, B);
    if (LDL) {
        for (int i = 0; i < n; ++i) {
            const T invd = T(1) / A(i, i);
            for (int j = 0; j < nrhs; ++j)
                B(i, j) = B(i, j) * invd;
        }
    }
    trsm_group<T>(true, upper, !upper, LDL, n, nrhs, T(1), A, B);
}

template <class T, bool LDL>
__global__ __launch_bounds__(kBlock) void chol_lane_kernel(bool factor, bool up, int n,
                                                           int nrhs, Operand<T> A,
                                                           Operand<T> B, int V, long nlanes,
                                                           int nm, int *info)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const LaneView<T> Av = lane_view(A, g, v, V);
    int lane_info = 0;
    const bool real = idx < nm;
    if (factor) chol_lane<T, LDL>(n, up ? Av.tr() : Av, (info && real) ? &lane_info : nullptr);
    if (info && real) info[idx] = lane_info;
    if (nrhs > 0 && lane_info == 0) chol_solve_lane<T, LDL>(up, n, nrhs, Av, lane_view(B, g, v, V));
}

template <class T>
__global__ __launch_bounds__(kBlock) void trsm_lane_kernel(bool left, bool up, bool tran,
                                                           bool unit, int m, int n, T alpha,
                                                           Operand<T> A, Operand<T> B, int V,
                                                           long nlanes)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    trsm_group<T>(left, up, tran, unit, m, n, alpha, lane_view(A, g, v, V), lane_view(B, g, v, V));
}

template <class T>
__global__ __launch_bounds__(kBlock) void zero_kernel(int rows, int cols, Operand<T> B, int V,
                                                      long nlanes)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const LaneView<T> Bv = lane_view(B, idx / V, (int)(idx % V), V);
    for (int j = 0; j < cols; ++j)
        for (int i = 0; i < rows; ++i)
            Bv(i, j) = T(0);
}

__global__ void clear_info_kernel(int *info, int nm)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < nm) info[idx] = 0;
}

/* ------------------------------------------------------------------ */
/* Householder QR: src/cbk_geqrf_compact.hpp, src/cbk_ormqr_compact.hpp */
/* ------------------------------------------------------------------ */

/* larf_block, slice by slice (src/cbk_ormqr_compact.hpp:68-105) */
template <class T>
__device__ void larf(int kk, int len, const LaneView<T> &A, T t, const LaneView<T> &C, int p0,
                     int p1)
{
    for (int p = p0; p < p1; ++p) {
        T w = C(kk, p);
        for (int i = kk + 1; i < len; ++i)
            w = fmaT(A(i, kk), C(i, p), w);
        w = w * t;
        C(kk, p) = C(kk, p) - w;
        for (int i = kk + 1; i < len; ++i)
            C(i, p) = fmaT(-A(i, kk), w, C(i, p));
    }
}

/* larfg_pack + geqrf_reflector (src/cbk_geqrf_compact.hpp:74-110) */
template <class T>
__device__ T geqrf_reflector(int kk, int m, const LaneView<T> &A, T *tau, int V)
{
    T tail = T(0);
    for (int i = kk + 1; i < m; ++i)
        tail = fmaT(A(i, kk), A(i, kk), tail);
    const T x0 = A(kk, kk);
    const T norm = sqrtT(fmaT(x0, x0, tail));
    const T beta = (x0 >= T(0)) ? -norm : norm;
    const bool has = tail > T(0);
    const T t = has ? (beta - x0) / beta : T(0);
    const T inv = has ? T(1) / (x0 - beta) : T(0);
    const T rdiag = has ? beta : x0;
    tau[(size_t)kk * V] = t;
    for (int i = kk + 1; i < m; ++i)
        A(i, kk) = A(i, kk) * inv;
    A(kk, kk) = rdiag;
    return t;
}

/* geqrf_compact_group (src/cbk_geqrf_compact.hpp:113-131) */
template <class T> __device__ void geqrf_lane(int m, int n, const LaneView<T> &A, T *tau, int V)
{
    const int k = m < n ? m : n;
    for (int kk = 0; kk < k; ++kk) {
        const T t = geqrf_reflector(kk, m, A, tau, V);
        larf(kk, m, A, t, A, kk + 1, n);
    }
}

/* ormqr_compact_group (src/cbk_ormqr_compact.hpp:109-124) */
template <class T>
__device__ void ormqr_lane(bool forward, int len, int npanel, int k, const LaneView<T> &A,
                           const T *tau, const LaneView<T> &C, int V)
{
    for (int s = 0; s < k; ++s) {
        const int kk = forward ? s : k - 1 - s;
        larf(kk, len, A, tau[(size_t)kk * V], C, 0, npanel);
    }
}

template <class T>
__global__ __launch_bounds__(kBlock) void geqrf_lane_kernel(int m, int n, Operand<T> A, T *tau,
                                                            int V, long nlanes)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const int k = m < n ? m : n;
    geqrf_lane<T>(m, n, lane_view(A, g, v, V), tau + (size_t)g * k * V + v, V);
}

/* ormqr_compact (src/cbk_ormqr_compact.hpp:131-156): dir = (left == tran) ?
 * Forward : Backward; side R sees C^T */
template <class T>
__global__ __launch_bounds__(kBlock) void ormqr_lane_kernel(bool left, bool tran, int m, int n,
                                                            int k, Operand<T> A, const T *tau,
                                                            Operand<T> C, int V, long nlanes)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    LaneView<T> Cv = lane_view(C, g, v, V);
    if (!left) Cv = Cv.tr();
    ormqr_lane<T>(left == tran, left ? m : n, left ? n : m, k, lane_view(A, g, v, V),
                  tau + (size_t)g * k * V + v, Cv, V);
}

/* orgqr_compact_group (src/cbk_orgqr_compact.hpp:51-85) */
template <class T>
__global__ __launch_bounds__(kBlock) void orgqr_lane_kernel(int m, int n, int k, Operand<T> A,
                                                            const T *tau_, int V, long nlanes)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const LaneView<T> Av = lane_view(A, g, v, V);
    const T *tau = tau_ + (size_t)g * k * V + v;
    for (int j = k; j < n; ++j) {
        for (int i = 0; i < m; ++i)
            Av(i, j) = T(0);
        Av(j, j) = T(1);
    }
    for (int kk = k - 1; kk >= 0; --kk) {
        const T t = tau[(size_t)kk * V];
        larf(kk, m, Av, t, Av, kk + 1, n);
        const T mt = -t;
        for (int i = kk + 1; i < m; ++i)
            Av(i, kk) = Av(i, kk) * mt;
        Av(kk, kk) = T(1) - t;
        for (int i = 0; i < kk; ++i)
            Av(i, kk) = T(0);
    }
}

/* gels_compact_group (src/cbk_gels_compact.hpp:84-106, 113-146) on one lane's
 * views; with info, a zero R(i,i) stops the lane where LAPACK dgels's dtrtrs
 * would. Returns that info code. */
template <class T>
__device__ int gels_one(bool tran, int m, int n, int nrhs, const LaneView<T> &Av,
                        const LaneView<T> &Bv, T *tau, int V, bool with_info)
{
    const bool tall = (m >= n);
    const int p = tall ? m : n, q = tall ? n : m;
    const bool overdet = (tall != tran);
    const LaneView<T> F = tall ? Av : Av.tr();
    geqrf_lane<T>(p, q, F, tau, V);
    int bad = 0;
    if (with_info) {
        for (int i = 0; i < q && !bad; ++i)
            if (F(i, i) == T(0)) bad = i + 1;
    }
    if (overdet) {
        ormqr_lane<T>(true, p, nrhs, q, F, tau, Bv, V);
        if (bad) return bad;
        trsm_group<T>(true, true, false, false, q, nrhs, T(1), F, Bv);
    }
    else {
        if (bad) return bad;
        trsm_group<T>(true, true, true, false, q, nrhs, T(1), F, Bv);
        for (int j = 0; j < nrhs; ++j)
            for (int i = q; i < p; ++i)
                Bv(i, j) = T(0);
        ormqr_lane<T>(false, p, nrhs, q, F, tau, Bv, V);
    }
    return bad;
}

template <class T>
__global__ __launch_bounds__(kBlock) void gels_lane_kernel(bool tran, int m, int n, int nrhs,
                                                           Operand<T> A, Operand<T> B, T *tau_,
                                                           int V, long nlanes, int nm,
                                                           int *info)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const int q = m < n ? m : n;
    const bool real = info && idx < nm;
    const int bad = gels_one<T>(tran, m, n, nrhs, lane_view(A, g, v, V), lane_view(B, g, v, V),
                                tau_ + (size_t)g * q * V + v, V, real);
    if (real) info[idx] = bad;
}

/* one problem of an uneven batch per lane, column-major with V = 1: the
 * general case of gels_batch (transposed or underdetermined problems, sizes
 * the register kernels do not hold, empty problems) */
template <class T>
__global__ __launch_bounds__(kBlock) void gels_batch_lane_kernel(BatchArgs<T> p)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (long)batch_count(p)) return;
    const BatchProblem d = p.desc[k];
    const LaneView<T> Av{p.a + d.a_off, 1, d.lda, 1};
    const LaneView<T> Bv{p.b + d.b_off, 1, d.ldb, 1};
    const int mx = d.m > d.n ? d.m : d.n, q = d.m < d.n ? d.m : d.n;
    int bad = 0;
    if (d.nrhs > 0 && mx > 0) {
        if (q == 0) {
            /* gels_compact with q = 0: B := 0 (src/cbk_gels_compact.hpp:127-131) */
            for (int j = 0; j < d.nrhs; ++j)
                for (int i = 0; i < mx; ++i)
                    Bv(i, j) = T(0);
        }
        else {
            bad = gels_one<T>(p.tran, d.m, d.n, d.nrhs, Av, Bv, p.tau + d.tau_off, 1, p.info != nullptr);
        }
    }
    if (p.info) p.info[d.idx] = bad;
}

/* ------------------------------------------------------------------ */
/* mixed precision posv (LAPACK dsposv), one lane per matrix          */
/* ------------------------------------------------------------------ */

/* The general path of dsposv_compact: the FP32 factor and the FP32 right-hand
 * side / correction live in the caller's workspace (compact, column-major,
 * ld = n, the lane's slot), A and B are read through their views, X is
 * written. Same algorithm and codes as the team kernel (mixed_team_impl.h):
 * refinement until ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n) for every
 * column, FP64 fallback in place otherwise. LDL: the unpivoted LDL^T of
 * sysvnp in both precisions (the FP32 factor fails on a zero or non-finite
 * pivot). */
template <bool LDL>
__global__ __launch_bounds__(kBlock) void mixed_posv_lane_kernel(
    bool up, int n, int nrhs, Operand<double> A, Operand<double> B, Operand<double> X,
    float *work, int V, long nlanes, int nm, int itermax, int *iter_out, int *info_out)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nlanes) return;
    const long g = idx / V;
    const int v = (int)(idx % V);
    const LaneView<double> Av0 = lane_view(A, g, v, V);
    const LaneView<double> Al = up ? Av0.tr() : Av0; /* lower view */
    const LaneView<double> Bv = lane_view(B, g, v, V), Xv = lane_view(X, g, v, V);
    float *wg = work + (size_t)g * V * (size_t)(n * n + n * nrhs) + v;
    const LaneView<float> S{wg, 1, n, V};                         /* FP32 factor */
    const LaneView<float> W{wg + (size_t)n * n * V, 1, n, V};     /* FP32 rhs */
    auto sym = [&](int i, int c) { return i >= c ? Al(i, c) : Al(c, i); };
    constexpr double kFltMax = 3.4028234663852886e38;

    double anrm = 0.0;
    bool over = false;
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) {
            const double x = __builtin_fabs(sym(i, c));
            s += x;
            over |= x > kFltMax;
        }
        anrm = fmax(anrm, s);
        for (int q = 0; q < nrhs; ++q) over |= __builtin_fabs(Bv(i, q)) > kFltMax;
    }
    const double cte = anrm * 0x1p-53 * __builtin_sqrt((double)n);
    int it_code = 0;
    bool done = false;
    if (over) {
        it_code = -2;
    }
    else {
        for (int c = 0; c < n; ++c)
            for (int i = c; i < n; ++i)
                S(i, c) = (float)Al(i, c);
        int finfo = 0;
        chol_lane<float, LDL>(n, S, &finfo);
        if (LDL) {
            for (int j = 0; j < n && finfo == 0; ++j)
                if (!__builtin_isfinite(S(j, j))) finfo = j + 1;
        }
        if (finfo != 0) {
            it_code = -3;
        }
        else {
            for (int q = 0; q < nrhs; ++q)
                for (int i = 0; i < n; ++i)
                    W(i, q) = (float)Bv(i, q);
            chol_solve_lane<float, LDL>(false, n, nrhs, S, W);
            for (int q = 0; q < nrhs; ++q)
                for (int i = 0; i < n; ++i)
                    Xv(i, q) = (double)W(i, q);
            for (int it = 0;; ++it) {
                bool conv = true, rover = false;
                for (int q = 0; q < nrhs; ++q) {
                    double rn = 0.0, xn = 0.0;
                    for (int i = 0; i < n; ++i) {
                        double r = Bv(i, q);
                        for (int c = 0; c < n; ++c)
                            r = fmaT(-sym(i, c), Xv(c, q), r);
                        const double ar = __builtin_fabs(r), ax = __builtin_fabs(Xv(i, q));
                        rn = (ar > rn || ar != ar) ? ar : rn;
                        xn = (ax > xn || ax != ax) ? ax : xn;
                        rover |= ar > kFltMax;
                        W(i, q) = (float)r;
                    }
                    if (!(rn <= xn * cte)) conv = false;
                }
                if (conv) {
                    it_code = it;
                    done = true;
                    break;
                }
                if (it == itermax) {
                    it_code = -itermax - 1;
                    break;
                }
                if (rover) {
                    it_code = -2;
                    break;
                }
                chol_solve_lane<float, LDL>(false, n, nrhs, S, W);
                for (int q = 0; q < nrhs; ++q)
                    for (int i = 0; i < n; ++i)
                        Xv(i, q) = Xv(i, q) + (double)W(i, q);
            }
        }
    }
    int dinfo = 0;
    if (!done) {
        chol_lane<double, LDL>(n, Al, &dinfo);
        if (dinfo == 0) {
            for (int q = 0; q < nrhs; ++q)
                for (int i = 0; i < n; ++i)
                    Xv(i, q) = Bv(i, q);
            chol_solve_lane<double, LDL>(up, n, nrhs, Av0, Xv);
        }
    }
    if (idx < nm) {
        iter_out[idx] = it_code;
        info_out[idx] = dinfo;
    }
}

inline long nlanes_of(int V, int nm) { return (long)((nm + V - 1) / V) * V; }
inline unsigned nblocks(long nlanes) { return (unsigned)((nlanes + kBlock - 1) / kBlock); }

} /* namespace */

template <class T>
hipError_t launch_chol_lane(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                            Operand<T> B, int V, int nm, int *info)
{
    const long nl = nlanes_of(V, nm);
    if (ldl)
        hipLaunchKernelGGL((chol_lane_kernel<T, true>), dim3(nblocks(nl)), dim3(kBlock), 0,
                           current_stream(), factor, up, n, nrhs, A, B, V, nl, nm, info);
    else
        hipLaunchKernelGGL((chol_lane_kernel<T, false>), dim3(nblocks(nl)), dim3(kBlock), 0,
                           current_stream(), factor, up, n, nrhs, A, B, V, nl, nm, info);
    return hipGetLastError();
}

template <class T>
hipError_t launch_trsm_lane(bool left, bool up, bool tran, bool unit, int m, int n, T alpha,
                            Operand<T> A, Operand<T> B, int V, int nm)
{
    const long nl = nlanes_of(V, nm);
    hipLaunchKernelGGL((trsm_lane_kernel<T>), dim3(nblocks(nl)), dim3(kBlock), 0,
                       current_stream(), left, up, tran, unit, m, n, alpha, A, B, V, nl);
    return hipGetLastError();
}

template <class T> hipError_t launch_zero(int rows, int cols, Operand<T> B, int V, int nm)
{
    const long nl = nlanes_of(V, nm);
    hipLaunchKernelGGL((zero_kernel<T>), dim3(nblocks(nl)), dim3(kBlock), 0, current_stream(),
                       rows, cols, B, V, nl);
    return hipGetLastError();
}

hipError_t launch_clear_info(int *info, int nm)
{
    if (nm <= 0 || !info) return hipSuccess;
    hipLaunchKernelGGL(clear_info_kernel, dim3((nm + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       current_stream(), info, nm);
    return hipGetLastError();
}

template <class T>
hipError_t launch_geqrf_lane(int m, int n, Operand<T> A, T *tau, int V, int nm)
{
    const long nl = nlanes_of(V, nm);
    hipLaunchKernelGGL((geqrf_lane_kernel<T>), dim3(nblocks(nl)), dim3(kBlock), 0,
                       current_stream(), m, n, A, tau, V, nl);
    return hipGetLastError();
}

template <class T>
hipError_t launch_ormqr_lane(bool left, bool tran, int m, int n, int k, Operand<T> A,
                             const T *tau, Operand<T> C, int V, int nm)
{
    const long nl = nlanes_of(V, nm);
    hipLaunchKernelGGL((ormqr_lane_kernel<T>), dim3(nblocks(nl)), dim3(kBlock), 0,
                       current_stream(), left, tran, m, n, k, A, tau, C, V, nl);
    return hipGetLastError();
}

template <class T>
hipError_t launch_orgqr_lane(int m, int n, int k, Operand<T> A, const T *tau, int V, int nm)
{
    const long nl = nlanes_of(V, nm);
    hipLaunchKernelGGL((orgqr_lane_kernel<T>), dim3(nblocks(nl)), dim3(kBlock), 0,
                       current_stream(), m, n, k, A, tau, V, nl);
    return hipGetLastError();
}

template <class T> hipError_t launch_gels_batch_lane(const BatchArgs<T> &args, hipStream_t s)
{
    if (args.count <= 0) return hipSuccess;
    hipLaunchKernelGGL((gels_batch_lane_kernel<T>), dim3(nblocks(args.count)), dim3(kBlock), 0, s,
                       args);
    return hipGetLastError();
}

__global__ void clear_u32_kernel(unsigned *p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void fill_int_kernel(int *p, int n, int value)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

hipError_t launch_fill_int(int *p, int n, int value)
{
    if (n <= 0 || !p) return hipSuccess;
    hipLaunchKernelGGL(fill_int_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       current_stream(), p, n, value);
    return hipGetLastError();
}

hipError_t launch_clear_u32(unsigned *p, int n, hipStream_t s)
{
    hipLaunchKernelGGL(clear_u32_kernel, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, p, n);
    return hipGetLastError();
}

template <class T>
hipError_t launch_gels_lane(bool tran, int m, int n, int nrhs, Operand<T> A, Operand<T> B,
                            T *tau, int V, int nm, int *info)
{
    const long nl = nlanes_of(V, nm);
    hipLaunchKernelGGL((gels_lane_kernel<T>), dim3(nblocks(nl)), dim3(kBlock), 0,
                       current_stream(), tran, m, n, nrhs, A, B, tau, V, nl, nm, info);
    return hipGetLastError();
}

hipError_t launch_mixed_posv_lane(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, float *work, int V,
                                  int nm, int itermax, int *iter, int *info)
{
    const long nl = nlanes_of(V, nm);
    if (ldl)
        hipLaunchKernelGGL(mixed_posv_lane_kernel<true>, dim3(nblocks(nl)), dim3(kBlock), 0,
                           current_stream(), up, n, nrhs, A, B, X, work, V, nl, nm, itermax, iter,
                           info);
    else
        hipLaunchKernelGGL(mixed_posv_lane_kernel<false>, dim3(nblocks(nl)), dim3(kBlock), 0,
                           current_stream(), up, n, nrhs, A, B, X, work, V, nl, nm, itermax, iter,
                           info);
    return hipGetLastError();
}

#define CBK_INSTANTIATE_LANE(T)                                                          \
    template hipError_t launch_chol_lane<T>(bool, bool, bool, int, int, Operand<T>,       \
                                            Operand<T>, int, int, int *);                 \
    template hipError_t launch_trsm_lane<T>(bool, bool, bool, bool, int, int, T,          \
                                            Operand<T>, Operand<T>, int, int);            \
    template hipError_t launch_zero<T>(int, int, Operand<T>, int, int);                   \
    template hipError_t launch_geqrf_lane<T>(int, int, Operand<T>, T *, int, int);        \
    template hipError_t launch_ormqr_lane<T>(bool, bool, int, int, int, Operand<T>,       \
                                             const T *, Operand<T>, int, int);            \
    template hipError_t launch_orgqr_lane<T>(int, int, int, Operand<T>, const T *, int,   \
                                             int);                                        \
    template hipError_t launch_gels_batch_lane<T>(const BatchArgs<T> &, hipStream_t);    \
    template hipError_t launch_gels_lane<T>(bool, int, int, int, Operand<T>, Operand<T>,  \
                                            T *, int, int, int *);

CBK_INSTANTIATE_LANE(double)
CBK_INSTANTIATE_LANE(float)

} /* namespace cbk */
