/* cbk_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A CPU restatement of the reference compact kernels, one lane (one matrix) at
 * a time, in the reference's exact operation order. Only tests/, the smoke
 * check and bench.py's CPU-baseline leg load it; the shipped HIP library never
 * does. Every function cites the reference routine it restates (paths relative
 * to the reference repository).
 *
 * The reference evaluates V lanes at once with GNU vector types; lanes never
 * interact (src/cbk_common.hpp:61-97), so a per-lane scalar loop performs the
 * same IEEE operations on every element. Where the reference's compiler, at
 * -ffp-contract=fast with an FMA-capable -march, fuses `x -= a*b` / `x += a*b`
 * into one rounding, this file calls fma(); it is itself compiled with
 * contraction off so its bits do not depend on the flags.
 */

#include "cbk_oracle.h"

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace {

/* ---------------------------------------------------------------- */
/* argument helpers: src/cbk.cpp:46-59                              */
/* ---------------------------------------------------------------- */

inline bool opt(char c, char upper) { return c == upper || c == upper + ('a' - 'A'); }
inline int max1(int x) { return x < 1 ? 1 : x; }
/* widened from {2,4,8,16} (src/cbk.cpp:56-59) to the GPU library's set */
inline bool vlen_ok(int V)
{
    return V == 2 || V == 4 || V == 8 || V == 16 || V == 32 || V == 64;
}

/* element (i,j) of one lane: data[(i*si + j*sj)*V] (src/cbk_common.hpp:156-177) */
template <class T> struct LV {
    T *p;
    long si, sj;
    long V;
    T &operator()(long i, long j) const { return p[(i * si + j * sj) * V]; }
    LV tr() const { return {p, sj, si, V}; }
};

/* make_view (src/cbk_common.hpp:182-193): column-major si=1, sj=ld */
template <class T> LV<T> make_view(T *p, bool row, long ld, long V)
{
    return {p, row ? ld : 1, row ? 1 : ld, V};
}

/* group_stride (src/cbk_common.hpp:210-214) */
inline size_t group_stride(bool row, long ld, long rows, long cols, long V)
{
    return (size_t)ld * (size_t)(row ? rows : cols) * (size_t)V;
}

/* every lane of every (padded) group, as for_each_group (src/cbk_common.hpp:277-297) */
template <class F> void for_each_lane(int nm, int V, F &&f)
{
    const long ng = (nm + V - 1) / V;
    const long nl = ng * V;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < nl; ++idx)
        f(idx / V, idx % V);
}

template <class T> inline T fmaT(T a, T b, T c) { return std::fma(a, b, c); }

/* ---------------------------------------------------------------- */
/* LDL^T / Cholesky machinery: src/cbk_sytrfnp_compact.hpp          */
/* ---------------------------------------------------------------- */

constexpr int NB = 8; /* potrf_nb / sytrfnp_nb: src/cbk_potrf_compact.hpp:71-74 */

/* sytrfnp_update_block, one column c (src/cbk_sytrfnp_compact.hpp:96-112):
 * A(i,c) -= A(i,j) * (A(c,j) * A(j,j)) for i >= c, the weight parked at A(j,j) */
template <class T> void update_col(long j, long n, const LV<T> &A, long c)
{
    const T w = A(c, j) * A(j, j);
    for (long i = c; i < n; ++i)
        A(i, c) = fmaT(-A(i, j), w, A(i, c));
}

/* sytrfnp_trailing_update<WithD> (src/cbk_sytrfnp_compact.hpp:132-209): every
 * element of columns k1..jend-1, rows c..n-1, accumulates -A(i,k)*w(c,k) over
 * k = k0..k1-1 in order, w = A(c,k)*A(k,k) with D, A(c,k) without. The tile
 * shapes only group independent elements, so the per-element order is this. */
template <bool WithD, class T>
void trailing_update(long n, long jend, const LV<T> &A, long k0, long k1)
{
    for (long c = k1; c < jend; ++c)
        for (long i = c; i < n; ++i) {
            T acc = A(i, c);
            for (long k = k0; k < k1; ++k) {
                const T w = WithD ? A(c, k) * A(k, k) : A(c, k);
                acc = fmaT(-A(i, k), w, acc);
            }
            A(i, c) = acc;
        }
}

/* potrf_compact_panel (src/cbk_potrf_compact.hpp:79-105); info: first pivot
 * a <= 0 or NaN (LAPACK dpotf2's test), recorded without changing the bits */
template <class T> void potrf_panel(long j0, long j1, long n, const LV<T> &A, int *info)
{
    for (long j = j0; j < j1; ++j) {
        const T a = A(j, j);
        if (info && *info == 0 && !(a > T(0))) *info = (int)j + 1;
        const T inva = T(1) / a;
        A(j, j) = inva;
        for (long jj = j + 1; jj < j1; ++jj)
            update_col(j, n, A, jj);
        const T d = std::sqrt(a);
        A(j, j) = d;
        const T invd = d * inva;
        for (long i = j + 1; i < n; ++i)
            A(i, j) = A(i, j) * invd;
    }
}

/* potrf_compact_block (src/cbk_potrf_compact.hpp:112-124) */
template <class T> void potrf_block(long j0, long j1, long n, const LV<T> &A, int *info)
{
    if (j1 - j0 <= NB) {
        potrf_panel(j0, j1, n, A, info);
        return;
    }
    const long npanels = (j1 - j0 + NB - 1) / NB;
    const long mid = j0 + (npanels / 2) * NB;
    potrf_block(j0, mid, n, A, info);
    trailing_update<false>(n, j1, A, j0, mid);
    potrf_block(mid, j1, n, A, info);
}

/* sytrfnp_compact_panel (src/cbk_sytrfnp_compact.hpp:216-235); info: first d == 0 */
template <class T> void sytrfnp_panel(long j0, long j1, long n, const LV<T> &A, int *info)
{
    for (long j = j0; j < j1; ++j) {
        const T d = A(j, j);
        if (info && *info == 0 && d == T(0)) *info = (int)j + 1;
        const T invd = T(1) / d;
        A(j, j) = invd;
        for (long jj = j + 1; jj < j1; ++jj)
            update_col(j, n, A, jj);
        A(j, j) = d;
        for (long i = j + 1; i < n; ++i)
            A(i, j) = A(i, j) * invd;
    }
}

/* sytrfnp_compact_block (src/cbk_sytrfnp_compact.hpp:242-265) */
template <class T> void sytrfnp_block(long j0, long j1, long n, const LV<T> &A, int *info)
{
    if (j1 - j0 <= NB) {
        sytrfnp_panel(j0, j1, n, A, info);
        return;
    }
    const long npanels = (j1 - j0 + NB - 1) / NB;
    const long mid = j0 + (npanels / 2) * NB;
    sytrfnp_block(j0, mid, n, A, info);
    trailing_update<true>(n, j1, A, j0, mid);
    sytrfnp_block(mid, j1, n, A, info);
}

/* ---------------------------------------------------------------- */
/* trsm: src/cbk_trsm_compact.hpp                                   */
/* ---------------------------------------------------------------- */

/* trsm_axpy_col (src/cbk_trsm_compact.hpp:49-67): one column, A column-major
 * (element (i,k) at a[(k*lda + i)*V]), b the column (element i at b[i*V]) */
template <class T>
void trsm_axpy_col(bool upper, bool unit, long m, const T *a, long lda, T *b, long V,
                   T alpha)
{
    for (long i = 0; i < m; ++i)
        b[i * V] = b[i * V] * alpha;
    for (long t = 0; t < m; ++t) {
        const long kk = upper ? m - 1 - t : t;
        const T *ak = a + kk * lda * V;
        const T xk = unit ? b[kk * V] : b[kk * V] / ak[kk * V];
        if (!unit) b[kk * V] = xk;
        const long lo = upper ? 0 : kk + 1, hi = upper ? kk : m;
        for (long i = lo; i < hi; ++i)
            b[i * V] = fmaT(-ak[i * V], xk, b[i * V]);
    }
}

/* trsm_dot_block for one column (src/cbk_trsm_compact.hpp:73-104); JB columns
 * of a block are independent, so the per-column sequence is the same */
template <class T>
void trsm_dot_col(bool upper, bool tran, bool unit, long m, const T *a, long lda, T *b,
                  long V, T alpha)
{
    const bool back = (upper != tran);
    for (long t = 0; t < m; ++t) {
        const long i = back ? m - 1 - t : t;
        T w = b[i * V] * alpha;
        const long lo = back ? i + 1 : 0, hi = back ? m : i;
        for (long l = lo; l < hi; ++l) {
            const T av = tran ? a[(i * lda + l) * V] : a[(l * lda + i) * V];
            w = fmaT(-av, b[l * V], w);
        }
        if (!unit) w = w / a[(i * lda + i) * V];
        b[i * V] = w;
    }
}

/* trsm_left_dot_tb (src/cbk_trsm_compact.hpp:110-145): RHS blocks of 4, then 2,
 * then a single leftover column, which takes the axpy form when !tran */
template <class T>
void trsm_left_dot(bool upper, bool tran, bool unit, long m, long n, T alpha, const T *a,
                   long lda, T *b, long ldb, long V)
{
    const bool single_tail = (n % 4 == 1) || (n % 4 == 3);
    for (long j = 0; j < n; ++j) {
        T *bj = b + j * ldb * V;
        if (single_tail && j == n - 1 && !tran)
            trsm_axpy_col(upper, unit, m, a, lda, bj, V, alpha);
        else
            trsm_dot_col(upper, tran, unit, m, a, lda, bj, V, alpha);
    }
}

/* trsm_compact_group_strided (src/cbk_trsm_compact.hpp:179-228) */
template <class T>
void trsm_strided(bool left, bool upper, bool tran, bool unit, long m, long n, T alpha,
                  const LV<T> &A, const LV<T> &B)
{
    if (left) {
        const bool back = (upper != tran);
        for (long j = 0; j < n; ++j)
            for (long t = 0; t < m; ++t) {
                const long i = back ? m - 1 - t : t;
                T w = B(i, j) * alpha;
                const long lo = back ? i + 1 : 0, hi = back ? m : i;
                for (long l = lo; l < hi; ++l)
                    w = fmaT(-(tran ? A(l, i) : A(i, l)), B(l, j), w);
                B(i, j) = unit ? w : w / A(i, i);
            }
    }
    else {
        const bool fwd = (upper != tran);
        for (long t = 0; t < n; ++t) {
            const long j = fwd ? t : n - 1 - t;
            for (long i = 0; i < m; ++i)
                B(i, j) = B(i, j) * alpha;
            const long lo = fwd ? 0 : j + 1, hi = fwd ? j : n;
            for (long l = lo; l < hi; ++l) {
                const T c = tran ? A(j, l) : A(l, j);
                for (long i = 0; i < m; ++i)
                    B(i, j) = fmaT(-c, B(i, l), B(i, j));
            }
            if (!unit) {
                const T d = A(j, j);
                for (long i = 0; i < m; ++i)
                    B(i, j) = B(i, j) / d;
            }
        }
    }
}

/* trsm_compact_group, the view router (src/cbk_trsm_compact.hpp:239-254) */
template <class T>
void trsm_group(bool left, bool upper, bool tran, bool unit, long m, long n, T alpha,
                const LV<T> &A, const LV<T> &B)
{
    if (left && B.si == 1 && A.si == 1)
        trsm_left_dot(upper, tran, unit, m, n, alpha, A.p, A.sj, B.p, B.sj, A.V);
    else if (left && B.si == 1 && A.sj == 1)
        trsm_left_dot(!upper, !tran, unit, m, n, alpha, A.p, A.si, B.p, B.sj, A.V);
    else
        trsm_strided(left, upper, tran, unit, m, n, alpha, A, B);
}

/* zero_block (src/cbk_common.hpp:301-307) */
template <class T> void zero_block(const LV<T> &B, long i0, long i1, long ncols)
{
    for (long j = 0; j < ncols; ++j)
        for (long i = i0; i < i1; ++i)
            B(i, j) = T(0);
}

/* potrs_compact_group (src/cbk_potrs_compact.hpp:46-58) */
template <class T>
void potrs_lane(bool row, bool upper, long n, long nrhs, T *a, long lda, T *b, long ldb,
                long V)
{
    const LV<T> A = make_view(a, row, lda, V), B = make_view(b, row, ldb, V);
    trsm_group(true, upper, upper, false, n, nrhs, T(1), A, B);
    trsm_group(true, upper, !upper, false, n, nrhs, T(1), A, B);
}

/* sytrsnp_compact_group (src/cbk_sytrsnp_compact.hpp:52-78) */
template <class T>
void sytrsnp_lane(bool row, bool upper, long n, long nrhs, T *a, long lda, T *b, long ldb,
                  long V)
{
    const LV<T> A = make_view(a, row, lda, V), B = make_view(b, row, ldb, V);
    trsm_group(true, upper, upper, true, n, nrhs, T(1), A, B);
    for (long i = 0; i < n; ++i) {
        const T invd = T(1) / A(i, i);
        for (long j = 0; j < nrhs; ++j)
            B(i, j) = B(i, j) * invd;
    }
    trsm_group(true, upper, !upper, true, n, nrhs, T(1), A, B);
}

/* ---------------------------------------------------------------- */
/* Householder QR: src/cbk_geqrf_compact.hpp, src/cbk_ormqr_compact.hpp */
/* ---------------------------------------------------------------- */

/* larf_block for one slice (src/cbk_ormqr_compact.hpp:68-91); slices of a
 * block are independent */
template <class T> void larf(long kk, long len, const LV<T> &A, T t, const LV<T> &C, long p0, long p1)
{
    for (long p = p0; p < p1; ++p) {
        T w = C(kk, p);
        for (long i = kk + 1; i < len; ++i)
            w = fmaT(A(i, kk), C(i, p), w);
        w = w * t;
        C(kk, p) = C(kk, p) - w;
        for (long i = kk + 1; i < len; ++i)
            C(i, p) = fmaT(-A(i, kk), w, C(i, p));
    }
}

/* larfg_pack + geqrf_reflector (src/cbk_geqrf_compact.hpp:74-110) */
template <class T> T geqrf_reflector(long kk, long m, const LV<T> &A, T *tau, long V)
{
    T tail = T(0);
    for (long i = kk + 1; i < m; ++i)
        tail = fmaT(A(i, kk), A(i, kk), tail);
    const T x0 = A(kk, kk);
    const T norm = std::sqrt(fmaT(x0, x0, tail));
    const T beta = (x0 >= T(0)) ? -norm : norm;
    const bool has = tail > T(0);
    const T t = has ? (beta - x0) / beta : T(0);
    const T inv = has ? T(1) / (x0 - beta) : T(0);
    const T rdiag = has ? beta : x0;
    tau[kk * V] = t;
    for (long i = kk + 1; i < m; ++i)
        A(i, kk) = A(i, kk) * inv;
    A(kk, kk) = rdiag;
    return t;
}

/* geqrf_compact_group (src/cbk_geqrf_compact.hpp:113-131) */
template <class T> void geqrf_lane(long m, long n, const LV<T> &A, T *tau, long V)
{
    const long k = m < n ? m : n;
    for (long kk = 0; kk < k; ++kk) {
        const T t = geqrf_reflector(kk, m, A, tau, V);
        larf(kk, m, A, t, A, kk + 1, n);
    }
}

/* ormqr_compact_group (src/cbk_ormqr_compact.hpp:109-124) */
template <class T>
void ormqr_lane(bool forward, long len, long npanel, long k, const LV<T> &A, const T *tau,
                const LV<T> &C, long V)
{
    for (long s = 0; s < k; ++s) {
        const long kk = forward ? s : k - 1 - s;
        larf(kk, len, A, tau[kk * V], C, 0, npanel);
    }
}

/* orgqr_compact_group (src/cbk_orgqr_compact.hpp:51-85) */
template <class T> void orgqr_lane(long m, long n, long k, const LV<T> &A, const T *tau, long V)
{
    for (long j = k; j < n; ++j) {
        for (long i = 0; i < m; ++i)
            A(i, j) = T(0);
        A(j, j) = T(1);
    }
    for (long kk = k - 1; kk >= 0; --kk) {
        const T t = tau[kk * V];
        larf(kk, m, A, t, A, kk + 1, n);
        const T mt = -t;
        for (long i = kk + 1; i < m; ++i)
            A(i, kk) = A(i, kk) * mt;
        A(kk, kk) = T(1) - t;
        for (long i = 0; i < kk; ++i)
            A(i, kk) = T(0);
    }
}

/* ---------------------------------------------------------------- */
/* the entry points: validation as src/cbk.cpp:61-320               */
/* ---------------------------------------------------------------- */

template <class T> int geqrf(char layout, int m, int n, T *ap, int ldap, T *taup, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (m < 0) return -2;
    if (n < 0) return -3;
    if (ldap < max1(row ? n : m)) return -5;
    if (!vlen_ok(V)) return -7;
    if (nm < 0) return -8;
    if (m == 0 || n == 0 || nm == 0) return 0;
    const long k = m < n ? m : n;
    const size_t sa = group_stride(row, ldap, m, n, V), st = (size_t)k * V;
    for_each_lane(nm, V, [&](long g, long v) {
        geqrf_lane<T>(m, n, make_view(ap + g * sa + v, row, ldap, V), taup + g * st + v, V);
    });
    return 0;
}

template <class T>
int ormqr(char layout, char side, char trans, int m, int n, int k, const T *ap, int ldap,
          const T *taup, T *cp, int ldcp, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool left = opt(side, 'L'), right = opt(side, 'R');
    const int nq = left ? m : n;
    if (!col && !row) return -1;
    if (!left && !right) return -2;
    if (!opt(trans, 'T') && !opt(trans, 'N')) return -3;
    if (m < 0) return -4;
    if (n < 0) return -5;
    if (k < 0 || k > nq) return -6;
    if (ldap < max1(row ? k : nq)) return -8;
    if (ldcp < max1(row ? n : m)) return -11;
    if (!vlen_ok(V)) return -12;
    if (nm < 0) return -13;
    if (m == 0 || n == 0 || k == 0 || nm == 0) return 0;
    /* src/cbk_ormqr_compact.hpp:131-156 */
    const bool tran = opt(trans, 'T');
    const bool forward = (left == tran);
    const long len = left ? m : n, npanel = left ? n : m;
    const size_t sa = group_stride(row, ldap, len, k, V), st = (size_t)k * V,
                 sc = group_stride(row, ldcp, m, n, V);
    for_each_lane(nm, V, [&](long g, long v) {
        LV<T> C = make_view(cp + g * sc + v, row, ldcp, V);
        if (!left) C = C.tr();
        const LV<T> A = make_view(const_cast<T *>(ap) + g * sa + v, row, ldap, V);
        ormqr_lane<T>(forward, len, npanel, k, A, taup + g * st + v, C, V);
    });
    return 0;
}

template <class T>
int orgqr(char layout, int m, int n, int k, T *ap, int ldap, const T *taup, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (m < 0) return -2;
    if (n < 0 || n > m) return -3;
    if (k < 0 || k > n) return -4;
    if (ldap < max1(row ? n : m)) return -6;
    if (!vlen_ok(V)) return -8;
    if (nm < 0) return -9;
    if (m == 0 || n == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, m, n, V), st = (size_t)k * V;
    for_each_lane(nm, V, [&](long g, long v) {
        orgqr_lane<T>(m, n, k, make_view(ap + g * sa + v, row, ldap, V), taup + g * st + v, V);
    });
    return 0;
}

template <class T>
int potrf(char layout, char uplo, int n, T *ap, int ldap, int V, int nm, int *info)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool lo = opt(uplo, 'L'), up = opt(uplo, 'U');
    if (!col && !row) return -1;
    if (!lo && !up) return -2;
    if (n < 0) return -3;
    if (ldap < max1(n)) return -5;
    if (!vlen_ok(V)) return -6;
    if (nm < 0) return -7;
    if (info)
        for (int i = 0; i < nm; ++i)
            info[i] = 0;
    if (n == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V);
    for_each_lane(nm, V, [&](long g, long v) {
        LV<T> A = make_view(ap + g * sa + v, row, ldap, V);
        if (up) A = A.tr(); /* make_lower_view, src/cbk_common.hpp:201-206 */
        int lane_info = 0;
        potrf_block<T>(0, n, n, A, &lane_info);
        if (info && g * V + v < nm) info[g * V + v] = lane_info;
    });
    return 0;
}

template <class T>
int sytrfnp(char layout, char uplo, int n, T *ap, int ldap, int V, int nm, int *info)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool lo = opt(uplo, 'L'), up = opt(uplo, 'U');
    if (!col && !row) return -1;
    if (!lo && !up) return -2;
    if (n < 0) return -3;
    if (ldap < max1(n)) return -5;
    if (!vlen_ok(V)) return -6;
    if (nm < 0) return -7;
    if (info)
        for (int i = 0; i < nm; ++i)
            info[i] = 0;
    if (n == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V);
    for_each_lane(nm, V, [&](long g, long v) {
        LV<T> A = make_view(ap + g * sa + v, row, ldap, V);
        if (up) A = A.tr();
        int lane_info = 0;
        sytrfnp_block<T>(0, n, n, A, &lane_info);
        if (info && g * V + v < nm) info[g * V + v] = lane_info;
    });
    return 0;
}

/* sym_solve_args (src/cbk.cpp:173-188) */
inline int sym_solve_args(char layout, char uplo, int n, int nrhs, int ldap, int ldbp,
                          int V, int nm, bool &row, bool &up)
{
    const bool col = opt(layout, 'C'), lo = opt(uplo, 'L');
    row = opt(layout, 'R');
    up = opt(uplo, 'U');
    if (!col && !row) return -1;
    if (!lo && !up) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldap < max1(n)) return -6;
    if (ldbp < max1(row ? nrhs : n)) return -8;
    if (!vlen_ok(V)) return -9;
    if (nm < 0) return -10;
    return 0;
}

template <class T>
int potrs(char layout, char uplo, int n, int nrhs, const T *ap, int ldap, T *bp, int ldbp,
          int V, int nm)
{
    bool row, up;
    if (int e = sym_solve_args(layout, uplo, n, nrhs, ldap, ldbp, V, nm, row, up)) return e;
    if (n == 0 || nrhs == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V), sb = group_stride(row, ldbp, n, nrhs, V);
    for_each_lane(nm, V, [&](long g, long v) {
        potrs_lane<T>(row, up, n, nrhs, const_cast<T *>(ap) + g * sa + v, ldap,
                      bp + g * sb + v, ldbp, V);
    });
    return 0;
}

/* posv_compact (src/cbk_posv_compact.hpp:37-64); with info, a failed lane keeps B
 * (LAPACK dposv skips dpotrs) */
template <class T>
int posv(char layout, char uplo, int n, int nrhs, T *ap, int ldap, T *bp, int ldbp, int V,
         int nm, int *info)
{
    bool row, up;
    if (int e = sym_solve_args(layout, uplo, n, nrhs, ldap, ldbp, V, nm, row, up)) return e;
    if (info)
        for (int i = 0; i < nm; ++i)
            info[i] = 0;
    if (n == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V), sb = group_stride(row, ldbp, n, nrhs, V);
    for_each_lane(nm, V, [&](long g, long v) {
        T *a = ap + g * sa + v;
        LV<T> A = make_view(a, row, ldap, V);
        if (up) A = A.tr();
        int lane_info = 0;
        potrf_block<T>(0, n, n, A, &lane_info);
        const bool real = g * V + v < nm;
        if (info && real) info[g * V + v] = lane_info;
        if (nrhs > 0 && !(info && real && lane_info != 0))
            potrs_lane<T>(row, up, n, nrhs, a, ldap, bp + g * sb + v, ldbp, V);
    });
    return 0;
}

template <class T>
int sytrsnp(char layout, char uplo, int n, int nrhs, const T *ap, int ldap, T *bp,
            int ldbp, int V, int nm)
{
    bool row, up;
    if (int e = sym_solve_args(layout, uplo, n, nrhs, ldap, ldbp, V, nm, row, up)) return e;
    if (n == 0 || nrhs == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V), sb = group_stride(row, ldbp, n, nrhs, V);
    for_each_lane(nm, V, [&](long g, long v) {
        sytrsnp_lane<T>(row, up, n, nrhs, const_cast<T *>(ap) + g * sa + v, ldap,
                        bp + g * sb + v, ldbp, V);
    });
    return 0;
}

/* sysvnp_compact (src/cbk_sysvnp_compact.hpp:37-64) */
template <class T>
int sysvnp(char layout, char uplo, int n, int nrhs, T *ap, int ldap, T *bp, int ldbp,
           int V, int nm, int *info)
{
    bool row, up;
    if (int e = sym_solve_args(layout, uplo, n, nrhs, ldap, ldbp, V, nm, row, up)) return e;
    if (info)
        for (int i = 0; i < nm; ++i)
            info[i] = 0;
    if (n == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V), sb = group_stride(row, ldbp, n, nrhs, V);
    for_each_lane(nm, V, [&](long g, long v) {
        T *a = ap + g * sa + v;
        LV<T> A = make_view(a, row, ldap, V);
        if (up) A = A.tr();
        int lane_info = 0;
        sytrfnp_block<T>(0, n, n, A, &lane_info);
        const bool real = g * V + v < nm;
        if (info && real) info[g * V + v] = lane_info;
        if (nrhs > 0 && !(info && real && lane_info != 0))
            sytrsnp_lane<T>(row, up, n, nrhs, a, ldap, bp + g * sb + v, ldbp, V);
    });
    return 0;
}

template <class T>
int trsm(char layout, char side, char uplo, char transa, char diag, int m, int n, T alpha,
         const T *ap, int ldap, T *bp, int ldbp, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool left = opt(side, 'L'), right = opt(side, 'R');
    const bool up = opt(uplo, 'U'), lo = opt(uplo, 'L');
    const bool tran = opt(transa, 'T') || opt(transa, 'C');
    const bool unit = opt(diag, 'U');
    const int s = left ? m : n;
    if (!col && !row) return -1;
    if (!left && !right) return -2;
    if (!up && !lo) return -3;
    if (!tran && !opt(transa, 'N')) return -4;
    if (!unit && !opt(diag, 'N')) return -5;
    if (m < 0) return -6;
    if (n < 0) return -7;
    if (ldap < max1(s)) return -10;
    if (ldbp < max1(row ? n : m)) return -12;
    if (!vlen_ok(V)) return -13;
    if (nm < 0) return -14;
    if (m == 0 || n == 0 || nm == 0) return 0;
    /* src/cbk_trsm_compact.hpp:271-299 */
    const size_t sa = group_stride(row, ldap, s, s, V), sb = group_stride(row, ldbp, m, n, V);
    if (alpha == T(0)) {
        for_each_lane(nm, V, [&](long g, long v) {
            zero_block(make_view(bp + g * sb + v, row, ldbp, V), 0, m, n);
        });
        return 0;
    }
    for_each_lane(nm, V, [&](long g, long v) {
        trsm_group<T>(left, up, tran, unit, m, n, alpha,
                      make_view(const_cast<T *>(ap) + g * sa + v, row, ldap, V),
                      make_view(bp + g * sb + v, row, ldbp, V));
    });
    return 0;
}

/* gels_compact_group (src/cbk_gels_compact.hpp:84-106); with info (LAPACK dgels
 * via dtrtrs): least squares stops after Q^T B, minimum norm before touching B */
template <class T>
void gels_lane(bool overdet, long p, long q, long nrhs, const LV<T> &F, T *tau,
               const LV<T> &B, long V, int *info)
{
    geqrf_lane<T>(p, q, F, tau, V);
    int bad = 0;
    if (info) {
        for (long i = 0; i < q && !bad; ++i)
            if (F(i, i) == T(0)) bad = (int)i + 1;
        *info = bad;
    }
    if (overdet) {
        ormqr_lane<T>(true, p, nrhs, q, F, tau, B, V);
        if (bad) return;
        trsm_group<T>(true, true, false, false, q, nrhs, T(1), F, B);
    }
    else {
        if (bad) return;
        trsm_group<T>(true, true, true, false, q, nrhs, T(1), F, B);
        zero_block(B, q, p, nrhs);
        ormqr_lane<T>(false, p, nrhs, q, F, tau, B, V);
    }
}

template <class T>
int gels(char layout, char trans, int m, int n, int nrhs, T *ap, int ldap, T *bp, int ldbp,
         T *taup, int V, int nm, int *info)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const int mx = m > n ? m : n;
    if (!col && !row) return -1;
    if (!opt(trans, 'N') && !opt(trans, 'T') && !opt(trans, 'C')) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (nrhs < 0) return -5;
    if (ldap < max1(row ? n : m)) return -7;
    if (ldbp < max1(row ? nrhs : mx)) return -9;
    if (!vlen_ok(V)) return -11;
    if (nm < 0) return -12;
    if (info)
        for (int i = 0; i < nm; ++i)
            info[i] = 0;
    if (nrhs == 0 || nm == 0 || mx == 0) return 0;
    /* src/cbk_gels_compact.hpp:113-146 */
    const bool tran = opt(trans, 'T') || opt(trans, 'C');
    const bool tall = (m >= n);
    const long p = tall ? m : n, q = tall ? n : m;
    const bool overdet = (tall != tran);
    const size_t sa = group_stride(row, ldap, m, n, V), sb = group_stride(row, ldbp, p, nrhs, V),
                 st = (size_t)q * V;
    if (q == 0) {
        for_each_lane(nm, V, [&](long g, long v) {
            zero_block(make_view(bp + g * sb + v, row, ldbp, V), 0, p, nrhs);
        });
        return 0;
    }
    for_each_lane(nm, V, [&](long g, long v) {
        const LV<T> A = make_view(ap + g * sa + v, row, ldap, V);
        int lane_info = 0;
        const bool real = g * V + v < nm;
        gels_lane<T>(overdet, p, q, nrhs, tall ? A : A.tr(), taup + g * st + v,
                     make_view(bp + g * sb + v, row, ldbp, V), V,
                     (info && real) ? &lane_info : nullptr);
        if (info && real) info[g * V + v] = lane_info;
    });
    return 0;
}

/* ---------------------------------------------------------------- */
/* LAPACK dsposv, per matrix (no reference counterpart)             */
/* ---------------------------------------------------------------- */

/* r := b - A x for the symmetric A stored in the lower triangle of view L
 * (element (i,j), i >= j), summed over j ascending with one fma per term */
void sym_residual(long n, long nrhs, const LV<double> &A, const LV<double> &X,
                  const LV<double> &B, std::vector<double> &R)
{
    for (long c = 0; c < nrhs; ++c)
        for (long i = 0; i < n; ++i) {
            double r = B(i, c);
            for (long j = 0; j < n; ++j) {
                const double a = (i >= j) ? A(i, j) : A(j, i);
                r = std::fma(-a, X(j, c), r);
            }
            R[c * n + i] = r;
        }
}

} /* namespace */

extern "C" {

int oracle_get_version(void) { return 200; }

#define ORACLE_DEFS(T, p)                                                                \
    int oracle_##p##geqrf_compact(char layout, int m, int n, T *ap, int ldap, T *taup,   \
                                  int V, int nm)                                         \
    {                                                                                    \
        return geqrf<T>(layout, m, n, ap, ldap, taup, V, nm);                            \
    }                                                                                    \
    int oracle_##p##ormqr_compact(char layout, char side, char trans, int m, int n,      \
                                  int k, const T *ap, int ldap, const T *taup, T *cp,    \
                                  int ldcp, int V, int nm)                               \
    {                                                                                    \
        return ormqr<T>(layout, side, trans, m, n, k, ap, ldap, taup, cp, ldcp, V, nm);  \
    }                                                                                    \
    int oracle_##p##orgqr_compact(char layout, int m, int n, int k, T *ap, int ldap,     \
                                  const T *taup, int V, int nm)                          \
    {                                                                                    \
        return orgqr<T>(layout, m, n, k, ap, ldap, taup, V, nm);                         \
    }                                                                                    \
    int oracle_##p##potrf_compact(char layout, char uplo, int n, T *ap, int ldap, int V, \
                                  int nm)                                                \
    {                                                                                    \
        return potrf<T>(layout, uplo, n, ap, ldap, V, nm, nullptr);                      \
    }                                                                                    \
    int oracle_##p##potrs_compact(char layout, char uplo, int n, int nrhs, const T *ap,  \
                                  int ldap, T *bp, int ldbp, int V, int nm)              \
    {                                                                                    \
        return potrs<T>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm);               \
    }                                                                                    \
    int oracle_##p##posv_compact(char layout, char uplo, int n, int nrhs, T *ap,         \
                                 int ldap, T *bp, int ldbp, int V, int nm)               \
    {                                                                                    \
        return posv<T>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm, nullptr);       \
    }                                                                                    \
    int oracle_##p##sytrfnp_compact(char layout, char uplo, int n, T *ap, int ldap,      \
                                    int V, int nm)                                       \
    {                                                                                    \
        return sytrfnp<T>(layout, uplo, n, ap, ldap, V, nm, nullptr);                    \
    }                                                                                    \
    int oracle_##p##sytrsnp_compact(char layout, char uplo, int n, int nrhs,             \
                                    const T *ap, int ldap, T *bp, int ldbp, int V,       \
                                    int nm)                                              \
    {                                                                                    \
        return sytrsnp<T>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm);             \
    }                                                                                    \
    int oracle_##p##sysvnp_compact(char layout, char uplo, int n, int nrhs, T *ap,       \
                                   int ldap, T *bp, int ldbp, int V, int nm)             \
    {                                                                                    \
        return sysvnp<T>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm, nullptr);     \
    }                                                                                    \
    int oracle_##p##trsm_compact(char layout, char side, char uplo, char transa,         \
                                 char diag, int m, int n, T alpha, const T *ap,          \
                                 int ldap, T *bp, int ldbp, int V, int nm)               \
    {                                                                                    \
        return trsm<T>(layout, side, uplo, transa, diag, m, n, alpha, ap, ldap, bp,      \
                       ldbp, V, nm);                                                     \
    }                                                                                    \
    int oracle_##p##gels_compact(char layout, char trans, int m, int n, int nrhs, T *ap, \
                                 int ldap, T *bp, int ldbp, T *taup, int V, int nm)      \
    {                                                                                    \
        return gels<T>(layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, taup, V, nm,       \
                       nullptr);                                                         \
    }                                                                                    \
    int oracle_##p##potrf_compact_info(char layout, char uplo, int n, T *ap, int ldap,   \
                                       int V, int nm, int *info)                         \
    {                                                                                    \
        return potrf<T>(layout, uplo, n, ap, ldap, V, nm, info);                         \
    }                                                                                    \
    int oracle_##p##posv_compact_info(char layout, char uplo, int n, int nrhs, T *ap,    \
                                      int ldap, T *bp, int ldbp, int V, int nm,          \
                                      int *info)                                         \
    {                                                                                    \
        return posv<T>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm, info);          \
    }                                                                                    \
    int oracle_##p##sytrfnp_compact_info(char layout, char uplo, int n, T *ap, int ldap, \
                                         int V, int nm, int *info)                       \
    {                                                                                    \
        return sytrfnp<T>(layout, uplo, n, ap, ldap, V, nm, info);                       \
    }                                                                                    \
    int oracle_##p##sysvnp_compact_info(char layout, char uplo, int n, int nrhs, T *ap,  \
                                        int ldap, T *bp, int ldbp, int V, int nm,        \
                                        int *info)                                       \
    {                                                                                    \
        return sysvnp<T>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm, info);        \
    }                                                                                    \
    int oracle_##p##gels_compact_info(char layout, char trans, int m, int n, int nrhs,   \
                                      T *ap, int ldap, T *bp, int ldbp, T *taup, int V,  \
                                      int nm, int *info)                                 \
    {                                                                                    \
        return gels<T>(layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, taup, V, nm, info);\
    }

ORACLE_DEFS(double, d)
ORACLE_DEFS(float, s)
#undef ORACLE_DEFS

} /* extern "C" */

namespace {
/* the reference's gels on one column-major matrix: the compact view with one
 * slot (V = 1), tau on the stack; the quick returns of gels<T> per problem */
template <class T>
int gels_batch(char trans, int nprob, const oracle_gels_problem *desc, T *ap, T *bp, int *info)
{
    if (!opt(trans, 'N') && !opt(trans, 'T') && !opt(trans, 'C')) return -1;
    if (nprob < 0) return -2;
    const bool tran = !opt(trans, 'N');
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < nprob; ++k) {
        const oracle_gels_problem &d = desc[k];
        const long mx = d.m > d.n ? d.m : d.n;
        const bool tall = d.m >= d.n;
        const long p = tall ? d.m : d.n, q = tall ? d.n : d.m;
        const LV<T> A = make_view(ap + d.a_off, false, d.lda, 1);
        const LV<T> B = make_view(bp + d.b_off, false, d.ldb, 1);
        int bad = 0;
        if (d.nrhs > 0 && mx > 0) {
            if (q == 0) {
                zero_block(B, 0, p, d.nrhs);
            }
            else {
                std::vector<T> tau(q);
                gels_lane<T>(tall != tran, p, q, d.nrhs, tall ? A : A.tr(), tau.data(), B, 1,
                             info ? &bad : nullptr);
            }
        }
        if (info) info[k] = bad;
    }
    return 0;
}
/* the NaN-propagating max of the convergence tests */
inline double nanmax(double a, double b) { return (a > b || a != a) ? a : b; }

/* One column of the mixed least-squares refinement (dsgels_compact,
 * dsgels_batch): x holds x_0 on entry (n values), the refined solution on
 * exit; F's upper triangle is the single precision R. Per step
 *     r = b - A x,  g = A^T r              (double, rows ascending)
 *     R^T y = fl32(g) right-looking, R h = y by row dots   (float)
 *     x += h
 * until ||h||_inf <= ||x||_inf 2^-53 sqrt(n), or the correction stopped
 * halving once below ||x||_inf 2^-40 (true); a correction that does not
 * contract (a NaN included) or itermax steps give false. steps: the largest
 * step index reached. */
bool refine_column(long m, long n, const LV<double> &A, const LV<double> &B, long q,
                   const LV<float> &F, std::vector<double> &x, int itermax, int &steps)
{
    const double cte = std::ldexp(1.0, -53) * std::sqrt((double)n), floor_ = std::ldexp(1.0, -40);
    std::vector<double> gr(n);
    std::vector<float> h(n);
    double prev = INFINITY;
    for (int it = 0; it <= itermax; ++it) {
        for (long c = 0; c < n; ++c)
            gr[c] = 0.0;
        for (long i = 0; i < m; ++i) {
            double r = B(i, q);
            for (long c = 0; c < n; ++c)
                r = std::fma(-x[c], A(i, c), r);
            for (long c = 0; c < n; ++c)
                gr[c] = std::fma(A(i, c), r, gr[c]);
        }
        for (long c = 0; c < n; ++c)
            h[c] = (float)gr[c];
        for (long k = 0; k < n; ++k) {
            h[k] = h[k] / F(k, k);
            for (long i = k + 1; i < n; ++i)
                h[i] = std::fma(-F(k, i), h[k], h[i]);
        }
        for (long i = n - 1; i >= 0; --i) {
            float s = h[i];
            for (long k = i + 1; k < n; ++k)
                s = std::fma(-F(i, k), h[k], s);
            h[i] = s / F(i, i);
        }
        double dn = 0.0, xn = 0.0;
        for (long i = 0; i < n; ++i) {
            dn = nanmax(dn, std::fabs((double)h[i]));
            x[i] = x[i] + (double)h[i];
        }
        for (long i = 0; i < n; ++i)
            xn = nanmax(xn, std::fabs(x[i]));
        steps = it > steps ? it : steps;
        if (dn <= xn * cte || (it > 0 && dn > 0.5 * prev && dn <= xn * floor_)) return true;
        if ((!(dn < prev) && it > 0) || it == itermax) return false;
        prev = dn;
    }
    return false;
}

/* R^-1 of the upper triangle of F (n x n, float), column j by back
 * substitution: z_j = 1 / R(j,j); for i = j-1 down to 0,
 * z_i = -(sum_{k=i+1..j} R(i,k) z_k) / R(i,i), k ascending. Column-major,
 * zeros below the diagonal. */
std::vector<float> tri_inverse(long n, const LV<float> &F)
{
    std::vector<float> Ri((size_t)n * n, 0.0f);
    for (long j = 0; j < n; ++j) {
        float *z = &Ri[(size_t)j * n];
        z[j] = 1.0f / F(j, j);
        for (long i = j - 1; i >= 0; --i) {
            float s = 0.0f;
            for (long k = i + 1; k <= j; ++k)
                s = std::fma(F(i, k), z[k], s);
            z[i] = -s / F(i, i);
        }
    }
    return Ri;
}

/* refine_column with the preconditioner applied through an explicit
 * inverse: y = R^-T fl32(g) (y_j = sum_{k<=j} Ri(k,j) g_k, k ascending), then
 * h = R^-1 y (h_i = sum_{k>=i} Ri(i,k) y_k, k ascending); the rest as
 * refine_column. The library's dsgels_batch refinement. */
bool refine_column_inv(long m, long n, const LV<double> &A, const LV<double> &B, long q,
                       const std::vector<float> &Ri, std::vector<double> &x, int itermax,
                       int &steps)
{
    const double cte = std::ldexp(1.0, -53) * std::sqrt((double)n), floor_ = std::ldexp(1.0, -40);
    std::vector<double> gr(n);
    std::vector<float> g(n), y(n);
    double prev = INFINITY;
    for (int it = 0; it <= itermax; ++it) {
        for (long c = 0; c < n; ++c)
            gr[c] = 0.0;
        for (long i = 0; i < m; ++i) {
            double r = B(i, q);
            for (long c = 0; c < n; ++c)
                r = std::fma(-x[c], A(i, c), r);
            for (long c = 0; c < n; ++c)
                gr[c] = std::fma(A(i, c), r, gr[c]);
        }
        for (long c = 0; c < n; ++c)
            g[c] = (float)gr[c];
        for (long j = 0; j < n; ++j) {
            float s = 0.0f;
            for (long k = 0; k <= j; ++k)
                s = std::fma(Ri[(size_t)j * n + k], g[k], s);
            y[j] = s;
        }
        double dn = 0.0, xn = 0.0;
        for (long i = 0; i < n; ++i) {
            float s = 0.0f;
            for (long k = i; k < n; ++k)
                s = std::fma(Ri[(size_t)k * n + i], y[k], s);
            dn = nanmax(dn, std::fabs((double)s));
            x[i] = x[i] + (double)s;
        }
        for (long i = 0; i < n; ++i)
            xn = nanmax(xn, std::fabs(x[i]));
        steps = it > steps ? it : steps;
        if (dn <= xn * cte || (it > 0 && dn > 0.5 * prev && dn <= xn * floor_)) return true;
        if ((!(dn < prev) && it > 0) || it == itermax) return false;
        prev = dn;
    }
    return false;
}

} /* namespace */

extern "C" {

/* Mixed-precision least squares (no reference counterpart): the single
 * precision Householder QR (this file's geqrf_lane, the reference's geqr2 in
 * float) preconditions refinement on the normal equations with double
 * residuals:  x = 0;  r = b - A x, g = A^T r (double);  R^T R d = g (float
 * sweeps);  x += d (double); per column until ||d||_inf <= ||x||_inf *
 * 2^-53 * sqrt(n), or the correction stops halving below ||x||_inf * 2^-40;
 * a zero or non-finite R(i,i) of the float factor, a correction that does not
 * contract, or 30 steps send the matrix to the double gels (A then holds its
 * QR, X the first n rows of the solution; info as gels_compact_info). iter:
 * refinement steps (max over the columns), -1 no single precision path for
 * the shape (n > 32, or V < 4), -31 the double path ran. Shapes: trans 'N',
 * n <= m <= 64, n + nrhs <= 64. */
int oracle_dsgels_compact(char layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
                          const double *bp, int ldbp, double *xp, int ldxp, int *iter,
                          int *info, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (!opt(trans, 'N')) return -2;
    if (m < 0 || m > 64) return -3;
    if (n < 0 || n > m) return -4;
    if (nrhs < 0 || n + nrhs > 64) return -5;
    if (ldap < max1(row ? n : m)) return -7;
    if (ldbp < max1(row ? nrhs : m)) return -9;
    if (ldxp < max1(row ? nrhs : n)) return -11;
    if (!vlen_ok(V)) return -14;
    if (nm < 0) return -15;
    for (int i = 0; i < nm; ++i) {
        iter[i] = 0;
        if (info) info[i] = 0;
    }
    if (n == 0 || nrhs == 0 || nm == 0) return 0;
    const bool mixed = V >= 4 && n >= 1 && n <= 32;
    const size_t sa = group_stride(row, ldap, m, n, V), sb = group_stride(row, ldbp, m, nrhs, V),
                 sx = group_stride(row, ldxp, n, nrhs, V);
    const int itermax = 30;
    for_each_lane(nm, V, [&](long g, long v) {
        const long idx = g * V + v;
        if (idx >= nm) return;
        const LV<double> A = make_view(ap + g * sa + v, row, ldap, V);
        const LV<double> B = make_view(const_cast<double *>(bp) + g * sb + v, row, ldbp, V);
        const LV<double> X = make_view(xp + g * sx + v, row, ldxp, V);
        bool ok = mixed;
        int steps = 0;
        if (mixed) {
            std::vector<float> F((size_t)m * n), tau(n);
            const LV<float> Fv{F.data(), 1, m, 1};
            for (long j = 0; j < n; ++j)
                for (long i = 0; i < m; ++i)
                    F[j * m + i] = (float)A(i, j);
            geqrf_lane<float>(m, n, Fv, tau.data(), 1);
            for (long i = 0; i < n && ok; ++i)
                ok = Fv(i, i) != 0.f && std::isfinite(Fv(i, i));
            for (long q = 0; q < nrhs && ok; ++q) {
                std::vector<double> x(n, 0.0);
                ok = refine_column(m, n, A, B, q, Fv, x, itermax, steps);
                if (!ok) break;
                for (long i = 0; i < n; ++i)
                    X(i, q) = x[i];
            }
        }
        if (ok) {
            iter[idx] = steps;
            return;
        }
        iter[idx] = mixed ? -itermax - 1 : -1;
        /* the double gels on one-slot copies of A and B (the trsm kernels take
         * one slot stride for both operands), A's factor copied back */
        std::vector<double> D((size_t)m * n), W((size_t)m * nrhs), tau(n);
        const LV<double> Dv{D.data(), 1, m, 1}, Wv{W.data(), 1, m, 1};
        for (long j = 0; j < n; ++j)
            for (long i = 0; i < m; ++i)
                D[j * m + i] = A(i, j);
        for (long q = 0; q < nrhs; ++q)
            for (long i = 0; i < m; ++i)
                W[q * m + i] = B(i, q);
        int bad = 0;
        gels_lane<double>(true, m, n, nrhs, Dv, tau.data(), Wv, 1, info ? &bad : nullptr);
        if (info) info[idx] = bad;
        for (long j = 0; j < n; ++j)
            for (long i = 0; i < m; ++i)
                A(i, j) = D[j * m + i];
        for (long q = 0; q < nrhs; ++q)
            for (long i = 0; i < n; ++i)
                X(i, q) = W[q * m + i];
    });
    return 0;
}

int oracle_dgels_batch(char trans, int nprob, const oracle_gels_problem *desc, double *ap,
                       double *bp, int *info)
{
    return gels_batch<double>(trans, nprob, desc, ap, bp, info);
}

int oracle_sgels_batch(char trans, int nprob, const oracle_gels_problem *desc, float *ap,
                       float *bp, int *info)
{
    return gels_batch<float>(trans, nprob, desc, ap, bp, info);
}

/* The mixed-precision uneven batch (the library's dsgels_batch; no reference
 * counterpart). A problem with trans 'N', 1 <= n <= m <= 64, n <= 32 and
 * 1 <= nrhs <= 8: the single precision gels (this file's gels_lane in float,
 * the reference's gels_compact_group, src/cbk_gels_compact.hpp:84-106) on
 * float copies of A and B gives R and x_0; each column is then refined from
 * x_0 as dsgels_compact's, with R^-1 formed once (refine_column_inv). A problem that does not converge
 * (or has a zero or non-finite R(i,i)) gets iter = -31, every other problem
 * iter = -1, and both are solved by the double gels in place on A and on X
 * := B. X is laid out as B. */
int oracle_dsgels_batch(char trans, int nprob, const oracle_gels_problem *desc, double *ap,
                        const double *bp, double *xp, int *iter, int *info)
{
    if (!opt(trans, 'N') && !opt(trans, 'T') && !opt(trans, 'C')) return -1;
    if (nprob < 0) return -2;
    const bool tran = !opt(trans, 'N');
    const int itermax = 30;
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < nprob; ++k) {
        const oracle_gels_problem &d = desc[k];
        const long m = d.m, n = d.n, nrhs = d.nrhs;
        const LV<double> A = make_view(ap + d.a_off, false, d.lda, 1);
        const LV<double> B = make_view(const_cast<double *>(bp) + d.b_off, false, d.ldb, 1);
        const LV<double> X = make_view(xp + d.b_off, false, d.ldb, 1);
        const bool mixed = !tran && n >= 1 && n <= m && m <= 64 && n <= 32 && nrhs >= 1 && nrhs <= 8;
        bool ok = mixed;
        int steps = 0;
        info[k] = 0;
        if (mixed) {
            std::vector<float> F((size_t)m * n), S((size_t)m * nrhs), tau(n);
            const LV<float> Fv{F.data(), 1, m, 1}, Sv{S.data(), 1, m, 1};
            for (long j = 0; j < n; ++j)
                for (long i = 0; i < m; ++i)
                    F[j * m + i] = (float)A(i, j);
            for (long q = 0; q < nrhs; ++q)
                for (long i = 0; i < m; ++i)
                    S[q * m + i] = (float)B(i, q);
            gels_lane<float>(true, m, n, nrhs, Fv, tau.data(), Sv, 1, nullptr);
            for (long i = 0; i < n && ok; ++i)
                ok = Fv(i, i) != 0.f && std::isfinite(Fv(i, i));
            std::vector<std::vector<double>> xs(nrhs, std::vector<double>(n));
            std::vector<float> Ri;
            if (ok) Ri = tri_inverse(n, Fv);
            for (long q = 0; q < nrhs && ok; ++q) {
                for (long i = 0; i < n; ++i)
                    xs[q][i] = (double)Sv(i, q);
                ok = refine_column_inv(m, n, A, B, q, Ri, xs[q], itermax, steps);
            }
            if (ok) {
                for (long q = 0; q < nrhs; ++q)
                    for (long i = 0; i < n; ++i)
                        X(i, q) = xs[q][i];
                iter[k] = steps;
                continue;
            }
        }
        iter[k] = mixed ? -itermax - 1 : -1;
        const long mx = m > n ? m : n;
        for (long q = 0; q < nrhs; ++q)
            for (long i = 0; i < mx; ++i)
                X(i, q) = B(i, q);
        const bool tall = m >= n;
        const long p = tall ? m : n, qq = tall ? n : m;
        int bad = 0;
        if (nrhs > 0 && mx > 0) {
            if (qq == 0) {
                zero_block(X, 0, p, nrhs);
            }
            else {
                std::vector<double> tau(qq);
                gels_lane<double>(tall != tran, p, qq, nrhs, tall ? A : A.tr(), tau.data(), X, 1,
                                  &bad);
            }
        }
        info[k] = bad;
    }
    return 0;
}

} /* extern "C" */

namespace {

/* ---------------------------------------------------------------- */
/* dense <-> compact: tests/test_compact_util.hpp:475-531           */
/* ---------------------------------------------------------------- */

/* pack_compact / unpack_compact of the reference's test utilities, with the
 * dense side's layout and leading dimension free and either a strided batch
 * or an array of matrix pointers; the library's cbk_?pack_compact family. A
 * padded slot (idx >= nm) carries pad on its diagonal, zero elsewhere. */
template <class T> struct DenseBatch {
    const T *a;
    const T *const *aptr;
    T *w;
    T *const *wptr;
    size_t stride;
    long si, sj;
    const T *src(long idx) const { return a ? a + idx * stride : aptr[idx]; }
    T *dst(long idx) const { return w ? w + idx * stride : wptr[idx]; }
};

inline int log2v(int V)
{
    for (int s = 1; s <= 6; ++s)
        if (V == (1 << s)) return s;
    return -1;
}

template <class T>
void pack_loop(bool row, int rows, int cols, const DenseBatch<T> &D, T *ap, int ldap, T pad,
               int V, int nm)
{
    const long ng = ((long)nm + V - 1) / V;
    const size_t gs = group_stride(row, ldap, rows, cols, V);
    for (long g = 0; g < ng; ++g) {
        const LV<T> P = make_view(ap + g * gs, row, ldap, V);
        for (long v = 0; v < V; ++v) {
            const long idx = g * V + v;
            for (long j = 0; j < cols; ++j)
                for (long i = 0; i < rows; ++i)
                    P.p[v + (i * P.si + j * P.sj) * V] =
                        idx < nm ? D.src(idx)[i * D.si + j * D.sj] : (i == j ? pad : T(0));
        }
    }
}

template <class T>
void unpack_loop(bool row, int rows, int cols, const T *ap, int ldap, const DenseBatch<T> &D,
                 int V, int nm)
{
    const long ng = ((long)nm + V - 1) / V;
    const size_t gs = group_stride(row, ldap, rows, cols, V);
    for (long g = 0; g < ng; ++g) {
        const LV<T> P = make_view(const_cast<T *>(ap) + g * gs, row, ldap, V);
        for (long v = 0; v < V && g * V + v < nm; ++v)
            for (long j = 0; j < cols; ++j)
                for (long i = 0; i < rows; ++i)
                    D.dst(g * V + v)[i * D.si + j * D.sj] = P.p[v + (i * P.si + j * P.sj) * V];
    }
}

template <class T>
int pack_strided(char layout, int rows, int cols, const T *a, char alayout, int lda,
                 size_t stride_a, T *ap, int ldap, T pad, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool acol = opt(alayout, 'C'), arow = opt(alayout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && !a) return -4;
    if (!acol && !arow) return -5;
    if (lda < max1(arow ? cols : rows)) return -6;
    if (nm > 1 && stride_a < (size_t)lda * (arow ? rows : cols)) return -7;
    if (any && !ap) return -8;
    if (ldap < max1(row ? cols : rows)) return -9;
    if (log2v(V) < 0) return -11;
    if (nm < 0) return -12;
    if (!any) return 0;
    DenseBatch<T> D{a, nullptr, nullptr, nullptr, stride_a, arow ? lda : 1, arow ? 1 : lda};
    pack_loop(row, rows, cols, D, ap, ldap, pad, V, nm);
    return 0;
}

template <class T>
int unpack_strided(char layout, int rows, int cols, const T *ap, int ldap, T *a, char alayout,
                   int lda, size_t stride_a, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool acol = opt(alayout, 'C'), arow = opt(alayout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && !ap) return -4;
    if (ldap < max1(row ? cols : rows)) return -5;
    if (any && !a) return -6;
    if (!acol && !arow) return -7;
    if (lda < max1(arow ? cols : rows)) return -8;
    if (nm > 1 && stride_a < (size_t)lda * (arow ? rows : cols)) return -9;
    if (log2v(V) < 0) return -10;
    if (nm < 0) return -11;
    if (!any) return 0;
    DenseBatch<T> D{nullptr, nullptr, a, nullptr, stride_a, arow ? lda : 1, arow ? 1 : lda};
    unpack_loop(row, rows, cols, ap, ldap, D, V, nm);
    return 0;
}

template <class T>
int gepack(char layout, int rows, int cols, const T *const *a, int lda, T *ap, int ldap, int V,
           int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && !a) return -4;
    if (lda < max1(row ? cols : rows)) return -5;
    if (any && !ap) return -6;
    if (ldap < max1(row ? cols : rows)) return -7;
    if (log2v(V) < 0) return -8;
    if (nm < 0) return -9;
    if (!any) return 0;
    DenseBatch<T> D{nullptr, a, nullptr, nullptr, 0, row ? lda : 1, row ? 1 : lda};
    pack_loop(row, rows, cols, D, ap, ldap, T(1), V, nm);
    return 0;
}

template <class T>
int geunpack(char layout, int rows, int cols, T *const *a, int lda, const T *ap, int ldap, int V,
             int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && !a) return -4;
    if (lda < max1(row ? cols : rows)) return -5;
    if (any && !ap) return -6;
    if (ldap < max1(row ? cols : rows)) return -7;
    if (log2v(V) < 0) return -8;
    if (nm < 0) return -9;
    if (!any) return 0;
    DenseBatch<T> D{nullptr, nullptr, nullptr, a, 0, row ? lda : 1, row ? 1 : lda};
    unpack_loop(row, rows, cols, ap, ldap, D, V, nm);
    return 0;
}

} /* namespace */

extern "C" {

#define ORACLE_PACK(T, p)                                                                     \
    int oracle_cbk_##p##pack_compact(char layout, int rows, int cols, const T *a,             \
                                     char alayout, int lda, size_t stride_a, T *ap, int ldap, \
                                     T pad_diag, int V, int nm)                               \
    {                                                                                         \
        return pack_strided<T>(layout, rows, cols, a, alayout, lda, stride_a, ap, ldap,       \
                               pad_diag, V, nm);                                              \
    }                                                                                         \
    int oracle_cbk_##p##unpack_compact(char layout, int rows, int cols, const T *ap,          \
                                       int ldap, T *a, char alayout, int lda,                 \
                                       size_t stride_a, int V, int nm)                        \
    {                                                                                         \
        return unpack_strided<T>(layout, rows, cols, ap, ldap, a, alayout, lda, stride_a, V,  \
                                 nm);                                                         \
    }                                                                                         \
    int oracle_cbk_##p##gepack_compact(char layout, int rows, int cols, const T *const *a,    \
                                       int lda, T *ap, int ldap, int V, int nm)               \
    {                                                                                         \
        return gepack<T>(layout, rows, cols, a, lda, ap, ldap, V, nm);                        \
    }                                                                                         \
    int oracle_cbk_##p##geunpack_compact(char layout, int rows, int cols, T *const *a,        \
                                         int lda, const T *ap, int ldap, int V, int nm)       \
    {                                                                                         \
        return geunpack<T>(layout, rows, cols, a, lda, ap, ldap, V, nm);                      \
    }

ORACLE_PACK(double, d)
ORACLE_PACK(float, s)
#undef ORACLE_PACK

} /* extern "C" */

namespace {

/* LAPACK 3.12 dsposv.f, per matrix, on the compact layout; LDL: the same
 * algorithm with the unpivoted LDL^T of sysvnp in both precisions. The inner
 * single precision factor and solves are this file's restatement of the
 * reference's spotrf/spotrs (ssytrfnp/ssytrsnp) lane kernels. Per-matrix
 * codes follow dsposv's ITER: 0..30 the refinement steps taken; -2 B, A or a
 * residual overflows single precision (dlag2s/dlat2s, tested on B first);
 * -3 the single factor failed (LDL: also a non-finite D); -31 refinement did
 * not converge in 30 steps. In the negative cases the double dposv (dsysvnp)
 * ran on A in place. A residual or iterate holding a NaN never passes the
 * convergence test. */
template <bool LDL>
int mixed_sv(char layout, char uplo, int n, int nrhs, double *ap, int ldap, const double *bp,
             int ldbp, double *xp, int ldxp, int *iter, int *info, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool lo = opt(uplo, 'L'), up = opt(uplo, 'U');
    if (!col && !row) return -1;
    if (!lo && !up) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldap < max1(n)) return -6;
    if (ldbp < max1(row ? nrhs : n)) return -8;
    if (ldxp < max1(row ? nrhs : n)) return -10;
    if (!vlen_ok(V)) return -15;
    if (nm < 0) return -16;
    for (int i = 0; i < nm; ++i) {
        iter[i] = 0;
        info[i] = 0;
    }
    if (n == 0 || nrhs == 0 || nm == 0) return 0;
    const size_t sa = group_stride(row, ldap, n, n, V), sb = group_stride(row, ldbp, n, nrhs, V),
                 sx = group_stride(row, ldxp, n, nrhs, V);
    const double eps = std::ldexp(1.0, -53); /* dlamch('Epsilon') */
    const double cte_scale = eps * std::sqrt((double)n);
    const double rmax = (double)std::numeric_limits<float>::max(); /* slamch('O') */
    const int itermax = 30;
    for_each_lane(nm, V, [&](long g, long v) {
        const long idx = g * V + v;
        if (idx >= nm) return;
        LV<double> A = make_view(ap + g * sa + v, row, ldap, V);
        if (up) A = A.tr();
        const LV<double> B = make_view(const_cast<double *>(bp) + g * sb + v, row, ldbp, V);
        const LV<double> X = make_view(xp + g * sx + v, row, ldxp, V);
        auto over = [&](double x) { return x < -rmax || x > rmax; };
        /* ||A||_inf of the symmetric matrix (dlansy 'I') */
        double anrm = 0;
        for (long i = 0; i < n; ++i) {
            double s = 0;
            for (long j = 0; j < n; ++j)
                s += std::fabs(i >= j ? A(i, j) : A(j, i));
            anrm = s > anrm ? s : anrm;
        }
        const double cte = anrm * cte_scale;
        bool b_over = false, a_over = false;
        for (long c = 0; c < nrhs; ++c)
            for (long i = 0; i < n; ++i)
                b_over |= over(B(i, c));
        for (long j = 0; j < n; ++j)
            for (long i = j; i < n; ++i)
                a_over |= over(A(i, j));
        /* single copies: column-major n x n (lower) and n x nrhs, V = 1 */
        std::vector<float> SA((size_t)n * n, 0.f), SX((size_t)n * nrhs);
        const LV<float> SAv{SA.data(), 1, n, 1};
        std::vector<double> R((size_t)n * nrhs);
        int it = 0;
        bool done = false;
        if (b_over || a_over) {
            it = -2;
        }
        else {
            for (long j = 0; j < n; ++j)
                for (long i = j; i < n; ++i)
                    SA[j * n + i] = (float)A(i, j);
            int finfo = 0;
            if (LDL) {
                sytrfnp_block<float>(0, n, n, SAv, &finfo);
                for (long j = 0; j < n && finfo == 0; ++j)
                    if (!std::isfinite(SA[j * n + j])) finfo = (int)j + 1;
            }
            else {
                potrf_block<float>(0, n, n, SAv, &finfo);
            }
            auto solve32 = [&]() {
                if (LDL)
                    sytrsnp_lane<float>(false, false, n, nrhs, SA.data(), n, SX.data(), n, 1);
                else
                    potrs_lane<float>(false, false, n, nrhs, SA.data(), n, SX.data(), n, 1);
            };
            if (finfo != 0) {
                it = -3;
            }
            else {
                for (long c = 0; c < nrhs; ++c)
                    for (long i = 0; i < n; ++i)
                        SX[c * n + i] = (float)B(i, c);
                solve32();
                for (long c = 0; c < nrhs; ++c)
                    for (long i = 0; i < n; ++i)
                        X(i, c) = (double)SX[c * n + i];
                for (int step = 0;; ++step) {
                    sym_residual(n, nrhs, A, X, B, R);
                    bool conv = true;
                    for (long c = 0; c < nrhs; ++c) {
                        double xn = 0, rn = 0;
                        for (long i = 0; i < n; ++i) {
                            xn = std::fmax(xn, std::fabs(X(i, c)));
                            rn = std::fmax(rn, std::fabs(R[c * n + i]));
                            if (std::isnan(X(i, c)) || std::isnan(R[c * n + i])) conv = false;
                        }
                        if (!(rn <= xn * cte)) conv = false;
                    }
                    if (conv) {
                        it = step;
                        done = true;
                        break;
                    }
                    if (step == itermax) {
                        it = -itermax - 1;
                        break;
                    }
                    bool r_over = false;
                    for (size_t q = 0; q < R.size(); ++q) r_over |= over(R[q]);
                    if (r_over) {
                        it = -2;
                        break;
                    }
                    for (size_t q = 0; q < R.size(); ++q)
                        SX[q] = (float)R[q];
                    solve32();
                    for (long c = 0; c < nrhs; ++c)
                        for (long i = 0; i < n; ++i)
                            X(i, c) = X(i, c) + (double)SX[c * n + i];
                }
            }
        }
        if (!done) {
            /* fallback: the double dposv (dsysvnp) on A, X := B first */
            int dinfo = 0;
            if (LDL)
                sytrfnp_block<double>(0, n, n, A, &dinfo);
            else
                potrf_block<double>(0, n, n, A, &dinfo);
            info[idx] = dinfo;
            if (dinfo == 0) {
                for (long c = 0; c < nrhs; ++c)
                    for (long i = 0; i < n; ++i)
                        X(i, c) = B(i, c);
                if (LDL)
                    sytrsnp_lane<double>(row, up, n, nrhs, ap + g * sa + v, ldap,
                                         xp + g * sx + v, ldxp, V);
                else
                    potrs_lane<double>(row, up, n, nrhs, ap + g * sa + v, ldap, xp + g * sx + v,
                                       ldxp, V);
            }
        }
        iter[idx] = it;
    });
    return 0;
}

} /* namespace */

extern "C" {

int oracle_dsposv_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                          const double *bp, int ldbp, double *xp, int ldxp, void *work,
                          size_t lwork, int *iter, int *info, int V, int nm)
{
    (void)work;
    (void)lwork;
    return mixed_sv<false>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, iter, info, V,
                           nm);
}

int oracle_dssysvnp_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                            const double *bp, int ldbp, double *xp, int ldxp, void *work,
                            size_t lwork, int *iter, int *info, int V, int nm)
{
    (void)work;
    (void)lwork;
    return mixed_sv<true>(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, iter, info, V,
                          nm);
}

} /* extern "C" */
