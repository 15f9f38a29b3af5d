/* compat.hip -- the MKL-style adapter (include/cbk_compat.h): the MKL enum
 * and format arguments decoded as the reference's src/cbk_compat.cpp:43-250
 * decodes them, then forwarded to this library's include/cbk.h routines. */

#include "cbk.h"
#include "cbk_compat.h"
#include "cbk_internal.h"

namespace cbk {
namespace {

/* V = register bytes / sizeof(T) for MKL's three formats
 * (include/cbk_compat.h:276-286 of the reference), 64 for CBK_COMPACT_V64 */
int vlen_for(int format, int size)
{
    switch (format) {
    case CBK_MKL_COMPACT_SSE: return 16 / size;
    case CBK_MKL_COMPACT_AVX: return 32 / size;
    case CBK_MKL_COMPACT_AVX512: return 64 / size;
    case CBK_COMPACT_V64: return 64;
    default: return 0;
    }
}

char layout_of(int layout) { return layout == CBK_MKL_ROW_MAJOR ? 'R' : 'C'; }
char uplo_of(int uplo) { return uplo == CBK_MKL_UPPER ? 'U' : 'L'; }
char tran_of(char trans)
{
    return (trans == 'T' || trans == 't' || trans == 'C' || trans == 'c') ? 'T' : 'N';
}

template <class T> __global__ void put_scalar(T *p, T v) { *p = v; }

/* the lwork = -1 answer, written to work[0] on the device */
template <class T> int answer_query(T *work, double value)
{
    hipLaunchKernelGGL(put_scalar<T>, dim3(1), dim3(1), 0, current_stream(), work, (T)value);
    return hipGetLastError() == hipSuccess ? 0 : CBK_ERR_LAUNCH;
}

template <class T> constexpr bool is_d() { return sizeof(T) == 8; }

template <class T>
void geqrf(int layout, int m, int n, T *ap, int ldap, T *taup, T *work, int lwork, int *info,
           int format, int nm)
{
    *info = 0;
    if (lwork == -1) {
        *info = answer_query(work, 1.0);
        return;
    }
    const int V = vlen_for(format, sizeof(T));
    if (V == 0) {
        *info = -1;
        return;
    }
    const char L = layout_of(layout);
    *info = is_d<T>() ? dgeqrf_compact(L, m, n, (double *)ap, ldap, (double *)taup, V, nm)
                      : sgeqrf_compact(L, m, n, (float *)ap, ldap, (float *)taup, V, nm);
}

template <class T>
void ormqr(int layout, char side, char trans, int m, int n, int k, const T *ap, int ldap,
           const T *taup, T *cp, int ldcp, T *work, int lwork, int *info, int format, int nm)
{
    *info = 0;
    if (lwork == -1) {
        *info = answer_query(work, 1.0);
        return;
    }
    const int V = vlen_for(format, sizeof(T));
    if (V == 0) {
        *info = -1;
        return;
    }
    const char L = layout_of(layout), S = (side == 'L' || side == 'l') ? 'L' : 'R',
               Tr = tran_of(trans);
    *info = is_d<T>() ? dormqr_compact(L, S, Tr, m, n, k, (const double *)ap, ldap,
                                       (const double *)taup, (double *)cp, ldcp, V, nm)
                      : sormqr_compact(L, S, Tr, m, n, k, (const float *)ap, ldap,
                                       (const float *)taup, (float *)cp, ldcp, V, nm);
}

template <class T>
void orgqr(int layout, int m, int n, int k, T *ap, int ldap, const T *taup, T *work, int lwork,
           int *info, int format, int nm)
{
    *info = 0;
    if (lwork == -1) {
        *info = answer_query(work, 1.0);
        return;
    }
    const int V = vlen_for(format, sizeof(T));
    if (V == 0) {
        *info = -1;
        return;
    }
    const char L = layout_of(layout);
    *info = is_d<T>() ? dorgqr_compact(L, m, n, k, (double *)ap, ldap, (const double *)taup, V, nm)
                      : sorgqr_compact(L, m, n, k, (float *)ap, ldap, (const float *)taup, V, nm);
}

/* potrf and sytrfnp: fac(layout, uplo, n, ap, ldap, V, nm) */
template <class T, class F>
void factor(F &&fac, int layout, int uplo, int n, T *ap, int ldap, int *info, int format, int nm)
{
    *info = 0;
    const int V = vlen_for(format, sizeof(T));
    if (V == 0) {
        *info = -1;
        return;
    }
    *info = fac(layout_of(layout), uplo_of(uplo), n, ap, ldap, V, nm);
}

/* potrs, posv, sytrsnp, sysvnp: sol(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm) */
template <class T, class A, class F>
void solve(F &&sol, int layout, int uplo, int n, int nrhs, A *ap, int ldap, T *bp, int ldbp,
           int *info, int format, int nm)
{
    *info = 0;
    const int V = vlen_for(format, sizeof(T));
    if (V == 0) {
        *info = -1;
        return;
    }
    *info = sol(layout_of(layout), uplo_of(uplo), n, nrhs, ap, ldap, bp, ldbp, V, nm);
}

template <class T>
void trsm(int layout, int side, int uplo, int transa, int diag, int m, int n, T alpha,
          const T *ap, int ldap, T *bp, int ldbp, int format, int nm)
{
    const int V = vlen_for(format, sizeof(T));
    if (V == 0) return; /* no info: an unrecognized format selects nothing */
    const char L = layout_of(layout), S = side == CBK_MKL_LEFT ? 'L' : 'R', U = uplo_of(uplo),
               Tr = transa != CBK_MKL_NOTRANS ? 'T' : 'N', D = diag == CBK_MKL_UNIT ? 'U' : 'N';
    if constexpr (is_d<T>())
        dtrsm_compact(L, S, U, Tr, D, m, n, alpha, ap, ldap, bp, ldbp, V, nm);
    else
        strsm_compact(L, S, U, Tr, D, m, n, alpha, ap, ldap, bp, ldbp, V, nm);
}

template <class T>
void gels(int layout, char trans, int m, int n, int nrhs, T *ap, int ldap, T *bp, int ldbp,
          T *work, int lwork, int *info, int format, int nm)
{
    const int V = vlen_for(format, sizeof(T));
    if (lwork == -1) {
        /* the tau scratch: one compact slot of min(m,n) per matrix */
        if (V == 0) {
            *info = -1;
            return;
        }
        const double groups = nm > 0 ? (double)((nm + V - 1) / V) : 0.0;
        const double need = (double)(m < n ? m : n) * V * groups;
        *info = answer_query(work, need > 1.0 ? need : 1.0);
        return;
    }
    *info = 0;
    if (V == 0) {
        *info = -1;
        return;
    }
    const char L = layout_of(layout), Tr = tran_of(trans);
    *info = is_d<T>() ? dgels_compact(L, Tr, m, n, nrhs, (double *)ap, ldap, (double *)bp, ldbp,
                                      (double *)work, V, nm)
                      : sgels_compact(L, Tr, m, n, nrhs, (float *)ap, ldap, (float *)bp, ldbp,
                                      (float *)work, V, nm);
}

} /* namespace */
} /* namespace cbk */

using namespace cbk;

extern "C" {

int cbk_compact_vlen(int format, int fp64) { return vlen_for(format, fp64 ? 8 : 4); }

#define CBK_COMPAT_ENTRY_POINTS(T, p)                                                              \
    void cbk_##p##geqrf_compact(int layout, int m, int n, T *ap, int ldap, T *taup, T *work,       \
                                int lwork, int *info, int format, int nm)                          \
    {                                                                                              \
        geqrf<T>(layout, m, n, ap, ldap, taup, work, lwork, info, format, nm);                     \
    }                                                                                              \
    void cbk_##p##ormqr_compact(int layout, char side, char trans, int m, int n, int k,            \
                                const T *ap, int ldap, const T *taup, T *cp, int ldcp, T *work,    \
                                int lwork, int *info, int format, int nm)                          \
    {                                                                                              \
        ormqr<T>(layout, side, trans, m, n, k, ap, ldap, taup, cp, ldcp, work, lwork, info,        \
                 format, nm);                                                                      \
    }                                                                                              \
    void cbk_##p##orgqr_compact(int layout, int m, int n, int k, T *ap, int ldap, const T *taup,   \
                                T *work, int lwork, int *info, int format, int nm)                 \
    {                                                                                              \
        orgqr<T>(layout, m, n, k, ap, ldap, taup, work, lwork, info, format, nm);                  \
    }                                                                                              \
    void cbk_##p##potrf_compact(int layout, int uplo, int n, T *ap, int ldap, int *info,           \
                                int format, int nm)                                                \
    {                                                                                              \
        factor<T>(p##potrf_compact, layout, uplo, n, ap, ldap, info, format, nm);                  \
    }                                                                                              \
    void cbk_##p##sytrfnp_compact(int layout, int uplo, int n, T *ap, int ldap, int *info,         \
                                  int format, int nm)                                              \
    {                                                                                              \
        factor<T>(p##sytrfnp_compact, layout, uplo, n, ap, ldap, info, format, nm);                \
    }                                                                                              \
    void cbk_##p##potrs_compact(int layout, int uplo, int n, int nrhs, const T *ap, int ldap,      \
                                T *bp, int ldbp, int *info, int format, int nm)                    \
    {                                                                                              \
        solve<T, const T>(p##potrs_compact, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, info,       \
                          format, nm);                                                             \
    }                                                                                              \
    void cbk_##p##posv_compact(int layout, int uplo, int n, int nrhs, T *ap, int ldap, T *bp,      \
                               int ldbp, int *info, int format, int nm)                            \
    {                                                                                              \
        solve<T, T>(p##posv_compact, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, info, format, nm); \
    }                                                                                              \
    void cbk_##p##sytrsnp_compact(int layout, int uplo, int n, int nrhs, const T *ap, int ldap,    \
                                  T *bp, int ldbp, int *info, int format, int nm)                  \
    {                                                                                              \
        solve<T, const T>(p##sytrsnp_compact, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, info,     \
                          format, nm);                                                             \
    }                                                                                              \
    void cbk_##p##sysvnp_compact(int layout, int uplo, int n, int nrhs, T *ap, int ldap, T *bp,    \
                                 int ldbp, int *info, int format, int nm)                          \
    {                                                                                              \
        solve<T, T>(p##sysvnp_compact, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, info, format,    \
                    nm);                                                                           \
    }                                                                                              \
    void cbk_##p##trsm_compact(int layout, int side, int uplo, int transa, int diag, int m, int n, \
                               T alpha, const T *ap, int ldap, T *bp, int ldbp, int format,        \
                               int nm)                                                             \
    {                                                                                              \
        trsm<T>(layout, side, uplo, transa, diag, m, n, alpha, ap, ldap, bp, ldbp, format, nm);    \
    }                                                                                              \
    void cbk_##p##gels_compact(int layout, char trans, int m, int n, int nrhs, T *ap, int ldap,    \
                               T *bp, int ldbp, T *work, int lwork, int *info, int format,         \
                               int nm)                                                             \
    {                                                                                              \
        gels<T>(layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, work, lwork, info, format, nm);     \
    }

CBK_COMPAT_ENTRY_POINTS(double, d)
CBK_COMPAT_ENTRY_POINTS(float, s)
#undef CBK_COMPAT_ENTRY_POINTS

} /* extern "C" */
