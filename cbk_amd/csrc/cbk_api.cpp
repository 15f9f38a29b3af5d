/* cbk_api.cpp -- the C ABI (include/cbk.h, include/cbk_ext.h).
 *
 * Host side of the boundary: LAPACK-style argument validation in argument
 * order and the quick returns of the reference's src/cbk.cpp:41-391 (the same
 * codes, with V widened to {2,4,8,16,32,64}), then a stream-ordered launch of
 * the kernel that serves the shape: a team kernel for the tuned shapes
 * (dispatch.cpp), the lane kernel for everything else.
 */

#include "cbk.h"
#include "cbk_ext.h"
#include "cbk_internal.h"
#include "dispatch.h"

#include <hip/hip_runtime.h>

namespace cbk {

namespace {
thread_local hipStream_t t_stream = nullptr;
thread_local int t_policy = CBK_POLICY_AUTO;
thread_local hipError_t t_last_error = hipSuccess;
}

hipStream_t current_stream() { return t_stream; }
int current_policy() { return t_policy; }

int launch_status(hipError_t e)
{
    if (e == hipSuccess) return 0;
    t_last_error = e;
    return CBK_ERR_LAUNCH;
}

namespace {

/* src/cbk.cpp:46-59 */
inline bool opt(char c, char upper) { return c == upper || c == upper + ('a' - 'A'); }
inline int max1(int x) { return x < 1 ? 1 : x; }
inline bool vlen_ok(int V)
{
    return V == 2 || V == 4 || V == 8 || V == 16 || V == 32 || V == 64;
}

inline int status(hipError_t e) { return launch_status(e); }

/* group_stride (src/cbk_common.hpp:210-214) */
inline size_t gstride(bool row, int ld, int rows, int cols, int V)
{
    return (size_t)ld * (size_t)(row ? rows : cols) * (size_t)V;
}

template <class T> Operand<T> operand(T *p, int ld, bool row, int rows, int cols, int V)
{
    return {p, ld, row, gstride(row, ld, rows, cols, V)};
}

/* src/cbk.cpp:61-78 */
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
    return status(dispatch_geqrf<T>(m, n, operand(ap, ldap, row, m, n, V), taup, V, nm));
}

/* src/cbk.cpp:80-105 */
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
    /* A is the (ldap, k) reflector batch: group stride ldap*k*V column-major
     * (src/cbk_ormqr_compact.hpp:144) */
    const Operand<T> A = operand(const_cast<T *>(ap), ldap, row, nq, k, V);
    const Operand<T> C = operand(cp, ldcp, row, m, n, V);
    const bool fit = offsets_fit(A, nq, k, V) && offsets_fit(C, m, n, V);
    if (fit && current_policy() != CBK_POLICY_LANE && ormqr_reg_supports(nq, sizeof(T) == 8))
        return status(launch_ormqr_reg<T>(left, opt(trans, 'T'), m, n, k, A, taup, C, V, nm));
    if (fit && current_policy() != CBK_POLICY_LANE && V >= 4 && ormqr_team_supports(nq))
        return status(launch_ormqr_team<T>(left, opt(trans, 'T'), m, n, k, A, taup, C, V, nm));
    if (fit && current_policy() != CBK_POLICY_LANE && apply_grid_supports(nq))
        return status(launch_ormqr_grid<T>(left, opt(trans, 'T'), m, n, k, A, taup, C, V, nm));
    return status(launch_ormqr_lane<T>(left, opt(trans, 'T'), m, n, k, A, taup, C, V, nm));
}

/* src/cbk.cpp:107-127 */
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
    if (m == 0 || n == 0 || nm == 0) return 0; /* k == 0 is not empty */
    const Operand<T> A = operand(ap, ldap, row, m, n, V);
    const bool fit = offsets_fit(A, m, n, V) && current_policy() != CBK_POLICY_LANE;
    if (fit && qr_reg_supports(m, n, sizeof(T) == 8))
        return status(launch_orgqr_reg<T>(m, n, k, A, taup, V, nm));
    if (fit && apply_grid_supports(m))
        return status(launch_orgqr_grid<T>(m, n, k, A, taup, V, nm));
    return status(launch_orgqr_lane<T>(m, n, k, A, taup, V, nm));
}

/* src/cbk.cpp:129-167: potrf and sytrfnp share their codes */
inline int sym_factor_args(char layout, char uplo, int n, int ldap, int V, int nm, bool &row,
                           bool &up)
{
    const bool col = opt(layout, 'C'), lo = opt(uplo, 'L');
    row = opt(layout, 'R');
    up = opt(uplo, 'U');
    if (!col && !row) return -1;
    if (!lo && !up) return -2;
    if (n < 0) return -3;
    if (ldap < max1(n)) return -5;
    if (!vlen_ok(V)) return -6;
    if (nm < 0) return -7;
    return 0;
}

/* src/cbk.cpp:173-188 */
inline int sym_solve_args(char layout, char uplo, int n, int nrhs, int ldap, int ldbp, int V,
                          int nm, bool &row, bool &up)
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

/* potrf / sytrfnp (src/cbk.cpp:129-167), with optional per-matrix info */
template <class T>
int sym_factor(bool ldl, char layout, char uplo, int n, T *ap, int ldap, int V, int nm,
               int *info)
{
    bool row, up;
    if (int e = sym_factor_args(layout, uplo, n, ldap, V, nm, row, up)) return e;
    if (n == 0 || nm == 0) {
        return info ? status(launch_clear_info(info, nm)) : 0;
    }
    const Operand<T> A = operand(ap, ldap, row, n, n, V);
    const Operand<T> B = {nullptr, 1, row, 0};
    return status(dispatch_sym<T>(ldl, true, up, n, 0, A, B, V, nm, info));
}

/* potrs / sytrsnp (src/cbk.cpp:190-205, 226-241) */
template <class T>
int sym_solve(bool ldl, char layout, char uplo, int n, int nrhs, const T *ap, int ldap, T *bp,
              int ldbp, int V, int nm)
{
    bool row, up;
    if (int e = sym_solve_args(layout, uplo, n, nrhs, ldap, ldbp, V, nm, row, up)) return e;
    if (n == 0 || nrhs == 0 || nm == 0) return 0;
    const Operand<T> A = operand(const_cast<T *>(ap), ldap, row, n, n, V);
    const Operand<T> B = operand(bp, ldbp, row, n, nrhs, V);
    return status(dispatch_sym<T>(ldl, false, up, n, nrhs, A, B, V, nm, nullptr));
}

/* posv / sysvnp (src/cbk.cpp:207-224, 243-260): nrhs == 0 still factors */
template <class T>
int sym_sv(bool ldl, char layout, char uplo, int n, int nrhs, T *ap, int ldap, T *bp,
           int ldbp, int V, int nm, int *info)
{
    bool row, up;
    if (int e = sym_solve_args(layout, uplo, n, nrhs, ldap, ldbp, V, nm, row, up)) return e;
    if (n == 0 || nm == 0) {
        return info ? status(launch_clear_info(info, nm)) : 0;
    }
    const Operand<T> A = operand(ap, ldap, row, n, n, V);
    const Operand<T> B = operand(bp, ldbp, row, n, nrhs, V);
    return status(dispatch_sym<T>(ldl, true, up, n, nrhs, A, B, V, nm, info));
}

/* src/cbk.cpp:262-293 */
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
    /* -8 alpha: every value is legal */
    if (ldap < max1(s)) return -10;
    if (ldbp < max1(row ? n : m)) return -12;
    if (!vlen_ok(V)) return -13;
    if (nm < 0) return -14;
    if (m == 0 || n == 0 || nm == 0) return 0;
    const Operand<T> B = operand(bp, ldbp, row, m, n, V);
    /* alpha == 0: B := 0, A not referenced (src/cbk_trsm_compact.hpp:286-289) */
    if (alpha == T(0)) return status(launch_zero<T>(m, n, B, V, nm));
    const Operand<T> A = operand(const_cast<T *>(ap), ldap, row, s, s, V);
    return status(dispatch_trsm<T>(left, up, tran, unit, m, n, alpha, A, B, V, nm));
}

/* src/cbk.cpp:295-320 */
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
    if (nrhs == 0 || nm == 0 || mx == 0) {
        return info ? status(launch_clear_info(info, nm)) : 0;
    }
    const bool tran = opt(trans, 'T') || opt(trans, 'C');
    const Operand<T> B = operand(bp, ldbp, row, mx, nrhs, V);
    if ((m < n ? m : n) == 0) {
        /* X = 0 for an empty op(A) (src/cbk_gels_compact.hpp:128-133) */
        hipError_t e = launch_zero<T>(mx, nrhs, B, V, nm);
        if (e == hipSuccess && info) e = launch_clear_info(info, nm);
        return status(e);
    }
    const Operand<T> A = operand(ap, ldap, row, m, n, V);
    return status(dispatch_gels<T>(tran, m, n, nrhs, A, B, taup, V, nm, info));
}

constexpr int kItermax = 30; /* LAPACK dsposv's ITERMAX */

/* the fused kernel serves the call: its shapes, V >= 4, and operands whose
 * 32-bit offsets fit (offsets_fit) */
bool mixed_uses_team(bool row, int n, int nrhs, int ldap, int ldbp, int ldxp, int V)
{
    return current_policy() == CBK_POLICY_AUTO && V >= 4 &&
           (mixed_team_supports(n) || quad_mixed_supports(n) || mixed_rows_supports(n)) &&
           offsets_fit_bytes(gstride(row, ldap, n, n, V), ldap, row, n, n, V, 8) &&
           offsets_fit_bytes(gstride(row, ldbp, n, nrhs, V), ldbp, row, n, nrhs, V, 8) &&
           offsets_fit_bytes(gstride(row, ldxp, n, nrhs, V), ldxp, row, n, nrhs, V, 8);
}

size_t mixed_worksize(bool row, int n, int nrhs, int ldap, int ldbp, int ldxp, int V, int nm)
{
    if (mixed_uses_team(row, n, nrhs, ldap, ldbp, ldxp, V)) return 0;
    const size_t groups = (size_t)((nm + V - 1) / V);
    return groups * V * ((size_t)n * n + (size_t)n * nrhs) * sizeof(float);
}

/* dsposv_compact and dssysvnp_compact (include/cbk_ext.h): the FP32 factor
 * (Cholesky, or ldl: the unpivoted LDL^T) refined in FP64, the FP64 factor on
 * the matrices it fails */
int mixed_sv(bool ldl, char layout, char uplo, int n, int nrhs, double *ap, int ldap,
             const double *bp, int ldbp, double *xp, int ldxp, void *work, size_t lwork,
             int *iter, int *info, int V, int nm)
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
    if (vlen_ok(V) && nm >= 0 && lwork < mixed_worksize(row, n, nrhs, ldap, ldbp, ldxp, V, nm))
        return -12;
    if (nm > 0 && iter == nullptr) return -13;
    if (nm > 0 && info == nullptr) return -14;
    if (!vlen_ok(V)) return -15;
    if (nm < 0) return -16;
    if (n == 0 || nrhs == 0 || nm == 0) {
        hipError_t e = launch_clear_info(iter, nm);
        if (e == hipSuccess) e = launch_clear_info(info, nm);
        return status(e);
    }
    const Operand<double> A = operand(ap, ldap, row, n, n, V);
    const Operand<double> B = operand(const_cast<double *>(bp), ldbp, row, n, nrhs, V);
    const Operand<double> X = operand(xp, ldxp, row, n, nrhs, V);
    if (mixed_uses_team(row, n, nrhs, ldap, ldbp, ldxp, V)) {
        if (quad_mixed_supports(n))
            return status(
                launch_mixed_posv_quad(ldl, up, n, nrhs, A, B, X, V, nm, kItermax, iter, info));
        if (mixed_rows_supports(n))
            return status(
                launch_mixed_posv_rows(ldl, up, n, nrhs, A, B, X, V, nm, kItermax, iter, info));
        return status(
            launch_mixed_posv_team(ldl, up, n, nrhs, A, B, X, V, nm, kItermax, iter, info));
    }
    return status(launch_mixed_posv_lane(ldl, up, n, nrhs, A, B, X, static_cast<float *>(work),
                                         V, nm, kItermax, iter, info));
}

/* dsgels_compact (include/cbk_ext.h): the FP32-preconditioned refinement of
 * mixed_gels_impl.h where a team kernel serves the shape, then the FP64
 * wavefront gels on the matrices it marked (iter < 0), B read only */
int dsgels(char layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
           const double *bp, int ldbp, double *xp, int ldxp, int *iter, int *info, int V, int nm)
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
    if (nm > 0 && iter == nullptr) return -12;
    if (!vlen_ok(V)) return -14;
    if (nm < 0) return -15;
    hipError_t e = launch_clear_info(info, nm);
    if (n == 0 || nrhs == 0 || nm == 0) {
        if (e == hipSuccess) e = launch_clear_info(iter, nm);
        return status(e);
    }
    const Operand<double> A = operand(ap, ldap, row, m, n, V);
    const Operand<double> B = operand(const_cast<double *>(bp), ldbp, row, m, nrhs, V);
    const Operand<double> X = operand(xp, ldxp, row, n, nrhs, V);
    if (e == hipSuccess) {
        if (current_policy() != CBK_POLICY_LANE && V >= 4 && mixed_gels_supports(m, n, nrhs) &&
            offsets_fit(A, m, n, V) && offsets_fit(B, m, nrhs, V) && offsets_fit(X, n, nrhs, V))
            e = launch_mixed_gels_team(m, n, nrhs, A, B, X, V, nm, kItermax, iter);
        else
            e = launch_fill_int(iter, nm, -1);
    }
    if (e == hipSuccess) e = launch_qr_wave_masked(m, n, nrhs, A, B, X, V, nm, iter, info);
    return status(e);
}

} /* namespace */
} /* namespace cbk */

using namespace cbk;

extern "C" {

int cbk_get_version(void) { return CBK_VERSION; }

int cbk_set_stream(void *stream)
{
    t_stream = static_cast<hipStream_t>(stream);
    return 0;
}

void *cbk_get_stream(void) { return static_cast<void *>(t_stream); }

int cbk_synchronize(void) { return status(hipStreamSynchronize(t_stream)); }

const char *cbk_last_error_string(void) { return hipGetErrorString(t_last_error); }

int cbk_set_policy(int policy)
{
    const int prev = t_policy;
    if (policy == CBK_POLICY_AUTO || policy == CBK_POLICY_LANE) t_policy = policy;
    return prev;
}

#define CBK_ENTRY_POINTS(T, p)                                                           \
    int p##geqrf_compact(char layout, int m, int n, T *ap, int ldap, T *taup, int V,     \
                         int nm)                                                         \
    {                                                                                    \
        return geqrf<T>(layout, m, n, ap, ldap, taup, V, nm);                            \
    }                                                                                    \
    int p##ormqr_compact(char layout, char side, char trans, int m, int n, int k,        \
                         const T *ap, int ldap, const T *taup, T *cp, int ldcp, int V,   \
                         int nm)                                                         \
    {                                                                                    \
        return ormqr<T>(layout, side, trans, m, n, k, ap, ldap, taup, cp, ldcp, V, nm);  \
    }                                                                                    \
    int p##orgqr_compact(char layout, int m, int n, int k, T *ap, int ldap,              \
                         const T *taup, int V, int nm)                                   \
    {                                                                                    \
        return orgqr<T>(layout, m, n, k, ap, ldap, taup, V, nm);                         \
    }                                                                                    \
    int p##potrf_compact(char layout, char uplo, int n, T *ap, int ldap, int V, int nm)  \
    {                                                                                    \
        return sym_factor<T>(false, layout, uplo, n, ap, ldap, V, nm, nullptr);          \
    }                                                                                    \
    int p##potrs_compact(char layout, char uplo, int n, int nrhs, const T *ap, int ldap, \
                         T *bp, int ldbp, int V, int nm)                                 \
    {                                                                                    \
        return sym_solve<T>(false, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm);    \
    }                                                                                    \
    int p##posv_compact(char layout, char uplo, int n, int nrhs, T *ap, int ldap, T *bp, \
                        int ldbp, int V, int nm)                                         \
    {                                                                                    \
        return sym_sv<T>(false, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm,        \
                         nullptr);                                                       \
    }                                                                                    \
    int p##sytrfnp_compact(char layout, char uplo, int n, T *ap, int ldap, int V,        \
                           int nm)                                                       \
    {                                                                                    \
        return sym_factor<T>(true, layout, uplo, n, ap, ldap, V, nm, nullptr);           \
    }                                                                                    \
    int p##sytrsnp_compact(char layout, char uplo, int n, int nrhs, const T *ap,         \
                           int ldap, T *bp, int ldbp, int V, int nm)                     \
    {                                                                                    \
        return sym_solve<T>(true, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm);     \
    }                                                                                    \
    int p##sysvnp_compact(char layout, char uplo, int n, int nrhs, T *ap, int ldap,      \
                          T *bp, int ldbp, int V, int nm)                                \
    {                                                                                    \
        return sym_sv<T>(true, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm,         \
                         nullptr);                                                       \
    }                                                                                    \
    int p##trsm_compact(char layout, char side, char uplo, char transa, char diag,       \
                        int m, int n, T alpha, const T *ap, int ldap, T *bp, int ldbp,   \
                        int V, int nm)                                                   \
    {                                                                                    \
        return trsm<T>(layout, side, uplo, transa, diag, m, n, alpha, ap, ldap, bp,      \
                       ldbp, V, nm);                                                     \
    }                                                                                    \
    int p##gels_compact(char layout, char trans, int m, int n, int nrhs, T *ap,          \
                        int ldap, T *bp, int ldbp, T *taup, int V, int nm)               \
    {                                                                                    \
        return gels<T>(layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, taup, V, nm,       \
                       nullptr);                                                         \
    }                                                                                    \
    int p##potrf_compact_info(char layout, char uplo, int n, T *ap, int ldap, int V,     \
                              int nm, int *info)                                         \
    {                                                                                    \
        return sym_factor<T>(false, layout, uplo, n, ap, ldap, V, nm, info);             \
    }                                                                                    \
    int p##posv_compact_info(char layout, char uplo, int n, int nrhs, T *ap, int ldap,   \
                             T *bp, int ldbp, int V, int nm, int *info)                  \
    {                                                                                    \
        return sym_sv<T>(false, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm, info); \
    }                                                                                    \
    int p##sytrfnp_compact_info(char layout, char uplo, int n, T *ap, int ldap, int V,   \
                                int nm, int *info)                                       \
    {                                                                                    \
        return sym_factor<T>(true, layout, uplo, n, ap, ldap, V, nm, info);              \
    }                                                                                    \
    int p##sysvnp_compact_info(char layout, char uplo, int n, int nrhs, T *ap, int ldap, \
                               T *bp, int ldbp, int V, int nm, int *info)                \
    {                                                                                    \
        return sym_sv<T>(true, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, V, nm, info);  \
    }                                                                                    \
    int p##gels_compact_info(char layout, char trans, int m, int n, int nrhs, T *ap,     \
                             int ldap, T *bp, int ldbp, T *taup, int V, int nm,          \
                             int *info)                                                  \
    {                                                                                    \
        return gels<T>(layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, taup, V, nm,       \
                       info);                                                            \
    }

int dsposv_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                   const double *bp, int ldbp, double *xp, int ldxp, void *work,
                   size_t lwork, int *iter, int *info, int V, int nm)
{
    return mixed_sv(false, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, work, lwork,
                    iter, info, V, nm);
}

int dssysvnp_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                     const double *bp, int ldbp, double *xp, int ldxp, void *work,
                     size_t lwork, int *iter, int *info, int V, int nm)
{
    return mixed_sv(true, layout, uplo, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, work, lwork,
                    iter, info, V, nm);
}

int dsgels_compact(char layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
                   const double *bp, int ldbp, double *xp, int ldxp, int *iter, int *info,
                   int V, int nm)
{
    return dsgels(layout, trans, m, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, iter, info, V, nm);
}

size_t dsposv_compact_worksize(char layout, char uplo, int n, int nrhs, int ldap, int ldbp,
                               int ldxp, int V, int nm)
{
    (void)uplo;
    if (n < 0 || nrhs < 0 || nm < 0 || !vlen_ok(V)) return 0;
    return mixed_worksize(opt(layout, 'R'), n, nrhs, ldap, ldbp, ldxp, V, nm);
}

size_t dssysvnp_compact_worksize(char layout, char uplo, int n, int nrhs, int ldap, int ldbp,
                                 int ldxp, int V, int nm)
{
    return dsposv_compact_worksize(layout, uplo, n, nrhs, ldap, ldbp, ldxp, V, nm);
}

CBK_ENTRY_POINTS(double, d)
CBK_ENTRY_POINTS(float, s)
#undef CBK_ENTRY_POINTS

} /* extern "C" */
