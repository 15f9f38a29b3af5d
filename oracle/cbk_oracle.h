/* cbk_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The CPU restatement of the reference compact kernels (ivan-pi/cbk), used by
 * tests/ and by bench.py's CPU-baseline leg as the checker of the HIP library.
 * It is never linked into, loaded by, or called from the shipped library.
 *
 * Same argument lists as include/cbk.h (reference include/cbk.h:77-351), over
 * HOST pointers, prefixed oracle_. V may be any of 2, 4, 8, 16, 32, 64 (the
 * reference accepts 2..16, src/cbk.cpp:56-59). Results are bit-identical to the
 * reference compiled with FMA contraction (oracle/Makefile `ref`), which the
 * tests check; every multiply-add the reference's compiler contracts is spelt
 * out here as fma().
 */
#ifndef CBK_ORACLE_H
#define CBK_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

int oracle_get_version(void);

#define ORACLE_DECLS(T, p)                                                               \
    int oracle_##p##geqrf_compact(char layout, int m, int n, T *ap, int ldap, T *taup,   \
                                  int V, int nm);                                        \
    int oracle_##p##ormqr_compact(char layout, char side, char trans, int m, int n,      \
                                  int k, const T *ap, int ldap, const T *taup, T *cp,    \
                                  int ldcp, int V, int nm);                              \
    int oracle_##p##orgqr_compact(char layout, int m, int n, int k, T *ap, int ldap,     \
                                  const T *taup, int V, int nm);                         \
    int oracle_##p##potrf_compact(char layout, char uplo, int n, T *ap, int ldap, int V, \
                                  int nm);                                               \
    int oracle_##p##potrs_compact(char layout, char uplo, int n, int nrhs, const T *ap,  \
                                  int ldap, T *bp, int ldbp, int V, int nm);             \
    int oracle_##p##posv_compact(char layout, char uplo, int n, int nrhs, T *ap,         \
                                 int ldap, T *bp, int ldbp, int V, int nm);              \
    int oracle_##p##sytrfnp_compact(char layout, char uplo, int n, T *ap, int ldap,      \
                                    int V, int nm);                                      \
    int oracle_##p##sytrsnp_compact(char layout, char uplo, int n, int nrhs,             \
                                    const T *ap, int ldap, T *bp, int ldbp, int V,       \
                                    int nm);                                             \
    int oracle_##p##sysvnp_compact(char layout, char uplo, int n, int nrhs, T *ap,       \
                                   int ldap, T *bp, int ldbp, int V, int nm);            \
    int oracle_##p##trsm_compact(char layout, char side, char uplo, char transa,         \
                                 char diag, int m, int n, T alpha, const T *ap,          \
                                 int ldap, T *bp, int ldbp, int V, int nm);              \
    int oracle_##p##gels_compact(char layout, char trans, int m, int n, int nrhs, T *ap, \
                                 int ldap, T *bp, int ldbp, T *taup, int V, int nm);     \
    /* per-matrix LAPACK info variants (no reference counterpart): the same       \
     * arithmetic as the plain call plus info[idx] for idx < nm -- potrf/posv: the  \
     * first j (1-based) whose pivot is <= 0 or NaN; sytrfnp/sysvnp: the first     \
     * exact zero pivot; gels: the first exact zero diagonal of R. A lane with      \
     * info > 0 keeps B as LAPACK leaves it: posv/sysvnp and minimum-norm gels      \
     * unchanged, least-squares gels holding Q^T B. */                             \
    int oracle_##p##potrf_compact_info(char layout, char uplo, int n, T *ap, int ldap,  \
                                       int V, int nm, int *info);                       \
    int oracle_##p##posv_compact_info(char layout, char uplo, int n, int nrhs, T *ap,   \
                                      int ldap, T *bp, int ldbp, int V, int nm,         \
                                      int *info);                                       \
    int oracle_##p##sytrfnp_compact_info(char layout, char uplo, int n, T *ap,          \
                                         int ldap, int V, int nm, int *info);           \
    int oracle_##p##sysvnp_compact_info(char layout, char uplo, int n, int nrhs, T *ap, \
                                        int ldap, T *bp, int ldbp, int V, int nm,       \
                                        int *info);                                     \
    int oracle_##p##gels_compact_info(char layout, char trans, int m, int n, int nrhs,  \
                                      T *ap, int ldap, T *bp, int ldbp, T *taup, int V, \
                                      int nm, int *info);

ORACLE_DECLS(double, d)
ORACLE_DECLS(float, s)
#undef ORACLE_DECLS

/* LAPACK dsposv restated per matrix on the compact layout (no reference
 * counterpart; LAPACK 3.12 dsposv.f): single-precision Cholesky, refinement with
 * double residuals, ITERMAX = 30, convergence when ||r||_inf <= ||x||_inf *
 * ||A||_inf * eps * sqrt(n) for every column, fallback to the double-precision
 * dposv otherwise (-2 when B, A or a residual overflows single precision, -3
 * when the single factor fails, -31 without convergence). xp receives X; ap
 * keeps A unless the fallback ran, in which case it holds the double Cholesky
 * factor. iter and info are per matrix. */
int oracle_dsposv_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                          const double *bp, int ldbp, double *xp, int ldxp, void *work,
                          size_t lwork, int *iter, int *info, int V, int nm);

/* Dense <-> compact conversion (the reference's test utilities pack_compact /
 * unpack_compact, tests/test_compact_util.hpp:475-531, with the dense layout,
 * lda and batch stride free); host pointers, the library's cbk_?pack_compact
 * family and arguments. The gepack forms take an array of nm matrix pointers
 * in the compact layout's own layout and pad with the identity. */
int oracle_cbk_dpack_compact(char layout, int rows, int cols, const double *a, char alayout,
                             int lda, size_t stride_a, double *ap, int ldap, double pad_diag,
                             int V, int nm);
int oracle_cbk_spack_compact(char layout, int rows, int cols, const float *a, char alayout,
                             int lda, size_t stride_a, float *ap, int ldap, float pad_diag,
                             int V, int nm);
int oracle_cbk_dunpack_compact(char layout, int rows, int cols, const double *ap, int ldap,
                               double *a, char alayout, int lda, size_t stride_a, int V, int nm);
int oracle_cbk_sunpack_compact(char layout, int rows, int cols, const float *ap, int ldap,
                               float *a, char alayout, int lda, size_t stride_a, int V, int nm);
int oracle_cbk_dgepack_compact(char layout, int rows, int cols, const double *const *a, int lda,
                               double *ap, int ldap, int V, int nm);
int oracle_cbk_sgepack_compact(char layout, int rows, int cols, const float *const *a, int lda,
                               float *ap, int ldap, int V, int nm);
int oracle_cbk_dgeunpack_compact(char layout, int rows, int cols, double *const *a, int lda,
                                 const double *ap, int ldap, int V, int nm);
int oracle_cbk_sgeunpack_compact(char layout, int rows, int cols, float *const *a, int lda,
                                 const float *ap, int ldap, int V, int nm);

/* dsposv's algorithm with the unpivoted LDL^T of sysvnp in both precisions
 * (no reference counterpart, no LAPACK routine): the single factor also fails
 * on a non-finite D; the fallback is the double dsysvnp, after which ap holds
 * the double LDL^T factor. The library's dssysvnp_compact. */
int oracle_dssysvnp_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                            const double *bp, int ldbp, double *xp, int ldxp, void *work,
                            size_t lwork, int *iter, int *info, int V, int nm);

/* Mixed-precision least squares on the compact layout (no reference
 * counterpart): single precision QR as the preconditioner of refinement on the
 * normal equations with double residuals, double gels as the fallback; the
 * library's dsgels_compact. Shapes: trans 'N', n <= m <= 64, n + nrhs <= 64. */
int oracle_dsgels_compact(char layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
                          const double *bp, int ldbp, double *xp, int ldxp, int *iter,
                          int *info, int V, int nm);

/* One uneven batch (no reference counterpart): problem k is gels_compact on
 * the single column-major matrix A_k (m x n, lda) at ap + a_off and B_k
 * (max(m,n) x nrhs, ldb) at bp + b_off, i.e. the reference's
 * gels_compact_group (src/cbk_gels_compact.hpp:84-146) with a one-matrix
 * view; info[k] as the _info variants. Same struct layout as the library's
 * cbk_gels_problem. */
typedef struct oracle_gels_problem {
    int m, n, nrhs, lda, ldb;
    long long a_off, b_off;
} oracle_gels_problem;
int oracle_dgels_batch(char trans, int nprob, const oracle_gels_problem *desc, double *ap,
                       double *bp, int *info);
int oracle_sgels_batch(char trans, int nprob, const oracle_gels_problem *desc, float *ap,
                       float *bp, int *info);

/* The library's dsgels_batch restated per problem (no reference
 * counterpart): single precision gels, then double refinement of its
 * solution on the seminormal equations (dsgels_compact's per-column scheme),
 * the double gels in place on A and X := B for the problems that do not
 * converge (iter = -31) or that have no single precision path (-1). X is
 * laid out as B; iter and info per problem. */
int oracle_dsgels_batch(char trans, int nprob, const oracle_gels_problem *desc, double *ap,
                        const double *bp, double *xp, int *iter, int *info);

#ifdef __cplusplus
}
#endif

#endif /* CBK_ORACLE_H */
