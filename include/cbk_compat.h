#ifndef CBK_COMPAT_H
#define CBK_COMPAT_H

/* cbk_compat.h -- the MKL-style interface of the compact routines, on the GPU.
 *
 * The reference's include/cbk_compat.h:99-250 (implemented in its
 * src/cbk_compat.cpp:43-250) gives MKL compact users the routines under
 * MKL's own argument conventions: MKL_LAYOUT / MKL_UPLO / MKL_SIDE /
 * MKL_TRANSPOSE / MKL_DIAG enums, an MKL_COMPACT_PACK format that fixes the
 * interleave width V, work / lwork where LAPACK has them, and a scalar info.
 * The same 22 routines are declared here, argument for argument, as a thin
 * adapter over the include/cbk.h entry points of this library (same kernels,
 * same results).
 *
 * Conventions (as the reference's adapter, except where said):
 *   - ap, bp, cp, taup and work are DEVICE pointers (compact buffers in HBM);
 *     info is a HOST pointer to one status word. Calls are asynchronous on
 *     the cbk stream (cbk_set_stream), like the rest of the library.
 *   - The enum arguments take MKL's values (mkl_types.h): a caller that
 *     includes MKL's headers passes MKL_ROW_MAJOR, MKL_LOWER, ... unchanged;
 *     the CBK_MKL_* constants below are the same numbers. MKL_INT is int
 *     here (MKL's LP64 interface).
 *   - format: MKL_COMPACT_SSE / AVX / AVX512 give V = 16 / 32 / 64 bytes over
 *     sizeof(scalar) (FP64: V = 2 / 4 / 8; FP32: 4 / 8 / 16), as MKL packs
 *     them. CBK_COMPACT_V64 (a value MKL does not use) selects V = 64, the
 *     layout this library's kernels are tuned for (one wavefront lane per
 *     matrix). info = -1 for any other format, and nothing runs.
 *   - Option decoding: layout MKL_ROW_MAJOR is row-major, anything else
 *     column-major; uplo MKL_UPPER is upper, anything else lower; side
 *     MKL_LEFT (trsm) or 'L' / 'l' (ormqr) is left, anything else right; a
 *     transpose is MKL_TRANS / MKL_CONJTRANS (trsm) or 'T', 't', 'C', 'c'
 *     (ormqr, gels), anything else none; diag MKL_UNIT is unit.
 *   - info (host): 0 on success, -1 for an unrecognized format. Where the
 *     reference's adapter runs its kernels on arguments it does not check,
 *     this one forwards to the checked cbk.h routine: an inconsistent
 *     dimension or leading dimension is reported as that routine's -j code
 *     (its argument order) and nothing runs; a failed launch as
 *     CBK_ERR_LAUNCH. trsm has no info; such a call is then a no-op.
 *   - Workspace: geqrf, ormqr and orgqr take work / lwork and need none; with
 *     lwork = -1 they write 1 to work[0] (on the device) and return. gels uses
 *     work as its tau scratch (one compact tau slot per matrix, left holding
 *     the factorization's tau); lwork = -1 writes max(1, min(m,n) V
 *     ceil(nm/V)) to work[0].
 *   - Quick returns as the reference: nrhs = 0 still factors for posv and
 *     sysvnp, orgqr with k = 0 writes I, trsm with alpha = 0 sets B := 0,
 *     gels with min(m,n) = 0 sets B := 0.
 */

#ifdef __cplusplus
extern "C" {
#endif

/* MKL's enum values (mkl_types.h) */
#define CBK_MKL_ROW_MAJOR 101
#define CBK_MKL_COL_MAJOR 102
#define CBK_MKL_NOTRANS 111
#define CBK_MKL_TRANS 112
#define CBK_MKL_CONJTRANS 113
#define CBK_MKL_UPPER 121
#define CBK_MKL_LOWER 122
#define CBK_MKL_NONUNIT 131
#define CBK_MKL_UNIT 132
#define CBK_MKL_LEFT 141
#define CBK_MKL_RIGHT 142
#define CBK_MKL_COMPACT_SSE 181
#define CBK_MKL_COMPACT_AVX 182
#define CBK_MKL_COMPACT_AVX512 183
/* no MKL counterpart: V = 64 */
#define CBK_COMPACT_V64 190

/* reference include/cbk_compat.h:99-105 (src/cbk_compat.cpp:58-71) */
void cbk_dgeqrf_compact(int layout, int m, int n, double *ap, int ldap, double *taup,
                        double *work, int lwork, int *info, int format, int nm);
void cbk_sgeqrf_compact(int layout, int m, int n, float *ap, int ldap, float *taup, float *work,
                        int lwork, int *info, int format, int nm);

/* reference include/cbk_compat.h:110-118 (src/cbk_compat.cpp:73-88) */
void cbk_dormqr_compact(int layout, char side, char trans, int m, int n, int k,
                        const double *ap, int ldap, const double *taup, double *cp, int ldcp,
                        double *work, int lwork, int *info, int format, int nm);
void cbk_sormqr_compact(int layout, char side, char trans, int m, int n, int k, const float *ap,
                        int ldap, const float *taup, float *cp, int ldcp, float *work,
                        int lwork, int *info, int format, int nm);

/* reference include/cbk_compat.h:125-131 (src/cbk_compat.cpp:90-106) */
void cbk_dorgqr_compact(int layout, int m, int n, int k, double *ap, int ldap,
                        const double *taup, double *work, int lwork, int *info, int format,
                        int nm);
void cbk_sorgqr_compact(int layout, int m, int n, int k, float *ap, int ldap, const float *taup,
                        float *work, int lwork, int *info, int format, int nm);

/* reference include/cbk_compat.h:138-142 (src/cbk_compat.cpp:108-121) */
void cbk_dpotrf_compact(int layout, int uplo, int n, double *ap, int ldap, int *info,
                        int format, int nm);
void cbk_spotrf_compact(int layout, int uplo, int n, float *ap, int ldap, int *info, int format,
                        int nm);

/* reference include/cbk_compat.h:150-156 (src/cbk_compat.cpp:123-137) */
void cbk_dpotrs_compact(int layout, int uplo, int n, int nrhs, const double *ap, int ldap,
                        double *bp, int ldbp, int *info, int format, int nm);
void cbk_spotrs_compact(int layout, int uplo, int n, int nrhs, const float *ap, int ldap,
                        float *bp, int ldbp, int *info, int format, int nm);

/* reference include/cbk_compat.h:166-172 (src/cbk_compat.cpp:139-154) */
void cbk_dposv_compact(int layout, int uplo, int n, int nrhs, double *ap, int ldap, double *bp,
                       int ldbp, int *info, int format, int nm);
void cbk_sposv_compact(int layout, int uplo, int n, int nrhs, float *ap, int ldap, float *bp,
                       int ldbp, int *info, int format, int nm);

/* reference include/cbk_compat.h:182-188 (src/cbk_compat.cpp:156-169) */
void cbk_dsytrfnp_compact(int layout, int uplo, int n, double *ap, int ldap, int *info,
                          int format, int nm);
void cbk_ssytrfnp_compact(int layout, int uplo, int n, float *ap, int ldap, int *info,
                          int format, int nm);

/* reference include/cbk_compat.h:195-201 (src/cbk_compat.cpp:171-185) */
void cbk_dsytrsnp_compact(int layout, int uplo, int n, int nrhs, const double *ap, int ldap,
                          double *bp, int ldbp, int *info, int format, int nm);
void cbk_ssytrsnp_compact(int layout, int uplo, int n, int nrhs, const float *ap, int ldap,
                          float *bp, int ldbp, int *info, int format, int nm);

/* reference include/cbk_compat.h:211-217 (src/cbk_compat.cpp:187-203) */
void cbk_dsysvnp_compact(int layout, int uplo, int n, int nrhs, double *ap, int ldap,
                         double *bp, int ldbp, int *info, int format, int nm);
void cbk_ssysvnp_compact(int layout, int uplo, int n, int nrhs, float *ap, int ldap, float *bp,
                         int ldbp, int *info, int format, int nm);

/* reference include/cbk_compat.h:225-233 (src/cbk_compat.cpp:205-223) */
void cbk_dtrsm_compact(int layout, int side, int uplo, int transa, int diag, int m, int n,
                       double alpha, const double *ap, int ldap, double *bp, int ldbp,
                       int format, int nm);
void cbk_strsm_compact(int layout, int side, int uplo, int transa, int diag, int m, int n,
                       float alpha, const float *ap, int ldap, float *bp, int ldbp, int format,
                       int nm);

/* reference include/cbk_compat.h:244-250 (src/cbk_compat.cpp:225-250) */
void cbk_dgels_compact(int layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
                       double *bp, int ldbp, double *work, int lwork, int *info, int format,
                       int nm);
void cbk_sgels_compact(int layout, char trans, int m, int n, int nrhs, float *ap, int ldap,
                       float *bp, int ldbp, float *work, int lwork, int *info, int format,
                       int nm);

/* V of a format for an FP64 (fp64 != 0) or FP32 scalar; 0 if unrecognized
 * (the reference's cbk::detail::vlen_for_format, include/cbk_compat.h:276-286) */
int cbk_compact_vlen(int format, int fp64);

#ifdef __cplusplus
}
#endif

#endif /* CBK_COMPAT_H */
