#ifndef CBK_EXT_H
#define CBK_EXT_H

/* cbk_ext.h -- entry points with no counterpart in the reference's cbk.h.
 *
 * Same conventions as cbk.h: device pointers, compact layout, V in
 * {2,4,8,16,32,64}, -j for the first illegal argument in argument order,
 * CBK_ERR_LAUNCH when a launch fails, asynchronous on the cbk stream.
 */

#include "cbk.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stream ------------------------------------------------------------
 * The HIP stream (a hipStream_t passed as void*) that this thread's calls
 * launch on; NULL is the null stream. Thread-local: each host thread keeps its
 * own. cbk_synchronize waits for that stream. */
int cbk_set_stream(void *stream);
void *cbk_get_stream(void);
int cbk_synchronize(void);

/* the HIP runtime's message for the last CBK_ERR_LAUNCH of this thread */
const char *cbk_last_error_string(void);

/* ---- kernel selection ---------------------------------------------------
 * CBK_POLICY_AUTO (default) lets the dispatcher pick the fastest kernel for the
 * shape (team kernels for the tuned orders, the lane kernel otherwise);
 * CBK_POLICY_LANE forces the general one-lane-per-matrix kernels, whose
 * results are bit-identical to the reference kernels for every routine.
 * Thread-local, like the stream. Returns the previous policy. */
#define CBK_POLICY_AUTO 0
#define CBK_POLICY_LANE 1
int cbk_set_policy(int policy);

/* ---- per-matrix info ---------------------------------------------------
 * The plain routine's arguments plus `info`, a device int[nm]: 0, or the LAPACK
 * code of that matrix. The factors are bit-identical to the plain call.
 *   potrf/posv   : j (1-based) of the first pivot that is <= 0 or NaN (dpotrf);
 *                  posv leaves that matrix's B unmodified (dposv skips dpotrs)
 *   sytrfnp/sysvnp: j of the first exact zero pivot D(j); sysvnp leaves B
 *   gels         : i of the first exact zero diagonal R(i,i) (dgels/dtrtrs);
 *                  least squares leaves Q^T B in B, minimum norm leaves B
 * Argument codes: those of the plain routine; info itself is not inspected. */
int dpotrf_compact_info(char layout, char uplo, int n, double *ap, int ldap, int V, int nm,
                        int *info);
int spotrf_compact_info(char layout, char uplo, int n, float *ap, int ldap, int V, int nm,
                        int *info);
int dposv_compact_info(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                       double *bp, int ldbp, int V, int nm, int *info);
int sposv_compact_info(char layout, char uplo, int n, int nrhs, float *ap, int ldap,
                       float *bp, int ldbp, int V, int nm, int *info);
int dsytrfnp_compact_info(char layout, char uplo, int n, double *ap, int ldap, int V,
                          int nm, int *info);
int ssytrfnp_compact_info(char layout, char uplo, int n, float *ap, int ldap, int V,
                          int nm, int *info);
int dsysvnp_compact_info(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                         double *bp, int ldbp, int V, int nm, int *info);
int ssysvnp_compact_info(char layout, char uplo, int n, int nrhs, float *ap, int ldap,
                         float *bp, int ldbp, int V, int nm, int *info);
int dgels_compact_info(char layout, char trans, int m, int n, int nrhs, double *ap,
                       int ldap, double *bp, int ldbp, double *taup, int V, int nm,
                       int *info);
int sgels_compact_info(char layout, char trans, int m, int n, int nrhs, float *ap,
                       int ldap, float *bp, int ldbp, float *taup, int V, int nm,
                       int *info);

/* ---- mixed precision ----------------------------------------------------
 * dsposv_compact: the compact, per-matrix form of LAPACK dsposv. Factor in
 * FP32, solve, then refine X with residuals r = B - A X and updates X += d
 * computed in FP64 until every column satisfies
 *     ||r||_inf <= ||x||_inf * ||A||_inf * eps * sqrt(n)   (eps = 2^-53)
 * or itermax = 30 steps pass; a matrix whose FP32 factor fails, whose data
 * overflow FP32, or whose refinement does not converge is solved by the FP64
 * Cholesky instead. Per matrix (device int[nm] each):
 *     iter  >= 0 refinement steps taken (A untouched);
 *           -2 overflow, -3 FP32 factor failed, -31 no convergence: the FP64
 *           path ran and A holds its Cholesky factor;
 *     info  0, or the FP64 dpotrf code of a matrix that is not SPD (X then
 *           unwritten).
 * B is read-only; X (n x nrhs, ldxp, same layout) receives the solution.
 * work: device scratch of dsposv_compact_worksize(...) bytes (0 -- and work
 * may be NULL -- when the fused kernel serves the shape: n in {16, 24, 32},
 * nrhs <= 4, V >= 4; see that function).
 * Codes: -1 layout -2 uplo -3 n -4 nrhs -6 ldap -8 ldbp -10 ldxp -12 lwork
 *        -15 V -16 nm */
int dsposv_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                   const double *bp, int ldbp, double *xp, int ldxp, void *work,
                   size_t lwork, int *iter, int *info, int V, int nm);

/* bytes of workspace dsposv_compact needs for this call shape (0 if none) */
size_t dsposv_compact_worksize(char layout, char uplo, int n, int nrhs, int V, int nm);

#ifdef __cplusplus
}
#endif

#endif /* CBK_EXT_H */
