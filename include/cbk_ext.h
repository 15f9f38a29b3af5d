#ifndef CBK_EXT_H
#define CBK_EXT_H

/* cbk_ext.h -- entry points with no counterpart in the reference's cbk.h.
 *
 * Same conventions as cbk.h: device pointers, compact layout, V in
 * {2,4,8,16,32,64}, -j for the first illegal argument in argument order,
 * CBK_ERR_LAUNCH when a launch fails, asynchronous on the cbk stream.
 */

#include "cbk.h"

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

#ifdef __cplusplus
}
#endif

#endif /* CBK_EXT_H */
