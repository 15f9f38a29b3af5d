#ifndef CBK_H
#define CBK_H

/* cbk.h -- compact batch kernels on one MI355X (gfx950).
 *
 * The drop-in for the portable C API of ivan-pi/cbk (reference include/cbk.h):
 * the same 22 entry points, names, argument lists, option characters,
 * LAPACK-style -j argument codes and quick-return rules, over DEVICE-resident
 * compact buffers. Each declaration below cites the reference declaration it
 * replaces.
 *
 * Compact layout (reference include/cbk.h:22-29); group g = idx/V, slot v = idx%V:
 *   A_v(i,j)  = ap [ g*ldap*ncol*V + (j*ldap + i)*V + v ]   (column-major)
 *   tau_v(kk) = taup[ g*k*V         +  kk*V          + v ]
 *   B_v(i,j)  = bp [ g*ldbp*nrhs*V  + (j*ldbp + i)*V + v ]
 * row-major swaps i and j. V is 2, 4, 8, 16, 32 or 64 (the reference accepts
 * 2..16); V = 64 gives one 512-byte coalesced access per element per wave.
 * Buffers hold ceil(nm/V) whole groups; padded slots are computed like real
 * ones (fill them with the identity, as the reference expects).
 *
 * Differences from the reference, all on the host side of the boundary:
 *   - every pointer is a device pointer (hipMalloc / torch CUDA tensor);
 *   - calls are asynchronous, ordered on the stream set by cbk_set_stream
 *     (cbk_ext.h; default: the null stream). Argument errors return -j before
 *     anything is launched; a launch failure returns CBK_ERR_LAUNCH;
 *   - per-matrix info, mixed precision and uneven batches are extra entry
 *     points in cbk_ext.h.
 * The numerical results are those of the reference kernels: the same
 * operations in the same order, one GPU lane (or a team of lanes) per matrix.
 */

#define CBK_VERSION_MAJOR 0
#define CBK_VERSION_MINOR 2
#define CBK_VERSION_PATCH 0
#define CBK_VERSION (CBK_VERSION_MAJOR * 10000 + CBK_VERSION_MINOR * 100 + CBK_VERSION_PATCH)

/* returned when the HIP runtime rejects a launch (outside every -j range) */
#define CBK_ERR_LAUNCH (-1000)

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/cbk.h:73-75 */
int cbk_get_version(void);

/* QR factorization A = QR; reference include/cbk.h:77-96.
 * -1 layout -2 m -3 n -5 ldap -7 V -8 nm */
int dgeqrf_compact(char layout, int m, int n, double *ap, int ldap, double *taup, int V,
                   int nm);
int sgeqrf_compact(char layout, int m, int n, float *ap, int ldap, float *taup, int V,
                   int nm);

/* C := op(Q) C or C op(Q); reference include/cbk.h:98-130.
 * -1 layout -2 side -3 trans (T or N only) -4 m -5 n -6 k -8 ldap -11 ldcp -12 V -13 nm */
int dormqr_compact(char layout, char side, char trans, int m, int n, int k,
                   const double *ap, int ldap, const double *taup, double *cp, int ldcp,
                   int V, int nm);
int sormqr_compact(char layout, char side, char trans, int m, int n, int k,
                   const float *ap, int ldap, const float *taup, float *cp, int ldcp,
                   int V, int nm);

/* Form Q in place; reference include/cbk.h:132-157. k = 0 still writes I.
 * -1 layout -2 m -3 n (<0 or >m) -4 k (<0 or >n) -6 ldap -8 V -9 nm */
int dorgqr_compact(char layout, int m, int n, int k, double *ap, int ldap,
                   const double *taup, int V, int nm);
int sorgqr_compact(char layout, int m, int n, int k, float *ap, int ldap,
                   const float *taup, int V, int nm);

/* Cholesky A = L L^T or U^T U; reference include/cbk.h:159-181.
 * -1 layout -2 uplo -3 n -5 ldap -6 V -7 nm */
int dpotrf_compact(char layout, char uplo, int n, double *ap, int ldap, int V, int nm);
int spotrf_compact(char layout, char uplo, int n, float *ap, int ldap, int V, int nm);

/* Solve from a Cholesky factor; reference include/cbk.h:183-208.
 * -1 layout -2 uplo -3 n -4 nrhs -6 ldap -8 ldbp -9 V -10 nm */
int dpotrs_compact(char layout, char uplo, int n, int nrhs, const double *ap, int ldap,
                   double *bp, int ldbp, int V, int nm);
int spotrs_compact(char layout, char uplo, int n, int nrhs, const float *ap, int ldap,
                   float *bp, int ldbp, int V, int nm);

/* Fused Cholesky factor + solve; reference include/cbk.h:210-222. nrhs = 0
 * still factors. Codes as ?potrs_compact. */
int dposv_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                  double *bp, int ldbp, int V, int nm);
int sposv_compact(char layout, char uplo, int n, int nrhs, float *ap, int ldap, float *bp,
                  int ldbp, int V, int nm);

/* Unpivoted LDL^T (upper: A = U^T D U); reference include/cbk.h:224-249.
 * -1 layout -2 uplo -3 n -5 ldap -6 V -7 nm */
int dsytrfnp_compact(char layout, char uplo, int n, double *ap, int ldap, int V, int nm);
int ssytrfnp_compact(char layout, char uplo, int n, float *ap, int ldap, int V, int nm);

/* Solve from an unpivoted LDL^T factor; reference include/cbk.h:251-274.
 * Codes as ?potrs_compact. */
int dsytrsnp_compact(char layout, char uplo, int n, int nrhs, const double *ap, int ldap,
                     double *bp, int ldbp, int V, int nm);
int ssytrsnp_compact(char layout, char uplo, int n, int nrhs, const float *ap, int ldap,
                     float *bp, int ldbp, int V, int nm);

/* Fused LDL^T factor + solve; reference include/cbk.h:276-288. */
int dsysvnp_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                    double *bp, int ldbp, int V, int nm);
int ssysvnp_compact(char layout, char uplo, int n, int nrhs, float *ap, int ldap,
                    float *bp, int ldbp, int V, int nm);

/* Triangular solve op(A) X = alpha B or X op(A) = alpha B; reference
 * include/cbk.h:290-318. alpha = 0 sets B := 0 without reading A.
 * -1 layout -2 side -3 uplo -4 transa -5 diag -6 m -7 n -10 ldap -12 ldbp -13 V -14 nm */
int dtrsm_compact(char layout, char side, char uplo, char transa, char diag, int m, int n,
                  double alpha, const double *ap, int ldap, double *bp, int ldbp, int V,
                  int nm);
int strsm_compact(char layout, char side, char uplo, char transa, char diag, int m, int n,
                  float alpha, const float *ap, int ldap, float *bp, int ldbp, int V,
                  int nm);

/* Least squares / minimum norm op(A) X = B; reference include/cbk.h:320-351.
 * min(m,n) = 0 sets B := 0; nrhs = 0 computes nothing.
 * -1 layout -2 trans -3 m -4 n -5 nrhs -7 ldap -9 ldbp -11 V -12 nm */
int dgels_compact(char layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
                  double *bp, int ldbp, double *taup, int V, int nm);
int sgels_compact(char layout, char trans, int m, int n, int nrhs, float *ap, int ldap,
                  float *bp, int ldbp, float *taup, int V, int nm);

#ifdef __cplusplus
}
#endif

#endif /* CBK_H */
