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
 *     info  0, or the FP64 dpotrf code of a matrix that is not SPD (X of
 *           that matrix then unspecified: as LAPACK dsposv, it may hold an
 *           unfinished iterate).
 * B is read-only; X (n x nrhs, ldxp, same layout) receives the solution.
 * iter and info are required when nm > 0.
 * work: device scratch of dsposv_compact_worksize(...) bytes (0 -- and work
 * may be NULL -- when a fused kernel serves the call; see that function).
 * Codes: -1 layout -2 uplo -3 n -4 nrhs -6 ldap -8 ldbp -10 ldxp -12 lwork
 *        -13 iter NULL -14 info NULL -15 V -16 nm */
int dsposv_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                   const double *bp, int ldbp, double *xp, int ldxp, void *work,
                   size_t lwork, int *iter, int *info, int V, int nm);

/* bytes of workspace dsposv_compact needs for this call (0 if none): 0 when
 * a fused kernel serves it (1 <= n <= 64, V >= 4, and leading
 * dimensions small enough for the kernels' 32-bit offsets -- below about
 * 2^31 / (64 * 8 * n) at V = 64), else groups * V * (n*n + n*nrhs) * 4 */
size_t dsposv_compact_worksize(char layout, char uplo, int n, int nrhs, int ldap, int ldbp,
                               int ldxp, int V, int nm);

/* dssysvnp_compact: dsposv_compact's contract with the unpivoted LDL^T of
 * sysvnp_compact in place of the Cholesky factor. No LAPACK routine has this
 * contract; the reference's sysvnp runs in one precision throughout
 * (include/cbk.h:274-279).
 * The FP32 LDL^T fails (iter = -3) on a zero or non-finite D; the FP64
 * fallback is dsysvnp_compact, A then holding its factor (unit L strictly
 * below the diagonal, D on it, bit-identical to dsytrfnp_compact), and info
 * its code: j (1-based) of the first exact zero pivot D(j) of the FP64
 * factor (X of that matrix unspecified). Same codes and workspace as
 * dsposv_compact; dssysvnp_compact_worksize equals dsposv_compact_worksize. */
int dssysvnp_compact(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                     const double *bp, int ldbp, double *xp, int ldxp, void *work,
                     size_t lwork, int *iter, int *info, int V, int nm);
size_t dssysvnp_compact_worksize(char layout, char uplo, int n, int nrhs, int ldap, int ldbp,
                                 int ldxp, int V, int nm);

/* ---- no device allocation ------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: dsposv_compact_noalloc is dsposv_compact with the
 * workspace taken from a static device array of the library instead of a
 * caller's buffer (CBK_NOALLOC_SCRATCH_BYTES per GPU, held by each GPU the
 * library runs on until the process ends). It calls no hipMalloc or hipFree.
 * A call that would need more than that (the lane path: groups * V *
 * (n*n + n*nrhs) * 4 bytes, e.g. n = 20, nrhs = 5 up to 16777 matrices)
 * returns CBK_ERR_SCRATCH before any launch; shapes the fused kernels serve
 * need none. Every call on a GPU shares that GPU's array: calls that overlap
 * on one GPU (several streams or host threads at once) race and must be
 * ordered by the caller. The array holds no state between calls.
 * Codes: those of dsposv_compact for the same arguments (-11 iter NULL,
 * -12 info NULL, -13 V, -14 nm, as their positions here), CBK_ERR_SCRATCH. */
#define CBK_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
#define CBK_ERR_SCRATCH (-1001)
int dsposv_compact_noalloc(char layout, char uplo, int n, int nrhs, double *ap, int ldap,
                           const double *bp, int ldbp, double *xp, int ldxp, int *iter,
                           int *info, int V, int nm);
/* the static array's size, CBK_NOALLOC_SCRATCH_BYTES */
size_t cbk_noalloc_scratch_bytes(void);

/* dsgels_compact: mixed-precision least squares min ||B - A X|| for every
 * matrix (trans 'N', n <= m <= 64, n + nrhs <= 64, full column rank). No LAPACK
 * routine has this contract; it follows dsposv's. The FP32 Householder QR of
 * A preconditions iterative refinement on the normal equations, whose
 * residuals r = B - A X and A^T r are computed in FP64 (X_0 = 0, then X += d
 * with R^T R d = A^T r solved in FP32), per column until
 *     ||d||_inf <= ||x||_inf * eps * sqrt(n)      (eps = 2^-53)
 * or until the correction stops halving once below ||x||_inf * 2^-40 (the
 * rounding floor of A^T r for an inconsistent system). Per matrix:
 *     iter  >= 0 refinement steps (the most any column took; A untouched);
 *           -1 the shape has no FP32 kernel (n > 32, or V < 4; the FP32 QR
 *           runs in the smallest of the 16/32/64-row by 8/16/32-column
 *           buckets that holds m x n) and -31 refinement failed (a zero or
 *           non-finite diagonal of the FP32 R, a correction that does not
 *           contract, or 30 steps): the FP64 gels ran and A holds its QR
 *           (reflectors below the diagonal, as gels_compact);
 *     info  0, or (FP64 path) the 1-based index of the first exact zero of
 *           R's diagonal, X then holding (Q^T B)(0:n).
 * B (m x nrhs, ldbp) is read only; X (n x nrhs, ldxp, same layout) receives
 * the solution. iter is required (device int[nm]) when nm > 0; info may be
 * NULL. Matrices that refine never reach the FP64 pass, whose launch then
 * only scans iter.
 * Codes: -1 layout -2 trans -3 m -4 n -5 nrhs -7 ldap -9 ldbp -11 ldxp
 *        -12 iter NULL -14 V -15 nm */
int dsgels_compact(char layout, char trans, int m, int n, int nrhs, double *ap, int ldap,
                   const double *bp, int ldbp, double *xp, int ldxp, int *iter, int *info,
                   int V, int nm);

/* ---- uneven batches -------------------------------------------------------
 * No reference counterpart. A batch of independent least-squares problems
 * min || B_k - op(A_k) X_k || of different sizes, each stored column-major on
 * its own (not compact): problem k's A is m x n at ap + a_off (lda >=
 * max(1, m)), its B is max(m, n) x nrhs at bp + b_off (ldb >= max(1, m, n)).
 * Per problem the result is that of gels_compact on that one matrix: A
 * receives its QR (LQ) factorization, the first n (m) rows of B the solution,
 * for a least-squares problem the remaining rows Q^T B; info[k] = 0, or the
 * 1-based index of the first exact zero on R's diagonal (LAPACK dgels), B
 * then left as Q^T B. Offsets are in elements.
 *
 * Usage: plan once per batch shape on the host, query the workspace, then
 * call dgels_batch / sgels_batch as often as wanted. The plan sorts the
 * problems into size classes, orders each class longest first, picks each
 * class's kernel from the tuning table (cbk_autotune), and records all of
 * it as one hipGraph: a counter-clearing kernel followed by one persistent
 * work-queue kernel per class, the classes running concurrently. The graph
 * is launched on the caller's stream; it is re-recorded only when the
 * pointers passed to the call change. The first call with a given work
 * buffer copies the plan's problem table into it; keep work untouched
 * between calls. A plan is not thread-safe: one host thread at a time. */
typedef struct cbk_gels_problem {
    int m, n, nrhs, lda, ldb;
    long long a_off, b_off;
} cbk_gels_problem;

typedef struct cbk_batch_plan *cbk_batch_plan_t;

/* prec 'D' (dgels_batch), 'S' (sgels_batch) or 'M' (dsgels_batch); trans
 * 'N', 'T' or 'C' as gels_compact. Codes: -1 prec, -2 trans, -3 nprob < 0,
 * -4 desc NULL, -5 a descriptor is invalid (its index goes to *bad if bad is
 * not NULL), -6 plan NULL. On success *plan is a new plan (host memory
 * only). */
int cbk_gels_batch_plan(char prec, char trans, int nprob, const cbk_gels_problem *desc,
                        cbk_batch_plan_t *plan, int *bad);

/* device bytes of workspace the plan's calls need */
size_t cbk_batch_plan_worksize(cbk_batch_plan_t plan);

/* problems per kernel class (count[CBK_BATCH_NCLASS]); returns the number
 * of classes. Classes: 0-2 one problem per 16-lane team (n + nrhs <= 16;
 * rows <= 16, 32, 64), 3-4 the wavefront kernels (n + nrhs <= 64; rows
 * <= 32, 64; two problems per wavefront when n + nrhs <= 32), 5 one problem per wavefront with A and B in separate registers
 * (n + nrhs > 64, rows <= 64, nrhs <= 64), 6 the general one-lane-per-problem
 * kernel (transposed or underdetermined problems, more than 64 rows, empty
 * problems). */
#define CBK_BATCH_NCLASS 7
int cbk_batch_plan_classes(cbk_batch_plan_t plan, int *count);

/* Codes: -1 plan NULL or of the other precision, -2 ap, -3 bp, -5 work NULL
 * with a nonzero worksize, -6 lwork too small; CBK_ERR_LAUNCH on a HIP error. */
int dgels_batch(cbk_batch_plan_t plan, double *ap, double *bp, int *info, void *work,
                size_t lwork);
int sgels_batch(cbk_batch_plan_t plan, float *ap, float *bp, int *info, void *work,
                size_t lwork);

/* dsgels_batch: the batch in mixed precision (a plan made with prec 'M').
 * A problem with trans 'N', 1 <= n <= m <= 64, n <= 32 and 1 <= nrhs <= 8 is
 * solved by the FP32 batch gels (on FP32 copies of A and B in the workspace)
 * and FP64 refinement of that solution on the seminormal equations, the
 * per-column scheme of dsgels_compact: r = B - A X and A^T r in FP64,
 * R^T R D = A^T r with the FP32 R (through R^-1, formed once), X += D, until
 * ||D||_inf <= ||X||_inf eps sqrt(n) (eps = 2^-53) or the correction stops
 * halving once below ||X||_inf 2^-40. A
 * problem that does not converge (a correction that does not contract, a
 * NaN, a zero or non-finite diagonal of the FP32 R, or 30 steps), and every
 * problem outside that range, is solved by the FP64 batch gels instead,
 * in place on A and X (those outside the range concurrently with the FP32
 * work). Everything runs as one recorded graph.
 * A: read (left as it was for a refined problem; the FP64 QR for the rest,
 * as dgels_batch). B: read only. X (xp): laid out as B (the same b_off and
 * ldb, so a buffer the size of B); its first n rows of each problem receive
 * the solution (the rows below are unspecified). Per problem (device ints):
 *     iter[k] >= 0 refinement steps (the most any column took);
 *             -1 outside the FP32 path's range, -31 refinement failed:
 *             the FP64 path ran;
 *     info[k] 0, or dgels_batch's code for a problem the FP64 path solved.
 * Codes: -1 plan NULL or not a mixed plan, -2 ap, -3 bp, -4 xp, -5 iter,
 * -6 info NULL, -7 work NULL with a nonzero worksize, -8 lwork too small;
 * CBK_ERR_LAUNCH on a HIP error. The workspace holds the FP32 copies of A
 * and B (as far as the largest offset of a refined problem reaches). */
int dsgels_batch(cbk_batch_plan_t plan, double *ap, const double *bp, double *xp, int *iter,
                 int *info, void *work, size_t lwork);

int cbk_batch_plan_destroy(cbk_batch_plan_t plan);

/* ---- tuning ----------------------------------------------------------------
 * The batch plans choose, per row bucket (rows <= 16, <= 32, <= 64), between
 * the team kernel (one problem per 16 lanes, n + nrhs <= 16 columns) and the
 * wavefront kernel (one problem per 64 lanes) by the largest column count
 * narrow_max_cols[b] the team kernel should take, and size each persistent
 * grid to blocks_per_cu[c] blocks of 256 threads per compute unit.
 * cbk_autotune measures both on the current device with synthetic problems
 * (allocating and freeing its own device memory), then times each class's
 * kernel at a typical shape with its persistent grid at half, once and twice
 * its occupancy per compute unit (blocks_per_cu: the fastest, 0 for the
 * occupancy itself unless another is more than 2 % faster), and installs the
 * result for plans created afterwards on this device. */
typedef struct cbk_tuning {
    int narrow_max_cols[3];
    int blocks_per_cu[CBK_BATCH_NCLASS];
} cbk_tuning;

int cbk_autotune(void);
int cbk_get_tuning(cbk_tuning *t);
int cbk_set_tuning(const cbk_tuning *t);

/* ---- kernel families of the compact routines -----------------------------
 * A compact routine runs each shape on one of several kernel families, which
 * differ in how many lanes one matrix gets (how its tile is laid over the
 * wavefront). The dispatcher's built-in order was measured on MI355X; the
 * tuner below measures the families that can serve one shape on the current
 * device and installs the fastest for it (per device; the table is consulted
 * under CBK_POLICY_AUTO, before the built-in order). */
#define CBK_FAMILY_DEFAULT 0  /* the built-in order (no table entry) */
#define CBK_FAMILY_REGISTER 1 /* one matrix per lane, held in registers */
#define CBK_FAMILY_QUAD 2     /* one matrix per 4 lanes */
#define CBK_FAMILY_TEAM 3     /* one matrix per 16 lanes, rows in lanes */
#define CBK_FAMILY_COLUMNS 4  /* one matrix per 16 lanes, columns in lanes (QR) */
#define CBK_FAMILY_WAVE 5     /* one matrix per wavefront (QR) */
#define CBK_FAMILY_LANE 6     /* the general one-lane-per-matrix kernel */
#define CBK_FAMILY_GRID 7     /* lane grids: QR two problems per wavefront (n <= 32), one above; posv / sysvnp 33..64 */

/* Measure and install the fastest family for one shape. routine: "posv"
 * (also potrf / potrs, whose nrhs is 0 / the call's), "sysvnp" (sytrfnp /
 * sytrsnp), "trsm" (side L: A m x m with uplo and trans, B m x n), "gels"
 * (trans 'N', m >= n, nrhs) or "geqrf" (m >= n). The routine runs nm matrices
 * (0: a count that fills the device, at most 2^17) of synthetic data
 * (diagonally dominant symmetric, well-conditioned triangular, or uniform
 * (-1,1)), column-major with ld = rows, best of 3 timed calls per family.
 * V: the interleave width to tune at; 0 tunes each of 16, 32 and 64 and
 * reports in *vbest the one with the most matrices per second. *family
 * receives the family installed (for *vbest when V = 0). Allocates and frees
 * its own device memory; synchronous. Codes: -1 prec, -2 routine, -3 uplo or
 * trans, -5 m, -6 n, -7 nrhs, -8 V, -9 nm, CBK_ERR_LAUNCH. */
int cbk_autotune_compact(char prec, const char *routine, char uplo, char trans, int m, int n,
                         int nrhs, int V, int nm, int *family, int *vbest);

/* Install (family > 0) or remove (CBK_FAMILY_DEFAULT) the family of one
 * shape, same key as cbk_autotune_compact; a family that cannot serve the
 * shape is ignored at dispatch. Codes as cbk_autotune_compact, -10 family. */
int cbk_set_kernel_family(char prec, const char *routine, char uplo, char trans, int m, int n,
                          int nrhs, int V, int family);
/* the family installed for the shape, CBK_FAMILY_DEFAULT if none (or a
 * negative code as above) */
int cbk_get_kernel_family(char prec, const char *routine, char uplo, char trans, int m, int n,
                          int nrhs, int V);

/* ---- dense <-> compact ------------------------------------------------------
 * Conversion on the device between nm dense matrices (rows x cols) and the
 * compact layout: element (i, j) of matrix g*V+v at
 * ap[g*gstride + (i*si + j*sj)*V + v] with (si, sj) = (1, ldap) for layout 'C',
 * (ldap, 1) for 'R', and gstride = ldap*cols*V ('C') or ldap*rows*V ('R').
 * Packing writes the slots of the last group beyond nm with pad_diag on the
 * diagonal and zero elsewhere -- 1 gives the identity padding the compact
 * routines expect, 0 the padding of a tau batch -- as the reference's test
 * utilities pack_compact / unpack_compact (tests/test_compact_util.hpp:475-531,
 * which fix the dense side to column-major with ld = rows; here its layout,
 * alayout, and lda are free). Unpacking writes only the nm real matrices.
 * Entries of ap outside the rows x cols block of each leading dimension are
 * never read or written. Asynchronous on the cbk stream; HBM-bound (each value
 * read once and written once).
 *
 * Strided dense batch: matrix idx at a + idx*stride_a (scalars), element
 * (i, j) at i + j*lda (alayout 'C') or i*lda + j ('R').
 * Codes (pack): -1 layout -2 rows -3 cols -4 a NULL -5 alayout -6 lda
 *   -7 stride_a < lda * (cols or rows) with nm > 1 -8 ap NULL -9 ldap -11 V -12 nm
 * Codes (unpack): -1 layout -2 rows -3 cols -4 ap NULL -5 ldap -6 a NULL
 *   -7 alayout -8 lda -9 stride_a -10 V -11 nm */
int cbk_dpack_compact(char layout, int rows, int cols, const double *a, char alayout, int lda,
                      size_t stride_a, double *ap, int ldap, double pad_diag, int V, int nm);
int cbk_spack_compact(char layout, int rows, int cols, const float *a, char alayout, int lda,
                      size_t stride_a, float *ap, int ldap, float pad_diag, int V, int nm);
int cbk_dunpack_compact(char layout, int rows, int cols, const double *ap, int ldap, double *a,
                        char alayout, int lda, size_t stride_a, int V, int nm);
int cbk_sunpack_compact(char layout, int rows, int cols, const float *ap, int ldap, float *a,
                        char alayout, int lda, size_t stride_a, int V, int nm);

/* MKL's mkl_?gepack_compact / mkl_?geunpack_compact argument order: a is a
 * device array of nm device pointers, one per dense matrix, in the same
 * layout as ap; packing pads with the identity.
 * Codes: -1 layout -2 rows -3 cols -4 a NULL -5 lda -6 ap NULL -7 ldap -8 V -9 nm */
int cbk_dgepack_compact(char layout, int rows, int cols, const double *const *a, int lda,
                        double *ap, int ldap, int V, int nm);
int cbk_sgepack_compact(char layout, int rows, int cols, const float *const *a, int lda,
                        float *ap, int ldap, int V, int nm);
int cbk_dgeunpack_compact(char layout, int rows, int cols, double *const *a, int lda,
                          const double *ap, int ldap, int V, int nm);
int cbk_sgeunpack_compact(char layout, int rows, int cols, float *const *a, int lda,
                          const float *ap, int ldap, int V, int nm);

#ifdef __cplusplus
}
#endif

#endif /* CBK_EXT_H */
