/* cbk_internal.h -- launch interface between the host dispatch (cbk_api.cpp)
 * and the HIP kernels. Not installed; not part of the C ABI. */
#ifndef CBK_INTERNAL_H
#define CBK_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstddef>

#include "offsets.h"

namespace cbk {

/* the stream of the calling host thread (cbk_set_stream) */
hipStream_t current_stream();

/* the kernel policy of the calling host thread (cbk_set_policy) */
int current_policy();

/* 0 for hipSuccess; otherwise records e for cbk_last_error_string and
 * returns CBK_ERR_LAUNCH */
int launch_status(hipError_t e);

/* One operand of a compact batch: base pointer, leading dimension, layout and
 * scalars per group (group_stride of the reference, src/cbk_common.hpp:210). */
template <class T> struct Operand {
    T *p;
    int ld;
    bool row;
    size_t gstride;
};

template <class T> inline bool offsets_fit(const Operand<T> &o, int rows, int cols, int V)
{
    return o.p == nullptr || offsets_fit_bytes(o.gstride, o.ld, o.row, rows, cols, V, sizeof(T));
}

/* ---- generic lane kernels: one lane = one matrix, any V, any layout ---- */

/* potrf / sytrfnp, optionally followed by potrs / sytrsnp on B (nrhs > 0).
 * ldl selects sytrfnp. factor=false runs the solve alone (potrs/sytrsnp on an
 * already-factored A). info (may be null): per-matrix code; with info, a
 * failed matrix skips its solve. */
template <class T>
hipError_t launch_chol_lane(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                            Operand<T> B, int V, int nm, int *info);

/* trsm (any side/uplo/trans/diag/layout); alpha == 0 zeroes B */
template <class T>
hipError_t launch_trsm_lane(bool left, bool up, bool tran, bool unit, int m, int n, T alpha,
                            Operand<T> A, Operand<T> B, int V, int nm);

template <class T>
hipError_t launch_geqrf_lane(int m, int n, Operand<T> A, T *tau, int V, int nm);

template <class T>
hipError_t launch_ormqr_lane(bool left, bool tran, int m, int n, int k, Operand<T> A,
                             const T *tau, Operand<T> C, int V, int nm);

template <class T>
hipError_t launch_orgqr_lane(int m, int n, int k, Operand<T> A, const T *tau, int V, int nm);

template <class T>
hipError_t launch_gels_lane(bool tran, int m, int n, int nrhs, Operand<T> A, Operand<T> B,
                            T *tau, int V, int nm, int *info);

/* mixed-precision posv (ldl: sysvnp), lane path: work holds groups*V*(n*n +
 * n*nrhs) floats */
hipError_t launch_mixed_posv_lane(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, float *work, int V,
                                  int nm, int itermax, int *iter, int *info);

/* ---- kernel families chosen per shape (tune_compact.cpp) ----
 * The dispatcher's built-in order can be overridden per shape by a family
 * (CBK_FAMILY_* of cbk_ext.h) that cbk_autotune_compact measured fastest or
 * cbk_set_kernel_family installed; 0 when the table has no entry. Keys: the
 * routine class, the precision, the dimensions, nrhs and V. */
enum TuneClass { kTuneChol = 0, kTuneLdl = 1, kTuneTrsmFwd = 2, kTuneTrsmBwd = 3, kTuneGels = 4,
                 kTuneGeqrf = 5 };
int tuned_family(int cls, bool fp64, int p, int q, int nrhs, int V);

/* ---- team kernels: one matrix per 16-lane team, held in registers ---- */

/* orders n <= 16 in FP64, n <= 24 in FP32: the register lane kernel
 * (reg_chol.hip), same contract as launch_chol_lane */
bool reg_chol_supports(int n, bool fp64);
template <class T>
hipError_t launch_reg_chol(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                           Operand<T> B, int V, int nm, int *info);

/* orders served by the potrf/posv/sytrfnp/sysvnp quad kernel (one matrix per
 * 4 lanes, quad_sym.hip): 24 and 32 in both precisions; any V; same contract as
 * launch_chol_lane. only: if set, the call touches only the matrices with
 * only[mat] < 0 (their A, B and info; the others are left as they are). */
bool quad_sym_supports(int n, bool fp64);
template <class T>
hipError_t launch_quad_sym(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                           Operand<T> B, int V, int nm, int *info, const int *only = nullptr);

/* orders served by the potrf/posv/sytrfnp/sysvnp team kernel */
bool sym_team_supports(int n);

/* same contract as launch_chol_lane */
template <class T>
hipError_t launch_sym_team(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                           Operand<T> B, int V, int nm, int *info);

/* trsm, side L, with one lane per matrix and op(A)'s triangle in registers
 * (trsm_reg.hip): m <= 16, and m <= 24 in FP32. strided: every column by row
 * dots, never the axpy form the reference's side-L router takes for the last
 * of an odd column count (the side-R solves run through here transposed) */
bool trsm_reg_supports(bool left, int m, bool fp64);
template <class T>
hipError_t launch_trsm_reg(bool up, bool tran, bool unit, int m, int n, T alpha, Operand<T> A,
                           Operand<T> B, int V, int nm, bool strided = false);

/* trsm, side L, forward sweep (op(A) lower) with one matrix per 4 lanes
 * (quad_trsm.hip): FP64 m = 24, FP32 m = 32 */
bool quad_trsm_supports(bool left, bool up, bool tran, int m, bool fp64);
template <class T>
hipError_t launch_quad_trsm(bool tran, bool unit, int m, int n, T alpha, Operand<T> A,
                            Operand<T> B, int V, int nm);

/* trsm, side L, forward sweep (op(A) lower), m in {16, 24, 32} */
bool trsm_team_supports(bool left, bool up, bool tran, int m);
template <class T>
hipError_t launch_trsm_team(bool up, bool tran, bool unit, int m, int n, T alpha, Operand<T> A,
                            Operand<T> B, int V, int nm);

/* ormqr with one lane per matrix and C in registers, reflector length nq (m
 * for side L, n for side R) in {8, 16, 24}, and 32 in FP32 (ormqr_reg.hip) */
bool ormqr_reg_supports(int nq, bool fp64);
template <class T>
hipError_t launch_ormqr_reg(bool left, bool tran, int m, int n, int k, Operand<T> A,
                            const T *tau, Operand<T> C, int V, int nm);

/* ormqr with reflector length nq (m for side L, n for side R) in {16, 32} */
bool ormqr_team_supports(int nq);
template <class T>
hipError_t launch_ormqr_team(bool left, bool tran, int m, int n, int k, Operand<T> A,
                             const T *tau, Operand<T> C, int V, int nm);

/* ormqr / orgqr on the 8 x 8 lane grid, one problem per wavefront
 * (qr_apply64.hip): reflector lengths (orgqr: m) up to 64 */
bool apply_grid_supports(int nq);
template <class T>
hipError_t launch_ormqr_grid(bool left, bool tran, int m, int n, int k, Operand<T> A,
                             const T *tau, Operand<T> C, int V, int nm);
template <class T>
hipError_t launch_orgqr_grid(int m, int n, int k, Operand<T> A, const T *tau, int V, int nm);

/* orders served by the fused mixed posv / sysvnp team kernel (any nrhs, in
 * blocks of 4 right-hand sides) */
bool mixed_team_supports(int n);

/* orders served by the fused mixed posv / sysvnp quad kernel (quad_mixed.hip,
 * one matrix per 4 lanes), followed by the masked FP64 quad pass */
bool quad_mixed_supports(int n);
hipError_t launch_mixed_posv_quad(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, int V, int nm,
                                  int itermax, int *iter, int *info);

/* mixed posv (ldl: sysvnp), team path (no workspace); A is the operand as
 * passed (uplo via up) */
hipError_t launch_mixed_posv_team(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, int V, int nm,
                                  int itermax, int *iter, int *info);

/* QR / overdetermined least squares with one lane per matrix and the tall view
 * F (p x q) in registers (qr_reg.hip): p <= 16 with q <= 8, p <= 32 with
 * q <= 4 (q <= 8 in FP32); any V, any right-hand side count */
bool qr_reg_supports(int p, int q, bool fp64);
template <class T>
hipError_t launch_qr_reg(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                         Operand<T> B, int V, int nm, int *info);
/* orgqr of the same shapes (m x n = p x q), any k <= n */
template <class T>
hipError_t launch_orgqr_reg(int m, int n, int k, Operand<T> A, const T *tau, int V, int nm);

/* QR / overdetermined least squares with the columns of [F | B] in the 16
 * lanes of a team (qr_cols.hip): p <= 32, q + nrhs <= 16, V >= 4 */
bool qr_cols_supports(int p, int q, int nrhs);
template <class T>
hipError_t launch_qr_cols(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info);

/* shapes served by the QR team kernel: the tall view p x q (p >= q) */
bool qr_team_supports(int p, int q, int nrhs, bool fp64);

/* geqrf (B.p == null) or overdetermined gels on the tall view of A
 * (transposed: the view is A^T); tau receives q scalars per matrix */
template <class T>
hipError_t launch_qr_team(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info);

/* ---- uneven batches (batch_gels_impl.h, batch_plan.cpp) ---- */

/* one problem of an uneven batch as the kernels read it (device table) */
struct BatchProblem {
    int m, n, nrhs, lda, ldb;
    int idx;            /* index in the caller's batch: info[idx] */
    long long a_off, b_off;
    long long tau_off;  /* general kernel: its tau in the workspace */
};

/* the kernels of one precision for one size class: one persistent grid over
 * `count` problems starting at `desc`, taking work from `counter` (zero at
 * launch) */
template <class T> struct BatchArgs {
    T *a, *b;
    int *info;
    const BatchProblem *desc;
    unsigned *counter;
    int count;
    T *tau;   /* general kernel only */
    bool tran;
    /* row strides in elements (1 for the column-major problems of a batch) */
    int a_rs, b_rs;
    int ninfo; /* info[idx] is written for idx < ninfo */
    /* compact mode (desc == null, counter == null, wavefront kernel): problem k
     * is cm x cn with cnrhs right-hand sides, element (i, c) of its A at
     * a + (k / cV) * a_gs + k % cV + i * a_rs + c * a_cs (B likewise); the
     * reflector scalars go to ctau[(k / cV) * cn * cV + kk * cV + k % cV] */
    int cm, cn, cnrhs, cV, a_cs, b_cs;
    long long a_gs, b_gs;
    T *ctau;
    /* compact mode, optional: only problems k < ninfo with mask[k] < 0 are
     * solved; with x set, B is read only and the first cn rows of the
     * solution go to x (element (i, j) at x + (k / cV) * x_gs + k % cV +
     * i * x_rs + j * x_cs) */
    const int *mask;
    T *x;
    int x_rs, x_cs;
    long long x_gs;
    /* optional: the number of problems at desc is *dcount (filled on the
     * device before the launch; count is then the capacity the grid is
     * sized for) */
    const unsigned *dcount;
};

/* the problems a batch launch takes: *dcount when set, else count */
template <class T> __device__ __forceinline__ unsigned batch_count(const BatchArgs<T> &p)
{
    return p.dcount ? *p.dcount : (unsigned)p.count;
}

/* classes: 0-2 team kernel with 16/32/64 row registers, 3-4 wavefront kernel
 * with 32/64, 5 the two-array wavefront kernel (n + nrhs > 64), 6 the general
 * lane kernel */
constexpr int kBatchClasses = 7;
constexpr int kBatchGeneral = 6;
/* kernels behind the classes: 0-6 as the classes; 9 / 10 the wavefront
 * classes 3 / 4's problems with n <= 32 and nrhs <= 8 on the lane-grid kernel
 * (qr_grid.hip), two per wavefront; 11 the wider problems of classes 4 and 5
 * with nrhs <= 8 on the 8 x 8 lane grid (qr_grid64.hip), one per wavefront;
 * 7 / 8 the classes 3 / 4's other problems with n + nrhs <= 32, two per
 * wavefront */
constexpr int kBatchKernels = 12;
template <class T>
hipError_t launch_gels_batch(int cls, const BatchArgs<T> &args, unsigned blocks, hipStream_t s);
template <class T>
hipError_t launch_gels_batch_lane(const BatchArgs<T> &args, hipStream_t s);
/* batch kernels 9 and 10 (qr_grid.hip) */
template <class T>
hipError_t launch_qr_grid_batch(int kernel, const BatchArgs<T> &args, unsigned blocks, hipStream_t s);
template <class T> hipError_t qr_grid_batch_occupancy(int kernel, int *blocks);
/* batch kernel 11 (qr_grid64.hip) */
template <class T>
hipError_t launch_qr_grid64_batch(const BatchArgs<T> &args, unsigned blocks, hipStream_t s);
template <class T> hipError_t qr_grid64_batch_occupancy(int *blocks);
/* gels on compact operands through the wavefront kernel: the tall view p x q
 * (transposed: A^T), p <= 64, q + nrhs <= 64, nrhs >= 1, any V */
bool qr_wave_supports(int p, int q, int nrhs);
/* gels / geqrf on lane grids (qr_grid.hip, qr_grid64.hip): p <= 64, q <= 64,
 * nrhs <= 8; two problems per wavefront on 4 x 8 lanes each for q <= 32, one
 * on 8 x 8 lanes above */
bool qr_grid_supports(int p, int q, int nrhs);
/* potrf / posv / potrs and the LDL^T routines at orders 33..64, one matrix per
 * wavefront on an 8 x 8 lane grid (chol_grid.hip) */
bool chol_grid_supports(int n);
/* trsm side L, forward sweep, orders 33..64, on the same grid */
bool trsm_grid_supports(bool left, bool up, bool tran, int m);
template <class T>
hipError_t launch_trsm_grid(bool tran, bool unit, int m, int n, T alpha, Operand<T> A, Operand<T> B,
                            int V, int nm);
/* mask: if set, only the matrices k < nm with mask[k] < 0 (the FP64 pass
 * after the mixed kernel of mixed_rows.hip) */
template <class T>
hipError_t launch_chol_grid(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                            Operand<T> B, int V, int nm, int *info, const int *mask = nullptr);
/* dsposv / dssysvnp at orders 33..64, one matrix per wavefront, lane i holding
 * row i (mixed_rows.hip); the FP64 pass on the matrices it marks follows */
bool mixed_rows_supports(int n);
hipError_t launch_mixed_posv_rows(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, int V, int nm,
                                  int itermax, int *iter, int *info);
template <class T>
hipError_t launch_qr_grid(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info);
template <class T>
hipError_t launch_qr_grid64(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                            Operand<T> B, int V, int nm, int *info);
template <class T>
hipError_t launch_qr_wave(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info);
/* the same kernel as the FP64 path of dsgels_compact: matrices k with
 * mask[k] < 0 only, B read only, X (q x nrhs) written; A receives the QR */
hipError_t launch_qr_wave_masked(int p, int q, int nrhs, Operand<double> A, Operand<double> B,
                                 Operand<double> X, int V, int nm, const int *mask, int *info);

/* ---- mixed-precision uneven batch (mixed_batch.hip, batch_plan.cpp) ---- */

/* a problem of the FP32 path: the BatchProblem fields plus the FP64 kernel
 * its fallback runs on */
struct MixedProblem {
    int m, n, nrhs, lda, ldb;
    int idx;
    long long a_off, b_off;
    int k64;
    int pad;
};

/* conv set: A32 / B32 := fl32(A / B) for the count problems at conv;
 * otherwise X := B (max(m, n) x nrhs) and iter := -1 for the count problems
 * at copy */
struct MixedPrepArgs {
    const double *a, *b;
    float *a32, *b32;
    double *x;
    const MixedProblem *conv;
    const BatchProblem *copy;
    int count;
    int *iter;
};

struct MixedRefineArgs {
    const double *a, *b;
    double *x;
    const float *a32, *b32;
    int *iter, *info;
    const MixedProblem *desc;
    int count;
    BatchProblem *fb;           /* the FP64 fallback lists */
    unsigned *fb_fill;          /* entries per list */
    int fb_begin[kBatchKernels]; /* where each kernel's list starts in fb */
    int itermax;
};

/* refinement kernels by shape: rows <= 32 / 64, columns <= 16 / 32,
 * right-hand sides <= 4 / 8 */
constexpr int kRefineBuckets = 8;
int mixed_refine_bucket(int m, int n, int nrhs);
hipError_t launch_mixed_refine(int bucket, const MixedRefineArgs &a, hipStream_t s);
hipError_t launch_mixed_prep(const MixedPrepArgs &a, hipStream_t s);

/* ---- mixed-precision least squares (mixed_gels_impl.h) ---- */
bool mixed_gels_supports(int p, int q, int nrhs);
hipError_t launch_mixed_gels_team(int p, int q, int nrhs, Operand<double> A, Operand<double> B,
                                  Operand<double> X, int V, int nm, int itermax, int *iter);
/* p[0:n] := value */
hipError_t launch_fill_int(int *p, int n, int value);
/* resident 256-thread blocks per compute unit of class cls's kernel */
template <class T> hipError_t gels_batch_occupancy(int cls, int *blocks);
hipError_t launch_clear_u32(unsigned *p, int n, hipStream_t s);

/* B(0:rows, 0:cols) := 0 in every group */
template <class T>
hipError_t launch_zero(int rows, int cols, Operand<T> B, int V, int nm);

/* info[0:nm] := 0 */
hipError_t launch_clear_info(int *info, int nm);

} /* namespace cbk */

#endif /* CBK_INTERNAL_H */
