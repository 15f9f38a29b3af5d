/* quad_chol.h -- Cholesky / LDL^T factor and solves on the quad layout
 * (quad_common.h): lane t of a quad holds rows t + 4k of the lower view,
 * a[k][c] = element (t + 4k, c), c < 4k + 4.
 *
 * quad_factor_panels: the reference's factorization exactly
 * (src/cbk_potrf_compact.hpp:79-124, src/cbk_sytrfnp_compact.hpp:96-265).
 * Per element it performs the operations of the right-looking per-pivot
 * order that lane_kernels.hip and reg_chol_impl.h use (pivot j from its
 * owner lane; the columns of j's 8-wide panel take the panel form on the
 * unscaled column with weight a(c,j) * (1/a), later columns the trailing
 * form on the scaled column, LDL^T times D(j)), in the same order, so the
 * factor is bit-identical to the reference's; the work is ordered by panels
 * so that A streams in and the factor streams out while it runs. The pivot's
 * own entry becomes scratch; the diagonal is kept apart (dg). In FP32 the
 * updates of whole 4x4 tiles can run as MFMA rank-1 updates (mfma_tile,
 * quad_mfma), which round as fma does, so the factor stays bit-identical.
 *
 * quad_solve: potrs / sytrsnp on NRB right-hand sides held as b[k][q] = B(t +
 * 4k, q). Forward L y = b right-looking per pivot l (y_l from its owner, the
 * update on the own rows: the reference's per-element order, so sytrsnp's unit
 * sweep and its reciprocal-then-multiply by D are bit-identical); backward
 * L^T x = y column-oriented from the bottom block: the 4 rows of block b of L
 * go through the quad's LDS region, so lane t reads L(l, 4m + t) for its own
 * rows. The backward sweep sums over l descending where the reference's
 * row-dot ascends, so X matches the reference to rounding.
 */
#ifndef CBK_QUAD_CHOL_H
#define CBK_QUAD_CHOL_H

#include "quad_common.h"

namespace cbk {
namespace quad {

using team::fmaT;
using team::sqrtT;

template <int N> constexpr int slots() { return N / 4; }

/* Whether the FP32 factor of order N takes its full-tile updates on MFMA
 * (mfma_tile): measured on 262144 matrices at V = 64, order 24 gains 1.5 to
 * 3.5 % (spotrf 151.5 -> 146.3 us, sposv nrhs 4 194.2 -> 188.7 us, the mixed
 * dsposv nrhs 4 781.5 -> 769.7 us), order 32 none (spotrf 251.0 -> 252.3 us,
 * sposv nrhs 4 305.1 -> 308.9 us); at either order the kernel is bound by
 * HBM or by its FP64 and solve work, not by the factor's fma's. */
template <class T, int N> constexpr bool quad_mfma() { return sizeof(T) == 4 && N == 24; }

typedef float mf32x4 __attribute__((ext_vector_type(4)));

/* Tile CB of one slot's row, row[4CB + i] = A(t + 4k, 4CB + i) in lane t,
 * plus the outer product x y^T, by one v_mfma_f32_4x4x1_16b_f32: its 16
 * blocks are the wavefront's quads, and register i of lane t of a block
 * receives C(i, t) + x_i y_t with x_i read from lane i of the quad and y_t
 * from lane t, so each lane gives its own values and nothing is broadcast.
 * The instruction rounds x_i y_t + C(i, t) once, as fma does, so it performs
 * the reference's per-element update bit for bit. */
template <int CB, int N>
__device__ __forceinline__ void mfma_tile(float (&row)[N], float x, float y)
{
    mf32x4 c = {row[4 * CB], row[4 * CB + 1], row[4 * CB + 2], row[4 * CB + 3]};
    c = __builtin_amdgcn_mfma_f32_4x4x1f32(x, y, c, 0, 0, 0);
    row[4 * CB] = c[0];
    row[4 * CB + 1] = c[1];
    row[4 * CB + 2] = c[2];
    row[4 * CB + 3] = c[3];
}

/* LDS elements of a quad's staging region for the backward sweep: 4 rows of
 * RS = N + 1 (the odd stride spreads a lane's row over the banks), the
 * region padded to 4 mod 32 elements */
template <int N> constexpr int stage_rs() { return N + 1; }
template <int N> constexpr int stage_elems() { return 4 * (N + 1); }

/* The factorization ordered by 8-wide panels (left-looking across panels,
 * right-looking within one): panel P's columns first receive the
 * trailing-form updates of every earlier pivot j < 8P in ascending order,
 * then the panel is factored pivot by pivot. Each element gets exactly the
 * updates of the right-looking order in the same sequence (the trailing form
 * of the earlier panels' pivots, then the panel form of its own panel's), but a panel needs
 * only its own columns of A and the final columns left of it, so the caller
 * can stream A in by panels and write the factor out by panels while the
 * next one is factored: load(P) is called before panel P - 1 is computed
 * (the loads of panel P overlap that work), store(P) after panel P is final.
 * N must be a multiple of 8. */
template <class T, int N, bool LDL, bool MF = false, class Load, class Store>
__device__ __forceinline__ int quad_factor_panels(T (&a)[N / 4][N], T (&dg)[N / 4], int t,
                                                  Load &&load, Store &&store)
{
    static_assert(N % 8 == 0, "whole panels");
    static_assert(!MF || sizeof(T) == 4, "the MFMA tile update is FP32");
    constexpr int K = N / 4;
    int linfo = 0;
    sfor<0, K>([&](auto k) { dg[k] = T(1); });
    load(std::integral_constant<int, 0>{});
    sfor<0, N / 8>([&](auto Pc) {
        constexpr int P = Pc;
        constexpr int c0 = 8 * P, c1 = 8 * P + 8;
        if constexpr (P + 1 < N / 8) load(std::integral_constant<int, P + 1>{});
        __builtin_amdgcn_sched_barrier(0); /* keep the next panel's loads ahead of this work */
        /* trailing form of the earlier pivots on this panel's columns */
        sfor<0, c0>([&](auto jc) {
            constexpr int j = jc;
            T dj = T(0);
            if constexpr (LDL) dj = qbc<j % 4>(dg[j / 4]);
            if constexpr (MF) {
                /* the panel's two column blocks, tile by tile */
                sfor<c0 / 4, c1 / 4>([&](auto cb) {
                    const T x = -(LDL ? a[cb][j] * dj : a[cb][j]);
                    sfor<cb, K>([&](auto k) { mfma_tile<cb>(a[k], x, a[k][j]); });
                });
            }
            else {
                sfor<c0, c1>([&](auto c) {
                    const T w = qbc<c % 4>(LDL ? a[c / 4][j] * dj : a[c / 4][j]);
                    sfor<c / 4, K>([&](auto k) { a[k][c] = fmaT(-w, a[k][j], a[k][c]); });
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        /* the panel, right-looking: panel form, then the column scaled */
        sfor<c0, c1>([&](auto jc) {
            constexpr int j = jc;
            constexpr int kj = j / 4, tj = j % 4;
            const T piv = qbc<tj>(a[kj][j]);
            if (linfo == 0 && (LDL ? (piv == T(0)) : !(piv > T(0)))) linfo = j + 1;
            const T inv = T(1) / piv;
            if constexpr (j + 1 < c1) {
                T wv[K];
                sfor<kj, K>([&](auto k) { wv[k] = a[k][j] * inv; });
                /* the columns of j's own block on the VALU; with MF, the
                 * panel's whole blocks right of it by tiles */
                constexpr int cv = MF ? cmin(c1, 4 * kj + 4) : c1;
                sfor<j + 1, cv>([&](auto c) {
                    const T w = qbc<c % 4>(wv[c / 4]);
                    sfor<c / 4, K>([&](auto k) { a[k][c] = fmaT(-w, a[k][j], a[k][c]); });
                });
                if constexpr (MF) {
                    sfor<cv / 4, c1 / 4>([&](auto cb) {
                        const T x = -wv[cb];
                        sfor<cb, K>([&](auto k) { mfma_tile<cb>(a[k], x, a[k][j]); });
                    });
                }
            }
            T s;
            if constexpr (LDL) {
                s = inv;
                dg[kj] = (t == tj) ? piv : dg[kj];
            }
            else {
                const T d = sqrtT(piv);
                s = d * inv;
                dg[kj] = (t == tj) ? d : dg[kj];
            }
            sfor<kj, K>([&](auto k) { a[k][j] = a[k][j] * s; });
            __builtin_amdgcn_sched_barrier(0);
        });
        store(std::integral_constant<int, P>{});
        __builtin_amdgcn_sched_barrier(0);
    });
    return linfo;
}

/* Ready a factor for quad_solve: dinv = 1/L(i,i) (Cholesky) or 1/D(i)
 * (sytrsnp's reciprocal, src/cbk_sytrsnp_compact.hpp:71), and the entries on
 * and right of the diagonal zeroed, so the sweeps need no row masks. */
template <class T, int N>
__device__ __forceinline__ void quad_prepare(T (&a)[N / 4][N], const T (&dg)[N / 4],
                                             T (&dinv)[N / 4], int t)
{
    constexpr int K = N / 4;
    sfor<0, K>([&](auto k) {
        dinv[k] = T(1) / dg[k];
        sfor<4 * k, 4 * k + 4>([&](auto c) { a[k][c] = (c < t + 4 * k) ? a[k][c] : T(0); });
    });
}

/* X := A^{-1} B on b[k][q] = B(t + 4k, q); L is the quad's staging region in
 * LDS (stage_elems<N>() elements), quad_prepare'd factor in a */
template <class T, int N, int NRB, bool LDL>
__device__ __forceinline__ void quad_solve(const T (&a)[N / 4][N], const T (&dinv)[N / 4],
                                           T (&b)[N / 4][NRB], T *L, int t)
{
    constexpr int K = N / 4;
    constexpr int RS = stage_rs<N>();
    /* forward: L y = b; step l: y_l = b_l (/ L(l,l)) from its owner, then the
     * rows below; the owner's own coefficient is zero */
    sfor<0, N>([&](auto lc) {
        constexpr int l = lc;
        constexpr int kl = l / 4, tl = l % 4;
        sfor<0, NRB>([&](auto q) {
            const T y = qbc<tl>(LDL ? b[kl][q] : b[kl][q] * dinv[kl]);
            sfor<kl, K>([&](auto k) { b[k][q] = fmaT(-a[k][l], y, b[k][q]); });
        });
    });
    /* Cholesky: y = b / L(i,i); sytrsnp: the reciprocal of D times b */
    sfor<0, K>([&](auto k) {
        sfor<0, NRB>([&](auto q) { b[k][q] = b[k][q] * dinv[k]; });
    });
    /* backward: L^T x = y from the bottom block; block kb's rows staged in LDS */
    T *Lt = L + t * RS;
    sfor_down<0, K>([&](auto kbc) {
        constexpr int kb = kbc;
        lds_sync(); /* every read of the previous block before the writes */
        sfor<0, 4 * kb + 4>([&](auto c) { Lt[c] = a[kb][c]; });
        lds_sync();
        sfor_down<4 * kb, 4 * kb + 4>([&](auto lc) {
            constexpr int l = lc;
            constexpr int tl = l % 4;
            T cf[kb + 1]; /* L(l, 4m + t), m <= kb */
            sfor<0, kb + 1>([&](auto m) { cf[m] = L[tl * RS + 4 * m + t]; });
            sfor<0, NRB>([&](auto q) {
                const T x = qbc<tl>(LDL ? b[kb][q] : b[kb][q] * dinv[kb]);
                sfor<0, kb + 1>([&](auto m) { b[m][q] = fmaT(-cf[m], x, b[m][q]); });
            });
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    if constexpr (!LDL) {
        sfor<0, K>([&](auto k) {
            sfor<0, NRB>([&](auto q) { b[k][q] = b[k][q] * dinv[k]; });
        });
    }
}

} /* namespace quad */
} /* namespace cbk */

#endif /* CBK_QUAD_CHOL_H */
