/* team_chol.h -- the register-resident Cholesky / LDL^T building blocks of
 * the team kernels (sym_team_impl.h, mixed_team_impl.h).
 *
 * A team is 16 lanes (one DPP row) holding one n x n matrix: lane t owns rows
 * t, t+16, ... of the lower view, a[r][c] = element (t+16r, c), c < 16(r+1).
 *
 * chol_factor: the reference's factorization exactly
 * (src/cbk_potrf_compact.hpp:79-124, src/cbk_sytrfnp_compact.hpp:96-265) in
 * right-looking per-pivot order (lane_kernels.hip explains why the operations
 * per element are the reference's, in its order): pivot j is broadcast from
 * its owner lane; columns of j's 8-wide panel take the panel form on the
 * unscaled column with weight 1/a, later columns the trailing form on the
 * scaled column. Entries above the diagonal ride along as scratch: they only
 * ever feed other scratch entries and are never stored.
 *
 * chol_prepare / chol_solve: forward substitution with L (right-looking; each
 * row accumulates in the reference's order), then backward substitution with
 * L^T, column-oriented over a copy of L in LDS so a lane reads L(l, i) for its
 * own rows i. For potrs the diagonal is folded into the factor once (columns
 * of L scaled by 1/L(i,i) for the forward sweep, rows for the backward one),
 * so a sweep step is one broadcast and one fma per row block; sytrsnp keeps
 * the reference's unit sweeps and reciprocal D scaling
 * (src/cbk_sytrsnp_compact.hpp:52-78). The backward sweep sums in descending
 * l where the reference's row-dot ascends, so X matches to rounding.
 */
#ifndef CBK_TEAM_CHOL_H
#define CBK_TEAM_CHOL_H

#include "team_common.h"

namespace cbk {
namespace team {

template <int N> constexpr int rows_per_lane() { return (N + TS - 1) / TS; }

/* LDS words a team needs for the L copy: the packed lower triangle
 * (L(i,c) at i(i+1)/2 + c), a 0.0 slot and a write sink */
template <int N> constexpr int chol_lds() { return N * (N + 1) / 2 + 2; }

/* Factor in registers. Returns the LAPACK-style code (first pivot <= 0 or NaN
 * for Cholesky, first exact zero for LDL^T, 1-based; 0 if none). dg receives
 * L(i,i) (Cholesky) or D(i) (LDL^T) for the lane's rows; the pivot entries of
 * a become scratch. */
/* the pivot j as every lane of the team sees it: a, 1/a and (Cholesky) sqrt(a) */
template <class T, bool LDL> struct Pivot {
    T piv, inv, d;
};
template <class T, int N, bool LDL, int J>
__device__ __forceinline__ Pivot<T, LDL> load_pivot(const T (&a)[rows_per_lane<N>()][N])
{
    Pivot<T, LDL> p;
    p.piv = bcast<J % TS>(a[J / TS][J]);
    p.inv = T(1) / p.piv;
    p.d = LDL ? T(0) : sqrtT(p.piv);
    return p;
}

/* Look-ahead: pivot j+1's divide and square root depend only on column j+1,
 * so that column is updated first and the next pivot's chain is issued
 * before the rest of pivot j's updates, which hide its latency. The
 * operations on every element, and their order per element, are those
 * described above. */
template <class T, int N, bool LDL>
__device__ __forceinline__ int chol_factor(T (&a)[rows_per_lane<N>()][N],
                                           T (&dg)[rows_per_lane<N>()], int t)
{
    constexpr int R = rows_per_lane<N>();
    int linfo = 0;
    sfor<0, R>([&](auto r) { dg[r] = T(1); });
    Pivot<T, LDL> cur = load_pivot<T, N, LDL, 0>(a);
    sfor<0, N>([&](auto jc) {
        constexpr int j = jc;
        constexpr int rj = j / TS, tj = j % TS;
        constexpr int pe = cmin((j / 8) * 8 + 8, N);
        const T piv = cur.piv, inv = cur.inv;
        if (linfo == 0 && (LDL ? (piv == T(0)) : !(piv > T(0)))) linfo = j + 1;
        Pivot<T, LDL> nxt = cur;
        /* panel form on columns [c0, c1): A(i,c) -= A(i,j) * (A(c,j) * (1/a)) */
        T wv[R];
        auto panel = [&](auto c0, auto c1) {
            sfor<c0, c1>([&](auto c) {
                sfor<0, R>([&](auto r) {
                    if constexpr (c < TS * (r + 1)) fnma_bcr<c % TS>(a[r][c], wv[c / TS], a[r][j]);
                });
            });
        };
        /* trailing form on columns [c0, c1): A(i,c) -= L(i,j) * w, w = L(c,j) or L(c,j)*D(j) */
        T wd[R];
        auto trailing = [&](auto c0, auto c1) {
            sfor<c0, c1>([&](auto c) {
                sfor<0, R>([&](auto r) {
                    if constexpr (c < TS * (r + 1))
                        fnma_bcr<c % TS>(a[r][c], LDL ? wd[c / TS] : a[c / TS][j], a[r][j]);
                });
            });
        };
        using I1 = std::integral_constant<int, j + 1>;
        using I2 = std::integral_constant<int, j + 2>;
        using IPE = std::integral_constant<int, pe>;
        using IPE1 = std::integral_constant<int, pe + 1>;
        using IN = std::integral_constant<int, N>;
        if constexpr (j + 1 < pe) {
            sfor<0, R>([&](auto r) {
                if constexpr (j < TS * (r + 1)) wv[r] = a[r][j] * inv;
            });
            dpp_ready_from<rj>(wv);
            panel(I1{}, I2{});
            nxt = load_pivot<T, N, LDL, j + 1>(a);
            panel(I2{}, IPE{});
        }
        /* scale the column below the pivot; the pivot's entry becomes scratch */
        T sc;
        if constexpr (LDL) {
            sc = inv;
            dg[rj] = (t == tj) ? piv : dg[rj];
        }
        else {
            sc = cur.d * inv;
            dg[rj] = (t == tj) ? cur.d : dg[rj];
        }
        sfor<0, R>([&](auto r) {
            if constexpr (j < TS * (r + 1)) a[r][j] = a[r][j] * sc;
        });
        if constexpr (pe < N) {
            if constexpr (LDL) {
                sfor<0, R>([&](auto r) {
                    if constexpr (j < TS * (r + 1)) wd[r] = a[r][j] * piv;
                });
                dpp_ready_from<rj>(wd);
            }
            else {
                dpp_ready_col<rj, j>(a);
            }
            if constexpr (j + 1 == pe) {
                trailing(IPE{}, IPE1{});
                nxt = load_pivot<T, N, LDL, j + 1>(a);
                trailing(IPE1{}, IN{});
            }
            else {
                trailing(IPE{}, IN{});
            }
        }
        cur = nxt;
        __builtin_amdgcn_sched_barrier(0); /* one pivot at a time: bounds live registers */
    });
    return linfo;
}

/* FP32 Cholesky for the mixed-precision solve, where the factor only has to
 * be an FP32-accurate preconditioner (refinement in FP64 removes its error):
 * plain right-looking per pivot, L(:,j) = A(:,j) * rsq(a) with the hardware
 * reciprocal square root, L(j,j) = a * rsq(a), then the trailing update with
 * the scaled column. Same look-ahead as chol_factor. Returns the first pivot
 * <= 0 or NaN (1-based), 0 if none. */
template <int N>
__device__ __forceinline__ int chol_factor_fast(float (&a)[rows_per_lane<N>()][N],
                                                float (&dg)[rows_per_lane<N>()], int t)
{
    constexpr int R = rows_per_lane<N>();
    int linfo = 0;
    sfor<0, R>([&](auto r) { dg[r] = 1.0f; });
    float piv = bcast<0>(a[0][0]);
    float rs = __builtin_amdgcn_rsqf(piv);
    sfor<0, N>([&](auto jc) {
        constexpr int j = jc;
        constexpr int rj = j / TS, tj = j % TS;
        if (linfo == 0 && !(piv > 0.0f)) linfo = j + 1;
        dg[rj] = (t == tj) ? piv * rs : dg[rj];
        sfor<0, R>([&](auto r) {
            if constexpr (j < TS * (r + 1)) a[r][j] = a[r][j] * rs;
        });
        if constexpr (j + 1 < N) dpp_ready_col<rj, j>(a);
        float npiv = piv, nrs = rs;
        sfor<j + 1, N>([&](auto c) {
            sfor<0, R>([&](auto r) {
                if constexpr (c < TS * (r + 1)) fnma_bcr<c % TS>(a[r][c], a[c / TS][j], a[r][j]);
            });
            if constexpr (c == j + 1) {
                npiv = bcast<c % TS>(a[c / TS][c]);
                nrs = __builtin_amdgcn_rsqf(npiv);
            }
        });
        piv = npiv;
        rs = nrs;
        __builtin_amdgcn_sched_barrier(0);
    });
    return linfo;
}

/* Ready a factor for chol_solve: dinv = 1/L(i,i) or 1/D(i) (sytrsnp's
 * reciprocal, src/cbk_sytrsnp_compact.hpp:71); the diagonal and everything
 * right of it zeroed (no row masks in the sweeps); L copied into the team's
 * LDS (rows pre-scaled by 1/L(i,i) for potrs); for potrs the columns of a
 * scaled by 1/L(c,c). */
template <class T, int N, bool LDL>
__device__ __forceinline__ void chol_prepare(T (&a)[rows_per_lane<N>()][N],
                                             const T (&dg)[rows_per_lane<N>()],
                                             T (&dinv)[rows_per_lane<N>()], T *L, int t)
{
    constexpr int R = rows_per_lane<N>();
    constexpr int ZERO = N * (N + 1) / 2, SINK = ZERO + 1;
    sfor<0, R>([&](auto r) { dinv[r] = T(1) / dg[r]; });
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        sfor<TS * r, cmin(TS * (r + 1), N)>([&](auto c) { a[r][c] = (c < i) ? a[r][c] : T(0); });
    });
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        sfor<0, cmin(TS * (r + 1), N)>([&](auto c) {
            const int at = ((i < N) && (c < TS * r || c < i)) ? i * (i + 1) / 2 + c : SINK;
            L[at] = LDL ? a[r][c] : a[r][c] * dinv[r];
        });
    });
    if (t == 0) L[ZERO] = T(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (!LDL) {
        sfor<0, N>([&](auto c) {
            const T s = bcast<c % TS>(dinv[c / TS]);
            sfor<0, R>([&](auto r) {
                if constexpr (c < TS * (r + 1)) a[r][c] = a[r][c] * s;
            });
        });
    }
}

/* X := A^{-1} B in place on b (the lane's rows, NRB right-hand sides) */
template <class T, int N, int NRB, bool LDL>
__device__ __forceinline__ void chol_solve(const T (&a)[rows_per_lane<N>()][N],
                                           const T (&dinv)[rows_per_lane<N>()],
                                           T (&b)[rows_per_lane<N>()][NRB], const T *L, int t)
{
    constexpr int R = rows_per_lane<N>();
    constexpr int ZERO = N * (N + 1) / 2;
    /* forward: L y = b */
    sfor<0, N>([&](auto ic) {
        constexpr int i = ic;
        constexpr int ri = i / TS, ti = i % TS;
        /* the pivot row's own update last: no DPP read may follow the write
         * of its source within the step */
        dpp_ready_from<0>(b[ri]);
        sfor<0, NRB>([&](auto q) {
            sfor<ri + 1, R>([&](auto r) { fnma_bcr<ti>(b[r][q], b[ri][q], a[r][i]); });
            fnma_bcr<ti>(b[ri][q], b[ri][q], a[ri][i]);
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    /* potrs: x = y / L(i,i); sytrsnp: w = D^{-1} z */
    sfor<0, R>([&](auto r) {
        sfor<0, NRB>([&](auto q) { b[r][q] = b[r][q] * dinv[r]; });
    });
    /* backward: L^T x = z, column-oriented from the bottom row; the LDS reads
     * of row l-1 are issued before step l's fmas, which hide their latency */
    auto load_row = [&](auto lc, T(&lv)[R]) {
        constexpr int l = lc;
        sfor<0, l / TS + 1>([&](auto r) {
            const int i = t + TS * r;
            lv[r] = L[(i < l) ? l * (l + 1) / 2 + i : ZERO];
        });
    };
    T lv[R];
    load_row(std::integral_constant<int, N - 1>{}, lv);
    sfor_down<0, N>([&](auto lc) {
        constexpr int l = lc;
        constexpr int rl = l / TS, tl = l % TS;
        T lvn[R];
        if constexpr (l > 0) load_row(std::integral_constant<int, l - 1>{}, lvn);
        dpp_ready_from<0>(b[rl]);
        sfor<0, NRB>([&](auto q) {
            sfor<0, rl>([&](auto r) { fnma_bcr<tl>(b[r][q], b[rl][q], lv[r]); });
            fnma_bcr<tl>(b[rl][q], b[rl][q], lv[rl]);
        });
        if constexpr (l > 0) {
            sfor<0, R>([&](auto r) { lv[r] = lvn[r]; });
        }
        __builtin_amdgcn_sched_barrier(0);
    });
    if constexpr (!LDL) {
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { b[r][q] = b[r][q] * dinv[r]; });
        });
    }
}

/* The FP32 Cholesky solve with both sweeps reading L from LDS, so the
 * factor's registers are free once it is prepared (the mixed kernel keeps
 * the FP64 A in registers for its residuals). One copy serves both sweeps:
 * the columns of L scaled by 1/L(c,c), each column c stored for the row
 * blocks from its own down (rows 16*(c/16) .. 16*R-1, zeros on and above the
 * diagonal and in rows past n), so every read is a lane's base plus a
 * compile-time offset and needs no mask:
 *   forward  L y = b as (unit lower L D^-1) v = b, v = D y: step i reads
 *            column i at the lane's own rows;
 *   backward L^T x = y as x = D^-1 y - (L D^-1)^T x, from v: step l reads
 *            row l across the lane's columns i = t + 16r (column i's slot l).
 * D = diag(L). */
template <int N> constexpr int cols_at(int c)
{
    int s = 0;
    for (int k = 0; k < c; ++k) s += TS * (rows_per_lane<N>() - k / TS);
    return s;
}
/* floats of the copy */
template <int N> constexpr int cols_lds() { return cols_at<N>(N); }

/* LDL: a holds the unit lower factor of an LDL^T and dg its D; the copy is
 * then L itself, and dsc = 1/D (the x = dsc * v step of chol_solve_cols).
 * Cholesky: the columns scaled by 1/L(c,c) and dsc = 1/L(i,i)^2. */
template <int N, bool LDL = false>
__device__ __forceinline__ void chol_prepare_cols(const float (&a)[rows_per_lane<N>()][N],
                                                  const float (&dg)[rows_per_lane<N>()],
                                                  float (&dsc)[rows_per_lane<N>()], float *L,
                                                  int t)
{
    constexpr int R = rows_per_lane<N>();
    float dinv[R];
    sfor<0, R>([&](auto r) {
        dinv[r] = 1.0f / dg[r];
        dsc[r] = LDL ? dinv[r] : dinv[r] * dinv[r];
    });
    sfor<0, N>([&](auto c) {
        const float dc = LDL ? 1.0f : bcast<c % TS>(dinv[c / TS]);
        sfor<c / TS, R>([&](auto r) {
            const int row = t + TS * r;
            L[cols_at<N>(c) + TS * (r - c / TS) + t] =
                (c < row && row < N) ? (LDL ? a[r][c] : a[r][c] * dc) : 0.0f;
        });
    });
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int N, int NRB>
__device__ __forceinline__ void chol_solve_cols(const float (&dsc)[rows_per_lane<N>()],
                                                float (&b)[rows_per_lane<N>()][NRB],
                                                const float *L, int t)
{
    constexpr int R = rows_per_lane<N>();
    constexpr int D = 4; /* forward coefficients are read D steps ahead */
    const float *Lt = L + t;
    /* forward: step i reads column i at rows t + 16r, r >= i/16 */
    auto load_col = [&](auto ic, float(&cv)[R]) {
        constexpr int i = ic;
        sfor<i / TS, R>([&](auto r) { cv[r] = Lt[cols_at<N>(i) + TS * (r - i / TS)]; });
    };
    float ring[D][R];
    sfor<0, cmin(D, N)>([&](auto ic) { load_col(ic, ring[ic]); });
    sfor<0, N>([&](auto ic) {
        constexpr int i = ic;
        constexpr int ri = i / TS, ti = i % TS;
        float cv[R];
        sfor<ri, R>([&](auto r) { cv[r] = ring[i % D][r]; });
        if constexpr (i + D < N) load_col(std::integral_constant<int, i + D>{}, ring[i % D]);
        dpp_ready_from<0>(b[ri]);
        sfor<0, NRB>([&](auto q) {
            sfor<ri + 1, R>([&](auto r) { fnma_bcr<ti>(b[r][q], b[ri][q], cv[r]); });
            fnma_bcr<ti>(b[ri][q], b[ri][q], cv[ri]);
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    /* x starts from D^-2 v (Cholesky; LDL^T: D^-1 v) */
    sfor<0, R>([&](auto r) {
        sfor<0, NRB>([&](auto q) { b[r][q] = b[r][q] * dsc[r]; });
    });
    /* backward: column i = t + 16r starts at cols_at(i) with row 16r */
    int base[R];
    sfor<0, R>([&](auto r) {
        int full = 0;
        sfor<0, r>([&](auto k) { full += TS * TS * (R - k); });
        base[r] = full + TS * (R - r) * t - TS * r;
    });
    auto load_row = [&](auto lc, float(&lv)[R]) {
        constexpr int l = lc;
        sfor<0, l / TS + 1>([&](auto r) { lv[r] = L[base[r] + l]; });
    };
    float lv[R];
    load_row(std::integral_constant<int, N - 1>{}, lv);
    sfor_down<0, N>([&](auto lc) {
        constexpr int l = lc;
        constexpr int rl = l / TS, tl = l % TS;
        float lvn[R];
        if constexpr (l > 0) load_row(std::integral_constant<int, l - 1>{}, lvn);
        dpp_ready_from<0>(b[rl]);
        sfor<0, NRB>([&](auto q) {
            sfor<0, rl>([&](auto r) { fnma_bcr<tl>(b[r][q], b[rl][q], lv[r]); });
            fnma_bcr<tl>(b[rl][q], b[rl][q], lv[rl]);
        });
        if constexpr (l > 0) {
            sfor<0, R>([&](auto r) { lv[r] = lvn[r]; });
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

/* max over the 16 lanes of the team, in every lane (mirror butterfly) */
template <class T> __device__ __forceinline__ T team_max(T x)
{
    x = fmax_nan(x, dpp<0xB1>(x));  /* quad_perm [1,0,3,2] */
    x = fmax_nan(x, dpp<0x4E>(x));  /* quad_perm [2,3,0,1] */
    x = fmax_nan(x, dpp<0x141>(x)); /* row_half_mirror */
    x = fmax_nan(x, dpp<0x140>(x)); /* row_mirror */
    return x;
}

} /* namespace team */
} /* namespace cbk */

#endif /* CBK_TEAM_CHOL_H */
