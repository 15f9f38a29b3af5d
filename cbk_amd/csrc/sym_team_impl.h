/* sym_team_impl.h -- Cholesky / LDL^T factor and solve, one matrix per 16-lane
 * team, the whole matrix in registers.
 *
 * Lane t of a team owns rows i = t, t+16, t+32, ... of the lower view of its
 * matrix (row block r holds row t+16r) and the same rows of B. A wavefront
 * carries 4 matrices; at V = 64 the 4 teams of a wave read 4 consecutive
 * matrices, 32 contiguous bytes per row of each element, and the 4 waves that
 * share a 128-byte line run back to back on the same CU.
 *
 * Factorization (potrf / sytrfnp): the reference's arithmetic exactly
 * (src/cbk_potrf_compact.hpp:79-124, src/cbk_sytrfnp_compact.hpp:96-265) in
 * the right-looking per-pivot order that lane_kernels.hip explains: pivot j is
 * broadcast from its owner lane; columns of j's 8-wide panel take the panel
 * form on the unscaled column with weight 1/a, later columns the trailing form
 * on the scaled column. The factor is bit-identical to the reference's.
 * Entries above the diagonal ride along as scratch: they only ever feed other
 * scratch entries and are never stored.
 *
 * Solve (potrs / sytrsnp): forward substitution with L, right-looking, then
 * backward substitution with L^T, column-oriented over a copy of L in LDS so a
 * lane reads L(l, i) for its own rows i. For potrs the diagonal is folded into
 * the triangular factor once (columns of L scaled by 1/L(i,i) for the forward
 * sweep, rows for the backward one), so each sweep step is one broadcast and
 * one fma per row block. X therefore agrees with the reference to rounding
 * (the test tolerance), not to the bit; sytrsnp's unit sweeps and reciprocal D
 * scaling (src/cbk_sytrsnp_compact.hpp:66-76) keep the reference's operations
 * but the backward sweep's summation order (l descending) differs.
 *
 * Addressing: a buffer resource per group; rows past n and entries that must
 * not be written get an offset beyond the resource's range, which the
 * hardware turns into a zero load / a dropped store -- no branches.
 */
#ifndef CBK_SYM_TEAM_IMPL_H
#define CBK_SYM_TEAM_IMPL_H

#include "cbk_internal.h"
#include "team_common.h"

namespace cbk {
namespace symteam {

using namespace team;

template <class T> struct SymArgs {
    T *a;              /* factor mode: A, overwritten by the factor; solve mode: the factor */
    bool store_factor; /* factor mode: write the factor back */
    int a_si, a_sj;
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long nmat; /* groups * V: every slot, padded ones included */
    int nm;
    int nrhs;
    int *info;
    bool factor;
};

constexpr int kTeamsPerBlock = 4; /* one wavefront per block */
constexpr unsigned kOOB = 0x80000000u; /* a byte offset past every group's range */

/* the byte range a group's resource covers: the rest of the buffer, capped so
 * that kOOB stays out of range */
__device__ __forceinline__ unsigned group_range(long groups_left, size_t gstride, unsigned sz)
{
    const unsigned long long left = (unsigned long long)groups_left * gstride * sz;
    return left < 0x7fffffffull ? (unsigned)left : 0x7fffffffu;
}

template <class T, int N> hipError_t launch_n(bool ldl, const SymArgs<T> &args);

template <class T, int N, int NRB, bool LDL>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void
sym_team_kernel(SymArgs<T> p)
{
    constexpr int R = (N + TS - 1) / TS;
    constexpr int TRI = N * (N + 1) / 2;
    constexpr int ZERO = TRI, SINK = TRI + 1; /* per-team LDS slots: a 0.0, a write sink */
    __shared__ T lds[kTeamsPerBlock * (TRI + 2)];

    const int t = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;
    const long mat0 = (long)blockIdx.x * kTeamsPerBlock; /* wave-uniform */
    const long mat = mat0 + team;
    const long g = mat0 / p.V;
    const long groups_left = p.nmat / p.V - g;
    const unsigned v = (unsigned)(mat0 % p.V) + team;
    const unsigned V = (unsigned)p.V;
    constexpr unsigned SZ = sizeof(T);
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc(p.a + g * p.a_gs, 0, group_range(groups_left, p.a_gs, SZ),
                                          0x00020000);
    const unsigned a_col = (unsigned)p.a_sj * V * SZ; /* byte stride of a column */
    unsigned a_row[R]; /* byte offset of (row t+16r, slot v); kOOB past row n-1 */
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        a_row[r] = (i < N) ? ((unsigned)(i * p.a_si) * V + v) * SZ : kOOB;
    });
    T *L = lds + team * (TRI + 2);

    /* a[r][c] = element (t+16r, c) of the lower view, c < 16(r+1) */
    T a[R][N];
    T dg[R]; /* the diagonal of the factor for the lane's rows: L(i,i) or D(i) */
    int linfo = 0;

    sfor<0, R>([&](auto r) {
        sfor<0, cmin(TS * (r + 1), N)>([&](auto c) { a[r][c] = bload(T(), ra, a_row[r], c * a_col); });
        dg[r] = T(1);
    });

    if (p.factor) {
        sfor<0, N>([&](auto jc) {
            constexpr int j = jc;
            constexpr int rj = j / TS, tj = j % TS;
            constexpr int pe = cmin((j / 8) * 8 + 8, N);
            const T piv = bcast<tj>(a[rj][j]);
            if (linfo == 0 && (LDL ? (piv == T(0)) : !(piv > T(0)))) linfo = j + 1;
            const T inv = T(1) / piv;
            if constexpr (j + 1 < pe) {
                /* panel form: A(i,c) -= A(i,j) * (A(c,j) * (1/a)) */
                T wv[R];
                sfor<0, R>([&](auto r) {
                    if constexpr (j < TS * (r + 1)) wv[r] = a[r][j] * inv;
                });
                sfor<j + 1, pe>([&](auto c) {
                    const T w = bcast<c % TS>(wv[c / TS]);
                    sfor<0, R>([&](auto r) {
                        if constexpr (c < TS * (r + 1)) a[r][c] = fmaT(-a[r][j], w, a[r][c]);
                    });
                });
            }
            /* scale the column below the pivot; the pivot's own entry becomes
             * scratch, its value lives in dg */
            T sc;
            if constexpr (LDL) {
                sc = inv;
                dg[rj] = (t == tj) ? piv : dg[rj];
            }
            else {
                const T d = sqrtT(piv);
                sc = d * inv;
                dg[rj] = (t == tj) ? d : dg[rj];
            }
            sfor<0, R>([&](auto r) {
                if constexpr (j < TS * (r + 1)) a[r][j] = a[r][j] * sc;
            });
            if constexpr (pe < N) {
                /* trailing form: A(i,c) -= L(i,j) * w, w = L(c,j) or L(c,j)*D(j) */
                T wd[R];
                if constexpr (LDL) {
                    sfor<0, R>([&](auto r) {
                        if constexpr (j < TS * (r + 1)) wd[r] = a[r][j] * piv;
                    });
                }
                sfor<pe, N>([&](auto c) {
                    const T w = LDL ? bcast<c % TS>(wd[c / TS]) : bcast<c % TS>(a[c / TS][j]);
                    sfor<0, R>([&](auto r) {
                        if constexpr (c < TS * (r + 1)) a[r][c] = fmaT(-a[r][j], w, a[r][c]);
                    });
                });
            }
            __builtin_amdgcn_sched_barrier(0); /* one pivot at a time */
        });
        if (p.info && mat < p.nm && t == 0) p.info[mat] = linfo;
        if (p.store_factor) {
            sfor<0, R>([&](auto r) {
                const int i = t + TS * r;
                sfor<0, cmin(TS * (r + 1), N)>([&](auto c) {
                    const unsigned off = (c < TS * r || c < i) ? a_row[r] : kOOB;
                    bstore(a[r][c], ra, off, c * a_col);
                });
                const unsigned doff = (a_row[r] == kOOB) ? kOOB : a_row[r] + (unsigned)i * a_col;
                bstore(dg[r], ra, doff, 0);
            });
        }
    }
    else {
        /* solve mode: the diagonal of the given factor */
        sfor<0, R>([&](auto r) {
            const int i = t + TS * r;
            const unsigned doff = (a_row[r] == kOOB) ? kOOB : a_row[r] + (unsigned)i * a_col;
            dg[r] = bload(T(), ra, doff, 0);
        });
    }

    /* with info, a failed matrix keeps its B (LAPACK dposv skips dpotrs) */
    if (p.nrhs == 0 || (p.info && mat < p.nm && linfo != 0)) return;

    /* 1/L(i,i) (potrf) or 1/D(i) (sytrsnp's reciprocal, src/cbk_sytrsnp_compact.hpp:71) */
    T dinv[R];
    sfor<0, R>([&](auto r) { dinv[r] = T(1) / dg[r]; });
    /* keep the strictly lower part: a row's diagonal and everything right of
     * it become 0, so the sweeps below need no row masks */
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        sfor<TS * r, cmin(TS * (r + 1), N)>([&](auto c) { a[r][c] = (c < i) ? a[r][c] : T(0); });
    });
    /* L into LDS for the backward sweep, packed by rows (L(i,c) at i(i+1)/2+c),
     * potrs rows pre-scaled by 1/L(i,i) */
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
        /* columns of L scaled by 1/L(c,c) for the forward sweep */
        sfor<0, N>([&](auto c) {
            const T s = bcast<c % TS>(dinv[c / TS]);
            sfor<0, R>([&](auto r) {
                if constexpr (c < TS * (r + 1)) a[r][c] = a[r][c] * s;
            });
        });
    }

    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc(p.b + g * p.b_gs, 0, group_range(groups_left, p.b_gs, SZ),
                                          0x00020000);
    const unsigned b_col = (unsigned)p.b_sj * V * SZ;
    unsigned b_row[R];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        b_row[r] = (i < N) ? ((unsigned)(i * p.b_si) * V + v) * SZ : kOOB;
    });
    for (int q0 = 0; q0 < p.nrhs; q0 += NRB) {
        T b[R][NRB];
        unsigned bcol[NRB];
        sfor<0, NRB>([&](auto q) { bcol[q] = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB; });
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { b[r][q] = bload(T(), rb, b_row[r], bcol[q]); });
        });
        /* forward: L y = b, right-looking; row k accumulates in the
         * reference's order */
        sfor<0, N>([&](auto ic) {
            constexpr int i = ic;
            constexpr int ri = i / TS, ti = i % TS;
            sfor<0, NRB>([&](auto q) {
                const T y = bcast<ti>(b[ri][q]);
                sfor<ri, R>([&](auto r) { b[r][q] = fmaT(-a[r][i], y, b[r][q]); });
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        /* potrs: x = y / L(i,i); sytrsnp: w = D^{-1} z */
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { b[r][q] = b[r][q] * dinv[r]; });
        });
        /* backward: L^T x = z, column-oriented from the bottom row */
        sfor_down<0, N>([&](auto lc) {
            constexpr int l = lc;
            constexpr int rl = l / TS, tl = l % TS;
            T lv[R];
            sfor<0, rl + 1>([&](auto r) {
                const int i = t + TS * r;
                lv[r] = L[(i < l) ? l * (l + 1) / 2 + i : ZERO];
            });
            sfor<0, NRB>([&](auto q) {
                const T z = bcast<tl>(b[rl][q]);
                sfor<0, rl + 1>([&](auto r) { b[r][q] = fmaT(-lv[r], z, b[r][q]); });
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        if constexpr (!LDL) {
            sfor<0, R>([&](auto r) {
                sfor<0, NRB>([&](auto q) { b[r][q] = b[r][q] * dinv[r]; });
            });
        }
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { bstore(b[r][q], rb, b_row[r], bcol[q]); });
        });
    }
}

/* one order: the NRB = 4 kernel for up to 4 right-hand sides, 8 above */
template <class T, int N> hipError_t launch_n(bool ldl, const SymArgs<T> &args)
{
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    const dim3 grid(blocks), block(64);
    if (args.nrhs > 4) {
        if (ldl)
            hipLaunchKernelGGL((sym_team_kernel<T, N, 8, true>), grid, block, 0, current_stream(), args);
        else
            hipLaunchKernelGGL((sym_team_kernel<T, N, 8, false>), grid, block, 0, current_stream(), args);
    }
    else {
        if (ldl)
            hipLaunchKernelGGL((sym_team_kernel<T, N, 4, true>), grid, block, 0, current_stream(), args);
        else
            hipLaunchKernelGGL((sym_team_kernel<T, N, 4, false>), grid, block, 0, current_stream(), args);
    }
    return hipGetLastError();
}

} /* namespace symteam */
} /* namespace cbk */

#endif /* CBK_SYM_TEAM_IMPL_H */
