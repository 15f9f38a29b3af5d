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
#include "team_chol.h"

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

/* 16 teams (4 wavefronts) per block: at V >= 16 a block's matrices fill whole
 * 128-byte lines of every element, so each line is fetched by one CU */
constexpr int kTeamsPerBlock = 16;

template <class T, int N> hipError_t launch_n(bool ldl, const SymArgs<T> &args);

template <class T, int N, int NRB, bool LDL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
sym_team_kernel(SymArgs<T> p)
{
    constexpr int R = rows_per_lane<N>();
    __shared__ T lds[kTeamsPerBlock * chol_lds<N>()];

    const int t = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;                    /* 0..15 in the block */
    const long mat0 = (long)blockIdx.x * kTeamsPerBlock + 4 * wave_in_block(); /* wave-uniform */
    if (mat0 >= p.nmat) return;                           /* whole wavefront */
    const long mat = mat0 + (team & 3);
    const long g = mat0 / p.V;
    const long groups_left = p.nmat / p.V - g;
    const unsigned v = (unsigned)(mat0 % p.V) + (team & 3);
    const unsigned V = (unsigned)p.V;
    constexpr unsigned SZ = sizeof(T);
    const __amdgpu_buffer_rsrc_t ra = group_rsrc(p.a + g * p.a_gs, group_range(groups_left, p.a_gs, SZ));
    const unsigned a_col = (unsigned)p.a_sj * V * SZ; /* byte stride of a column */
    unsigned a_row[R]; /* byte offset of (row t+16r, slot v); kOOB past row n-1 */
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        a_row[r] = (i < N) ? ((unsigned)(i * p.a_si) * V + v) * SZ : kOOB;
    });
    T *L = lds + team * chol_lds<N>();

    /* a[r][c] = element (t+16r, c) of the lower view, c < 16(r+1) */
    T a[R][N];
    T dg[R]; /* the diagonal of the factor for the lane's rows: L(i,i) or D(i) */
    int linfo = 0;
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        /* the lower triangle only: entries right of the diagonal read as 0 */
        sfor<0, cmin(TS * (r + 1), N)>([&](auto c) {
            a[r][c] = bload(T(), ra, (c < TS * r || c <= i) ? a_row[r] : kOOB, c * a_col);
        });
    });

    if (p.factor) {
        linfo = chol_factor<T, N, LDL>(a, dg, t);
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

    T dinv[R];
    chol_prepare<T, N, LDL>(a, dg, dinv, L, t);

    const __amdgpu_buffer_rsrc_t rb = group_rsrc(p.b + g * p.b_gs, group_range(groups_left, p.b_gs, SZ));
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
        chol_solve<T, N, NRB, LDL>(a, dinv, b, L, t);
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { bstore(b[r][q], rb, b_row[r], bcol[q]); });
        });
    }
}

/* one order: the NRB = 4 kernel for up to 4 right-hand sides, 8 above */
template <class T, int N> hipError_t launch_n(bool ldl, const SymArgs<T> &args)
{
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    const dim3 grid(blocks), block(256);
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
