/* trsm_team_impl.h -- B := alpha op(A)^-1 B (side L) with one matrix per 16-lane
 * team, for the forward sweeps: op(A) lower triangular (uplo L, no transpose,
 * or uplo U, transposed), orders 16, 24 and 32.
 *
 * Lane t owns rows t, t+16 of the lower view L = op(A) (strictly below the
 * diagonal; the diagonal apart) and the same rows of B, NRB right-hand sides at
 * a time. Per pivot k (ascending):
 *     x_k = b_k / L(k,k)            (non-unit; lane k)
 *     b_i -= L(i,k) x_k,  i > k     (x_k broadcast by DPP, fused into the fma)
 * after b := alpha b. Every element therefore receives the reference's
 * operations in its order -- the multiply by alpha, the fma's over l ascending,
 * then the divide (the row-dot form, src/cbk_trsm_compact.hpp:73-104, and the
 * axpy form of its single-column tail, :49-67, which is this same sweep) -- so
 * results are bit-identical to the reference's. The backward sweeps (op(A)
 * upper) sum in ascending l in the reference, which a right-looking sweep
 * cannot reproduce; they stay on the lane kernel.
 */
#ifndef CBK_TRSM_TEAM_IMPL_H
#define CBK_TRSM_TEAM_IMPL_H

#include "cbk_internal.h"
#include "team_chol.h"

namespace cbk {
namespace trsmteam {

using namespace team;

template <class T> struct TrsmArgs {
    const T *a; /* op(A): element (i, c) at (i*a_si + c*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long nmat;
    int nrhs;
    T alpha;
    bool unit;
};

constexpr int kTeamsPerBlock = 16;

template <class T, int N, int NRB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
trsm_team_kernel(TrsmArgs<T> p)
{
    constexpr int R = rows_per_lane<N>();
    constexpr unsigned SZ = sizeof(T);
    const int t = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;
    const long mat0 = (long)blockIdx.x * kTeamsPerBlock + 4 * wave_in_block(); /* wave-uniform */
    if (mat0 >= p.nmat) return;
    const long g = mat0 / p.V;
    const long groups_left = p.nmat / p.V - g;
    const unsigned v = (unsigned)(mat0 % p.V) + (team & 3);
    const unsigned V = (unsigned)p.V;
    const __amdgpu_buffer_rsrc_t ra = group_rsrc(p.a + g * p.a_gs, group_range(groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb = group_rsrc(p.b + g * p.b_gs, group_range(groups_left, p.b_gs, SZ));
    const unsigned a_col = (unsigned)p.a_sj * V * SZ, b_col = (unsigned)p.b_sj * V * SZ;
    unsigned a_row[R], b_row[R];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        a_row[r] = (i < N) ? ((unsigned)(i * p.a_si) * V + v) * SZ : kOOB;
        b_row[r] = (i < N) ? ((unsigned)(i * p.b_si) * V + v) * SZ : kOOB;
    });

    /* L strictly below the diagonal (zeros elsewhere: no row masks in the
     * sweep), its diagonal apart */
    T a[R][N], dg[R];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        sfor<0, cmin(TS * (r + 1), N)>([&](auto c) {
            a[r][c] = bload(T(), ra, (c < TS * r || c < i) ? a_row[r] : kOOB, c * a_col);
        });
        const unsigned doff = (a_row[r] == kOOB || p.unit) ? kOOB : a_row[r] + (unsigned)i * a_col;
        dg[r] = bload(T(), ra, doff, 0);
    });

    /* right-hand sides in blocks of NRB; the next block's loads are issued
     * before the current block's sweep, so their latency hides behind it */
    auto load_block = [&](int q0, T(&bb)[R][NRB]) {
        sfor<0, NRB>([&](auto q) {
            const unsigned col = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB;
            sfor<0, R>([&](auto r) { bb[r][q] = bload(T(), rb, b_row[r], col); });
        });
    };
    T bn[R][NRB];
    if (p.nrhs > 0) load_block(0, bn);
    for (int q0 = 0; q0 < p.nrhs; q0 += NRB) {
        unsigned bcol[NRB];
        sfor<0, NRB>([&](auto q) { bcol[q] = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB; });
        T b[R][NRB];
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { b[r][q] = bn[r][q] * p.alpha; });
        });
        if (q0 + NRB < p.nrhs) load_block(q0 + NRB, bn);
        sfor<0, N>([&](auto kc) {
            constexpr int k = kc;
            constexpr int rk = k / TS, tk = k % TS;
            if (!p.unit) {
                sfor<0, NRB>([&](auto q) {
                    b[rk][q] = (t == tk) ? b[rk][q] / dg[rk] : b[rk][q];
                });
            }
            dpp_ready_from<0>(b[rk]);
            sfor<0, NRB>([&](auto q) {
                /* rows i <= k of x_k's own row block keep their value: an fma
                 * with a zero coefficient would still turn a -0 into +0 */
                T own = b[rk][q];
                fnma_bcr<tk>(own, b[rk][q], a[rk][k]);
                sfor<rk + 1, R>([&](auto r) { fnma_bcr<tk>(b[r][q], b[rk][q], a[r][k]); });
                b[rk][q] = (t > tk) ? own : b[rk][q];
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { bstore(b[r][q], rb, b_row[r], bcol[q]); });
        });
    }
}

template <class T, int N> hipError_t launch_n(const TrsmArgs<T> &args)
{
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    if (args.nrhs <= 4)
        hipLaunchKernelGGL((trsm_team_kernel<T, N, 4>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((trsm_team_kernel<T, N, 8>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    return hipGetLastError();
}

} /* namespace trsmteam */
} /* namespace cbk */

#endif /* CBK_TRSM_TEAM_IMPL_H */
