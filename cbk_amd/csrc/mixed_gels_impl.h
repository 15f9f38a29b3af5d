/* mixed_gels_impl.h -- mixed-precision least squares (dsgels_compact), one
 * P x Q matrix per 16-lane team.
 *
 * The FP32 Householder QR of A (the team QR of qr_team_impl.h, on the FP32
 * rounding of A) gives R with R^T R = A^T A + E, ||E|| ~ eps32 ||A||^2. R then
 * preconditions iterative refinement on the normal equations, whose residuals
 * are computed in FP64 from the FP64 A and B read from HBM:
 *     x_0 = 0;  r = b - A x,  g = A^T r      (FP64, the only FP64 arithmetic)
 *               R^T R d = g                  (FP32: two triangular sweeps)
 *               x += d                       (FP64)
 * The fixed point is A^T (b - A x) = 0 in FP64, the least-squares solution, so
 * X reaches FP64 accuracy while the factorization and the sweeps run in FP32;
 * each step contracts the error by about kappa(A)^2 eps32. Per column the
 * loop stops when ||d||_inf <= ||x||_inf * 2^-53 * sqrt(Q) (the dsposv test on
 * the correction), or when the correction no longer halves and is already
 * below ||x||_inf * 2^-40 (the rounding floor of A^T r, which kappa^2 lifts
 * above 2^-53 for inconsistent systems); a matrix whose FP32 R has a zero or
 * non-finite diagonal, whose correction grows or stalls above that floor, or
 * that runs ITERMAX steps is marked for the FP64 path (iter < 0), which a
 * second launch (the wavefront gels kernel on those matrices only) runs.
 *
 * The right-hand sides go through the loop in blocks of NRB (2 or 4) in
 * lockstep, so one pass over A per step serves the whole block; a column that
 * has converged keeps its x while the others finish.
 *
 * Layout in the team: lane t owns rows t + 16 r of A (as the QR team kernel);
 * after the QR only R's rows (r < Q/16) are kept, FP32. The sweeps work on a
 * right-hand side replicated in every lane of the team (g, then h, then d):
 *   R^T h = g : h_k = g_k / R(k,k); g_i -= R(k,i) h_k, R(k,i) broadcast from
 *               row k's lane (fused DPP fma)
 *   R d = h   : d_i = (h_i - sum_{k>i} R(i,k) d_k) / R(i,i) in row i's lane,
 *               then broadcast
 * and x stays distributed like the rows (lane t: x_t, x_{t+16}).
 */
#ifndef CBK_MIXED_GELS_IMPL_H
#define CBK_MIXED_GELS_IMPL_H

#include "cbk_internal.h"
#include "team_common.h"
#include "team_chol.h"

namespace cbk {
namespace mixedgels {

using namespace team;

struct MixedGelsArgs {
    int m, n; /* the matrix: m <= P rows, n <= Q columns of the kernel's bucket */
    const double *a; /* m x n, element (i, c) at (i*a_si + c*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    const double *b; /* p x nrhs */
    int b_si, b_sj;
    size_t b_gs;
    double *x; /* q x nrhs */
    int x_si, x_sj;
    size_t x_gs;
    int V;
    long nmat;
    int nm;
    int nrhs;
    int itermax;
    int *iter;
};

/* Teams per block. 16 (4 wavefronts): at V = 64 the block's 16 matrices fill
 * whole 128-byte lines, so a line that A's re-reads go through serves one CU.
 * At 64 x 32 a wavefront takes a whole SIMD's registers; one-wavefront blocks
 * in the XCD-aware order (team_common.h) keep the 4 wavefronts of a line on
 * one XCD's L2 without tying their schedules together. */
template <int P> constexpr int teams_per_block() { return P >= 64 ? 4 : 16; }

template <class T> __device__ __forceinline__ T team_sum(T x)
{
    x = x + dpp<0xB1>(x);
    x = x + dpp<0x4E>(x);
    x = x + dpp<0x141>(x);
    x = x + dpp<0x140>(x);
    return x;
}

/* resident wavefronts per SIMD: one for the large buckets, and whenever
 * four right-hand sides go through the loop together */
template <int P, int Q, int NRB> constexpr int gels_waves() { return (P * Q >= 1024 || NRB > 2) ? 1 : 2; }

/* EXACT: the matrix is P x Q (m and n fold away); otherwise m <= P rows and
 * n <= Q columns at run time */
template <int P, int Q, int NRB, bool EXACT>
__global__ __launch_bounds__(teams_per_block<P>() * TS)
__attribute__((amdgpu_waves_per_eu(gels_waves<P, Q, NRB>(), gels_waves<P, Q, NRB>() == 1 ? 1 : 8))) void
mixed_gels_kernel(MixedGelsArgs p)
{
    const int m = EXACT ? P : p.m, n = EXACT ? Q : p.n;
    constexpr int R = P / TS;  /* row blocks of A */
    constexpr int RQ = (Q + TS - 1) / TS; /* row blocks of R */
    constexpr unsigned SZ = 8;
    const int t = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;
    const long mat0 = (long)(P >= 64 ? xcd_block(blockIdx.x) : blockIdx.x) * teams_per_block<P>() +
                      4 * wave_in_block();
    if (mat0 >= p.nmat) return;
    const long mat = mat0 + (team & 3);
    const long g = mat0 / p.V;
    const long groups_left = p.nmat / p.V - g;
    const unsigned v = (unsigned)(mat0 % p.V) + (team & 3);
    const unsigned V = (unsigned)p.V;
    const __amdgpu_buffer_rsrc_t ra = group_rsrc(p.a + g * p.a_gs, group_range(groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb = group_rsrc(p.b + g * p.b_gs, group_range(groups_left, p.b_gs, SZ));
    const __amdgpu_buffer_rsrc_t rx = group_rsrc(p.x + g * p.x_gs, group_range(groups_left, p.x_gs, SZ));
    const unsigned a_col = (unsigned)p.a_sj * V * SZ, b_col = (unsigned)p.b_sj * V * SZ,
                   x_col = (unsigned)p.x_sj * V * SZ;
    /* rows past m read as zero (they change neither R nor the residuals);
     * columns past n read as zero and get R(c,c) = 1, which decouples them */
    unsigned a_row[R], b_row[R], x_row[RQ];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        a_row[r] = (i < m) ? ((unsigned)(i * p.a_si) * V + v) * SZ : kOOB;
        b_row[r] = (i < m) ? ((unsigned)(i * p.b_si) * V + v) * SZ : kOOB;
    });
    sfor<0, RQ>([&](auto r) {
        const int i = t + TS * r;
        x_row[r] = (i < n) ? ((unsigned)(i * p.x_si) * V + v) * SZ : kOOB;
    });
    bool colin[Q];
    sfor<0, Q>([&](auto c) { colin[c] = c < n; });

    /* ---- FP32 QR of A (geqr2, qr_team_impl.h), R only; the load of A also
     * forms the first step's g = A^T b of the first block of right-hand
     * sides (x_0 = 0), so that step needs no pass of its own over A ---- */
    float rr[RQ][Q];
    bool ok = true;
    /* (two right-hand sides at most: with four, the extra registers cost more
     * than the pass they save) */
    constexpr bool FOLD = NRB <= 2;
    double gp0[FOLD ? NRB : 1][Q];
    {
        float f[R][Q];
        sfor<0, FOLD ? NRB : 1>([&](auto j) { sfor<0, Q>([&](auto c) { gp0[j][c] = 0.0; }); });
        sfor<0, R>([&](auto r) {
            double bv[NRB];
            if constexpr (FOLD) {
                sfor<0, NRB>([&](auto j) {
                    bv[j] = bload(0.0, rb, b_row[r], j < p.nrhs ? j * b_col : kOOB);
                });
            }
            sfor<0, Q>([&](auto c) {
                const double a = bload(0.0, ra, colin[c] ? a_row[r] : kOOB, c * a_col);
                f[r][c] = (float)a;
                if constexpr (FOLD) {
                    sfor<0, NRB>([&](auto j) { gp0[j][c] = __builtin_fma(a, bv[j], gp0[j][c]); });
                }
            });
        });
        sfor<0, Q>([&](auto kc) {
            constexpr int kk = kc;
            constexpr int rk = kk / TS, tk = kk % TS;
            float part = 0.f;
            sfor<rk, R>([&](auto r) {
                const int i = t + TS * r;
                const float xv = (i > kk) ? f[r][kk] : 0.f;
                part = __builtin_fmaf(xv, xv, part);
            });
            const float tail = team_sum(part);
            const float x0 = bcast<tk>(f[rk][kk]);
            const float norm = __builtin_sqrtf(__builtin_fmaf(x0, x0, tail));
            const float beta = (x0 >= 0.f) ? -norm : norm;
            const bool has = tail > 0.f;
            const float tau = has ? (beta - x0) / beta : 0.f;
            const float inv = has ? 1.f / (x0 - beta) : 0.f;
            const float rdiag = (kk < n) ? (has ? beta : x0) : 1.f;
            float vi[R];
            sfor<rk, R>([&](auto r) {
                const int i = t + TS * r;
                f[r][kk] = (i > kk) ? f[r][kk] * inv : (i == kk ? rdiag : f[r][kk]);
                vi[r] = (i > kk) ? f[r][kk] : (i == kk ? 1.f : 0.f);
            });
            sfor<kk + 1, Q>([&](auto c) {
                float pw = 0.f;
                sfor<rk, R>([&](auto r) { pw = __builtin_fmaf(vi[r], f[r][c], pw); });
                const float w = team_sum(pw) * tau;
                sfor<rk, R>([&](auto r) { f[r][c] = __builtin_fmaf(-vi[r], w, f[r][c]); });
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        /* keep R (rows < Q, upper triangle); a zero or non-finite diagonal
         * sends the matrix to the FP64 path */
        sfor<0, RQ>([&](auto r) {
            const int i = t + TS * r;
            sfor<0, Q>([&](auto c) { rr[r][c] = (c >= i) ? f[r][c] : 0.f; });
        });
        sfor<0, RQ>([&](auto r) { dpp_ready_from<0>(rr[r]); }); /* the sweeps' DPP sources */
        sfor<0, Q>([&](auto kc) {
            const float d = bcast<kc % TS>(rr[kc / TS][kc]);
            ok = ok && (d != 0.f) && __builtin_isfinite(d);
        });
    }

    int steps = 0;
    const double cte = 0x1p-53 * __builtin_sqrt((double)n), floor_ = 0x1p-40;
    /* NRB right-hand sides in lockstep: one pass over A per step serves them all */
    for (int q0 = 0; q0 < p.nrhs && ok; q0 += NRB) {
        double x[NRB][RQ], prev[NRB];
        bool done[NRB];
        unsigned bcol[NRB], xcol[NRB];
        sfor<0, NRB>([&](auto j) {
            sfor<0, RQ>([&](auto r) { x[j][r] = 0.0; });
            prev[j] = __builtin_inf();
            done[j] = q0 + j >= p.nrhs;
            bcol[j] = done[j] ? kOOB : (q0 + j) * b_col;
            xcol[j] = done[j] ? kOOB : (q0 + j) * x_col;
        });
        bool all = false;
        for (int it = 0; it <= p.itermax && ok && !all; ++it) {
            /* g = A^T (b - A x), FP64, one row block of A at a time */
            double gp[NRB][Q];
            const bool first = FOLD && it == 0 && q0 == 0;
            sfor<0, NRB>([&](auto j) {
                sfor<0, Q>([&](auto c) { gp[j][c] = first ? gp0[FOLD ? j : 0][c] : 0.0; });
            });
            sfor<0, NRB>([&](auto j) { dpp_ready_from<0>(x[j]); });
            if (!first) sfor<0, R>([&](auto r) {
                double arow[Q];
                sfor<0, Q>([&](auto c) { arow[c] = bload(0.0, ra, colin[c] ? a_row[r] : kOOB, c * a_col); });
                sfor<0, NRB>([&](auto j) {
                    double res = bload(0.0, rb, b_row[r], bcol[j]);
                    sfor<0, Q>([&](auto c) { fnma_bcr<c % TS>(res, x[j][c / TS], arow[c]); });
                    sfor<0, Q>([&](auto c) { gp[j][c] = __builtin_fma(arow[c], res, gp[j][c]); });
                });
                __builtin_amdgcn_sched_barrier(0);
            });
            all = true;
            sfor<0, NRB>([&](auto j) {
                float h[Q];
                sfor<0, Q>([&](auto c) { h[c] = (float)team_sum(gp[j][c]); });
                /* R^T h = g */
                sfor<0, Q>([&](auto kc) {
                    constexpr int k = kc;
                    const float dk = bcast<k % TS>(rr[k / TS][k]);
                    h[k] = h[k] / dk;
                    sfor<k + 1, Q>([&](auto i) { fnma_bcr<k % TS>(h[i], rr[k / TS][i], h[k]); });
                });
                /* R d = h, d overwriting h from the bottom */
                sfor_down<0, Q>([&](auto ic) {
                    constexpr int i = ic;
                    float s = h[i];
                    sfor<i + 1, Q>([&](auto k) { s = __builtin_fmaf(-rr[i / TS][k], h[k], s); });
                    s = s / rr[i / TS][i];
                    h[i] = bcast<i % TS>(s);
                });
                if (!done[j]) {
                    double dn = 0.0;
                    sfor<0, Q>([&](auto i) {
                        dn = fmax_nan(dn, (double)__builtin_fabsf(h[i]));
                        x[j][i / TS] = x[j][i / TS] + ((t == i % TS) ? (double)h[i] : 0.0);
                    });
                    double xn = 0.0;
                    sfor<0, RQ>([&](auto r) { xn = fmax_nan(xn, __builtin_fabs(x[j][r])); });
                    xn = team_max(xn);
                    steps = it > steps ? it : steps;
                    if (dn <= xn * cte || (it > 0 && dn > 0.5 * prev[j] && dn <= xn * floor_))
                        done[j] = true;
                    else if ((!(dn < prev[j]) && it > 0) || it == p.itermax)
                        ok = false; /* not contracting, or out of steps: FP64 path */
                    prev[j] = dn;
                }
                all = all && done[j];
            });
        }
        ok = ok && all;
        if (ok) {
            sfor<0, NRB>([&](auto j) {
                sfor<0, RQ>([&](auto r) { bstore(x[j][r], rx, x_row[r], xcol[j]); });
            });
        }
    }
    if (mat < p.nm && t == 0) p.iter[mat] = ok ? steps : -p.itermax - 1;
}

template <int P, int Q, bool EXACT> hipError_t launch_pq(const MixedGelsArgs &args)
{
    constexpr int TPB = teams_per_block<P>();
    unsigned blocks = (unsigned)((args.nmat + TPB - 1) / TPB);
    if (P >= 64) blocks = xcd_grid(blocks);
    if (args.nrhs <= 2)
        hipLaunchKernelGGL((mixed_gels_kernel<P, Q, 2, EXACT>), dim3(blocks), dim3(TPB * TS), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((mixed_gels_kernel<P, Q, 4, EXACT>), dim3(blocks), dim3(TPB * TS), 0,
                           current_stream(), args);
    return hipGetLastError();
}

} /* namespace mixedgels */
} /* namespace cbk */

#endif /* CBK_MIXED_GELS_IMPL_H */
