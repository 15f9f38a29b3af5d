/* mixed_team_impl.h -- mixed-precision SPD solve (LAPACK dsposv semantics),
 * one matrix per 16-lane team, fused in one pass over HBM.
 *
 * Per matrix: A (FP64) and B are read once; A stays in registers as full
 * symmetric rows (the upper half read from the mirrored lower storage), the
 * FP32 Cholesky factor and its LDS copy are built from it, X0 = the FP32
 * solve, then LAPACK dsposv's refinement: r = b - A x in FP64 (the only FP64
 * arithmetic), d = the FP32 solve of r, x += d in FP64, until every column
 * satisfies ||r||_inf <= ||x||_inf * ||A||_inf * eps * sqrt(n) (eps = 2^-53,
 * dlamch('E')) or ITERMAX steps pass. On success A is untouched and only X is
 * written: 8 n(n+1)/2 + 16 n nrhs bytes of compulsory traffic against the plain
 * dposv's 8 n(n+1) + 16 n nrhs. If the FP32 factor fails, a value overflows
 * FP32, or refinement does not converge, the matrix falls back to the FP64
 * Cholesky (team_chol.h) in place, A receives that factor, and iter records
 * why (-2 overflow, -3 FP32 factor failed, -(ITERMAX+1) no convergence), as
 * dsposv's ITER does.
 *
 * The FP32 factor (chol_factor_fast: hardware reciprocal square root) and the
 * FP32 solves are the team building blocks of team_chol.h; the solves read the
 * factor from the team's LDS (chol_solve_cols), so that during refinement the
 * registers hold the FP64 A, B, X and the residual only, and B is loaded after
 * the factor: two wavefronts fit a SIMD at n = 32, four at n = 16. The
 * symmetric residual sums A(i,c) x_c over c ascending with one FP64 fma each.
 */
#ifndef CBK_MIXED_TEAM_IMPL_H
#define CBK_MIXED_TEAM_IMPL_H

#include "cbk_internal.h"
#include "team_chol.h"

namespace cbk {
namespace mixedteam {

using namespace team;

struct MixedArgs {
    double *a; /* lower view */
    int a_si, a_sj;
    size_t a_gs;
    const double *b;
    int b_si, b_sj;
    size_t b_gs;
    double *x;
    int x_si, x_sj;
    size_t x_gs;
    int V;
    long nmat;
    int nm;
    int n; /* the order: N, or below N when padded */
    int nrhs;
    int itermax;
    int *iter;
    int *info;
};

/* 4 teams (one wavefront) per block: with few wavefronts per SIMD a 4-wave
 * block would hold its CU's slots until its slowest matrix converged */
constexpr int kTeamsPerBlock = 4;

template <int N> hipError_t launch_n(bool ldl, const MixedArgs &args);

/* whether c holds in any lane of the calling lane's team (a wave-wide ballot,
 * then the team's 16 bits) */
__device__ __forceinline__ bool team_any(bool c)
{
    const unsigned long long m = __builtin_amdgcn_ballot_w64(c);
    return ((m >> (threadIdx.x & 48u)) & 0xFFFFull) != 0;
}

/* The largest of each of 4 per-lane values over the team, in every lane,
 * NaN-propagating: a transposed butterfly -- lane pairs (t, t^1) exchange
 * the half of the values each does not keep, pairs (t, t^2) the quarter, then
 * rotations by 4 and 8 combine the four lanes that keep the same value --
 * and one broadcast per value back. 5 exchanges instead of 16. */
__device__ __forceinline__ void team_max4(double (&x)[4], int t)
{
    const bool b0 = (t & 1) != 0, b1 = (t & 2) != 0;
    double k0 = b0 ? x[2] : x[0], k1 = b0 ? x[3] : x[1];
    const double s0 = b0 ? x[0] : x[2], s1 = b0 ? x[1] : x[3];
    k0 = fmax_nan(k0, dpp<0xB1>(s0)); /* quad_perm [1,0,3,2] */
    k1 = fmax_nan(k1, dpp<0xB1>(s1));
    double k = b1 ? k1 : k0;
    const double s = b1 ? k0 : k1;
    k = fmax_nan(k, dpp<0x4E>(s));    /* quad_perm [2,3,0,1] */
    k = fmax_nan(k, dpp<0x124>(k));   /* row_ror:4 */
    k = fmax_nan(k, dpp<0x128>(k));   /* row_ror:8 */
    /* lane b0 + 2 b1 now holds value 2 b0 + b1 */
    x[0] = bcast<0>(k);
    x[1] = bcast<2>(k);
    x[2] = bcast<1>(k);
    x[3] = bcast<3>(k);
}

/* PAD: an order n < N solved as diag(A, I) of order N -- rows and columns
 * from n on load zero and a unit diagonal, nothing is stored there, and the
 * norms, the stopping test and its sqrt(n) use the n real rows. The factor's
 * entries of A and the residual sums of its rows are those of order n (the
 * padded columns follow the real ones and add exact zeros); the unpadded
 * instances keep N as a constant. */
template <int N, int NRB, bool LDL, bool PAD = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(N <= 16 ? 4 : 2, N <= 16 ? 4 : 2))) void
mixed_posv_kernel(MixedArgs p)
{
    static_assert(NRB == 4, "the convergence test reduces four columns at once");
    constexpr int R = rows_per_lane<N>();
    const int n = PAD ? p.n : N;
    /* a team's region: the FP32 column copy during refinement, the FP64
     * packed rows of chol_prepare on the fallback */
    constexpr int TW = chol_lds<N>() > (cols_lds<N>() + 1) / 2 ? chol_lds<N>() : (cols_lds<N>() + 1) / 2;
    __shared__ double lds[kTeamsPerBlock * TW];

    const int t = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;                    /* 0..15 in the block */
    const long mat0 = (long)xcd_block(blockIdx.x) * kTeamsPerBlock + 4 * wave_in_block(); /* wave-uniform */
    if (mat0 >= p.nmat) return;                           /* whole wavefront */
    const long mat = mat0 + (team & 3);
    const long g = mat0 / p.V;
    const long groups_left = p.nmat / p.V - g;
    const unsigned v = (unsigned)(mat0 % p.V) + (team & 3);
    const unsigned V = (unsigned)p.V;
    constexpr unsigned SZ = 8;
    const __amdgpu_buffer_rsrc_t ra = group_rsrc(p.a + g * p.a_gs, group_range(groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb = group_rsrc(p.b + g * p.b_gs, group_range(groups_left, p.b_gs, SZ));
    const __amdgpu_buffer_rsrc_t rx = group_rsrc(p.x + g * p.x_gs, group_range(groups_left, p.x_gs, SZ));
    const unsigned a_col = (unsigned)p.a_sj * V * SZ, a_rowb = (unsigned)p.a_si * V * SZ;
    const unsigned b_colb = (unsigned)p.b_sj * V * SZ, x_colb = (unsigned)p.x_sj * V * SZ;
    unsigned a_row[R], a_cbase[R], b_row[R], x_row[R];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        const bool in = i < n;
        a_row[r] = in ? ((unsigned)(i * p.a_si) * V + v) * SZ : kOOB;   /* (i, c): + c*a_col */
        a_cbase[r] = in ? ((unsigned)(i * p.a_sj) * V + v) * SZ : kOOB; /* (c, i): + c*a_rowb */
        b_row[r] = in ? ((unsigned)(i * p.b_si) * V + v) * SZ : kOOB;
        x_row[r] = in ? ((unsigned)(i * p.x_si) * V + v) * SZ : kOOB;
    });

    /* full symmetric rows of A in FP64: (i, c) from the lower view's (i, c)
     * when c <= i, from its (c, i) otherwise */
    double a64[R][N];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        sfor<0, N>([&](auto c) {
            const unsigned off = (a_row[r] == kOOB || (PAD && c >= n)) ? kOOB
                                 : (c <= i)                           ? a_row[r] + c * a_col
                                                                      : a_cbase[r] + c * a_rowb;
            a64[r][c] = bload(0.0, ra, off, 0);
            if constexpr (PAD && c >= TS * r && c < TS * (r + 1))
                a64[r][c] = (c >= n && c == i) ? 1.0 : a64[r][c];
        });
    });
    /* ||A||_inf (dlansy 'I') and the FP32 overflow test of A (dlat2s) */
    double rowsum[R];
    bool over = false;
    sfor<0, R>([&](auto r) {
        rowsum[r] = 0.0;
        sfor<0, N>([&](auto c) {
            rowsum[r] += __builtin_fabs(a64[r][c]);
            over |= __builtin_fabs(a64[r][c]) > 3.4028234663852886e38;
        });
    });
    double anrm = 0.0;
    sfor<0, R>([&](auto r) {
        if (t + TS * r < n) anrm = fmax_nan(anrm, rowsum[r]);
    });
    anrm = team_max(anrm);
    const double cte = anrm * 0x1p-53 * __builtin_sqrt((double)n);
    /* any lane of the team overflowing sends the whole matrix to FP64 */
    const bool a_over = team_any(over);

    int iter = 0;
    bool failed = false;
    float *L32 = reinterpret_cast<float *>(lds + team * TW);
    float dsc32[R];
    if (!a_over) {
        /* the FP32 factor: Cholesky with the hardware reciprocal square root,
         * or the unpivoted LDL^T (which fails on a zero or non-finite D) */
        float a32[R][N], dg32[R];
        sfor<0, R>([&](auto r) {
            sfor<0, cmin(TS * (r + 1), N)>([&](auto c) { a32[r][c] = (float)a64[r][c]; });
        });
        int finfo;
        if constexpr (LDL) {
            finfo = chol_factor<float, N, true>(a32, dg32, t);
            bool bad = false;
            sfor<0, R>([&](auto r) {
                if (t + TS * r < n) bad |= !__builtin_isfinite(dg32[r]);
            });
            if (team_any(bad)) finfo = 1;
        }
        else {
            finfo = chol_factor_fast<N>(a32, dg32, t);
        }
        if (finfo != 0) {
            iter = -3;
            failed = true;
        }
        else {
            chol_prepare_cols<N, LDL>(a32, dg32, dsc32, L32, t);
        }
    }
    else {
        iter = -2;
        failed = true;
    }

    /* right-hand sides in blocks of NRB, each refined until its columns pass
     * (iter: the most steps a block took); B is read after the factor, whose
     * registers it would otherwise share */
    for (int q0 = 0; q0 < p.nrhs; q0 += NRB) {
        unsigned bcol[NRB], xcol[NRB];
        sfor<0, NRB>([&](auto q) {
            bcol[q] = (q0 + q < p.nrhs) ? (q0 + q) * b_colb : kOOB;
            xcol[q] = (q0 + q < p.nrhs) ? (q0 + q) * x_colb : kOOB;
        });
        double b64[R][NRB];
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { b64[r][q] = bload(0.0, rb, b_row[r], bcol[q]); });
        });
        /* an overflowing B (dlag2s) sends the matrix to FP64 like an
         * overflowing A, whatever the factor did (dsposv tests B first) */
        bool b_over = false;
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { b_over |= __builtin_fabs(b64[r][q]) > 3.4028234663852886e38; });
        });
        if (team_any(b_over)) {
            iter = -2;
            failed = true;
        }
        if (failed) continue; /* the team's lanes alike */
        double x64[R][NRB];
        float w[R][NRB];
        sfor<0, R>([&](auto r) { sfor<0, NRB>([&](auto q) { w[r][q] = (float)b64[r][q]; }); });
        chol_solve_cols<N, NRB>(dsc32, w, L32, t);
        sfor<0, R>([&](auto r) { sfor<0, NRB>([&](auto q) { x64[r][q] = (double)w[r][q]; }); });
        int it_blk = 0;
        for (int it = 0;; ++it) {
            /* r = b - A x, FP64 */
            double res[R][NRB];
            sfor<0, R>([&](auto r) { sfor<0, NRB>([&](auto q) { res[r][q] = b64[r][q]; }); });
            sfor<0, R>([&](auto r) { dpp_ready_from<0>(x64[r]); });
            sfor<0, N>([&](auto c) {
                sfor<0, NRB>([&](auto q) {
                    sfor<0, R>([&](auto r) { fnma_bcr<c % TS>(res[r][q], x64[c / TS][q], a64[r][c]); });
                });
            });
            /* every column: ||r||_inf <= ||x||_inf * cte, tested per row
             * against the team's ||x||_inf (a NaN fails either way) */
            double xm[NRB];
            sfor<0, NRB>([&](auto q) {
                xm[q] = 0.0;
                sfor<0, R>([&](auto r) {
                    if (t + TS * r < n) xm[q] = fmax_nan(xm[q], __builtin_fabs(x64[r][q]));
                });
            });
            team_max4(xm, t);
            bool fail = false;
            sfor<0, NRB>([&](auto q) {
                sfor<0, R>([&](auto r) {
                    if (q0 + q < p.nrhs && t + TS * r < n)
                        fail = fail || !(__builtin_fabs(res[r][q]) <= xm[q] * cte);
                });
            });
            if (!team_any(fail)) {
                it_blk = it;
                break;
            }
            if (it == p.itermax) {
                iter = -p.itermax - 1;
                failed = true;
                break;
            }
            bool rover = false;
            sfor<0, R>([&](auto r) {
                sfor<0, NRB>([&](auto q) {
                    rover |= __builtin_fabs(res[r][q]) > 3.4028234663852886e38;
                    w[r][q] = (float)res[r][q];
                });
            });
            if (team_any(rover)) {
                iter = -2;
                failed = true;
                break;
            }
            chol_solve_cols<N, NRB>(dsc32, w, L32, t);
            sfor<0, R>([&](auto r) {
                sfor<0, NRB>([&](auto q) { x64[r][q] = x64[r][q] + (double)w[r][q]; });
            });
        }
        if (failed) continue;
        iter = it_blk > iter ? it_blk : iter;
        sfor<0, R>([&](auto r) {
            sfor<0, NRB>([&](auto q) { bstore(x64[r][q], rx, x_row[r], xcol[q]); });
        });
    }

    int dinfo = 0;
    if (__builtin_amdgcn_ballot_w64(failed) != 0) {
        /* FP64 fallback in place on the failed matrices: A receives the factor
         * (bit-identical to the reference's potrf / sytrfnp), X the solution */
        double a[R][N], dg[R], dinv[R];
        sfor<0, R>([&](auto r) {
            sfor<0, cmin(TS * (r + 1), N)>([&](auto c) { a[r][c] = a64[r][c]; });
        });
        dinfo = chol_factor<double, N, LDL>(a, dg, t);
        sfor<0, R>([&](auto r) {
            const int i = t + TS * r;
            sfor<0, cmin(TS * (r + 1), N)>([&](auto c) {
                const unsigned off = (failed && (c < TS * r || c < i)) ? a_row[r] : kOOB;
                bstore(a[r][c], ra, off, c * a_col);
            });
            const unsigned doff = (!failed || a_row[r] == kOOB) ? kOOB : a_row[r] + (unsigned)i * a_col;
            bstore(dg[r], ra, doff, 0);
        });
        chol_prepare<double, N, LDL>(a, dg, dinv, lds + team * TW, t);
        const bool put = failed && dinfo == 0;
        for (int q0 = 0; q0 < p.nrhs; q0 += NRB) {
            unsigned bcol[NRB], xcol[NRB];
            sfor<0, NRB>([&](auto q) {
                bcol[q] = (q0 + q < p.nrhs) ? (q0 + q) * b_colb : kOOB;
                xcol[q] = (q0 + q < p.nrhs) ? (q0 + q) * x_colb : kOOB;
            });
            double xb[R][NRB];
            sfor<0, R>([&](auto r) {
                sfor<0, NRB>([&](auto q) { xb[r][q] = bload(0.0, rb, b_row[r], bcol[q]); });
            });
            chol_solve<double, N, NRB, LDL>(a, dinv, xb, lds + team * TW, t);
            sfor<0, R>([&](auto r) {
                sfor<0, NRB>([&](auto q) { bstore(xb[r][q], rx, put ? x_row[r] : kOOB, xcol[q]); });
            });
        }
        if (!failed) dinfo = 0;
    }
    if (mat < p.nm && t == 0) {
        p.iter[mat] = iter;
        p.info[mat] = dinfo;
    }
}

template <int N, bool LDL, bool PAD> void launch_np(const MixedArgs &args)
{
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    hipLaunchKernelGGL((mixed_posv_kernel<N, 4, LDL, PAD>), dim3(xcd_grid(blocks)), dim3(64), 0,
                       current_stream(), args);
}

template <int N> hipError_t launch_n(bool ldl, const MixedArgs &args)
{
    const bool pad = args.n != N;
    if (ldl)
        pad ? launch_np<N, true, true>(args) : launch_np<N, true, false>(args);
    else
        pad ? launch_np<N, false, true>(args) : launch_np<N, false, false>(args);
    return hipGetLastError();
}

} /* namespace mixedteam */
} /* namespace cbk */

#endif /* CBK_MIXED_TEAM_IMPL_H */
