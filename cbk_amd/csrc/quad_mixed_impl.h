/* quad_mixed_impl.h -- mixed-precision posv / sysvnp (LAPACK dsposv
 * semantics) with one matrix per quad of lanes (quad_common.h), 16 matrices
 * per wavefront.
 *
 * Per matrix, A (FP64, the referenced triangle) is read once and B once per
 * refinement step (the first from HBM, the others from L2); on success A is
 * untouched and only X is written.
 *
 * Registers hold A in FP64 for the residual: lane t of a quad keeps rows
 * t + 4k up to the end of their diagonal 4x4 block (that block completed
 * from the mirrored entries), and every 4-column block m stored rotated by the
 * lane index -- slot 4m + j holds column 4m + ((t + j) mod 4). The symmetric
 * product r = b - A x then needs no per-lane selects: the row part takes
 * x(4m + (t + j) mod 4) from lane t + j of the quad, the column part (the
 * entries above the diagonal, read from the rows below it) sends its partial
 * sums for column 4m + (t + j) mod 4 to lane t + j; both are one DPP quad
 * rotation per value.
 *
 * The FP32 factor (right-looking Cholesky with the hardware reciprocal square
 * root, or the unpivoted LDL^T with the hardware reciprocal) is computed from
 * a second, unrotated read of the same lines and parked in the wavefront's
 * LDS; the sweeps read it from there and run as plain FP32 fma's at the full
 * FP32 rate, each pivot value reaching its quad by one DPP move. Refinement
 * per block of NRB right-hand sides: r = b - A x in FP64, the test
 * ||r||_inf <= ||x||_inf ||A||_inf eps sqrt(n) per column, d = the FP32 solve
 * of r, x += d in FP64, up to itermax steps. A matrix whose FP32 factor fails
 * (iter -3), whose data or a residual overflow FP32 (-2), or that does not
 * converge (-(itermax + 1)) has B copied to X and is left to the FP64 pass
 * that follows (the quad posv / sysvnp kernel masked by iter), which factors A
 * in place and solves X.
 */
#ifndef CBK_QUAD_MIXED_IMPL_H
#define CBK_QUAD_MIXED_IMPL_H

#include "cbk_internal.h"
#include "quad_chol.h"

namespace cbk {
namespace quadmixed {

using namespace quad;
using team::bload;
using team::bstore;
using team::fmax_nan;
using team::group_range;
using team::group_rsrc;

struct QuadMixedArgs {
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

/* right-hand sides per refinement block. The FP64 triangle of order 24 (84
 * doubles a lane) and two columns of x, r and the FP32 sweep fill the
 * wavefront's registers without scratch; at order 32 (144 doubles a lane)
 * the kernel spills even with one column, so that order stays on the 16-lane
 * team kernel. */
constexpr int nrb_of(int) { return 2; }

typedef float f32x4 __attribute__((ext_vector_type(4)));

/* value of lane (t + J) mod 4 of this lane's quad (quad_perm rotations) */
template <int J> constexpr int rot_ctrl() { return J == 1 ? 0x39 : J == 2 ? 0x4E : 0x93; }
template <int J> __device__ __forceinline__ double qrot(double x)
{
    if constexpr (J == 0) return x;
    else {
        int2 w = __builtin_bit_cast(int2, x);
        w.x = __builtin_amdgcn_mov_dpp(w.x, rot_ctrl<J>(), 0xF, 0xF, true);
        w.y = __builtin_amdgcn_mov_dpp(w.y, rot_ctrl<J>(), 0xF, 0xF, true);
        return __builtin_bit_cast(double, w);
    }
}

/* LDS floats of one quad's factor: slot k's 4 rows, each padded to 4k + 4
 * entries (8 K (K + 1) in all), plus 4 so that consecutive quads start 4
 * banks apart */
template <int N> constexpr int quad_lds() { return 8 * (N / 4) * (N / 4 + 1) + 4; }
/* first float of slot k's rows in a quad's region */
constexpr int slot_base(int k) { return 8 * k * (k + 1); }

/* r -= A x for the rotated symmetric storage described above; a[k][4m + j],
 * m <= k, x and r as [slot][rhs] */
template <int N, int NRB>
__device__ __forceinline__ void sym_residual(const double (&a)[N / 4][N], const double (&x)[N / 4][NRB],
                                             double (&r)[N / 4][NRB])
{
    constexpr int K = N / 4;
    sfor<0, K>([&](auto mc) {
        constexpr int m = mc;
        double xr[4][NRB];
        sfor<0, NRB>([&](auto q) {
            xr[0][q] = x[m][q];
            xr[1][q] = qrot<1>(x[m][q]);
            xr[2][q] = qrot<2>(x[m][q]);
            xr[3][q] = qrot<3>(x[m][q]);
        });
        /* row part: the rows of slots k >= m times column block m */
        sfor<m, K>([&](auto k) {
            sfor<0, 4>([&](auto j) {
                sfor<0, NRB>([&](auto q) { r[k][q] = __builtin_fma(-a[k][4 * m + j], xr[j][q], r[k][q]); });
            });
        });
        /* column part: block m's columns from the rows of slots k > m */
        if constexpr (m + 1 < K) {
            double P[4][NRB];
            sfor<0, 4>([&](auto j) {
                sfor<0, NRB>([&](auto q) { P[j][q] = a[m + 1][4 * m + j] * x[m + 1][q]; });
            });
            sfor<m + 2, K>([&](auto k) {
                sfor<0, 4>([&](auto j) {
                    sfor<0, NRB>([&](auto q) { P[j][q] = __builtin_fma(a[k][4 * m + j], x[k][q], P[j][q]); });
                });
            });
            /* lane t's column 4m + t: partial j of lane t - j */
            sfor<0, NRB>([&](auto q) {
                const double s = (P[0][q] + qrot<3>(P[1][q])) + (qrot<2>(P[2][q]) + qrot<1>(P[3][q]));
                r[m][q] = r[m][q] - s;
            });
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

/* row sums of |A| of the rows t + 4k (dlansy 'I' before the maximum) */
template <int N>
__device__ __forceinline__ void sym_abs_rowsum(const double (&a)[N / 4][N], double (&rs)[N / 4])
{
    constexpr int K = N / 4;
    sfor<0, K>([&](auto k) { rs[k] = 0.0; });
    sfor<0, K>([&](auto mc) {
        constexpr int m = mc;
        sfor<m, K>([&](auto k) {
            sfor<0, 4>([&](auto j) { rs[k] += __builtin_fabs(a[k][4 * m + j]); });
        });
        if constexpr (m + 1 < K) {
            double P[4] = {0.0, 0.0, 0.0, 0.0};
            sfor<m + 1, K>([&](auto k) {
                sfor<0, 4>([&](auto j) { P[j] += __builtin_fabs(a[k][4 * m + j]); });
            });
            rs[m] += (P[0] + qrot<3>(P[1])) + (qrot<2>(P[2]) + qrot<1>(P[3]));
        }
    });
}

/* FP32 right-looking factor of a[k][c] = (t + 4k, c), c < 4k + 4 (entries
 * right of the diagonal are scratch): Cholesky (column scaled by the hardware
 * reciprocal square root of the pivot) or LDL^T (by the hardware reciprocal of
 * D). On return column j holds L(., j) below the diagonal, dinv[k] is 1/L(i,i)
 * or 1/D(i) of the lane's row i = t + 4k; returns whether a pivot failed
 * (Cholesky: not > 0; LDL^T: zero or not finite), alike in the quad's lanes. */
template <int N, bool LDL, bool MF = false>
__device__ __forceinline__ bool factor32(float (&a)[N / 4][N], float (&dinv)[N / 4], int t)
{
    static_assert(!MF || N % 4 == 0, "whole tiles");
    constexpr int K = N / 4;
    bool bad = false;
    sfor<0, K>([&](auto k) { dinv[k] = 1.0f; });
    sfor<0, N>([&](auto jc) {
        constexpr int j = jc;
        constexpr int kj = j / 4, tj = j % 4;
        const float piv = qbc<tj>(a[kj][j]);
        float sc, dmul;
        if constexpr (LDL) {
            bad |= (piv == 0.0f) || !__builtin_isfinite(piv);
            sc = __builtin_amdgcn_rcpf(piv);
            dmul = piv; /* a(c, j) = L(c, j) D(j) */
        }
        else {
            bad |= !(piv > 0.0f);
            sc = __builtin_amdgcn_rsqf(piv);
            dmul = 1.0f;
        }
        dinv[kj] = (t == tj) ? sc : dinv[kj];
        sfor<kj, K>([&](auto k) { a[k][j] = a[k][j] * sc; });
        /* with MF, the column blocks right of j's own by MFMA tiles
         * (quad::mfma_tile; the same single rounding per element) */
        constexpr int cv = MF ? 4 * kj + 4 : N;
        sfor<j + 1, cv>([&](auto c) {
            const float w = LDL ? qbc<c % 4>(a[c / 4][j]) * dmul : qbc<c % 4>(a[c / 4][j]);
            sfor<c / 4, K>([&](auto k) { a[k][c] = __builtin_fmaf(-a[k][j], w, a[k][c]); });
        });
        if constexpr (MF) {
            sfor<kj + 1, K>([&](auto cb) {
                const float x = -(LDL ? a[cb][j] * dmul : a[cb][j]);
                sfor<cb, K>([&](auto k) { mfma_tile<cb>(a[k], x, a[k][j]); });
            });
        }
        __builtin_amdgcn_sched_barrier(0);
    });
    return bad;
}

/* X := A^{-1} B with the factor in the quad's LDS region Lq (slot k's row
 * t + 4k at slot_base(k) + t (4k + 4), zeros on and right of the diagonal) and
 * dinv as factor32 left it; b[k][q] = B(t + 4k, q) */
template <int N, int NRB, bool LDL>
__device__ __forceinline__ void solve32(float (&b)[N / 4][NRB], const float (&dinv)[N / 4],
                                        const float *Lq, int t)
{
    constexpr int K = N / 4;
    /* forward, by column blocks of 4: L(t + 4k, 4m .. 4m + 3) in one read */
    sfor<0, K>([&](auto mc) {
        constexpr int m = mc;
        f32x4 Lf[K];
        sfor<m, K>([&](auto k) {
            Lf[k] = *reinterpret_cast<const f32x4 *>(Lq + slot_base(k) + t * (4 * k + 4) + 4 * m);
        });
        sfor<0, 4>([&](auto s) {
            sfor<0, NRB>([&](auto q) {
                const float y = qbc<s>(LDL ? b[m][q] : b[m][q] * dinv[m]);
                sfor<m, K>([&](auto k) { b[k][q] = __builtin_fmaf(-Lf[k][(int)s], y, b[k][q]); });
            });
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    sfor<0, K>([&](auto k) {
        sfor<0, NRB>([&](auto q) { b[k][q] = b[k][q] * dinv[k]; });
    });
    /* backward: row l of L from its owner's LDS row, entries 4m + t */
    const float *Lt = Lq + t;
    sfor_down<0, K>([&](auto kbc) {
        constexpr int kb = kbc;
        sfor_down<0, 4>([&](auto sc) {
            constexpr int s = sc;
            float cf[kb + 1];
            sfor<0, kb + 1>([&](auto m) { cf[m] = Lt[slot_base(kb) + s * (4 * kb + 4) + 4 * m]; });
            sfor<0, NRB>([&](auto q) {
                const float xv = qbc<s>(LDL ? b[kb][q] : b[kb][q] * dinv[kb]);
                sfor<0, kb + 1>([&](auto m) { b[m][q] = __builtin_fmaf(-cf[m], xv, b[m][q]); });
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

/* PAD: an order n < N solved as diag(A, I) of order N -- rows and columns
 * from n on read as zero (the FP32 factor's diagonal there as 1), nothing is
 * stored there, and the stopping test's rows and sqrt(n) are the n real ones;
 * the padded rows keep b = x = r = 0 */
template <int N, int NRB, bool LDL, bool PAD = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void
quad_mixed_kernel(QuadMixedArgs p)
{
    constexpr int K = N / 4;
    constexpr unsigned SZ = 8;
    constexpr int QS = quad_lds<N>();
    constexpr float FMAX = 3.4028234663852886e38f;
    __shared__ float lds[4 * kMatsPerWave * QS];

    QuadPlace pl;
    if (!quad_place(p.nmat, p.V, pl)) return; /* whole wavefront */
    const int t = pl.t;
    const int n = PAD ? p.n : N;
    const unsigned VS = (unsigned)p.V * SZ;
    const __amdgpu_buffer_rsrc_t ra =
        group_rsrc(p.a + pl.w0 * p.a_gs, group_range(pl.groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb =
        group_rsrc(p.b + pl.w0 * p.b_gs, group_range(pl.groups_left, p.b_gs, SZ));
    const __amdgpu_buffer_rsrc_t rx =
        group_rsrc(p.x + pl.w0 * p.x_gs, group_range(pl.groups_left, p.x_gs, SZ));
    const unsigned aslot = pl.slot(p.a_gs, p.V, SZ);
    const unsigned bslot = pl.slot(p.b_gs, p.V, SZ);
    const unsigned xslot = pl.slot(p.x_gs, p.V, SZ);
    const unsigned a_col = (unsigned)p.a_sj * VS, a_rowb = (unsigned)p.a_si * VS;
    const unsigned b_col = (unsigned)p.b_sj * VS, x_col = (unsigned)p.x_sj * VS;
    unsigned b_row[K], x_row[K];
    sfor<0, K>([&](auto k) {
        const bool in = !PAD || t + 4 * k < n;
        b_row[k] = (bslot == kOOB || !in) ? kOOB : bslot + (unsigned)((t + 4 * k) * p.b_si) * VS;
        x_row[k] = (xslot == kOOB || !in) ? kOOB : xslot + (unsigned)((t + 4 * k) * p.x_si) * VS;
    });

    const bool live = aslot != kOOB;
    int iter = 0;
    bool failed = false;
    float dinv[K];
    float *Lq = lds + (threadIdx.x >> 2) * QS;
    {
        /* the FP32 factor from an unrotated read of the referenced triangle
         * (entries right of the diagonal read as 0), parked in LDS with zeros
         * on and right of the diagonal */
        float a32[K][N];
        unsigned arow[K];
        sfor<0, K>([&](auto k) {
            arow[k] = (live && (!PAD || t + 4 * k < n)) ? aslot + (unsigned)(t + 4 * k) * a_rowb : kOOB;
        });
        sfor<0, K>([&](auto k) {
            sfor<0, 4 * k + 4>([&](auto c) {
                a32[k][c] = (float)bload(0.0, ra, (c < 4 * k || c <= t + 4 * k) ? arow[k] : kOOB,
                                         (unsigned)c * a_col);
                if constexpr (PAD && c >= 4 * k)
                    a32[k][c] = (c >= n && c == t + 4 * k) ? 1.0f : a32[k][c];
            });
        });
        if (factor32<N, LDL, quad_mfma<float, N>()>(a32, dinv, t)) {
            iter = -3;
            failed = true;
        }
        sfor<0, K>([&](auto k) {
            sfor<0, k + 1>([&](auto m) {
                f32x4 v;
                sfor<0, 4>([&](auto s) {
                    v[(int)s] = (m < k || s < t) ? a32[k][4 * m + s] : 0.0f;
                });
                *reinterpret_cast<f32x4 *>(Lq + slot_base(k) + t * (4 * k + 4) + 4 * m) = v;
            });
        });
    }
    __builtin_amdgcn_sched_barrier(0); /* the FP64 loads stay behind the factor */

    /* A again (from L2 or the last-level cache), in FP64 and rotated, for the
     * residuals */
    double a64[K][N];
    {
        /* per rotation j: column (t + j) mod 4's offset, the mirror's row
         * offset, and whether the diagonal block takes the entry itself
         * (precomputed, so that the loads' offsets are selects) */
        const unsigned tcol = (unsigned)t * a_col;
        unsigned rc[4], mr[4];
        bool own[4];
        sfor<0, 4>([&](auto j) {
            const unsigned s = (unsigned)((t + j) & 3);
            rc[j] = s * a_col;
            mr[j] = s * a_rowb + tcol;
            own[j] = (int)s <= t;
        });
        sfor<0, K>([&](auto k) {
            const unsigned arow = aslot + (unsigned)(t + 4 * k) * a_rowb;
            const unsigned drow = aslot + (unsigned)(4 * k) * a_rowb;
            sfor<0, k + 1>([&](auto m) {
                sfor<0, 4>([&](auto j) {
                    unsigned off;
                    if constexpr (m < k) {
                        off = arow + rc[j];
                    }
                    else {
                        /* diagonal block: (t + 4k, s + 4k) for s <= t, else its mirror */
                        const unsigned o1 = arow + rc[j], o2 = drow + mr[j];
                        off = own[j] ? o1 : o2;
                    }
                    off = live ? off : kOOB;
                    if constexpr (PAD) /* row t + 4k and column 4m + (t + j) mod 4 (its mirror likewise) */
                        off = (t + 4 * k < n && 4 * m + ((t + j) & 3) < n) ? off : kOOB;
                    a64[k][4 * m + j] = bload(0.0, ra, off, (unsigned)(4 * m) * a_col);
                });
            });
        });
    }

    /* ||A||_inf (dlansy 'I') and the FP32 overflow test of A (dlat2s), which
     * decides before the factor's outcome */
    double rs[K];
    sym_abs_rowsum<N>(a64, rs);
    double anrm = 0.0;
    bool over = false;
    sfor<0, K>([&](auto k) {
        anrm = fmax_nan(anrm, rs[k]);
        sfor<0, 4 * k + 4>([&](auto c) { over |= __builtin_fabs(a64[k][c]) > (double)FMAX; });
    });
    anrm = quad_max(anrm);
    const double cte = anrm * 0x1p-53 * __builtin_sqrt((double)n);
    if (quad_any(over)) {
        iter = -2;
        failed = true;
    }
    lds_sync();

    /* right-hand sides in blocks of NRB, each refined until its columns pass
     * (iter: the most steps a block took) */
    for (int q0 = 0; q0 < p.nrhs && !failed; q0 += NRB) {
        unsigned bcol[NRB], xcol[NRB];
        sfor<0, NRB>([&](auto q) {
            bcol[q] = (q0 + q < p.nrhs) ? (unsigned)(q0 + q) * b_col : kOOB;
            xcol[q] = (q0 + q < p.nrhs) ? (unsigned)(q0 + q) * x_col : kOOB;
        });
        double r[K][NRB];
        sfor<0, K>([&](auto k) {
            sfor<0, NRB>([&](auto q) { r[k][q] = bload(0.0, rb, b_row[k], bcol[q]); });
        });
        /* an overflowing B (dlag2s) sends the matrix to FP64 */
        bool b_over = false;
        sfor<0, K>([&](auto k) {
            sfor<0, NRB>([&](auto q) { b_over |= __builtin_fabs(r[k][q]) > (double)FMAX; });
        });
        if (quad_any(b_over)) {
            iter = -2;
            failed = true;
            break;
        }
        float w[K][NRB];
        sfor<0, K>([&](auto k) { sfor<0, NRB>([&](auto q) { w[k][q] = (float)r[k][q]; }); });
        solve32<N, NRB, LDL>(w, dinv, Lq, t);
        double x[K][NRB];
        sfor<0, K>([&](auto k) { sfor<0, NRB>([&](auto q) { x[k][q] = (double)w[k][q]; }); });
        int it_blk = 0;
        for (int it = 0;; ++it) {
            sym_residual<N, NRB>(a64, x, r);
            /* every column: ||r||_inf <= ||x||_inf * cte (a NaN fails) */
            bool fail = false, rover = false;
            sfor<0, NRB>([&](auto q) {
                double xm = 0.0;
                sfor<0, K>([&](auto k) { xm = fmax_nan(xm, __builtin_fabs(x[k][q])); });
                xm = quad_max(xm);
                sfor<0, K>([&](auto k) {
                    fail = fail || (q0 + q < p.nrhs && !(__builtin_fabs(r[k][q]) <= xm * cte));
                    rover = rover || __builtin_fabs(r[k][q]) > (double)FMAX;
                });
            });
            if (!quad_any(fail)) {
                it_blk = it;
                break;
            }
            if (it == p.itermax) {
                iter = -p.itermax - 1;
                failed = true;
                break;
            }
            if (quad_any(rover)) {
                iter = -2;
                failed = true;
                break;
            }
            sfor<0, K>([&](auto k) { sfor<0, NRB>([&](auto q) { w[k][q] = (float)r[k][q]; }); });
            /* B again for the next residual (from L2), while the sweeps run */
            sfor<0, K>([&](auto k) {
                sfor<0, NRB>([&](auto q) { r[k][q] = bload(0.0, rb, b_row[k], bcol[q]); });
            });
            solve32<N, NRB, LDL>(w, dinv, Lq, t);
            sfor<0, K>([&](auto k) {
                sfor<0, NRB>([&](auto q) { x[k][q] = x[k][q] + (double)w[k][q]; });
            });
        }
        if (failed) break;
        iter = it_blk > iter ? it_blk : iter;
        sfor<0, K>([&](auto k) {
            sfor<0, NRB>([&](auto q) { bstore(x[k][q], rx, x_row[k], xcol[q]); });
        });
    }

    /* a failed matrix: X := B, for the FP64 pass */
    if (failed) {
        for (int q0 = 0; q0 < p.nrhs; ++q0) {
            sfor<0, K>([&](auto k) {
                bstore(bload(0.0, rb, b_row[k], (unsigned)q0 * b_col), rx, x_row[k], (unsigned)q0 * x_col);
            });
        }
    }
    if (pl.live && pl.mat < p.nm && t == 0) {
        p.iter[pl.mat] = iter;
        p.info[pl.mat] = 0;
    }
}

template <int N, int NRB, bool LDL, bool PAD> void launch_np(const QuadMixedArgs &args)
{
    const long waves = (args.nmat + kMatsPerWave - 1) / kMatsPerWave;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    hipLaunchKernelGGL((quad_mixed_kernel<N, NRB, LDL, PAD>), grid, block, 0, current_stream(), args);
}

template <int N, int NRB> hipError_t launch_n(bool ldl, const QuadMixedArgs &args)
{
    const bool pad = args.n != N;
    if (ldl)
        pad ? launch_np<N, NRB, true, true>(args) : launch_np<N, NRB, true, false>(args);
    else
        pad ? launch_np<N, NRB, false, true>(args) : launch_np<N, NRB, false, false>(args);
    return hipGetLastError();
}

} /* namespace quadmixed */
} /* namespace cbk */

#endif /* CBK_QUAD_MIXED_IMPL_H */
