/* mixed_rows.hip -- dsposv_compact and dssysvnp_compact (LAPACK dsposv
 * semantics) at orders 33..64: one matrix per wavefront, lane i holding row i.
 *
 * Per matrix: A's full symmetric row i (the lower view's (i, j) for j <= i,
 * its (j, i) above) stays in lane i's registers in FP64 for the residual; the
 * FP32 factor is right-looking on an FP32 copy of the row -- per pivot k the
 * pivot comes by v_readlane, lane i scales its own entry (hardware reciprocal
 * square root, or reciprocal for the unpivoted LDL^T) and the column reaches
 * every lane through LDS -- and is then parked in the wavefront's LDS rows.
 * The sweeps run column-oriented: forward, lane i takes L(i, k) from its own
 * LDS row and y_k by v_readlane; backward, lane k takes L(i, k) from row i.
 * Refinement per block of Q right-hand sides: r = b - A x in FP64 (x through
 * LDS, read as broadcasts), the test ||r||_inf <= ||x||_inf ||A||_inf eps
 * sqrt(n) per column, d = the FP32 solve of r, x += d in FP64, up to itermax
 * steps. A matrix whose FP32 factor fails (iter -3), whose data or a residual
 * overflow FP32 (-2), or that does not converge (-(itermax + 1)) gets X := B;
 * the FP64 posv / sysvnp of chol_grid.hip, masked by iter, then factors A in
 * place (bit-identical to the reference's potrf / sytrfnp) and solves X.
 * Rows past n load zero, are never pivots and are never stored.
 */
#include "cbk_internal.h"
#include "quad_common.h"
#include "team_common.h"

namespace cbk {
namespace mixedrows {

using team::bload;
using team::bstore;
using team::fmax_nan;
using team::group_range;
using team::group_rsrc;
using team::kOOB;
using team::sfor;
using team::sfor_down;

constexpr int NB = 64;      /* rows: one a lane */
constexpr int Q = 4;        /* right-hand sides per refinement block */
constexpr int LS = NB + 1;  /* LDS row stride of the FP32 factor (floats) */
constexpr int kWaves = 4;   /* matrices (wavefronts) per block */

struct RowsArgs {
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
    long long count; /* slots, padded ones included */
    int nm, n, nrhs, itermax;
    int *iter;
    int *info;
};

template <int L> __device__ __forceinline__ float lanef(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), L));
}

/* v of lane (this lane xor mask) */
__device__ __forceinline__ double xor_lane(double v, int lane, int mask)
{
    const int addr = (lane ^ mask) << 2;
    int2 w = __builtin_bit_cast(int2, v);
    w.x = __builtin_amdgcn_ds_bpermute(addr, w.x);
    w.y = __builtin_amdgcn_ds_bpermute(addr, w.y);
    return __builtin_bit_cast(double, w);
}

/* the wavefront's maximum, NaN-propagating, in every lane */
__device__ __forceinline__ double wave_max(double v, int lane)
{
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) v = fmax_nan(v, xor_lane(v, lane, m));
    return v;
}

__device__ __forceinline__ bool wave_any(bool c) { return __builtin_amdgcn_ballot_w64(c) != 0ull; }

template <class T> __device__ __forceinline__ T opaque(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

/* w := A^{-1} w with the factor in the wavefront's LDS rows Lw (row i at
 * i * LS: L(i, k) for k < i, zeros from the diagonal on) and dinv the lane's
 * 1 / L(i, i) (1 / D(i)). With the zeros no lane needs a mask: a lane's value
 * changes only at the steps before its own, where it is then read. */
template <bool LDL>
__device__ __forceinline__ void solve32(float (&w)[Q], float dinv, const float *Lw, int lane, int n)
{
    /* forward: y_k (Cholesky: w_k / L(k, k)) from lane k, then the rows
     * below; in chunks of 8 steps (a step past n changes nothing: its L
     * entries and its lane's w are zero), so that a chunk's LDS reads are
     * issued together */
    sfor<0, NB / 8>([&](auto cc) {
        if (8 * cc < n) {
            sfor<8 * cc, 8 * cc + 8>([&](auto kc) {
                constexpr int k = kc;
                const float lik = Lw[lane * LS + k];
                const float dk = LDL ? 1.0f : lanef<k>(dinv);
                sfor<0, Q>([&](auto q) {
                    const float yk = LDL ? lanef<k>(w[q]) : lanef<k>(w[q]) * dk;
                    w[q] = __builtin_fmaf(-lik, yk, w[q]);
                });
            });
        }
    });
    /* y (Cholesky), or z = D^-1 y (LDL^T) */
    sfor<0, Q>([&](auto q) { w[q] = w[q] * dinv; });
    /* backward: x_i (Cholesky: z_i / L(i, i)) from lane i, then the rows
     * above with L(i, k) from row i */
    sfor_down<0, NB / 8>([&](auto cc) {
        if (8 * cc < n) {
            sfor_down<8 * cc, 8 * cc + 8>([&](auto ic) {
                constexpr int i = ic;
                const float lik = Lw[i * LS + lane];
                const float di = LDL ? 1.0f : lanef<i>(dinv);
                sfor<0, Q>([&](auto q) {
                    const float xi = LDL ? lanef<i>(w[q]) : lanef<i>(w[q]) * di;
                    w[q] = __builtin_fmaf(-lik, xi, w[q]);
                });
            });
        }
    });
    if constexpr (!LDL) {
        sfor<0, Q>([&](auto q) { w[q] = w[q] * dinv; });
    }
}

template <bool LDL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
mixed_rows_kernel(RowsArgs p)
{
    constexpr unsigned SZ = 8;
    constexpr double FMAX = 3.4028234663852886e38;
    __shared__ float Ls[kWaves][NB * LS]; /* the FP32 factor's rows */
    __shared__ float col[kWaves][NB];     /* a factor column being broadcast */
    __shared__ float dvs[kWaves][NB];     /* 1 / L(k, k) (1 / D(k)) */
    __shared__ double xs[kWaves][NB * Q]; /* x being broadcast to the residual */

    const int lane = threadIdx.x & 63;
    const int wv = team::wave_in_block();
    /* the matrices of 4 consecutive logical blocks (16 consecutive slots) on
     * one XCD */
    const long long kq = (long long)team::xcd_block(blockIdx.x) * kWaves + wv;
    if (kq >= p.count) return; /* whole wavefront; no block-wide barriers below */
    const int n = p.n;
    const long long g = kq / p.V, v = kq % p.V;
    const long long groups = p.count / p.V;
    const __amdgpu_buffer_rsrc_t ra = group_rsrc(p.a + g * p.a_gs, group_range(groups - g, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb = group_rsrc(p.b + g * p.b_gs, group_range(groups - g, p.b_gs, SZ));
    const __amdgpu_buffer_rsrc_t rx = group_rsrc(p.x + g * p.x_gs, group_range(groups - g, p.x_gs, SZ));
    const unsigned VS = (unsigned)p.V * SZ;
    const unsigned ars = (unsigned)p.a_si * VS, acs = (unsigned)p.a_sj * VS;
    const unsigned bcs = (unsigned)p.b_sj * VS, xcs = (unsigned)p.x_sj * VS;
    const bool row_in = lane < n;
    const unsigned slot = (unsigned)(v * SZ);
    const unsigned b_row = row_in ? slot + (unsigned)lane * p.b_si * VS : kOOB;
    const unsigned x_row = row_in ? slot + (unsigned)lane * p.x_si * VS : kOOB;
    float *Lw = Ls[wv];
    float *cw = col[wv];
    float *dv = dvs[wv];
    double *xw = xs[wv];

    /* row i = lane of A: (i, j) for j <= i, (j, i) above; read once for the
     * norm, the overflow test and the FP32 copy, and once more in FP64 for the
     * residuals after the factor (then mostly from L2) */
    /* entry (i, j) of the symmetric matrix at (max(i, j), min(i, j)) of the
     * lower view: arithmetic, so that no per-entry lane mask stays live; ln is
     * the lane index, passed anew to the second read so that its offsets are
     * recomputed there rather than kept live across the factor */
    auto a_off = [&](int j, int ln) {
        const unsigned hi = (unsigned)(ln > j ? ln : j), lo = (unsigned)(ln < j ? ln : j);
        return (ln < n && j < n) ? slot + hi * ars + lo * acs : kOOB;
    };
    float a32[NB];
    double rs = 0.0;
    bool over = false;
    sfor<0, NB / 16>([&](auto h) {
        /* ||A||_inf (dlansy 'I') and the FP32 overflow test of A (dlat2s),
         * 16 entries at a time */
        sfor<16 * h, 16 * h + 16>([&](auto j) {
            const double aij = bload(0.0, ra, a_off(j, lane), 0);
            rs += __builtin_fabs(aij);
            over |= __builtin_fabs(aij) > FMAX;
            a32[j] = (float)aij;
        });
        __builtin_amdgcn_sched_barrier(0);
    });
    const double anrm = wave_max(rs, lane);
    const double cte = anrm * 0x1p-53 * __builtin_sqrt((double)n);

    int iter = 0;
    bool failed = false;
    float dinv = 1.0f;
    if (wave_any(over)) {
        iter = -2;
        failed = true;
    }
    else {
        /* the FP32 factor, right-looking; entries right of the diagonal are
         * scratch */
        bool bad = false;
        sfor<0, NB>([&](auto kc) {
            constexpr int k = kc;
            if (k < n) {
                const float piv = lanef<k>(a32[k]);
                float sc;
                if constexpr (LDL) {
                    bad |= (piv == 0.0f) || !__builtin_isfinite(piv);
                    sc = __builtin_amdgcn_rcpf(piv);
                }
                else {
                    bad |= !(piv > 0.0f);
                    sc = __builtin_amdgcn_rsqf(piv);
                }
                /* every lane scales its entry k (rows above k: scratch) */
                const float aik = a32[k];
                const float lik = aik * sc;
                a32[k] = lik;
                /* the update's other factor: L(j, k) (LDL^T: L(j, k) D(k), the
                 * unscaled entry); and 1 / L(k, k) (1 / D(k)) for lane k */
                cw[lane] = LDL ? aik : lik;
                dv[k] = sc;
                quad::lds_sync();
                sfor<k + 1, NB>([&](auto j) { a32[j] = __builtin_fmaf(-lik, cw[j], a32[j]); });
                quad::lds_sync();
            }
        });
        if (bad) {
            iter = -3;
            failed = true;
        }
        else {
            sfor<0, NB>([&](auto k) { Lw[lane * LS + k] = (k < lane) ? a32[k] : 0.0f; });
            quad::lds_sync();
            dinv = row_in ? dv[lane] : 1.0f; /* rows past n: never pivots */
        }
    }
    __builtin_amdgcn_sched_barrier(0); /* the FP64 loads stay behind the factor */
    double a64[NB];
    {
        const int ln = opaque(lane);
        sfor<0, NB>([&](auto j) { a64[j] = failed ? 0.0 : bload(0.0, ra, a_off(j, ln), 0); });
    }

    /* right-hand sides in blocks of Q, each refined until its columns pass;
     * every block's B is tested for FP32 overflow even after a failure, since
     * that test comes first in dsposv (-2 over any other code) */
    for (int q0 = 0; q0 < p.nrhs; q0 += Q) {
        unsigned bcol[Q], xcol[Q];
        sfor<0, Q>([&](auto q) {
            bcol[q] = (q0 + q < p.nrhs) ? b_row + (unsigned)(q0 + q) * bcs : kOOB;
            xcol[q] = (q0 + q < p.nrhs) ? x_row + (unsigned)(q0 + q) * xcs : kOOB;
            if (b_row == kOOB) bcol[q] = kOOB;
            if (x_row == kOOB) xcol[q] = kOOB;
        });
        double b64[Q];
        sfor<0, Q>([&](auto q) { b64[q] = bload(0.0, rb, bcol[q], 0); });
        bool b_over = false;
        sfor<0, Q>([&](auto q) { b_over |= __builtin_fabs(b64[q]) > FMAX; });
        if (wave_any(b_over)) {
            iter = -2;
            failed = true;
        }
        if (failed) continue;
        float w[Q];
        sfor<0, Q>([&](auto q) { w[q] = (float)b64[q]; });
        solve32<LDL>(w, dinv, Lw, lane, n);
        double x64[Q];
        sfor<0, Q>([&](auto q) { x64[q] = (double)w[q]; });
        int it_blk = 0;
        for (int it = 0;; ++it) {
            /* r = b - A x, FP64, columns ascending */
            sfor<0, Q>([&](auto q) { xw[lane * Q + q] = x64[q]; });
            quad::lds_sync();
            double r[Q];
            sfor<0, Q>([&](auto q) { r[q] = b64[q]; });
            sfor<0, NB / 8>([&](auto cc) {
                if (8 * cc < n) { /* entries past n are zero */
                    sfor<8 * cc, 8 * cc + 8>([&](auto j) {
                        sfor<0, Q>([&](auto q) { r[q] = __builtin_fma(-a64[j], xw[j * Q + q], r[q]); });
                    });
                }
            });
            quad::lds_sync();
            /* every column: ||r||_inf <= ||x||_inf * cte (a NaN fails) */
            bool fail = false, rover = false;
            sfor<0, Q>([&](auto q) {
                const double xm = wave_max(__builtin_fabs(x64[q]), lane);
                fail = fail || (q0 + q < p.nrhs && row_in && !(__builtin_fabs(r[q]) <= xm * cte));
                rover = rover || __builtin_fabs(r[q]) > FMAX;
            });
            if (!wave_any(fail)) {
                it_blk = it;
                break;
            }
            if (it == p.itermax) {
                iter = -p.itermax - 1;
                failed = true;
                break;
            }
            if (wave_any(rover)) {
                iter = -2;
                failed = true;
                break;
            }
            sfor<0, Q>([&](auto q) { w[q] = (float)r[q]; });
            solve32<LDL>(w, dinv, Lw, lane, n);
            sfor<0, Q>([&](auto q) { x64[q] = x64[q] + (double)w[q]; });
        }
        if (failed) continue;
        iter = it_blk > iter ? it_blk : iter;
        sfor<0, Q>([&](auto q) { bstore(x64[q], rx, xcol[q], 0); });
    }

    /* a failed matrix: X := B, for the FP64 pass */
    if (failed) {
        for (int q = 0; q < p.nrhs; ++q) {
            const unsigned bo = (b_row == kOOB) ? kOOB : b_row + (unsigned)q * bcs;
            const unsigned xo = (x_row == kOOB) ? kOOB : x_row + (unsigned)q * xcs;
            bstore(bload(0.0, rb, bo, 0), rx, xo, 0);
        }
    }
    if (kq < p.nm && lane == 0) {
        p.iter[kq] = iter;
        p.info[kq] = 0;
    }
}

} /* namespace mixedrows */

bool mixed_rows_supports(int n) { return n > 32 && n <= mixedrows::NB; }

hipError_t launch_mixed_posv_rows(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, int V, int nm,
                                  int itermax, int *iter, int *info)
{
    if (!mixed_rows_supports(n)) return hipErrorInvalidValue;
    mixedrows::RowsArgs p;
    p.a = A.p;
    /* the lower view: uplo 'U' reads A transposed (src/cbk_common.hpp:201-206) */
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (up) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    p.a_si = si;
    p.a_sj = sj;
    p.a_gs = A.gstride;
    p.b = B.p;
    p.b_si = B.row ? B.ld : 1;
    p.b_sj = B.row ? 1 : B.ld;
    p.b_gs = B.gstride;
    p.x = X.p;
    p.x_si = X.row ? X.ld : 1;
    p.x_sj = X.row ? 1 : X.ld;
    p.x_gs = X.gstride;
    p.V = V;
    p.count = (long long)((nm + V - 1) / V) * V;
    p.nm = nm;
    p.n = n;
    p.nrhs = nrhs;
    p.itermax = itermax;
    p.iter = iter;
    p.info = info;
    const unsigned blocks = team::xcd_grid((unsigned)((p.count + mixedrows::kWaves - 1) / mixedrows::kWaves));
    if (ldl)
        hipLaunchKernelGGL((mixedrows::mixed_rows_kernel<true>), dim3(blocks), dim3(256), 0, current_stream(), p);
    else
        hipLaunchKernelGGL((mixedrows::mixed_rows_kernel<false>), dim3(blocks), dim3(256), 0, current_stream(), p);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    /* the FP64 posv / sysvnp on the matrices marked iter < 0, in place on A and
     * on X (which holds B for them): A receives the factor, info its code */
    return launch_chol_grid<double>(ldl, true, up, n, nrhs, A, X, V, nm, info, iter);
}

} /* namespace cbk */
