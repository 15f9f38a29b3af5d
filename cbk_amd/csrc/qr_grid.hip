/* qr_grid.hip -- FP64 gels for tall problems up to 64 x 32 with up to 8
 * right-hand sides (the reference's geqrf + ormqr + trsm path of gels,
 * src/cbk_gels_compact.hpp:124-160, src/cbk_geqrf_compact.hpp:74-131,
 * src/cbk_ormqr_compact.hpp:68-105), two problems per wavefront.
 *
 * Lane grid. Each 32-lane half of a wavefront holds one problem as a 4 x 8
 * grid of lanes: lane (r, c) = lane bits [1:0] and [4:2] holds rows i = r + 4a
 * (a < 16) of the columns j = c + 8b (b < 4) of A, f[a][b], and rows r + 4a of
 * B's column c (c < nrhs), g[a]: 80 doubles a lane. Per reflector kk (rows
 * and slots compile-time):
 *   tail   the pivot column's quad (column lane kk % 8) sums its squares below
 *          row kk, then a quad_perm DPP sum over the 4 row lanes;
 *   larfg  in every lane from the quad's values (x0 = A(kk, kk) by a quad
 *          broadcast): tau = (beta - x0) / beta, the body scaled by
 *          1 / (x0 - beta), R(kk, kk) = beta (batch_gels_impl.h's larfg);
 *   v      the scaled body reaches the other column lanes of the half by
 *          ds_swizzle (bit mode: lane (r, c) reads lane (r, kk % 8)), no LDS
 *          memory and no barrier;
 *   apply  every lane forms the partial w_j = sum of its rows' v_i A(i, j) for
 *          its columns j > kk and its B column, a quad DPP sum completes it,
 *          then A(i, j) -= v_i tau w_j on its own rows.
 * The back substitution R X = (Q^T B)(0:n) walks l = n-1 .. 0 the same way,
 * R's column l swizzled from column lane l % 8 to the B lanes.
 *
 * Sums run per lane in row order and then over the quad, not in the
 * reference's single row order, so results agree with it to rounding, which
 * the tests bound (as for the wavefront kernel). Per problem and reflector
 * a wavefront issues about 2 (rows/4) (columns/8) fma's for the update, where
 * the one-problem-per-wavefront kernel issues 3 x rows with 34 of 64 lanes
 * busy.
 */
#include "batch_gels_impl.h"
#include "cbk_internal.h"
#include "quad_common.h"

namespace cbk {
namespace qrgrid {

using team::fmaT;
using team::sfor;
using team::sfor_down;
using quad::qbc;
using quad::qx;

constexpr int RS = 16;  /* row slots: rows r + 4a */
constexpr int CSA = 4;  /* A column slots: columns c + 8b */
constexpr int KMAX = 32;

struct GridArgs {
    double *a, *b, *tau;
    int *info;
    int a_rs, a_cs, b_rs, b_cs; /* element strides, V folded in */
    long long a_gs, b_gs;
    int V, m, n, nrhs, ninfo;
    long long count; /* slots, padded ones included */
};

/* value of lane (r, C) of this lane's 32-lane half (ds_swizzle bit mode:
 * and 0x03, or 4C) */
template <int C> __device__ __forceinline__ double colbc(double x)
{
    constexpr int pat = 0x03 | ((4 * C) << 5);
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_ds_swizzle(w.x, pat);
    w.y = __builtin_amdgcn_ds_swizzle(w.y, pat);
    return __builtin_bit_cast(double, w);
}

/* x, as a value the compiler cannot relate to earlier ones */
template <class T> __device__ __forceinline__ T opaque(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

/* the sum over the 4 row lanes of a quad, the same value in each */
__device__ __forceinline__ double quad_sum(double x)
{
    x = x + qx<1>(x);
    return x + qx<2>(x);
}

/* OR over the 32 lanes of the half */
__device__ __forceinline__ unsigned half_or(unsigned x)
{
    x |= (unsigned)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (0x01 << 10));
    x |= (unsigned)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (0x02 << 10));
    x |= (unsigned)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (0x04 << 10));
    x |= (unsigned)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (0x08 << 10));
    x |= (unsigned)__builtin_amdgcn_ds_swizzle((int)x, 0x1F | (0x10 << 10));
    return x;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
qr_grid_kernel(GridArgs p)
{
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5, c = (lane >> 2) & 7, r = lane & 3;
    /* blocks in XCD order: the 32 problems of 4 consecutive logical blocks
     * (a 256-byte run of each element at V >= 32) are read through one L2 */
    const long long wave = (long long)team::xcd_block(blockIdx.x) * 4 + team::wave_in_block();
    if (wave * 2 >= p.count) return; /* whole wavefront */
    const long long kq = wave * 2 + h;
    const bool live = kq < p.count;
    const int m = p.m, n = p.n, nrhs = p.nrhs;
    /* buffer resources based at the wavefront's first group (wave-uniform);
     * a lane's byte offset is its slot, row lane and column lane, the slot
     * indices a, b go to the scalar offset, and a masked access uses kOOB
     * (zero load, dropped store) */
    const long long w0 = __builtin_amdgcn_readfirstlane((int)((wave * 2) / p.V));
    const long long groups = p.count / p.V;
    const long long g = live ? kq / p.V : w0, v = live ? kq % p.V : 0;
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + w0 * p.a_gs, team::group_range(groups - w0, p.a_gs, 8));
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + w0 * p.b_gs, team::group_range(groups - w0, p.b_gs, 8));
    const unsigned ars = 8u * p.a_rs, acs = 8u * p.a_cs, brs = 8u * p.b_rs, bcs = 8u * p.b_cs;
    const unsigned aoff = (unsigned)(((g - w0) * p.a_gs + v) * 8) + r * ars + c * acs;
    const unsigned boff = (unsigned)(((g - w0) * p.b_gs + v) * 8) + r * brs + c * bcs;
    /* the masked offsets of element slot (a, b) from lane offsets lo (A) and
     * lb (B) and row/column lanes rr, cc; the stores at the end pass their
     * operands through opaque() so that the compiler recomputes these few
     * instructions instead of keeping the loads' 80 offsets live throughout */
    auto a_at = [&](unsigned lo, int rr, int cc, int a, int b) {
        return (live && rr + 4 * a < m && cc + 8 * b < n) ? lo : team::kOOB;
    };
    auto b_at = [&](unsigned lo, int rr, int cc, int a) {
        return (live && rr + 4 * a < m && cc < nrhs) ? lo : team::kOOB;
    };

    double f[RS][CSA], gb[RS];
    sfor<0, RS>([&](auto a) {
        sfor<0, CSA>([&](auto b) {
            f[a][b] = team::bload(0.0, ra, a_at(aoff, r, c, a, b), a * 4 * ars + b * 8 * acs);
        });
        gb[a] = team::bload(0.0, rb, b_at(boff, r, c, a), a * 4 * brs);
    });

    double tauv[CSA] = {0.0, 0.0, 0.0, 0.0}; /* tau of reflector c + 8b */

    /* reflector kk from column kk (slot kk / 8 of column lane kk % 8),
     * which has received every earlier reflector: its squares below row kk
     * (two partial sums a lane, then the quad), larfg, the body scaled in
     * place, R(kk, kk) and tau; returns tau (valid in column lane kk % 8) */
    auto gen = [&](auto kc) {
        constexpr int kk = kc;
        constexpr int ak = kk / 4, rk = kk % 4, bk = kk / 8, ck = kk % 8;
        double s0 = (r > rk) ? f[ak][bk] * f[ak][bk] : 0.0, s1 = 0.0;
        sfor<ak + 1, RS>([&](auto a) {
            if constexpr ((a - ak) % 2 == 1)
                s1 = fmaT(f[a][bk], f[a][bk], s1);
            else
                s0 = fmaT(f[a][bk], f[a][bk], s0);
        });
        const double s = quad_sum(s0 + s1);
        const batch::Larfg<double> hh = batch::larfg(qbc<rk>(f[ak][bk]), s);
        const bool own = c == ck;
        const double sc = own ? hh.inv : 1.0; /* x 1.0 is exact */
        sfor<ak + 1, RS>([&](auto a) { f[a][bk] = f[a][bk] * sc; });
        f[ak][bk] = own ? ((r > rk) ? f[ak][bk] * hh.inv : ((r == rk) ? hh.rdiag : f[ak][bk]))
                        : f[ak][bk];
        tauv[bk] = own ? hh.tau : tauv[bk];
        return hh.tau;
    };
    /* v of reflector kk on this lane's rows (v_kk = 1, rows above kk 0) and
     * tau, from column lane kk % 8 */
    double vv[RS];
    auto bcast = [&](auto kc, double tl) {
        constexpr int kk = kc;
        constexpr int ak = kk / 4, rk = kk % 4, bk = kk / 8, ck = kk % 8;
        sfor<ak + 1, RS>([&](auto a) { vv[a] = colbc<ck>(f[a][bk]); });
        const double v0 = colbc<ck>(f[ak][bk]);
        vv[ak] = (r > rk) ? v0 : ((r == rk) ? 1.0 : 0.0);
        return colbc<ck>(tl);
    };
    /* reflector kk on A's column slot b (b >= kk / 8): w = tau v^T A(:, j)
     * on this lane's rows, then over the quad (columns j <= kk get w = 0, the
     * wavefront kernel's masking); then A(i, j) -= v_i w_j on rows i >= kk,
     * the rows above kk in slot kk / 4 keeping their values. B's column is
     * slot CSA. */
    auto dot = [&](auto kc, auto bc, double tau) {
        constexpr int kk = kc, b = bc;
        constexpr int ak = kk / 4, bk = kk / 8, ck = kk % 8;
        double d;
        if constexpr (b < CSA) {
            d = vv[ak] * f[ak][b];
            sfor<ak + 1, RS>([&](auto a) { d = fmaT(vv[a], f[a][b], d); });
        }
        else {
            d = vv[ak] * gb[ak];
            sfor<ak + 1, RS>([&](auto a) { d = fmaT(vv[a], gb[a], d); });
        }
        d = quad_sum(d);
        if constexpr (b == CSA) return d * tau;
        return (b > bk || c > ck) ? d * tau : 0.0;
    };
    auto upd = [&](auto kc, auto bc, double w) {
        constexpr int kk = kc, b = bc;
        constexpr int ak = kk / 4, rk = kk % 4;
        if constexpr (b < CSA) {
            f[ak][b] = (r >= rk) ? fmaT(-vv[ak], w, f[ak][b]) : f[ak][b];
            sfor<ak + 1, RS>([&](auto a) { f[a][b] = fmaT(-vv[a], w, f[a][b]); });
        }
        else {
            gb[ak] = (r >= rk) ? fmaT(-vv[ak], w, gb[ak]) : gb[ak];
            sfor<ak + 1, RS>([&](auto a) { gb[a] = fmaT(-vv[a], w, gb[a]); });
        }
    };
    /* Look-ahead: step kk first finishes the slot that holds column kk + 1,
     * then generates reflector kk + 1 from it in the same basic block as the
     * updates of the other slots and of B, so that the scheduler overlaps its
     * sums, square root and divisions with them; its v is broadcast once
     * reflector kk's v is dead. Past
     * column n - 1 the generated reflector comes from a zero column (tau = 0,
     * nothing scaled) and is never stored. */
    double tau = 0.0;
    if (n > 0) tau = bcast(std::integral_constant<int, 0>{}, gen(std::integral_constant<int, 0>{}));
    sfor<0, KMAX>([&](auto kc) {
        constexpr int kk = kc;
        constexpr int bk = kk / 8, b1 = (kk + 1) / 8;
        if (kk < n) {
            if constexpr (kk + 1 < KMAX) {
                constexpr auto B1 = std::integral_constant<int, b1>{};
                upd(kc, B1, dot(kc, B1, tau));
                const double tn = gen(std::integral_constant<int, kk + 1>{});
                sfor<bk, CSA + 1>([&](auto b) {
                    if constexpr (b != b1) upd(kc, b, dot(kc, b, tau));
                });
                tau = bcast(std::integral_constant<int, kk + 1>{}, tn);
            }
            else {
                sfor<bk, CSA + 1>([&](auto b) { upd(kc, b, dot(kc, b, tau)); });
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    });

    /* A: R and the reflector bodies; tau (reflector c + 8b of the lane's
     * column lane, at ((g * n + kk) * V + v)) */
    const unsigned aoff2 = opaque(aoff), boff2 = opaque(boff);
    const int r2 = opaque(r), c2 = opaque(c);
    sfor<0, RS>([&](auto a) {
        sfor<0, CSA>([&](auto b) {
            team::bstore(f[a][b], ra, a_at(aoff2, r2, c2, a, b), a * 4 * ars + b * 8 * acs);
        });
    });
    if (p.tau) {
        const unsigned tV = 8u * p.V;
        const __amdgpu_buffer_rsrc_t rt =
            team::group_rsrc(p.tau + w0 * (long long)n * p.V,
                             team::group_range(groups - w0, (size_t)n * p.V, 8));
        const unsigned toff = (unsigned)(((g - w0) * n * p.V + v) * 8) + c * tV;
        sfor<0, CSA>([&](auto b) {
            const unsigned o = (live && r == 0 && c + 8 * b < n) ? toff : team::kOOB;
            team::bstore(tauv[b], rt, o, b * 8 * tV);
        });
    }

    /* info: the first exact zero on R's diagonal (LAPACK dtrtrs's test);
     * R(l, l), l = c + 8b, sits in row lane l % 4 = c % 4, slot l / 4 */
    unsigned zero = 0;
    sfor<0, CSA>([&](auto b) {
        sfor<0, 2>([&](auto hc) {
            constexpr int a = 2 * b + hc;
            const int l = c + 8 * b;
            if ((c >> 2) == hc && r == (c & 3) && l < n && f[a][b] == 0.0) zero |= 1u << l;
        });
    });
    zero = half_or(zero);
    const int bad = zero ? __builtin_ctz(zero) + 1 : 0;
    if (live && lane % 32 == 0 && p.info && kq < p.ninfo) p.info[kq] = bad;
    /* LAPACK dgels stops at a singular R (B keeps Q^T B); without info the
     * substitution runs as the reference's gels_compact */
    const bool solve = !(p.info && bad);

    /* R X = (Q^T B)(0:n), column-oriented from row n-1 up: x_l = b_l /
     * R(l, l), then b_i -= R(i, l) x_l for rows i < l. solve is uniform in
     * the half, and every swizzle and DPP source lies in the same half. */
    if (solve) {
        sfor_down<0, KMAX>([&](auto lc) {
            constexpr int l = lc;
            constexpr int al = l / 4, rl = l % 4, bl = l / 8, cl = l % 8;
            if (l < n) {
                double rc[al + 1]; /* R(r + 4a, l) from column lane l % 8 */
                sfor<0, al + 1>([&](auto a) { rc[a] = colbc<cl>(f[a][bl]); });
                const double x = qbc<rl>(gb[al]) / qbc<rl>(rc[al]);
                sfor<0, al>([&](auto a) { gb[a] = fmaT(-rc[a], x, gb[a]); });
                gb[al] = (r < rl) ? fmaT(-rc[al], x, gb[al]) : ((r == rl) ? x : gb[al]);
            }
            __builtin_amdgcn_sched_barrier(0);
        });
    }
    sfor<0, RS>([&](auto a) { team::bstore(gb[a], rb, b_at(boff2, r2, c2, a), a * 4 * brs); });
}

} /* namespace qrgrid */

bool qr_grid_supports(int p, int q, int nrhs) { return p <= 64 && q <= 32 && q <= p && nrhs <= 8; }

hipError_t launch_qr_grid(int p, int q, int nrhs, bool transposed, Operand<double> A, double *tau,
                          Operand<double> B, int V, int nm, int *info)
{
    qrgrid::GridArgs args;
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (transposed) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.a = A.p;
    args.b = B.p;
    args.tau = tau;
    args.info = info;
    args.a_rs = si * V;
    args.a_cs = sj * V;
    args.b_rs = (B.row ? B.ld : 1) * V;
    args.b_cs = (B.row ? 1 : B.ld) * V;
    args.a_gs = (long long)A.gstride;
    args.b_gs = (long long)B.gstride;
    args.V = V;
    args.m = p;
    args.n = q;
    args.nrhs = nrhs;
    args.ninfo = nm;
    args.count = (long long)((nm + V - 1) / V) * V; /* padded slots too, as the reference */
    const long long waves = (args.count + 1) / 2;
    const unsigned blocks = team::xcd_grid((unsigned)((waves + 3) / 4));
    hipLaunchKernelGGL(qrgrid::qr_grid_kernel, dim3(blocks), dim3(256), 0, current_stream(), args);
    return hipGetLastError();
}

} /* namespace cbk */
