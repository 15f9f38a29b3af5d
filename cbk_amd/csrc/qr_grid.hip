/* qr_grid.hip -- gels for tall problems up to 64 x 32 with up to 8
 * right-hand sides (the reference's geqrf + ormqr + trsm path of gels,
 * src/cbk_gels_compact.hpp:124-160, src/cbk_geqrf_compact.hpp:74-131,
 * src/cbk_ormqr_compact.hpp:68-105), two problems per wavefront: compact
 * FP64 operands (dgels_compact), and the uneven batches' problems of that size
 * in both precisions (persistent work-queue grids of gels_batch).
 *
 * Lane grid. Each 32-lane half of a wavefront holds one problem as a 4 x 8
 * grid of lanes: lane (r, c) = lane bits [1:0] and [4:2] holds rows i = r + 4a
 * (a < RS: 16 slots for up to 64 rows, 8 for up to 32) of the columns
 * j = c + 8b (b < 4) of A, f[a][b], and rows r + 4a of B's column c
 * (c < nrhs), gb[a]. Per reflector kk (rows and slots compile-time):
 *   gen    the pivot column's quad (column lane kk % 8) sums its squares below
 *          row kk, then a quad_perm DPP sum over the 4 row lanes; larfg in
 *          every lane from the quad's values (x0 = A(kk, kk) by a quad
 *          broadcast): tau = (beta - x0) / beta, the body scaled by
 *          1 / (x0 - beta), R(kk, kk) = beta (batch_gels_impl.h's larfg);
 *   bcast  the scaled body reaches the other column lanes of the half by
 *          ds_swizzle (bit mode: lane (r, c) reads lane (r, kk % 8)), no LDS
 *          memory and no barrier;
 *   apply  every lane forms the partial w_j = sum of its rows' v_i A(i, j) for
 *          its columns j > kk and its B column, a quad DPP sum completes it,
 *          then A(i, j) -= v_i tau w_j on its own rows.
 * Step kk finishes the slot holding column kk + 1 first and generates
 * reflector kk + 1 from it while the other slots are updated (look-ahead).
 * The back substitution R X = (Q^T B)(0:n) walks l = n-1 .. 0 the same way,
 * R's column l swizzled from column lane l % 8 to the B lanes.
 *
 * Sums run per lane in row order and then over the quad, not in the
 * reference's single row order, so results agree with it to rounding, which
 * the tests bound (as for the wavefront kernels). Per problem and reflector
 * a wavefront issues about 2 (rows/4) (columns/8) fma's for the update, where
 * the one-problem-per-wavefront kernel issues 3 x rows with 34 of 64 lanes
 * busy. Every swizzle and DPP source lies in the lane's own half, and every
 * branch that depends on a problem's sizes is uniform in the half, so the two
 * halves may hold problems of different sizes.
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

constexpr int CSA = 4;  /* A column slots: columns c + 8b */
constexpr int KMAX = 32;

template <class T> struct GridArgs {
    T *a, *b, *tau;
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
template <int C> __device__ __forceinline__ float colbc(float x)
{
    constexpr int pat = 0x03 | ((4 * C) << 5);
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, x), pat));
}

/* x, as a value the compiler cannot relate to earlier ones */
template <class T> __device__ __forceinline__ T opaque(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

/* the sum over the 4 row lanes of a quad, the same value in each */
template <class T> __device__ __forceinline__ T quad_sum(T x)
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

/* QR of the half's n columns (n <= 4 RS, rows past the problem zero), Q^T B
 * in gb as it goes; tauv[b] = tau of reflector c + 8b */
template <class T, int RS>
__device__ __forceinline__ void factor(T (&f)[RS][CSA], T (&gb)[RS], T (&tauv)[CSA], int r, int c,
                                       int n)
{
    /* reflector kk from column kk (slot kk / 8 of column lane kk % 8), which
     * has received every earlier reflector: its squares below row kk (two
     * partial sums a lane, then the quad), larfg, the body scaled in place,
     * R(kk, kk) and tau; returns tau (valid in column lane kk % 8) */
    auto gen = [&](auto kc) {
        constexpr int kk = kc;
        constexpr int ak = kk / 4, rk = kk % 4, bk = kk / 8, ck = kk % 8;
        T s0 = (r > rk) ? f[ak][bk] * f[ak][bk] : T(0), s1 = T(0);
        sfor<ak + 1, RS>([&](auto a) {
            if constexpr ((a - ak) % 2 == 1)
                s1 = fmaT(f[a][bk], f[a][bk], s1);
            else
                s0 = fmaT(f[a][bk], f[a][bk], s0);
        });
        const T s = quad_sum(s0 + s1);
        const batch::Larfg<T> hh = batch::larfg(qbc<rk>(f[ak][bk]), s);
        const bool own = c == ck;
        const T sc = own ? hh.inv : T(1); /* x 1 is exact */
        sfor<ak + 1, RS>([&](auto a) { f[a][bk] = f[a][bk] * sc; });
        f[ak][bk] = own ? ((r > rk) ? f[ak][bk] * hh.inv : ((r == rk) ? hh.rdiag : f[ak][bk]))
                        : f[ak][bk];
        tauv[bk] = own ? hh.tau : tauv[bk];
        return hh.tau;
    };
    /* v of reflector kk on this lane's rows (v_kk = 1, rows above kk 0) and
     * tau, from column lane kk % 8 */
    T vv[RS];
    auto bcast = [&](auto kc, T tl) {
        constexpr int kk = kc;
        constexpr int ak = kk / 4, rk = kk % 4, bk = kk / 8, ck = kk % 8;
        sfor<ak + 1, RS>([&](auto a) { vv[a] = colbc<ck>(f[a][bk]); });
        const T v0 = colbc<ck>(f[ak][bk]);
        vv[ak] = (r > rk) ? v0 : ((r == rk) ? T(1) : T(0));
        return colbc<ck>(tl);
    };
    /* reflector kk on column slot b (b >= kk / 8; slot CSA is B's column):
     * w = tau v^T A(:, j) on this lane's rows, then over the quad (columns
     * j <= kk get w = 0, the wavefront kernel's masking); then A(i, j) -=
     * v_i w_j on rows i >= kk, the rows above kk in slot kk / 4 keeping their
     * values */
    auto dot = [&](auto kc, auto bc, T tau) {
        constexpr int kk = kc, b = bc;
        constexpr int ak = kk / 4, bk = kk / 8, ck = kk % 8;
        T d;
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
        return (b > bk || c > ck) ? d * tau : T(0);
    };
    auto upd = [&](auto kc, auto bc, T w) {
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
     * reflector kk's v is dead. Past column n - 1 the generated reflector
     * comes from a zero column (tau = 0, nothing scaled) and is never
     * stored. */
    constexpr int K = KMAX < 4 * RS ? KMAX : 4 * RS;
    T tau = T(0);
    if (n > 0) tau = bcast(std::integral_constant<int, 0>{}, gen(std::integral_constant<int, 0>{}));
    sfor<0, K>([&](auto kc) {
        constexpr int kk = kc;
        constexpr int bk = kk / 8, b1 = (kk + 1) / 8;
        if (kk < n) {
            if constexpr (kk + 1 < K) {
                constexpr auto B1 = std::integral_constant<int, b1>{};
                upd(kc, B1, dot(kc, B1, tau));
                const T tn = gen(std::integral_constant<int, kk + 1>{});
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
}

/* info of the half's problem: the first exact zero on R's diagonal (LAPACK
 * dtrtrs's test); R(l, l), l = c + 8b, sits in row lane l % 4 = c % 4, slot
 * l / 4 */
template <class T, int RS>
__device__ __forceinline__ int first_zero_pivot(const T (&f)[RS][CSA], int r, int c, int n)
{
    unsigned zero = 0;
    sfor<0, CSA>([&](auto b) {
        sfor<0, 2>([&](auto hc) {
            constexpr int a = 2 * b + hc;
            if constexpr (a < RS) {
                const int l = c + 8 * b;
                if ((c >> 2) == hc && r == (c & 3) && l < n && f[a][b] == T(0)) zero |= 1u << l;
            }
        });
    });
    zero = half_or(zero);
    return zero ? __builtin_ctz(zero) + 1 : 0;
}

/* R X = (Q^T B)(0:n), column-oriented from row n-1 up: x_l = b_l / R(l, l),
 * then b_i -= R(i, l) x_l for rows i < l (the reference's trsm order) */
template <class T, int RS>
__device__ __forceinline__ void back_substitute(const T (&f)[RS][CSA], T (&gb)[RS], int r, int n)
{
    constexpr int K = KMAX < 4 * RS ? KMAX : 4 * RS;
    sfor_down<0, K>([&](auto lc) {
        constexpr int l = lc;
        constexpr int al = l / 4, rl = l % 4, bl = l / 8, cl = l % 8;
        if (l < n) {
            T rc[al + 1]; /* R(r + 4a, l) from column lane l % 8 */
            sfor<0, al + 1>([&](auto a) { rc[a] = colbc<cl>(f[a][bl]); });
            const T x = qbc<rl>(gb[al]) / qbc<rl>(rc[al]);
            sfor<0, al>([&](auto a) { gb[a] = fmaT(-rc[a], x, gb[a]); });
            gb[al] = (r < rl) ? fmaT(-rc[al], x, gb[al]) : ((r == rl) ? x : gb[al]);
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

/* compact operands, two consecutive slots per wavefront */
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, sizeof(T) == 8 ? 2 : 4))) void
qr_grid_kernel(GridArgs<T> p)
{
    constexpr int RS = 16;
    constexpr unsigned SZ = sizeof(T);
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
        team::group_rsrc(p.a + w0 * p.a_gs, team::group_range(groups - w0, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + w0 * p.b_gs, team::group_range(groups - w0, p.b_gs, SZ));
    const unsigned ars = SZ * p.a_rs, acs = SZ * p.a_cs, brs = SZ * p.b_rs, bcs = SZ * p.b_cs;
    const unsigned aoff = (unsigned)(((g - w0) * p.a_gs + v) * SZ) + r * ars + c * acs;
    const unsigned boff = (unsigned)(((g - w0) * p.b_gs + v) * SZ) + r * brs + c * bcs;
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

    T f[RS][CSA], gb[RS];
    sfor<0, RS>([&](auto a) {
        sfor<0, CSA>([&](auto b) {
            f[a][b] = team::bload(T(0), ra, a_at(aoff, r, c, a, b), a * 4 * ars + b * 8 * acs);
        });
        gb[a] = team::bload(T(0), rb, b_at(boff, r, c, a), a * 4 * brs);
    });
    T tauv[CSA] = {T(0), T(0), T(0), T(0)};
    factor<T, RS>(f, gb, tauv, r, c, n);

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
        const unsigned tV = SZ * p.V;
        const __amdgpu_buffer_rsrc_t rt =
            team::group_rsrc(p.tau + w0 * (long long)n * p.V,
                             team::group_range(groups - w0, (size_t)n * p.V, SZ));
        const unsigned toff = (unsigned)(((g - w0) * n * p.V + v) * SZ) + c * tV;
        sfor<0, CSA>([&](auto b) {
            const unsigned o = (live && r == 0 && c + 8 * b < n) ? toff : team::kOOB;
            team::bstore(tauv[b], rt, o, b * 8 * tV);
        });
    }

    const int bad = first_zero_pivot<T, RS>(f, r, c, n);
    if (live && lane % 32 == 0 && p.info && kq < p.ninfo) p.info[kq] = bad;
    /* LAPACK dgels stops at a singular R (B keeps Q^T B); without info the
     * substitution runs as the reference's gels_compact. bad is uniform in
     * the half. */
    if (nrhs > 0 && !(p.info && bad)) back_substitute<T, RS>(f, gb, r, n);
    sfor<0, RS>([&](auto a) { team::bstore(gb[a], rb, b_at(boff2, r2, c2, a), a * 4 * brs); });
}

/* the uneven batch (gels_batch): a persistent grid, each wavefront taking the
 * next two problems of the class queue (longest first) until it is empty;
 * each problem column-major on its own (row stride 1), lda / ldb its own */
template <class T, int RS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
qr_grid_batch_kernel(BatchArgs<T> p)
{
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5, c = (lane >> 2) & 7, r = lane & 3;
    const unsigned count = batch_count(p);
    for (;;) {
        unsigned k0 = 0;
        if (lane == 0) k0 = atomicAdd(p.counter, 2u);
        k0 = __builtin_amdgcn_readfirstlane(k0);
        if (k0 >= count) return;
        const unsigned kq = k0 + h;
        const bool live = kq < count;
        BatchProblem d{0, 0, 0, 1, 1, 0, 0, 0, 0};
        if (live) d = p.desc[kq];
        const int m = d.m, n = d.n, nrhs = d.nrhs;
        /* lane (r, c): column c + 8b of A from row r, column c of B */
        T *acol = p.a + d.a_off + (long long)c * d.lda + r;
        T *bcol = p.b + d.b_off + (long long)c * d.ldb + r;
        const long long astep = 8ll * d.lda;

        T f[RS][CSA], gb[RS];
        sfor<0, RS>([&](auto a) {
            sfor<0, CSA>([&](auto b) {
                f[a][b] = (r + 4 * a < m && c + 8 * b < n) ? acol[b * astep + 4 * a] : T(0);
            });
            gb[a] = (r + 4 * a < m && c < nrhs) ? bcol[4 * a] : T(0);
        });
        T tauv[CSA] = {T(0), T(0), T(0), T(0)};
        factor<T, RS>(f, gb, tauv, r, c, n);

        T *acol2 = opaque(acol), *bcol2 = opaque(bcol);
        const int r2 = opaque(r), c2 = opaque(c);
        sfor<0, RS>([&](auto a) {
            sfor<0, CSA>([&](auto b) {
                if (r2 + 4 * a < m && c2 + 8 * b < n) acol2[b * astep + 4 * a] = f[a][b];
            });
        });
        const int bad = first_zero_pivot<T, RS>(f, r, c, n);
        if (live && lane % 32 == 0 && p.info && d.idx < p.ninfo) p.info[d.idx] = bad;
        if (!(p.info && bad)) back_substitute<T, RS>(f, gb, r, n);
        sfor<0, RS>([&](auto a) {
            if (r2 + 4 * a < m && c2 < nrhs) bcol2[4 * a] = gb[a];
        });
    }
}

} /* namespace qrgrid */

/* gels with nrhs <= 8, or geqrf (nrhs = 0), of a tall p x q view: two
 * problems per wavefront for q <= 32, the 8 x 8 lane grid (qr_grid64.hip)
 * above */
bool qr_grid_supports(int p, int q, int nrhs)
{
    return p <= 64 && q >= 1 && q <= 64 && q <= p && nrhs >= 0 && nrhs <= 8;
}

template <class T>
hipError_t launch_qr_grid(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info)
{
    if (q > 32) return launch_qr_grid64<T>(p, q, nrhs, transposed, A, tau, B, V, nm, info);
    qrgrid::GridArgs<T> args;
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
    hipLaunchKernelGGL((qrgrid::qr_grid_kernel<T>), dim3(blocks), dim3(256), 0, current_stream(), args);
    return hipGetLastError();
}

template hipError_t launch_qr_grid<double>(int, int, int, bool, Operand<double>, double *,
                                           Operand<double>, int, int, int *);
template hipError_t launch_qr_grid<float>(int, int, int, bool, Operand<float>, float *,
                                          Operand<float>, int, int, int *);

/* batch kernels 9 (rows <= 32) and 10 (rows <= 64) */
template <class T>
hipError_t launch_qr_grid_batch(int kernel, const BatchArgs<T> &args, unsigned blocks, hipStream_t s)
{
    if (kernel == 9)
        hipLaunchKernelGGL((qrgrid::qr_grid_batch_kernel<T, 8>), dim3(blocks), dim3(256), 0, s, args);
    else
        hipLaunchKernelGGL((qrgrid::qr_grid_batch_kernel<T, 16>), dim3(blocks), dim3(256), 0, s, args);
    return hipGetLastError();
}

template <class T> hipError_t qr_grid_batch_occupancy(int kernel, int *blocks)
{
    return kernel == 9
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, qrgrid::qr_grid_batch_kernel<T, 8>, 256, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, qrgrid::qr_grid_batch_kernel<T, 16>, 256, 0);
}

template hipError_t launch_qr_grid_batch<double>(int, const BatchArgs<double> &, unsigned, hipStream_t);
template hipError_t launch_qr_grid_batch<float>(int, const BatchArgs<float> &, unsigned, hipStream_t);
template hipError_t qr_grid_batch_occupancy<double>(int, int *);
template hipError_t qr_grid_batch_occupancy<float>(int, int *);

} /* namespace cbk */
