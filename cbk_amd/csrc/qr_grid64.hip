/* qr_grid64.hip -- gels for the uneven batches' wide problems: m <= 64,
 * 32 < n <= 64 or n + nrhs > 64, nrhs <= 8 (the reference's geqrf + ormqr +
 * trsm path of gels, src/cbk_gels_compact.hpp:124-160,
 * src/cbk_geqrf_compact.hpp:74-131, src/cbk_ormqr_compact.hpp:68-105), one
 * problem per wavefront on an 8 x 8 grid of lanes.
 *
 * Lane (r, c) = lane bits [2:0] and [5:3] holds rows i = r + 8a (a < 8) of the
 * columns j = c + 8b (b < 8) of A, f[a][b], and rows r + 8a of B's column c
 * (c < nrhs), gb[a]: 72 values a lane. Per reflector kk (rows and slots
 * compile-time; the pivot row and column both sit in slot kk / 8 of lane
 * kk % 8):
 *   gen    column lane kk % 8 sums the squares below row kk (row order per
 *          lane, then over its 8 row lanes: quad_perm and row_half_mirror DPP
 *          adds); x0 = A(kk, kk) and the sum reach every lane by v_readlane,
 *          so larfg runs on wave-uniform values (batch_gels_impl.h's larfg);
 *   bcast  the scaled body goes from column lane kk % 8 to the others by
 *          ds_bpermute (the column lanes span both halves of the wavefront);
 *   apply  each lane's partial w_j over its rows, the 8-row-lane DPP sum,
 *          then A(i, j) -= v_i tau w_j on its own rows; B's column likewise.
 * Step kk first finishes the slot that holds column kk + 1 and generates
 * reflector kk + 1 from it while the other slots are updated (look-ahead).
 * R X = (Q^T B)(0:n) then runs column-oriented from row n-1 up.
 *
 * Sums run per lane in row order and then over the row lanes, not in the
 * reference's single row order, so results agree with it to rounding (the
 * tests bound it, as for the wavefront kernels). Where the one-problem
 * wavefront kernels run a 64-row fma chain per column lane and reflector, a
 * lane here runs at most 8 rows, and the wavefront's 64 lanes all hold data.
 */
#include "batch_gels_impl.h"
#include "cbk_internal.h"
#include "grid64_common.h"

namespace cbk {
namespace qrgrid64 {

/* QR of the problem's n columns (rows past the problem zero), Q^T B in gb as
 * it goes; tauv[b] = tau of reflector c + 8b */
template <class T>
__device__ __forceinline__ void factor(T (&f)[RS][CS], T (&gb)[RS], T (&tauv)[CS], int r, int c, int n,
                                       int raddr)
{
    auto gen = [&](auto kc) {
        constexpr int kk = kc;
        constexpr int ak = kk / 8, rk = kk % 8; /* row slot / lane; column slot / lane likewise */
        T s0 = (r > rk) ? f[ak][ak] * f[ak][ak] : T(0), s1 = T(0);
        sfor<ak + 1, RS>([&](auto a) {
            if constexpr ((a - ak) % 2 == 1)
                s1 = fmaT(f[a][ak], f[a][ak], s1);
            else
                s0 = fmaT(f[a][ak], f[a][ak], s0);
        });
        const T s = row_sum(s0 + s1);
        /* lane (rk, rk) holds A(kk, kk); lane (0, rk) the column's sum */
        const batch::Larfg<T> hh =
            batch::larfg(lane_value<rk + 8 * rk>(f[ak][ak]), lane_value<8 * rk>(s));
        const bool own = c == rk;
        const T sc = own ? hh.inv : T(1); /* x 1 is exact */
        sfor<ak + 1, RS>([&](auto a) { f[a][ak] = f[a][ak] * sc; });
        f[ak][ak] = own ? ((r > rk) ? f[ak][ak] * hh.inv : ((r == rk) ? hh.rdiag : f[ak][ak]))
                        : f[ak][ak];
        tauv[ak] = own ? hh.tau : tauv[ak];
        return hh.tau;
    };
    T vv[RS];
    auto bcast = [&](auto kc) {
        constexpr int kk = kc;
        constexpr int ak = kk / 8, rk = kk % 8;
        sfor<ak + 1, RS>([&](auto a) { vv[a] = colbc<rk>(f[a][ak], raddr); });
        const T v0 = colbc<rk>(f[ak][ak], raddr);
        vv[ak] = (r > rk) ? v0 : ((r == rk) ? T(1) : T(0));
    };
    auto dot = [&](auto kc, auto bc, T tau) {
        constexpr int kk = kc, b = bc;
        constexpr int ak = kk / 8, rk = kk % 8;
        T d;
        if constexpr (b < CS) {
            d = vv[ak] * f[ak][b];
            sfor<ak + 1, RS>([&](auto a) { d = fmaT(vv[a], f[a][b], d); });
        }
        else {
            d = vv[ak] * gb[ak];
            sfor<ak + 1, RS>([&](auto a) { d = fmaT(vv[a], gb[a], d); });
        }
        d = row_sum(d);
        if constexpr (b == CS) return d * tau;
        return (b > ak || c > rk) ? d * tau : T(0);
    };
    auto upd = [&](auto kc, auto bc, T w) {
        constexpr int kk = kc, b = bc;
        constexpr int ak = kk / 8, rk = kk % 8;
        if constexpr (b < CS) {
            f[ak][b] = (r >= rk) ? fmaT(-vv[ak], w, f[ak][b]) : f[ak][b];
            sfor<ak + 1, RS>([&](auto a) { f[a][b] = fmaT(-vv[a], w, f[a][b]); });
        }
        else {
            gb[ak] = (r >= rk) ? fmaT(-vv[ak], w, gb[ak]) : gb[ak];
            sfor<ak + 1, RS>([&](auto a) { gb[a] = fmaT(-vv[a], w, gb[a]); });
        }
    };
    /* look-ahead, as in qr_grid.hip: past column n - 1 the generated
     * reflector comes from a zero column (tau = 0) and is never stored */
    T tau = T(0);
    if (n > 0) {
        tau = gen(std::integral_constant<int, 0>{});
        bcast(std::integral_constant<int, 0>{});
    }
    sfor<0, KMAX>([&](auto kc) {
        constexpr int kk = kc;
        constexpr int bk = kk / 8, b1 = (kk + 1) / 8;
        if (kk < n) {
            if constexpr (kk + 1 < KMAX) {
                constexpr auto B1 = std::integral_constant<int, b1>{};
                upd(kc, B1, dot(kc, B1, tau));
                const T tn = gen(std::integral_constant<int, kk + 1>{});
                sfor<bk, CS + 1>([&](auto b) {
                    if constexpr (b != b1) upd(kc, b, dot(kc, b, tau));
                });
                bcast(std::integral_constant<int, kk + 1>{});
                tau = tn;
            }
            else {
                sfor<bk, CS + 1>([&](auto b) { upd(kc, b, dot(kc, b, tau)); });
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

/* the first exact zero on R's diagonal, + 1 (0: none): R(l, l), l = c + 8b,
 * sits in slot (b, b) of lane (c, c) = 9c */
template <class T> __device__ __forceinline__ int first_zero_pivot(const T (&f)[RS][CS], int r, int c, int n)
{
    int bad = 0;
    sfor<0, CS>([&](auto b) {
        const unsigned long long z =
            __builtin_amdgcn_ballot_w64(r == c && c + 8 * b < n && f[b][b] == T(0));
        if (bad == 0 && z != 0ull) bad = (int)(__builtin_ctzll(z) / 9) + 8 * b + 1;
    });
    return bad;
}

/* R X = (Q^T B)(0:n): x_l = b_l / R(l, l), then b_i -= R(i, l) x_l for rows
 * i < l (the reference's trsm order) */
template <class T>
__device__ __forceinline__ void back_substitute(const T (&f)[RS][CS], T (&gb)[RS], int r, int n, int raddr)
{
    sfor_down<0, KMAX>([&](auto lc) {
        constexpr int l = lc;
        constexpr int al = l / 8, rl = l % 8;
        if (l < n) {
            T rc[al + 1]; /* R(r + 8a, l) from column lane l % 8 */
            sfor<0, al + 1>([&](auto a) { rc[a] = colbc<rl>(f[a][al], raddr); });
            const T x = rowbc<rl>(gb[al]) / lane_value<rl + 8 * rl>(f[al][al]);
            sfor<0, al>([&](auto a) { gb[a] = fmaT(-rc[a], x, gb[a]); });
            gb[al] = (r < rl) ? fmaT(-rc[al], x, gb[al]) : ((r == rl) ? x : gb[al]);
        }
        __builtin_amdgcn_sched_barrier(0);
    });
}

/* a persistent grid over a class queue: each wavefront takes the next
 * problem until the queue is empty; each problem column-major on its own */
template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
qr_grid64_batch_kernel(BatchArgs<T> p)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 7, c = lane >> 3;
    const int raddr = 4 * r; /* ds_bpermute byte address of lane (r, 0) */
    const unsigned count = batch_count(p);
    for (;;) {
        unsigned kq = 0;
        if (lane == 0) kq = atomicAdd(p.counter, 1u);
        kq = __builtin_amdgcn_readfirstlane(kq);
        if (kq >= count) return;
        const BatchProblem d = p.desc[kq];
        const int m = d.m, n = d.n, nrhs = d.nrhs;
        T *acol = p.a + d.a_off + (long long)c * d.lda + r;
        T *bcol = p.b + d.b_off + (long long)c * d.ldb + r;
        const long long astep = 8ll * d.lda;

        T f[RS][CS], gb[RS];
        sfor<0, RS>([&](auto a) {
            sfor<0, CS>([&](auto b) {
                f[a][b] = (r + 8 * a < m && c + 8 * b < n) ? acol[b * astep + 8 * a] : T(0);
            });
            gb[a] = (r + 8 * a < m && c < nrhs) ? bcol[8 * a] : T(0);
        });
        T tauv[CS];
        sfor<0, CS>([&](auto b) { tauv[b] = T(0); });
        factor<T>(f, gb, tauv, r, c, n, raddr);

        T *acol2 = opaque(acol), *bcol2 = opaque(bcol);
        const int r2 = opaque(r), c2 = opaque(c);
        sfor<0, RS>([&](auto a) {
            sfor<0, CS>([&](auto b) {
                if (r2 + 8 * a < m && c2 + 8 * b < n) acol2[b * astep + 8 * a] = f[a][b];
            });
        });
        const int bad = first_zero_pivot<T>(f, r, c, n);
        if (lane == 0 && p.info && d.idx < p.ninfo) p.info[d.idx] = bad;
        if (!(p.info && bad)) back_substitute<T>(f, gb, r, n, raddr);
        sfor<0, RS>([&](auto a) {
            if (r2 + 8 * a < m && c2 < nrhs) bcol2[8 * a] = gb[a];
        });
    }
}

/* compact operands (gels_compact / geqrf_compact with 32 < q <= 64): one
 * problem per wavefront, the wavefronts of 4 consecutive logical blocks (16
 * consecutive slots, a 128-byte run of each element at V >= 16 in FP64) on
 * one XCD; buffer addressing as in qr_grid.hip */
template <class T> struct Grid64Args {
    T *a, *b, *tau;
    int *info;
    int a_rs, a_cs, b_rs, b_cs; /* element strides, V folded in */
    long long a_gs, b_gs;
    int V, m, n, nrhs, ninfo;
    long long count; /* slots, padded ones included */
};

template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
qr_grid64_kernel(Grid64Args<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int r = lane & 7, c = lane >> 3;
    const int raddr = 4 * r;
    const long long kq = (long long)team::xcd_block(blockIdx.x) * 4 + team::wave_in_block();
    if (kq >= p.count) return; /* whole wavefront */
    const int m = p.m, n = p.n, nrhs = p.nrhs;
    const long long g = kq / p.V, v = kq % p.V;
    const long long groups = p.count / p.V;
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + g * p.a_gs, team::group_range(groups - g, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + g * p.b_gs, team::group_range(groups - g, p.b_gs, SZ));
    const unsigned ars = SZ * p.a_rs, acs = SZ * p.a_cs, brs = SZ * p.b_rs, bcs = SZ * p.b_cs;
    const unsigned aoff = (unsigned)(v * SZ) + r * ars + c * acs;
    const unsigned boff = (unsigned)(v * SZ) + r * brs + c * bcs;
    auto a_at = [&](unsigned lo, int rr, int cc, int a, int b) {
        return (rr + 8 * a < m && cc + 8 * b < n) ? lo : team::kOOB;
    };
    auto b_at = [&](unsigned lo, int rr, int cc, int a) {
        return (rr + 8 * a < m && cc < nrhs) ? lo : team::kOOB;
    };
    T f[RS][CS], gb[RS];
    sfor<0, RS>([&](auto a) {
        sfor<0, CS>([&](auto b) {
            f[a][b] = team::bload(T(0), ra, a_at(aoff, r, c, a, b), a * 8 * ars + b * 8 * acs);
        });
        gb[a] = team::bload(T(0), rb, b_at(boff, r, c, a), a * 8 * brs);
    });
    T tauv[CS];
    sfor<0, CS>([&](auto b) { tauv[b] = T(0); });
    factor<T>(f, gb, tauv, r, c, n, raddr);

    const unsigned aoff2 = opaque(aoff), boff2 = opaque(boff);
    const int r2 = opaque(r), c2 = opaque(c);
    sfor<0, RS>([&](auto a) {
        sfor<0, CS>([&](auto b) {
            team::bstore(f[a][b], ra, a_at(aoff2, r2, c2, a, b), a * 8 * ars + b * 8 * acs);
        });
    });
    if (p.tau) {
        /* tau of reflector kk at ((g * n + kk) * V + v) */
        const unsigned tV = SZ * p.V;
        const __amdgpu_buffer_rsrc_t rt = team::group_rsrc(
            p.tau + g * (long long)n * p.V, team::group_range(groups - g, (size_t)n * p.V, SZ));
        const unsigned toff = (unsigned)(v * SZ) + c * tV;
        sfor<0, CS>([&](auto b) {
            team::bstore(tauv[b], rt, (r == 0 && c + 8 * b < n) ? toff : team::kOOB, b * 8 * tV);
        });
    }
    const int bad = first_zero_pivot<T>(f, r, c, n);
    if (lane == 0 && p.info && kq < p.ninfo) p.info[kq] = bad;
    if (nrhs > 0 && !(p.info && bad)) back_substitute<T>(f, gb, r, n, raddr);
    sfor<0, RS>([&](auto a) { team::bstore(gb[a], rb, b_at(boff2, r2, c2, a), a * 8 * brs); });
}

} /* namespace qrgrid64 */

template <class T>
hipError_t launch_qr_grid64(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                            Operand<T> B, int V, int nm, int *info)
{
    qrgrid64::Grid64Args<T> args;
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
    const unsigned blocks = team::xcd_grid((unsigned)((args.count + 3) / 4));
    hipLaunchKernelGGL((qrgrid64::qr_grid64_kernel<T>), dim3(blocks), dim3(256), 0, current_stream(), args);
    return hipGetLastError();
}

template hipError_t launch_qr_grid64<double>(int, int, int, bool, Operand<double>, double *,
                                             Operand<double>, int, int, int *);
template hipError_t launch_qr_grid64<float>(int, int, int, bool, Operand<float>, float *,
                                            Operand<float>, int, int, int *);

/* batch kernel 11 */
template <class T>
hipError_t launch_qr_grid64_batch(const BatchArgs<T> &args, unsigned blocks, hipStream_t s)
{
    hipLaunchKernelGGL((qrgrid64::qr_grid64_batch_kernel<T>), dim3(blocks), dim3(256), 0, s, args);
    return hipGetLastError();
}

template <class T> hipError_t qr_grid64_batch_occupancy(int *blocks)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, qrgrid64::qr_grid64_batch_kernel<T>, 256, 0);
}

template hipError_t launch_qr_grid64_batch<double>(const BatchArgs<double> &, unsigned, hipStream_t);
template hipError_t launch_qr_grid64_batch<float>(const BatchArgs<float> &, unsigned, hipStream_t);
template hipError_t qr_grid64_batch_occupancy<double>(int *);
template hipError_t qr_grid64_batch_occupancy<float>(int *);

} /* namespace cbk */
