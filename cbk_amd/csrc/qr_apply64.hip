/* qr_apply64.hip -- ormqr_compact and orgqr_compact for reflector lengths up
 * to 64 on the 8 x 8 lane grid of qr_grid64.hip, one problem per wavefront
 * (src/cbk_ormqr_compact.hpp:68-156, src/cbk_orgqr_compact.hpp:58-92).
 *
 * Lane (r, c) holds rows i = r + 8a (a < RS) of columns c + 8b. Per reflector
 * kk (rows and slots compile-time, kk < k a wave-uniform branch):
 *   bcast  column kk's body (rows > kk) goes from column lane kk % 8 to the
 *          others by ds_bpermute, with the implicit 1 at row kk and 0 above;
 *          tau(kk) reaches every lane by v_readlane;
 *   larf   each lane's partial w_j = C'(kk, j) + sum of its rows' v_i C'(i, j),
 *          the 8-row-lane DPP sum, w_j *= tau, then C'(i, j) -= v_i w_j on
 *          the lane's rows i >= kk.
 * ormqr: C' = C (side L) or C^T (side R), nq x ncols, the reflectors held in
 * registers for the whole problem and C' streamed through in 8-column chunks
 * (one column a column lane); the sweep runs forward when side L and trans T
 * or side R and trans N, backward otherwise (the LAPACK DORMQR table).
 * orgqr: A (m x n) in registers, columns k..n-1 seeded as unit vectors, then
 * for kk = k-1 .. 0 H(kk) applied to the columns after kk and column kk formed
 * as H(kk) e_kk: the body times -tau, 1 - tau on the diagonal, zeros above.
 *
 * The reference sums a reflector's rows in one ascending chain; here each lane
 * sums its rows in order and the row lanes are then added, so results agree
 * with it to rounding (the tests bound it, as for the other grid and team
 * kernels). Every other operation is the reference's, in its order.
 */
#include "grid64_common.h"

namespace cbk {
namespace qrgrid64 {

template <class T> struct ApplyArgs {
    const T *a;   /* reflectors (ormqr) / the matrix formed in place (orgqr) */
    T *c;         /* C' (ormqr) */
    const T *tau; /* k per problem, compact: tau(kk) at (g k + kk) V + v */
    int a_rs, a_cs, c_rs, c_cs; /* element strides, V folded in */
    long long a_gs, c_gs;
    int V, nq, k, ncols; /* orgqr: nq = m, ncols = n */
    long long count;     /* slots, padded ones included */
};

/* the reflector body of column kk: v_i for the lane's rows (rows > kk from
 * column lane kk % 8, 1 at kk, 0 above) */
template <class T, int RSN, int KK, int CSN>
__device__ __forceinline__ void reflector(const T (&f)[RSN][CSN], T (&vv)[RSN], int r, int raddr)
{
    constexpr int ak = KK / 8, rk = KK % 8;
    sfor<ak + 1, RSN>([&](auto a) { vv[a] = colbc<rk>(f[a][ak], raddr); });
    const T v0 = colbc<rk>(f[ak][ak], raddr);
    vv[ak] = (r > rk) ? v0 : ((r == rk) ? T(1) : T(0));
}

/* w = tau (x(kk) + sum_{i > kk} v_i x(i)) over the column, then x -= v w on
 * rows >= kk; skip: leave this lane's column as it is (w = 0) */
template <class T, int RSN, int KK, class X>
__device__ __forceinline__ void larf(const T (&vv)[RSN], X &&x, T tau, int r, bool keep)
{
    constexpr int ak = KK / 8, rk = KK % 8;
    T d = (r >= rk) ? vv[ak] * x(ak) : T(0);
    sfor<ak + 1, RSN>([&](auto a) { d = fmaT(vv[a], x(a), d); });
    d = row_sum(d);
    const T w = keep ? d * tau : T(0);
    x(ak) = (r >= rk) ? fmaT(-vv[ak], w, x(ak)) : x(ak);
    sfor<ak + 1, RSN>([&](auto a) { x(a) = fmaT(-vv[a], w, x(a)); });
}

template <class T, int RSN, int KS, bool FWD>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
ormqr_grid_kernel(ApplyArgs<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int r = lane & 7, c = lane >> 3;
    const int raddr = 4 * r;
    const long long kq = (long long)team::xcd_block(blockIdx.x) * 4 + team::wave_in_block();
    if (kq >= p.count) return; /* whole wavefront */
    const int nq = p.nq, k = p.k, ncols = p.ncols;
    const long long g = kq / p.V, v = kq % p.V;
    const long long groups = p.count / p.V;
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + g * p.a_gs, team::group_range(groups - g, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rc =
        team::group_rsrc(p.c + g * p.c_gs, team::group_range(groups - g, p.c_gs, SZ));
    const __amdgpu_buffer_rsrc_t rt = team::group_rsrc(
        p.tau + g * (long long)k * p.V, team::group_range(groups - g, (size_t)k * p.V, SZ));
    const unsigned ars = SZ * p.a_rs, acs = SZ * p.a_cs, crs = SZ * p.c_rs, ccs = SZ * p.c_cs;
    const unsigned aoff = (unsigned)(v * SZ) + r * ars + c * acs;
    const unsigned coff = (unsigned)(v * SZ) + r * crs + c * ccs;
    const unsigned tV = SZ * p.V, toff = (unsigned)(v * SZ) + c * tV;

    /* the bodies: rows below the diagonal of the k reflector columns */
    T va[RSN][KS];
    sfor<0, RSN>([&](auto a) {
        sfor<0, KS>([&](auto b) {
            const int i = r + 8 * a, j = c + 8 * b;
            va[a][b] = team::bload(T(0), ra, (i < nq && j < k && i > j) ? aoff : team::kOOB,
                                   a * 8 * ars + b * 8 * acs);
        });
    });
    T tv[KS];
    sfor<0, KS>([&](auto b) { tv[b] = team::bload(T(0), rt, (c + 8 * b < k) ? toff : team::kOOB, b * 8 * tV); });

    for (int q0 = 0; q0 < ncols; q0 += 8) {
        const bool col = q0 + c < ncols;
        const unsigned qs = (unsigned)q0 * ccs;
        T x[RSN];
        sfor<0, RSN>([&](auto a) {
            x[a] = team::bload(T(0), rc, (r + 8 * a < nq && col) ? coff : team::kOOB, a * 8 * crs + qs);
        });
        auto step = [&](auto kc) {
            constexpr int kk = kc;
            if (kk < k) {
                T vv[RSN];
                reflector<T, RSN, kk>(va, vv, r, raddr);
                const T tau = lane_value<8 * (kk % 8)>(tv[kk / 8]);
                larf<T, RSN, kk>(vv, [&](int a) -> T & { return x[a]; }, tau, r, true);
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        if constexpr (FWD)
            sfor<0, 8 * KS>(step);
        else
            sfor_down<0, 8 * KS>(step);
        const unsigned coff2 = opaque(coff);
        sfor<0, RSN>([&](auto a) {
            team::bstore(x[a], rc, (r + 8 * a < nq && col) ? coff2 : team::kOOB, a * 8 * crs + qs);
        });
    }
}

template <class T, int RSN, int CSN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
orgqr_grid_kernel(ApplyArgs<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int r = lane & 7, c = lane >> 3;
    const int raddr = 4 * r;
    const long long kq = (long long)team::xcd_block(blockIdx.x) * 4 + team::wave_in_block();
    if (kq >= p.count) return; /* whole wavefront */
    const int m = p.nq, k = p.k, n = p.ncols;
    const long long g = kq / p.V, v = kq % p.V;
    const long long groups = p.count / p.V;
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + g * p.a_gs, team::group_range(groups - g, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rt = team::group_rsrc(
        p.tau + g * (long long)k * p.V, team::group_range(groups - g, (size_t)k * p.V, SZ));
    const unsigned ars = SZ * p.a_rs, acs = SZ * p.a_cs;
    const unsigned aoff = (unsigned)(v * SZ) + r * ars + c * acs;
    const unsigned tV = SZ * p.V, toff = (unsigned)(v * SZ) + c * tV;

    /* only the bodies below the diagonal of columns < k are read; columns
     * k..n-1 start as unit vectors */
    T f[RSN][CSN];
    sfor<0, RSN>([&](auto a) {
        sfor<0, CSN>([&](auto b) {
            const int i = r + 8 * a, j = c + 8 * b;
            f[a][b] = team::bload(T(0), ra, (i < m && j < k && i > j) ? aoff : team::kOOB,
                                  a * 8 * ars + b * 8 * acs);
            if constexpr (a == b) f[a][b] = (r == c && j >= k) ? T(1) : f[a][b];
        });
    });
    T tv[CSN];
    sfor<0, CSN>([&](auto b) { tv[b] = team::bload(T(0), rt, (c + 8 * b < k) ? toff : team::kOOB, b * 8 * tV); });

    sfor_down<0, 8 * CSN>([&](auto kc) {
        constexpr int kk = kc;
        constexpr int ak = kk / 8, rk = kk % 8;
        if (kk < k) {
            T vv[RSN];
            reflector<T, RSN, kk>(f, vv, r, raddr);
            const T tau = lane_value<8 * rk>(tv[ak]);
            /* H(kk) on the columns after kk */
            sfor<ak, CSN>([&](auto b) {
                larf<T, RSN, kk>(vv, [&](int a) -> T & { return f[a][b]; }, tau, r, b > ak || c > rk);
            });
            /* column kk := H(kk) e_kk */
            const bool own = c == rk;
            const T mt = -tau;
            sfor<0, RSN>([&](auto a) {
                if constexpr (a < ak)
                    f[a][ak] = own ? T(0) : f[a][ak];
                else if constexpr (a == ak)
                    f[a][ak] = own ? ((r > rk) ? f[a][ak] * mt : ((r == rk) ? T(1) - tau : T(0))) : f[a][ak];
                else
                    f[a][ak] = own ? f[a][ak] * mt : f[a][ak];
            });
        }
        __builtin_amdgcn_sched_barrier(0);
    });

    const unsigned aoff2 = opaque(aoff);
    const int r2 = opaque(r), c2 = opaque(c);
    sfor<0, RSN>([&](auto a) {
        sfor<0, CSN>([&](auto b) {
            team::bstore(f[a][b], ra, (r2 + 8 * a < m && c2 + 8 * b < n) ? aoff2 : team::kOOB,
                         a * 8 * ars + b * 8 * acs);
        });
    });
}

template <class T> unsigned apply_blocks(const ApplyArgs<T> &args)
{
    return team::xcd_grid((unsigned)((args.count + 3) / 4));
}

template <class T, int RSN, int KS>
void launch_ormqr_rk(bool fwd, const ApplyArgs<T> &args)
{
    if (fwd)
        hipLaunchKernelGGL((ormqr_grid_kernel<T, RSN, KS, true>), dim3(apply_blocks(args)), dim3(256), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((ormqr_grid_kernel<T, RSN, KS, false>), dim3(apply_blocks(args)), dim3(256), 0,
                           current_stream(), args);
}

} /* namespace qrgrid64 */

/* reflector lengths up to 64; the register and team kernels keep the shorter
 * lengths they serve (the dispatch order) */
bool apply_grid_supports(int nq) { return nq >= 1 && nq <= 64; }

template <class T>
hipError_t launch_ormqr_grid(bool left, bool tran, int m, int n, int k, Operand<T> A, const T *tau,
                             Operand<T> C, int V, int nm)
{
    const int nq = left ? m : n;
    if (!apply_grid_supports(nq) || k > nq) return hipErrorInvalidValue;
    qrgrid64::ApplyArgs<T> args;
    args.a = A.p;
    args.c = C.p;
    args.tau = tau;
    args.a_rs = (A.row ? A.ld : 1) * V;
    args.a_cs = (A.row ? 1 : A.ld) * V;
    int crs = (C.row ? C.ld : 1) * V, ccs = (C.row ? 1 : C.ld) * V;
    if (!left) { /* C' = C^T */
        const int tmp = crs;
        crs = ccs;
        ccs = tmp;
    }
    args.c_rs = crs;
    args.c_cs = ccs;
    args.a_gs = (long long)A.gstride;
    args.c_gs = (long long)C.gstride;
    args.V = V;
    args.nq = nq;
    args.k = k;
    args.ncols = left ? n : m;
    args.count = (long long)((nm + V - 1) / V) * V; /* padded slots too, as the reference */
    const bool fwd = left == tran;
    if (nq <= 32)
        qrgrid64::launch_ormqr_rk<T, 4, 4>(fwd, args);
    else if (k <= 32)
        qrgrid64::launch_ormqr_rk<T, 8, 4>(fwd, args);
    else
        qrgrid64::launch_ormqr_rk<T, 8, 8>(fwd, args);
    return hipGetLastError();
}

template <class T>
hipError_t launch_orgqr_grid(int m, int n, int k, Operand<T> A, const T *tau, int V, int nm)
{
    if (!apply_grid_supports(m) || n > m || k > n) return hipErrorInvalidValue;
    qrgrid64::ApplyArgs<T> args;
    args.a = A.p;
    args.c = nullptr;
    args.tau = tau;
    args.a_rs = (A.row ? A.ld : 1) * V;
    args.a_cs = (A.row ? 1 : A.ld) * V;
    args.c_rs = args.c_cs = 0;
    args.a_gs = (long long)A.gstride;
    args.c_gs = 0;
    args.V = V;
    args.nq = m;
    args.k = k;
    args.ncols = n;
    args.count = (long long)((nm + V - 1) / V) * V;
    const unsigned blocks = qrgrid64::apply_blocks(args);
    if (m <= 32)
        hipLaunchKernelGGL((qrgrid64::orgqr_grid_kernel<T, 4, 4>), dim3(blocks), dim3(256), 0, current_stream(), args);
    else if (n <= 32)
        hipLaunchKernelGGL((qrgrid64::orgqr_grid_kernel<T, 8, 4>), dim3(blocks), dim3(256), 0, current_stream(), args);
    else
        hipLaunchKernelGGL((qrgrid64::orgqr_grid_kernel<T, 8, 8>), dim3(blocks), dim3(256), 0, current_stream(), args);
    return hipGetLastError();
}

template hipError_t launch_ormqr_grid<double>(bool, bool, int, int, int, Operand<double>, const double *,
                                              Operand<double>, int, int);
template hipError_t launch_ormqr_grid<float>(bool, bool, int, int, int, Operand<float>, const float *,
                                             Operand<float>, int, int);
template hipError_t launch_orgqr_grid<double>(int, int, int, Operand<double>, const double *, int, int);
template hipError_t launch_orgqr_grid<float>(int, int, int, Operand<float>, const float *, int, int);

} /* namespace cbk */
