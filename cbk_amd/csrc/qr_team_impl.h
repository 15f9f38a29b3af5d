/* qr_team_impl.h -- Householder QR (geqrf) and the least-squares solve (gels,
 * overdetermined case) with one p x q matrix per 16-lane team.
 *
 * Lane t owns rows t, t+16, ... of the tall view F (F = A, or A^T for gels
 * with trans and m < n) and of B: f[r][c] = F(t+16r, c). Per reflector kk
 * (the reference's geqr2, src/cbk_geqrf_compact.hpp:74-131, and larf,
 * src/cbk_ormqr_compact.hpp:68-105):
 *   tail  = sum_{i>kk} F(i,kk)^2           (team all-reduce)
 *   larfg: norm = sqrt(x0^2 + tail), beta = x0 >= 0 ? -norm : norm,
 *          tau = (beta - x0)/beta, body scale 1/(x0 - beta), R(kk,kk) = beta
 *          (tau = 0 and the column unchanged when tail = 0)
 *   apply H(kk) to the trailing columns and to B:
 *          w = F(kk,c) + sum_{i>kk} v_i F(i,c)  (team all-reduce), w *= tau,
 *          F(kk,c) -= w, F(i,c) -= v_i w
 * then R X = (Q^T B)(0:q) by back substitution, column-oriented over the rows
 * the lanes already hold (x_l = b_l / R(l,l), b_i -= R(i,l) x_l).
 *
 * The team sums (tail and w) are a butterfly over the 16 lanes -- every lane
 * ends with the same bits -- where the reference sums rows in ascending order,
 * and the substitution subtracts in descending l, so results match the
 * reference to rounding (the tests' tolerance), not to the bit.
 */
#ifndef CBK_QR_TEAM_IMPL_H
#define CBK_QR_TEAM_IMPL_H

#include "cbk_internal.h"
#include "team_common.h"

namespace cbk {
namespace qrteam {

using namespace team;

template <class T> struct QrArgs {
    T *a; /* the tall view F: element (i,c) at (i*si + c*sj)*V */
    int a_si, a_sj;
    size_t a_gs;
    T *tau; /* q per matrix, compact */
    T *b;   /* p x nrhs, may be null (geqrf) */
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long nmat;
    int nm;
    int nrhs;
    int *info; /* gels_info: first zero R(i,i), or null */
};

constexpr int kTeamsPerBlock = 16; /* 4 wavefronts: whole 128-byte lines per block */

/* sum over the 16 lanes of the row, the same bits in every lane */
template <class T> __device__ __forceinline__ T team_sum(T x)
{
    x = x + dpp<0xB1>(x);  /* quad_perm [1,0,3,2] */
    x = x + dpp<0x4E>(x);  /* quad_perm [2,3,0,1] */
    x = x + dpp<0x141>(x); /* row_half_mirror */
    x = x + dpp<0x140>(x); /* row_mirror */
    return x;
}

template <class T, int P, int Q, int NRB> hipError_t launch_pq(const QrArgs<T> &args);

template <class T, int P, int Q, int NRB>
__global__ __launch_bounds__(256) void qr_team_kernel(QrArgs<T> p)
{
    constexpr int R = (P + TS - 1) / TS;
    constexpr unsigned SZ = sizeof(T);
    const int t = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;
    const long mat0 = (long)blockIdx.x * kTeamsPerBlock + 4 * wave_in_block(); /* wave-uniform */
    if (mat0 >= p.nmat) return;
    const long mat = mat0 + (team & 3);
    const long g = mat0 / p.V;
    const long groups_left = p.nmat / p.V - g;
    const unsigned v = (unsigned)(mat0 % p.V) + (team & 3);
    const unsigned V = (unsigned)p.V;
    const __amdgpu_buffer_rsrc_t ra = group_rsrc(p.a + g * p.a_gs, group_range(groups_left, p.a_gs, SZ));
    const unsigned a_col = (unsigned)p.a_sj * V * SZ;
    unsigned a_row[R];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        a_row[r] = (i < P) ? ((unsigned)(i * p.a_si) * V + v) * SZ : kOOB;
    });

    T f[R][Q];
    sfor<0, R>([&](auto r) {
        sfor<0, Q>([&](auto c) { f[r][c] = bload(T(), ra, a_row[r], c * a_col); });
    });
    const bool solve = p.b != nullptr;
    T b[R][NRB];
    const __amdgpu_buffer_rsrc_t rb =
        group_rsrc(solve ? p.b + g * p.b_gs : p.a, solve ? group_range(groups_left, p.b_gs, SZ) : 0u);
    const unsigned b_col = (unsigned)p.b_sj * V * SZ;
    unsigned b_row[R], bcol[NRB];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        b_row[r] = (i < P && solve) ? ((unsigned)(i * p.b_si) * V + v) * SZ : kOOB;
    });
    sfor<0, NRB>([&](auto q) { bcol[q] = (q < p.nrhs) ? q * b_col : kOOB; });
    sfor<0, R>([&](auto r) {
        sfor<0, NRB>([&](auto q) { b[r][q] = bload(T(), rb, b_row[r], bcol[q]); });
    });

    T taus[(Q + TS - 1) / TS]; /* tau(kk) kept by lane kk % 16 */
    sfor<0, Q>([&](auto kc) {
        constexpr int kk = kc;
        constexpr int rk = kk / TS, tk = kk % TS;
        /* tail = sum_{i>kk} F(i,kk)^2 over the team */
        T part = T(0);
        sfor<rk, R>([&](auto r) {
            const int i = t + TS * r;
            const T x = (i > kk) ? f[r][kk] : T(0);
            part = fmaT(x, x, part);
        });
        const T tail = team_sum(part);
        const T x0 = bcast<tk>(f[rk][kk]);
        const T norm = sqrtT(fmaT(x0, x0, tail));
        const T beta = (x0 >= T(0)) ? -norm : norm;
        const bool has = tail > T(0);
        const T tau = has ? (beta - x0) / beta : T(0);
        const T inv = has ? T(1) / (x0 - beta) : T(0);
        const T rdiag = has ? beta : x0;
        if constexpr (kk < Q) taus[rk] = (t == tk) ? tau : taus[rk];
        /* the reflector body v_i = F(i,kk) / (x0 - beta) below the diagonal */
        sfor<rk, R>([&](auto r) {
            const int i = t + TS * r;
            f[r][kk] = (i > kk) ? f[r][kk] * inv : (i == kk ? rdiag : f[r][kk]);
        });
        /* apply H(kk) = I - tau v v^T (v_kk = 1) to columns kk+1.. and to B */
        T vi[R];
        sfor<rk, R>([&](auto r) {
            const int i = t + TS * r;
            vi[r] = (i > kk) ? f[r][kk] : (i == kk ? T(1) : T(0));
        });
        sfor<kk + 1, Q>([&](auto c) {
            T pw = T(0);
            sfor<rk, R>([&](auto r) { pw = fmaT(vi[r], f[r][c], pw); });
            const T w = team_sum(pw) * tau;
            sfor<rk, R>([&](auto r) { f[r][c] = fmaT(-vi[r], w, f[r][c]); });
        });
        if (solve) {
            sfor<0, NRB>([&](auto q) {
                T pw = T(0);
                sfor<rk, R>([&](auto r) { pw = fmaT(vi[r], b[r][q], pw); });
                const T w = team_sum(pw) * tau;
                sfor<rk, R>([&](auto r) { b[r][q] = fmaT(-vi[r], w, b[r][q]); });
            });
        }
        __builtin_amdgcn_sched_barrier(0);
    });

    /* A := (R, reflectors), tau */
    sfor<0, R>([&](auto r) {
        sfor<0, Q>([&](auto c) { bstore(f[r][c], ra, a_row[r], c * a_col); });
    });
    {
        const __amdgpu_buffer_rsrc_t rt = group_rsrc(p.tau + (size_t)g * Q * V,
                                                     group_range(groups_left, (size_t)Q * V, SZ));
        sfor<0, (Q + TS - 1) / TS>([&](auto r) {
            const int kk = t + TS * r;
            bstore(taus[r], rt, kk < Q ? ((unsigned)kk * V + v) * SZ : kOOB, 0);
        });
    }
    if (!solve) return;

    /* info: the first exact zero on R's diagonal (LAPACK dtrtrs's test) */
    int bad = 0;
    sfor<0, Q>([&](auto lc) {
        constexpr int l = lc;
        const T d = bcast<l % TS>(f[l / TS][l]);
        if (bad == 0 && d == T(0)) bad = l + 1;
    });
    if (p.info && mat < p.nm && t == 0) p.info[mat] = bad;
    if (!(p.info && bad)) {
        /* R X = (Q^T B)(0:q): x_l = b_l / R(l,l), then b_i -= R(i,l) x_l */
        sfor_down<0, Q>([&](auto lc) {
            constexpr int l = lc;
            constexpr int rl = l / TS, tl = l % TS;
            const T d = bcast<tl>(f[rl][l]);
            sfor<0, NRB>([&](auto q) {
                const T x = bcast<tl>(b[rl][q]) / d;
                b[rl][q] = (t == tl) ? x : b[rl][q];
                sfor<0, rl + 1>([&](auto r) {
                    const int i = t + TS * r;
                    const T nb = fmaT(-f[r][l], x, b[r][q]);
                    b[r][q] = (i < l) ? nb : b[r][q];
                });
            });
        });
    }
    sfor<0, R>([&](auto r) {
        sfor<0, NRB>([&](auto q) { bstore(b[r][q], rb, b_row[r], bcol[q]); });
    });
}

template <class T, int P, int Q, int NRB> hipError_t launch_pq(const QrArgs<T> &args)
{
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    hipLaunchKernelGGL((qr_team_kernel<T, P, Q, NRB>), dim3(blocks), dim3(256), 0,
                       current_stream(), args);
    return hipGetLastError();
}

} /* namespace qrteam */
} /* namespace cbk */

#endif /* CBK_QR_TEAM_IMPL_H */
