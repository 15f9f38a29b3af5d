/* ormqr_team.hip -- C := op(Q) C or C op(Q) (ormqr_compact) with one matrix per
 * 16-lane team, for reflector lengths nq in {16, 32}.
 *
 * The effective operand is C' = C (side L) or C^T (side R), nq x ncols; the
 * reflectors are applied forward when side L and trans T (or side R and N),
 * backward otherwise (src/cbk_ormqr_compact.hpp:131-156). Per reflector kk
 * (larf, src/cbk_ormqr_compact.hpp:68-105):
 *     w = C'(kk,:) + sum_{i>kk} V(i,kk) C'(i,:),  w *= tau(kk),
 *     C'(kk,:) -= w,  C'(i,:) -= V(i,kk) w
 * with lane t owning rows t, t+16 of V and C' and the sum a butterfly over the
 * team (every lane ends with the same bits), four columns of C' at a time.
 * The reference sums rows in ascending order, so results match it to rounding.
 */

#include "cbk_internal.h"
#include "team_common.h"

namespace cbk {
namespace {

using namespace team;

template <class T> struct OrmqrArgs {
    const T *a; /* reflectors: V(i, kk) at (i*a_si + kk*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    const T *tau; /* k per matrix, compact */
    T *c; /* C': element (i, j) at (i*c_si + j*c_sj)*V */
    int c_si, c_sj;
    size_t c_gs;
    int V;
    long nmat;
    int k, ncols;
    bool forward;
};

constexpr int kTeamsPerBlock = 16;
constexpr int NRB = 4;

template <class T> __device__ __forceinline__ T team_sum(T x)
{
    x = x + dpp<0xB1>(x);
    x = x + dpp<0x4E>(x);
    x = x + dpp<0x141>(x);
    x = x + dpp<0x140>(x);
    return x;
}

template <class T, int P>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
ormqr_team_kernel(OrmqrArgs<T> p)
{
    constexpr int R = P / TS;
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
    const __amdgpu_buffer_rsrc_t rc = group_rsrc(p.c + g * p.c_gs, group_range(groups_left, p.c_gs, SZ));
    const __amdgpu_buffer_rsrc_t rt = group_rsrc(p.tau + g * (size_t)p.k * V,
                                                 group_range(groups_left, (size_t)p.k * V, SZ));
    const unsigned a_col = (unsigned)p.a_sj * V * SZ, c_col = (unsigned)p.c_sj * V * SZ;
    unsigned a_row[R], c_row[R];
    sfor<0, R>([&](auto r) {
        const int i = t + TS * r;
        a_row[r] = ((unsigned)(i * p.a_si) * V + v) * SZ;
        c_row[r] = ((unsigned)(i * p.c_si) * V + v) * SZ;
    });
    /* v_i of each reflector: V(i, kk) below the diagonal, 1 on it, 0 above
     * (the diagonal of A is never read, as larf's implicit unit pivot) */
    T vr[R][P];
    T taus[P];
    sfor<0, P>([&](auto kk) {
        sfor<0, R>([&](auto r) {
            const int i = t + TS * r;
            const T x = bload(T(), ra, (kk < p.k && i > kk) ? a_row[r] : kOOB, kk * a_col);
            vr[r][kk] = (i == kk) ? T(1) : x;
        });
        taus[kk] = bload(T(), rt, kk < p.k ? ((unsigned)kk * V + v) * SZ : kOOB, 0);
    });
    for (int j0 = 0; j0 < p.ncols; j0 += NRB) {
        unsigned ccol[NRB];
        sfor<0, NRB>([&](auto q) { ccol[q] = (j0 + q < p.ncols) ? (j0 + q) * c_col : kOOB; });
        T c[R][NRB];
        sfor<0, R>([&](auto r) { sfor<0, NRB>([&](auto q) { c[r][q] = bload(T(), rc, c_row[r], ccol[q]); }); });
        auto apply = [&](auto kc) {
            constexpr int kk = kc;
            constexpr int rk = kk / TS;
            if (kk < p.k) {
                sfor<0, NRB>([&](auto q) {
                    T pw = T(0);
                    sfor<rk, R>([&](auto r) { pw = fmaT(vr[r][kk], c[r][q], pw); });
                    const T w = team_sum(pw) * taus[kk];
                    sfor<rk, R>([&](auto r) { c[r][q] = fmaT(-vr[r][kk], w, c[r][q]); });
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        if (p.forward)
            sfor<0, P>(apply);
        else
            sfor_down<0, P>(apply);
        sfor<0, R>([&](auto r) { sfor<0, NRB>([&](auto q) { bstore(c[r][q], rc, c_row[r], ccol[q]); }); });
    }
}

} /* namespace */

bool ormqr_team_supports(int nq) { return nq == 16 || nq == 32; }

template <class T>
hipError_t launch_ormqr_team(bool left, bool tran, int m, int n, int k, Operand<T> A,
                             const T *tau, Operand<T> C, int V, int nm)
{
    OrmqrArgs<T> args;
    args.a = A.p;
    args.a_si = A.row ? A.ld : 1;
    args.a_sj = A.row ? 1 : A.ld;
    args.a_gs = A.gstride;
    args.tau = tau;
    args.c = C.p;
    int si = C.row ? C.ld : 1, sj = C.row ? 1 : C.ld;
    if (!left) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.c_si = si;
    args.c_sj = sj;
    args.c_gs = C.gstride;
    args.V = V;
    args.nmat = (long)((nm + V - 1) / V) * V;
    args.k = k;
    args.ncols = left ? n : m;
    args.forward = (left == tran);
    const int nq = left ? m : n;
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    if (nq == 16)
        hipLaunchKernelGGL((ormqr_team_kernel<T, 16>), dim3(blocks), dim3(256), 0, current_stream(), args);
    else if (nq == 32)
        hipLaunchKernelGGL((ormqr_team_kernel<T, 32>), dim3(blocks), dim3(256), 0, current_stream(), args);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

template hipError_t launch_ormqr_team<double>(bool, bool, int, int, int, Operand<double>,
                                              const double *, Operand<double>, int, int);
template hipError_t launch_ormqr_team<float>(bool, bool, int, int, int, Operand<float>,
                                             const float *, Operand<float>, int, int);

} /* namespace cbk */
