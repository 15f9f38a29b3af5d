/* trsm_team.hip -- host launcher of the forward-sweep trsm team kernel
 * (trsm_team_impl.h); the kernels of each precision in trsm_team_<p>.hip */
#include "trsm_team_impl.h"

namespace cbk {

namespace trsmteam {
#define CBK_TRSM_TEAM_EXTERN(T)                                                          \
    extern template hipError_t launch_n<T, 16>(const TrsmArgs<T> &);                      \
    extern template hipError_t launch_n<T, 24>(const TrsmArgs<T> &);                      \
    extern template hipError_t launch_n<T, 32>(const TrsmArgs<T> &);
CBK_TRSM_TEAM_EXTERN(double)
CBK_TRSM_TEAM_EXTERN(float)
#undef CBK_TRSM_TEAM_EXTERN
} /* namespace trsmteam */

bool trsm_team_supports(bool left, bool up, bool tran, int m)
{
    return left && up == tran && (m == 16 || m == 24 || m == 32);
}

template <class T>
hipError_t launch_trsm_team(bool up, bool tran, bool unit, int m, int n, T alpha, Operand<T> A,
                            Operand<T> B, int V, int nm)
{
    (void)up;
    trsmteam::TrsmArgs<T> args;
    args.a = A.p;
    /* op(A): the transposed view when tran (src/cbk_common.hpp:156-177) */
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (tran) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.a_si = si;
    args.a_sj = sj;
    args.a_gs = A.gstride;
    args.b = B.p;
    args.b_si = B.row ? B.ld : 1;
    args.b_sj = B.row ? 1 : B.ld;
    args.b_gs = B.gstride;
    args.V = V;
    args.nmat = (long)((nm + V - 1) / V) * V;
    args.nrhs = n;
    args.alpha = alpha;
    args.unit = unit;
    switch (m) {
    case 16: return trsmteam::launch_n<T, 16>(args);
    case 24: return trsmteam::launch_n<T, 24>(args);
    case 32: return trsmteam::launch_n<T, 32>(args);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_trsm_team<double>(bool, bool, bool, int, int, double, Operand<double>,
                                             Operand<double>, int, int);
template hipError_t launch_trsm_team<float>(bool, bool, bool, int, int, float, Operand<float>,
                                            Operand<float>, int, int);

} /* namespace cbk */
