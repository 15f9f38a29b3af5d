/* sym_team.hip -- host launcher of the potrf/posv/sytrfnp/sysvnp team kernel
 * (sym_team_impl.h); one order per translation unit, sym_team_<p><n>.hip. */

#include "sym_team_impl.h"

namespace cbk {

/* instantiated in sym_team_<p><n>.hip, one order per translation unit */
namespace symteam {
#define CBK_SYM_TEAM_EXTERN(T)                                                           \
    extern template hipError_t launch_n<T, 16>(bool, const SymArgs<T> &);                 \
    extern template hipError_t launch_n<T, 24>(bool, const SymArgs<T> &);                 \
    extern template hipError_t launch_n<T, 32>(bool, const SymArgs<T> &);
CBK_SYM_TEAM_EXTERN(double)
CBK_SYM_TEAM_EXTERN(float)
#undef CBK_SYM_TEAM_EXTERN
} /* namespace symteam */

using symteam::SymArgs;

bool sym_team_supports(int n) { return n == 16 || n == 24 || n == 32; }

template <class T>
hipError_t launch_sym_team(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                           Operand<T> B, int V, int nm, int *info)
{
    SymArgs<T> args;
    args.a = A.p;
    args.store_factor = factor;
    /* the lower view: uplo 'U' reads A transposed (src/cbk_common.hpp:201-206) */
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (up) {
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
    args.nm = nm;
    args.nrhs = nrhs;
    args.info = info;
    args.factor = factor;
    switch (n) {
    case 16: return symteam::launch_n<T, 16>(ldl, args);
    case 24: return symteam::launch_n<T, 24>(ldl, args);
    case 32: return symteam::launch_n<T, 32>(ldl, args);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_sym_team<double>(bool, bool, bool, int, int, Operand<double>,
                                            Operand<double>, int, int, int *);
template hipError_t launch_sym_team<float>(bool, bool, bool, int, int, Operand<float>,
                                           Operand<float>, int, int, int *);

} /* namespace cbk */
