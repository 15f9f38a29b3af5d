/* mixed_team.hip -- host launcher of the fused mixed-precision posv / sysvnp
 * team kernel (mixed_team_impl.h); one order per translation unit,
 * mixed_team_<n>.hip. */

#include "mixed_team_impl.h"

namespace cbk {

namespace mixedteam {
extern template hipError_t launch_n<16>(bool, const MixedArgs &);
extern template hipError_t launch_n<24>(bool, const MixedArgs &);
extern template hipError_t launch_n<32>(bool, const MixedArgs &);
} /* namespace mixedteam */

/* orders 1..32: the kernels of orders 16, 24 and 32, any other order padded
 * with the identity up to the next of them */
bool mixed_team_supports(int n) { return n >= 1 && n <= 32; }

hipError_t launch_mixed_posv_team(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, int V, int nm,
                                  int itermax, int *iter, int *info)
{
    mixedteam::MixedArgs p;
    p.a = A.p;
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
    p.nmat = (long)((nm + V - 1) / V) * V;
    p.nm = nm;
    p.n = n;
    p.nrhs = nrhs;
    p.itermax = itermax;
    p.iter = iter;
    p.info = info;
    if (!mixed_team_supports(n)) return hipErrorInvalidValue;
    if (n <= 16) return mixedteam::launch_n<16>(ldl, p);
    if (n <= 24) return mixedteam::launch_n<24>(ldl, p);
    return mixedteam::launch_n<32>(ldl, p);
}

} /* namespace cbk */
