/* mixed_gels.hip -- host launcher of the mixed-precision least-squares team
 * kernel (mixed_gels_impl.h) */
#include "mixed_gels_impl.h"

namespace cbk {

bool mixed_gels_supports(int p, int q, int nrhs)
{
    return nrhs >= 1 && nrhs <= 8 &&
           ((p == 16 && q == 8) || (p == 32 && q == 16) || (p == 64 && q == 32));
}

hipError_t launch_mixed_gels_team(int p, int q, int nrhs, Operand<double> A, Operand<double> B,
                                  Operand<double> X, int V, int nm, int itermax, int *iter)
{
    mixedgels::MixedGelsArgs args;
    args.a = A.p;
    args.a_si = A.row ? A.ld : 1;
    args.a_sj = A.row ? 1 : A.ld;
    args.a_gs = A.gstride;
    args.b = B.p;
    args.b_si = B.row ? B.ld : 1;
    args.b_sj = B.row ? 1 : B.ld;
    args.b_gs = B.gstride;
    args.x = X.p;
    args.x_si = X.row ? X.ld : 1;
    args.x_sj = X.row ? 1 : X.ld;
    args.x_gs = X.gstride;
    args.V = V;
    args.nmat = (long)((nm + V - 1) / V) * V;
    args.nm = nm;
    args.nrhs = nrhs;
    args.itermax = itermax;
    args.iter = iter;
    if (p == 16 && q == 8) return mixedgels::launch_pq<16, 8>(args);
    if (p == 32 && q == 16) return mixedgels::launch_pq<32, 16>(args);
    if (p == 64 && q == 32) return mixedgels::launch_pq<64, 32>(args);
    return hipErrorInvalidValue;
}

} /* namespace cbk */
