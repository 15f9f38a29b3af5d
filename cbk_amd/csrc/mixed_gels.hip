/* mixed_gels.hip -- host launcher of the mixed-precision least-squares team
 * kernel (mixed_gels_impl.h): a matrix of m x n runs in the smallest
 * P x Q bucket that holds it, P in {16, 32, 64}, Q in {8, 16, 32}, Q <= P
 * (an instance of its own, with m and n folded, where the matrix fills the
 * bucket at 16 x 8, 32 x 16 and 64 x 32); the instances sit in
 * mixed_gels_<P>.hip and mixed_gels_64x32.hip */
#include "mixed_gels_impl.h"

namespace cbk {

namespace mixedgels {
extern template hipError_t launch_pq<16, 8, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<16, 16, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<32, 8, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<32, 16, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<32, 32, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<64, 8, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<64, 16, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<64, 32, false>(const MixedGelsArgs &);
extern template hipError_t launch_pq<16, 8, true>(const MixedGelsArgs &);
extern template hipError_t launch_pq<32, 16, true>(const MixedGelsArgs &);
extern template hipError_t launch_pq<64, 32, true>(const MixedGelsArgs &);
} /* namespace mixedgels */

bool mixed_gels_supports(int p, int q, int nrhs)
{
    return nrhs >= 1 && q >= 1 && q <= 32 && q <= p && p <= 64;
}

hipError_t launch_mixed_gels_team(int p, int q, int nrhs, Operand<double> A, Operand<double> B,
                                  Operand<double> X, int V, int nm, int itermax, int *iter)
{
    mixedgels::MixedGelsArgs args;
    args.m = p;
    args.n = q;
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
    const int P = p <= 16 ? 16 : p <= 32 ? 32 : 64;
    const int Q = q <= 8 ? 8 : q <= 16 ? 16 : 32;
    const bool exact = p == P && q == Q;
    switch (P * 100 + Q) {
    case 1608:
        return exact ? mixedgels::launch_pq<16, 8, true>(args)
                     : mixedgels::launch_pq<16, 8, false>(args);
    case 1616: return mixedgels::launch_pq<16, 16, false>(args);
    case 3208: return mixedgels::launch_pq<32, 8, false>(args);
    case 3216:
        return exact ? mixedgels::launch_pq<32, 16, true>(args)
                     : mixedgels::launch_pq<32, 16, false>(args);
    case 3232: return mixedgels::launch_pq<32, 32, false>(args);
    case 6408: return mixedgels::launch_pq<64, 8, false>(args);
    case 6416: return mixedgels::launch_pq<64, 16, false>(args);
    case 6432:
        return exact ? mixedgels::launch_pq<64, 32, true>(args)
                     : mixedgels::launch_pq<64, 32, false>(args);
    default: return hipErrorInvalidValue;
    }
}

} /* namespace cbk */
