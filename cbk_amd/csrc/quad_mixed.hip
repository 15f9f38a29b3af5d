/* quad_mixed.hip -- host launcher of the fused mixed-precision posv / sysvnp
 * quad kernel (quad_mixed_impl.h) and the masked FP64 quad pass after it;
 * one order per translation unit, quad_mixed_<n>.hip. */

#include "quad_mixed_impl.h"

namespace cbk {

namespace quadmixed {
extern template hipError_t launch_n<24, nrb_of(24)>(bool, const QuadMixedArgs &);
} /* namespace quadmixed */

/* orders 17..24: the kernel of order 24, an order below it padded with the
 * identity */
bool quad_mixed_supports(int n) { return n > 16 && n <= 24; }

hipError_t launch_mixed_posv_quad(bool ldl, bool up, int n, int nrhs, Operand<double> A,
                                  Operand<double> B, Operand<double> X, int V, int nm,
                                  int itermax, int *iter, int *info)
{
    quadmixed::QuadMixedArgs p;
    p.a = A.p;
    /* the lower view: uplo 'U' reads A transposed (src/cbk_common.hpp:201-206) */
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
    if (!quad_mixed_supports(n)) return hipErrorInvalidValue;
    hipError_t e = quadmixed::launch_n<24, quadmixed::nrb_of(24)>(ldl, p);
    if (e != hipSuccess) return e;
    /* the FP64 posv / sysvnp on the matrices marked iter < 0, in place on A and
     * on X (which holds B for them): A receives the factor, info its code */
    return launch_quad_sym<double>(ldl, true, up, n, nrhs, A, X, V, nm, info, iter);
}

} /* namespace cbk */
