/* quad_sym.hip -- host launcher of the potrf/posv/sytrfnp/sysvnp quad kernel
 * (quad_sym_impl.h); one order per translation unit, quad_sym_<p><n>.hip. */

#include "quad_sym_impl.h"

namespace cbk {

namespace quadsym {
extern template hipError_t launch_n<double, 24>(bool, const QuadSymArgs<double> &);
extern template hipError_t launch_n<double, 32>(bool, const QuadSymArgs<double> &);
extern template hipError_t launch_n<float, 24>(bool, const QuadSymArgs<float> &);
extern template hipError_t launch_n<float, 32>(bool, const QuadSymArgs<float> &);
} /* namespace quadsym */

/* the instances of order 24 and 32; an order between 16 and 32 runs on the
 * next one as diag(A, I) */
bool quad_sym_supports(int n, bool fp64)
{
    (void)fp64;
    return n > 16 && n <= 32;
}

template <class T>
hipError_t launch_quad_sym(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                           Operand<T> B, int V, int nm, int *info, const int *only)
{
    quadsym::QuadSymArgs<T> args;
    args.only = only;
    args.a = A.p;
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
    args.n = n;
    if (n > 16 && n <= 24) return quadsym::launch_n<T, 24>(ldl, args);
    if (n > 24 && n <= 32) return quadsym::launch_n<T, 32>(ldl, args);
    return hipErrorInvalidValue;
}

template hipError_t launch_quad_sym<double>(bool, bool, bool, int, int, Operand<double>,
                                            Operand<double>, int, int, int *, const int *);
template hipError_t launch_quad_sym<float>(bool, bool, bool, int, int, Operand<float>,
                                           Operand<float>, int, int, int *, const int *);

} /* namespace cbk */
