/* quad_trsm.hip -- host launcher and instances of the forward-sweep trsm quad
 * kernel (quad_trsm_impl.h): FP64 order 24, FP32 order 32. */
#include "quad_trsm_impl.h"

namespace cbk {

/* the FP64 instance of order 24 and the FP32 one of order 32; the orders
 * above the register kernel's (FP64 16, FP32 24) run on them as
 * diag(op(A), I) */
bool quad_trsm_supports(bool left, bool up, bool tran, int m, bool fp64)
{
    return left && up == tran && (fp64 ? (m > 16 && m <= 24) : (m > 24 && m <= 32));
}

template <class T>
hipError_t launch_quad_trsm(bool tran, bool unit, int m, int n, T alpha, Operand<T> A,
                            Operand<T> B, int V, int nm)
{
    quadtrsm::QuadTrsmArgs<T> args;
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
    args.m = m;
    if constexpr (sizeof(T) == 8) {
        if (m > 16 && m <= 24) return quadtrsm::launch_n<T, 24>(args);
    }
    else {
        if (m > 24 && m <= 32) return quadtrsm::launch_n<T, 32>(args);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_quad_trsm<double>(bool, bool, int, int, double, Operand<double>,
                                             Operand<double>, int, int);
template hipError_t launch_quad_trsm<float>(bool, bool, int, int, float, Operand<float>,
                                            Operand<float>, int, int);

} /* namespace cbk */
