/* qr_reg_d.hip -- FP64 register geqrf / overdetermined gels, and the shapes the register QR kernels serve (kernels in qr_reg_impl.h) */

#include "qr_reg_impl.h"

namespace cbk {

/* F's p x q (p >= q) fits a bucket held in a lane's registers without spills */
bool qr_reg_supports(int p, int q, bool fp64)
{
    if (q < 1 || p < q) return false;
    return (p <= 16 && q <= 8) || (p <= 32 && q <= (fp64 ? 4 : 8));
}

template hipError_t launch_qr_reg<double>(int, int, int, bool, Operand<double>, double *,
                                          Operand<double>, int, int, int *);

} /* namespace cbk */
