/* orgqr_reg_d.hip -- FP64 register orgqr (kernels in qr_reg_impl.h) */

#include "qr_reg_impl.h"

namespace cbk {

template hipError_t launch_orgqr_reg<double>(int, int, int, Operand<double>, const double *, int,
                                             int);

} /* namespace cbk */
