/* orgqr_reg_s.hip -- FP32 register orgqr (kernels in qr_reg_impl.h) */

#include "qr_reg_impl.h"

namespace cbk {

template hipError_t launch_orgqr_reg<float>(int, int, int, Operand<float>, const float *, int, int);

} /* namespace cbk */
