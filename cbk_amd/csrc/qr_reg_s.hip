/* qr_reg_s.hip -- FP32 register geqrf / overdetermined gels (kernels in qr_reg_impl.h) */

#include "qr_reg_impl.h"

namespace cbk {

template hipError_t launch_qr_reg<float>(int, int, int, bool, Operand<float>, float *,
                                         Operand<float>, int, int, int *);

} /* namespace cbk */
