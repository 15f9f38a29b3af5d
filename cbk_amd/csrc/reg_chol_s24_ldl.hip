/* reg_chol_s24_ldl.hip -- register LDL^T instances, FP32, orders 17 to 24 (kernel in reg_chol_impl.h) */

#include "reg_chol_impl.h"

namespace cbk {

template hipError_t launch_reg_chol_n<float, 24, true>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);

} /* namespace cbk */
