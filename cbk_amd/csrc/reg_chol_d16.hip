/* reg_chol_d16.hip -- register Cholesky instances, FP64, orders 13 to 16 (kernel in reg_chol_impl.h) */

#include "reg_chol_impl.h"

namespace cbk {

template hipError_t launch_reg_chol_n<double, 16, false>(bool, bool, int, int, Operand<double>,
                                                   Operand<double>, int, int, int *);

} /* namespace cbk */
