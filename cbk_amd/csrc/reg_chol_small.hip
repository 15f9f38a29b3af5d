/* reg_chol_small.hip -- register Cholesky / LDL^T instances for orders up to 12 (FP64) and 16 (FP32) (kernel in reg_chol_impl.h) */

#include "reg_chol_impl.h"

namespace cbk {

template hipError_t launch_reg_chol_n<double, 8, false>(bool, bool, int, int, Operand<double>,
                                                   Operand<double>, int, int, int *);
template hipError_t launch_reg_chol_n<double, 8, true>(bool, bool, int, int, Operand<double>,
                                                   Operand<double>, int, int, int *);
template hipError_t launch_reg_chol_n<double, 12, false>(bool, bool, int, int, Operand<double>,
                                                   Operand<double>, int, int, int *);
template hipError_t launch_reg_chol_n<double, 12, true>(bool, bool, int, int, Operand<double>,
                                                   Operand<double>, int, int, int *);
template hipError_t launch_reg_chol_n<float, 8, false>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);
template hipError_t launch_reg_chol_n<float, 8, true>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);
template hipError_t launch_reg_chol_n<float, 12, false>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);
template hipError_t launch_reg_chol_n<float, 12, true>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);
template hipError_t launch_reg_chol_n<float, 16, false>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);
template hipError_t launch_reg_chol_n<float, 16, true>(bool, bool, int, int, Operand<float>,
                                                   Operand<float>, int, int, int *);

} /* namespace cbk */
