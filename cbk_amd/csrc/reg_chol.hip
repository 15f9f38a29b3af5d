/* reg_chol.hip -- dispatch of the register Cholesky / LDL^T lane kernel
 * (reg_chol_impl.h) to its order bounds 8, 12, 16 (and 24 in FP32); the
 * instances live in reg_chol_small.hip, reg_chol_d16*.hip, reg_chol_s24*.hip. */

#include "cbk_internal.h"

#include <hip/hip_runtime.h>

namespace cbk {

template <class T, int NMAX, bool LDL>
hipError_t launch_reg_chol_n(bool factor, bool up, int n, int nrhs, Operand<T> A, Operand<T> B,
                             int V, int nm, int *info);

bool reg_chol_supports(int n, bool fp64) { return n >= 1 && n <= (fp64 ? 16 : 24); }

template <class T>
hipError_t launch_reg_chol(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                           Operand<T> B, int V, int nm, int *info)
{
    if (n <= 8)
        return ldl ? launch_reg_chol_n<T, 8, true>(factor, up, n, nrhs, A, B, V, nm, info)
                   : launch_reg_chol_n<T, 8, false>(factor, up, n, nrhs, A, B, V, nm, info);
    if (n <= 12)
        return ldl ? launch_reg_chol_n<T, 12, true>(factor, up, n, nrhs, A, B, V, nm, info)
                   : launch_reg_chol_n<T, 12, false>(factor, up, n, nrhs, A, B, V, nm, info);
    if (n <= 16)
        return ldl ? launch_reg_chol_n<T, 16, true>(factor, up, n, nrhs, A, B, V, nm, info)
                   : launch_reg_chol_n<T, 16, false>(factor, up, n, nrhs, A, B, V, nm, info);
    if constexpr (sizeof(T) == 4) {
        if (n <= 24)
            return ldl ? launch_reg_chol_n<T, 24, true>(factor, up, n, nrhs, A, B, V, nm, info)
                       : launch_reg_chol_n<T, 24, false>(factor, up, n, nrhs, A, B, V, nm, info);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_reg_chol<double>(bool, bool, bool, int, int, Operand<double>,
                                            Operand<double>, int, int, int *);
template hipError_t launch_reg_chol<float>(bool, bool, bool, int, int, Operand<float>,
                                           Operand<float>, int, int, int *);

} /* namespace cbk */
