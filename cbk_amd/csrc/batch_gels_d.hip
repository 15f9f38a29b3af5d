/* the uneven-batch gels kernels, double, and their compact-operand launcher */
#include "batch_gels_impl.h"

namespace cbk {
template hipError_t launch_gels_batch<double>(int, const BatchArgs<double> &, unsigned, hipStream_t);
template hipError_t gels_batch_occupancy<double>(int, int *);
template hipError_t launch_qr_wave<double>(int, int, int, bool, Operand<double>, double *, Operand<double>,
                                       int, int, int *);
} /* namespace cbk */
