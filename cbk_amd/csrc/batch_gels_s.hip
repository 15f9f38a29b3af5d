/* the uneven-batch gels kernels, float, and their compact-operand launcher */
#include "batch_gels_impl.h"

namespace cbk {
template hipError_t launch_gels_batch<float>(int, const BatchArgs<float> &, unsigned, hipStream_t);
template hipError_t gels_batch_occupancy<float>(int, int *);
template hipError_t launch_qr_wave<float>(int, int, int, bool, Operand<float>, float *, Operand<float>,
                                       int, int, int *);
} /* namespace cbk */
