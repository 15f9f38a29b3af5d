/* the uneven-batch gels kernels, float */
#include "batch_gels_impl.h"

namespace cbk {
template hipError_t launch_gels_batch<float>(int, const BatchArgs<float> &, unsigned, hipStream_t);
template hipError_t gels_batch_occupancy<float>(int, int *);
} /* namespace cbk */
