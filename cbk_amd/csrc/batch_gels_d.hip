/* the uneven-batch gels kernels, double */
#include "batch_gels_impl.h"

namespace cbk {
template hipError_t launch_gels_batch<double>(int, const BatchArgs<double> &, unsigned, hipStream_t);
template hipError_t gels_batch_occupancy<double>(int, int *);
} /* namespace cbk */
