/* the quad potrf/posv/sytrfnp/sysvnp kernel, float, order 24 */
#include "quad_sym_impl.h"

namespace cbk {
namespace quadsym {
template hipError_t launch_n<float, 24>(bool, const QuadSymArgs<float> &);
} /* namespace quadsym */
} /* namespace cbk */
