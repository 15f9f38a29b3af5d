/* the quad potrf/posv/sytrfnp/sysvnp kernel, double, order 24 */
#include "quad_sym_impl.h"

namespace cbk {
namespace quadsym {
template hipError_t launch_n<double, 24>(bool, const QuadSymArgs<double> &);
} /* namespace quadsym */
} /* namespace cbk */
