/* the fused mixed-precision posv / sysvnp quad kernel of order 24 */
#include "quad_mixed_impl.h"

namespace cbk {
namespace quadmixed {
template hipError_t launch_n<24, nrb_of(24)>(bool, const QuadMixedArgs &);
}
} /* namespace cbk */
