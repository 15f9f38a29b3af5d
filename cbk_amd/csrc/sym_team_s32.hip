/* the potrf/posv/sytrfnp/sysvnp team kernels of order 32, float */
#include "sym_team_impl.h"

namespace cbk {
namespace symteam {
template hipError_t launch_n<float, 32>(bool, const SymArgs<float> &);
}
} /* namespace cbk */
