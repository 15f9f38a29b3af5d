/* the potrf/posv/sytrfnp/sysvnp team kernels of order 24, double */
#include "sym_team_impl.h"

namespace cbk {
namespace symteam {
template hipError_t launch_n<double, 24>(bool, const SymArgs<double> &);
}
} /* namespace cbk */
