/* the forward-sweep trsm team kernels, double */
#include "trsm_team_impl.h"

namespace cbk {
namespace trsmteam {
template hipError_t launch_n<double, 16>(const TrsmArgs<double> &);
template hipError_t launch_n<double, 24>(const TrsmArgs<double> &);
template hipError_t launch_n<double, 32>(const TrsmArgs<double> &);
} /* namespace trsmteam */
} /* namespace cbk */
