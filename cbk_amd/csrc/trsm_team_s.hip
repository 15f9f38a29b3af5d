/* the forward-sweep trsm team kernels, float */
#include "trsm_team_impl.h"

namespace cbk {
namespace trsmteam {
template hipError_t launch_n<float, 16>(const TrsmArgs<float> &);
template hipError_t launch_n<float, 24>(const TrsmArgs<float> &);
template hipError_t launch_n<float, 32>(const TrsmArgs<float> &);
} /* namespace trsmteam */
} /* namespace cbk */
