/* the fused mixed-precision posv team kernel of order 32 */
#include "mixed_team_impl.h"

namespace cbk {
namespace mixedteam {
template hipError_t launch_n<32>(bool, const MixedArgs &);
}
} /* namespace cbk */
