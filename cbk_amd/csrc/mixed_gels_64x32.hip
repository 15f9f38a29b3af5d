/* the mixed-precision least-squares team kernel, the 64 x 32 bucket */
#include "mixed_gels_impl.h"

namespace cbk {
namespace mixedgels {
template hipError_t launch_pq<64, 32, true>(const MixedGelsArgs &);
template hipError_t launch_pq<64, 32, false>(const MixedGelsArgs &);
}
} /* namespace cbk */
