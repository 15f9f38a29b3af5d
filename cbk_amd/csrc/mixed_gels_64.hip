/* the mixed-precision least-squares team kernel, buckets of 64 rows */
#include "mixed_gels_impl.h"

namespace cbk {
namespace mixedgels {
template hipError_t launch_pq<64, 8, false>(const MixedGelsArgs &);
template hipError_t launch_pq<64, 16, false>(const MixedGelsArgs &);
}
} /* namespace cbk */
