/* the mixed-precision least-squares team kernel, buckets of 16 rows */
#include "mixed_gels_impl.h"

namespace cbk {
namespace mixedgels {
template hipError_t launch_pq<16, 8, false>(const MixedGelsArgs &);
template hipError_t launch_pq<16, 8, true>(const MixedGelsArgs &);
template hipError_t launch_pq<16, 16, false>(const MixedGelsArgs &);
}
} /* namespace cbk */
