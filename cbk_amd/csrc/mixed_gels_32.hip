/* the mixed-precision least-squares team kernel, buckets of 32 rows */
#include "mixed_gels_impl.h"

namespace cbk {
namespace mixedgels {
template hipError_t launch_pq<32, 8, false>(const MixedGelsArgs &);
template hipError_t launch_pq<32, 16, false>(const MixedGelsArgs &);
template hipError_t launch_pq<32, 16, true>(const MixedGelsArgs &);
template hipError_t launch_pq<32, 32, false>(const MixedGelsArgs &);
}
} /* namespace cbk */
