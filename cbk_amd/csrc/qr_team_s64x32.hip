/* the QR / least-squares team kernel for 64 x 32 matrices, float */
#include "qr_team_impl.h"

namespace cbk {
namespace qrteam {
template hipError_t launch_pq<float, 64, 32, 2>(const QrArgs<float> &);
}
} /* namespace cbk */
