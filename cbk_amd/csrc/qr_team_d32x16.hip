/* the QR / least-squares team kernel for 32 x 16 matrices, double */
#include "qr_team_impl.h"

namespace cbk {
namespace qrteam {
template hipError_t launch_pq<double, 32, 16, 2>(const QrArgs<double> &);
}
} /* namespace cbk */
