/* the QR / least-squares team kernel for 16 x 8 matrices, double */
#include "qr_team_impl.h"

namespace cbk {
namespace qrteam {
template hipError_t launch_pq<double, 16, 8, 2>(const QrArgs<double> &);
}
} /* namespace cbk */
