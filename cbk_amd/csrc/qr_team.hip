/* qr_team.hip -- host launcher of the QR / least-squares team kernel
 * (qr_team_impl.h); one shape per translation unit, qr_team_<p><m>x<n>.hip. */

#include "qr_team_impl.h"

namespace cbk {

namespace qrteam {
#define CBK_QR_TEAM_EXTERN(T)                                                            \
    extern template hipError_t launch_pq<T, 16, 8, 2>(const QrArgs<T> &);                 \
    extern template hipError_t launch_pq<T, 32, 16, 2>(const QrArgs<T> &);
CBK_QR_TEAM_EXTERN(double)
CBK_QR_TEAM_EXTERN(float)
#undef CBK_QR_TEAM_EXTERN
/* 64 x 32 in FP32 only: in FP64 the 64 x 32 panel does not fit a team's registers */
extern template hipError_t launch_pq<float, 64, 32, 2>(const QrArgs<float> &);
} /* namespace qrteam */

bool qr_wave_supports(int p, int q, int nrhs) { return p <= 64 && nrhs >= 1 && q + nrhs <= 64; }

bool qr_team_supports(int p, int q, int nrhs, bool fp64)
{
    return nrhs <= 2 && ((p == 16 && q == 8) || (p == 32 && q == 16) ||
                         (!fp64 && p == 64 && q == 32));
}

template <class T>
hipError_t launch_qr_team(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info)
{
    qrteam::QrArgs<T> args;
    args.a = A.p;
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (transposed) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.a_si = si;
    args.a_sj = sj;
    args.a_gs = A.gstride;
    args.tau = tau;
    args.b = B.p;
    args.b_si = B.row ? B.ld : 1;
    args.b_sj = B.row ? 1 : B.ld;
    args.b_gs = B.gstride;
    args.V = V;
    args.nmat = (long)((nm + V - 1) / V) * V;
    args.nm = nm;
    args.nrhs = nrhs;
    args.info = info;
    if (p == 16 && q == 8) return qrteam::launch_pq<T, 16, 8, 2>(args);
    if (p == 32 && q == 16) return qrteam::launch_pq<T, 32, 16, 2>(args);
    if constexpr (sizeof(T) == 4) {
        if (p == 64 && q == 32) return qrteam::launch_pq<T, 64, 32, 2>(args);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_qr_team<double>(int, int, int, bool, Operand<double>, double *,
                                           Operand<double>, int, int, int *);
template hipError_t launch_qr_team<float>(int, int, int, bool, Operand<float>, float *,
                                          Operand<float>, int, int, int *);

} /* namespace cbk */
