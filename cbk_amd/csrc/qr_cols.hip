/* qr_cols.hip -- Householder QR (geqrf) and the overdetermined least-squares
 * solve (gels) with one p x q matrix per 16-lane team and the COLUMNS of
 * [F | B] in the lanes: lane c holds column c (c < q + nrhs <= 16), all p <= P
 * rows of it in registers. F = A, or A^T for gels with trans and m < n.
 *
 * Per reflector kk (the reference's geqr2 and larf,
 * src/cbk_geqrf_compact.hpp:74-131, src/cbk_ormqr_compact.hpp:68-105):
 *   tail = sum_{i>kk} F(i,kk)^2 in lane kk, rows ascending;
 *   larfg: norm = sqrt(x0^2 + tail), beta = x0 >= 0 ? -norm : norm,
 *          tau = (beta - x0)/beta, body scale 1/(x0 - beta), R(kk,kk) = beta
 *          (tau = 0 and the column unchanged when tail = 0);
 *   every later column c: w = F(kk,c) + sum_{i>kk} v_i F(i,c), rows ascending,
 *          w *= tau, F(kk,c) -= w, F(i,c) -= v_i w;
 * v_i lives in lane kk and reaches the other lanes as the DPP row broadcast
 * source of the fma itself. Every sum is a chain of fmas in the reference's
 * row order (the row-oriented team kernel, qr_team_impl.h, sums with a
 * butterfly instead), and there are no cross-lane reductions at all: per
 * reflector about 4 (p - kk) VALU instructions for the whole wavefront.
 * Then R X = (Q^T B)(0:q) column-oriented in the B lanes (x_l = b_l / R(l,l),
 * b_i -= R(i,l) x_l, l descending), as the batch team kernel does. Results
 * agree with the reference to rounding.
 *
 * A wavefront holds 4 consecutive matrices of one group (V >= 4); a block of
 * 4 wavefronts covers the 16 slots of a 128-byte line of every element.
 */

#include "qr_team_impl.h"

namespace cbk {
namespace qrcols {

using namespace team;
using qrteam::QrArgs;

constexpr int kTeamsPerBlock = 16;

/* larfg_pack (src/cbk_geqrf_compact.hpp:74-90): no rescaling, direct sqrt */
template <class T> struct Larfg {
    T tau, inv, rdiag;
};
template <class T> __device__ __forceinline__ Larfg<T> larfg(T x0, T tail)
{
    Larfg<T> h;
    const T norm = sqrtT(fmaT(x0, x0, tail));
    const T beta = (x0 >= T(0)) ? -norm : norm;
    const bool has = tail > T(0);
    h.tau = has ? (beta - x0) / beta : T(0);
    h.inv = has ? T(1) / (x0 - beta) : T(0);
    h.rdiag = has ? beta : x0;
    return h;
}

template <class T, int P>
__global__ __launch_bounds__(256) void qr_cols_kernel(QrArgs<T> p, int pp, int q)
{
    const int c = threadIdx.x & (TS - 1);
    const int team = threadIdx.x / TS;
    const long mat0 = (long)blockIdx.x * kTeamsPerBlock + 4 * wave_in_block(); /* wave-uniform */
    if (mat0 >= p.nmat) return;
    const long mat = mat0 + (team & 3);
    const long g = mat0 / p.V;
    const size_t v = (size_t)(mat0 % p.V) + (team & 3);
    const size_t V = (size_t)p.V;
    const bool solve = p.b != nullptr;
    const int w = q + (solve ? p.nrhs : 0);
    const bool isA = c < q, isB = c >= q && c < w;
    T *col = isA ? p.a + g * p.a_gs + (size_t)c * p.a_sj * V + v
                 : (isB ? p.b + g * p.b_gs + (size_t)(c - q) * p.b_sj * V + v : p.a);
    const size_t rs = (size_t)(isA ? p.a_si : p.b_si) * V;

    T f[P];
    sfor<0, P>([&](auto i) { f[i] = ((isA || isB) && i < pp) ? col[i * rs] : T(0); });

    T tau = T(0); /* lane kk keeps tau(kk) */
    sfor<0, TS>([&](auto kc) {
        constexpr int kk = kc;
        if (kk < q) {
            T tl = T(0);
            sfor<kk + 1, P>([&](auto i) { tl = fmaT(f[i], f[i], tl); });
            const Larfg<T> h = larfg(bcast<kk>(f[kk]), bcast<kk>(tl));
            if (c == kk) {
                sfor<kk + 1, P>([&](auto i) { f[i] = f[i] * h.inv; });
                f[kk] = h.rdiag;
                tau = h.tau;
            }
            /* the broadcast sources f[kk+1..P-1] of lane kk are final here */
            dpp_ready_from<kk + 1>(f);
            T wv = f[kk];
            sfor<kk + 1, P>([&](auto i) { fma_bcr<kk>(wv, f[i], f[i]); });
            wv = (c > kk) ? wv * h.tau : T(0);
            f[kk] = f[kk] - wv;
            sfor<kk + 1, P>([&](auto i) { fnma_bcr<kk>(f[i], f[i], wv); });
        }
        __builtin_amdgcn_sched_barrier(0);
    });

    /* A := (R, reflectors), tau */
    if (isA) {
        sfor<0, P>([&](auto i) {
            if (i < pp) col[i * rs] = f[i];
        });
        p.tau[(size_t)g * q * V + (size_t)c * V + v] = tau;
    }
    if (!solve) return;

    /* info: the first exact zero on R's diagonal (LAPACK dtrtrs's test) */
    int bad = 0;
    sfor<0, TS>([&](auto lc) {
        constexpr int l = lc;
        const T dl = bcast<l>(f[l]);
        if (l < q && bad == 0 && dl == T(0)) bad = l + 1;
    });
    if (p.info && mat < p.nm && c == 0) p.info[mat] = bad;
    if (!(p.info && bad)) {
        sfor_down<0, TS>([&](auto lc) {
            constexpr int l = lc;
            if (l < q) {
                const T dl = bcast<l>(f[l]);
                const T x = isB ? f[l] / dl : T(0);
                f[l] = isB ? x : f[l];
                sfor<0, l>([&](auto i) { fnma_bc<l>(f[i], f[i], x); });
            }
        });
    }
    if (isB) {
        sfor<0, P>([&](auto i) {
            if (i < pp) col[i * rs] = f[i];
        });
    }
}

template <class T, int P> hipError_t launch_p(const QrArgs<T> &args, int pp, int q)
{
    const unsigned blocks = (unsigned)((args.nmat + kTeamsPerBlock - 1) / kTeamsPerBlock);
    hipLaunchKernelGGL((qr_cols_kernel<T, P>), dim3(blocks), dim3(256), 0, current_stream(), args,
                       pp, q);
    return hipGetLastError();
}

} /* namespace qrcols */

bool qr_cols_supports(int p, int q, int nrhs) { return p <= 32 && q >= 1 && q + nrhs <= 16; }

template <class T>
hipError_t launch_qr_cols(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
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
    if (p <= 16) return qrcols::launch_p<T, 16>(args, p, q);
    return qrcols::launch_p<T, 32>(args, p, q);
}

template hipError_t launch_qr_cols<double>(int, int, int, bool, Operand<double>, double *,
                                           Operand<double>, int, int, int *);
template hipError_t launch_qr_cols<float>(int, int, int, bool, Operand<float>, float *,
                                          Operand<float>, int, int, int *);

} /* namespace cbk */
