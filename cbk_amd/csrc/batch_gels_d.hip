/* the uneven-batch gels kernels, double, their compact-operand launchers,
 * and the FP64 path of dsgels_compact */
#include "batch_gels_impl.h"

namespace cbk {
template hipError_t launch_gels_batch<double>(int, const BatchArgs<double> &, unsigned, hipStream_t);
template hipError_t gels_batch_occupancy<double>(int, int *);
template hipError_t launch_qr_wave<double>(int, int, int, bool, Operand<double>, double *,
                                           Operand<double>, int, int, int *);

hipError_t launch_qr_wave_masked(int p, int q, int nrhs, Operand<double> A, Operand<double> B,
                                 Operand<double> X, int V, int nm, const int *mask, int *info)
{
    BatchArgs<double> args;
    args.a = A.p;
    args.b = B.p;
    args.info = info;
    args.desc = nullptr;
    args.counter = nullptr;
    args.count = nm;
    args.tau = nullptr;
    args.tran = false;
    args.a_rs = (A.row ? A.ld : 1) * V;
    args.a_cs = (A.row ? 1 : A.ld) * V;
    args.b_rs = (B.row ? B.ld : 1) * V;
    args.b_cs = (B.row ? 1 : B.ld) * V;
    args.ninfo = nm;
    args.cm = p;
    args.cn = q;
    args.cnrhs = nrhs;
    args.cV = V;
    args.a_gs = (long long)A.gstride;
    args.b_gs = (long long)B.gstride;
    args.ctau = nullptr;
    args.mask = mask;
    args.x = X.p;
    args.x_rs = (X.row ? X.ld : 1) * V;
    args.x_cs = (X.row ? 1 : X.ld) * V;
    args.x_gs = (long long)X.gstride;
    args.dcount = nullptr;
    /* persistent: one wavefront per 64-entry chunk of the mask */
    const unsigned blocks = (unsigned)(((nm + 63) / 64 + 3) / 4);
    if (p <= 32)
        hipLaunchKernelGGL((batch::gels_wave_kernel<double, 32>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((batch::gels_wave_kernel<double, 64>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    return hipGetLastError();
}

} /* namespace cbk */
