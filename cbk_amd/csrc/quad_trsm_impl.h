/* quad_trsm_impl.h -- B := alpha op(A)^-1 B (side L) for the forward sweeps
 * (op(A) lower: uplo L, no transpose, or uplo U, transposed) with one matrix
 * per quad of lanes (quad_common.h), 16 matrices per wavefront.
 *
 * Lane t of a quad holds rows t + 4k of op(A)'s strict lower triangle (its
 * diagonal apart) and the same rows of B, NRB right-hand sides at a time. Per
 * pivot l (ascending), after b := alpha b:
 *     x_l = b_l / L(l,l)            (non-unit; the owner lane only)
 *     b_i -= L(i,l) x_l,  i > l     (x_l from its owner by one DPP move)
 * Every element so receives the reference's operations in its order -- the
 * multiply by alpha, the fma's over l ascending, then the divide (the row-dot
 * form, src/cbk_trsm_compact.hpp:73-104, and the axpy form of its
 * single-column tail, :49-67, which is this same sweep) -- and the rows of
 * the pivot's own slot at or above it keep their value untouched (an fma
 * with a zero coefficient would turn a -0 into +0), so results are
 * bit-identical to the reference's, signed zeros included. The next block of
 * B is loaded before the current block's sweep.
 */
#ifndef CBK_QUAD_TRSM_IMPL_H
#define CBK_QUAD_TRSM_IMPL_H

#include "cbk_internal.h"
#include "quad_common.h"

namespace cbk {
namespace quadtrsm {

using namespace quad;
using team::bload;
using team::bstore;
using team::fmaT;
using team::group_range;
using team::group_rsrc;

template <class T> struct QuadTrsmArgs {
    const T *a; /* op(A): element (i, c) at (i*a_si + c*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long nmat;
    int nrhs;
    T alpha;
    bool unit;
    int m; /* the order, m <= N: rows and columns past it read as the identity */
};

template <class T, int N, int NRB, bool PREFETCH, bool UNIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
quad_trsm_kernel(QuadTrsmArgs<T> p)
{
    constexpr int K = N / 4;
    constexpr unsigned SZ = sizeof(T);
    QuadPlace pl;
    if (!quad_place(p.nmat, p.V, pl)) return;
    const int t = pl.t;
    const unsigned VS = (unsigned)p.V * SZ;
    const __amdgpu_buffer_rsrc_t ra =
        group_rsrc(p.a + pl.w0 * p.a_gs, group_range(pl.groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb =
        group_rsrc(p.b + pl.w0 * p.b_gs, group_range(pl.groups_left, p.b_gs, SZ));
    const unsigned aslot = pl.slot(p.a_gs, p.V, SZ), bslot = pl.slot(p.b_gs, p.V, SZ);
    const unsigned a_col = (unsigned)p.a_sj * VS, b_col = (unsigned)p.b_sj * VS;
    unsigned a_row[K], b_row[K];
    sfor<0, K>([&](auto k) {
        const int i = t + 4 * k;
        a_row[k] = (aslot == kOOB || i >= p.m) ? kOOB : aslot + (unsigned)(i * p.a_si) * VS;
        b_row[k] = (bslot == kOOB || i >= p.m) ? kOOB : bslot + (unsigned)(i * p.b_si) * VS;
    });

    /* op(A) strictly below the diagonal (zeros elsewhere), its diagonal apart;
     * an order m below N runs as diag(op(A), I) with B's rows past m zero and
     * never stored: the padded pivots reach only padded rows, so the rows
     * below m get exactly the order-m operations */
    T a[K][N], dg[K];
    sfor<0, K>([&](auto k) {
        const int i = t + 4 * k;
        sfor<0, 4 * k + 4>([&](auto c) {
            a[k][c] = bload(T(), ra, (c < p.m && (c < 4 * k || c < i)) ? a_row[k] : kOOB, c * a_col);
        });
        if constexpr (!UNIT) {
            const unsigned doff = (a_row[k] == kOOB) ? kOOB : a_row[k] + (unsigned)i * a_col;
            dg[k] = (i >= p.m) ? T(1) : bload(T(), ra, doff, 0);
        }
    });

    auto load_block = [&](int q0, T(&bb)[K][NRB]) {
        sfor<0, NRB>([&](auto q) {
            const unsigned col = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB;
            sfor<0, K>([&](auto k) { bb[k][q] = bload(T(), rb, b_row[k], col); });
        });
    };
    T bn[K][NRB];
    if (PREFETCH && p.nrhs > 0) load_block(0, bn);
    for (int q0 = 0; q0 < p.nrhs; q0 += NRB) {
        T b[K][NRB];
        if (!PREFETCH) load_block(q0, bn);
        sfor<0, K>([&](auto k) { sfor<0, NRB>([&](auto q) { b[k][q] = bn[k][q] * p.alpha; }); });
        if (PREFETCH && q0 + NRB < p.nrhs) load_block(q0 + NRB, bn);
        sfor<0, N>([&](auto lc) {
            constexpr int l = lc;
            constexpr int kl = l / 4, tl = l % 4;
            if constexpr (!UNIT) {
                sfor<0, NRB>([&](auto q) {
                    b[kl][q] = (t == tl) ? b[kl][q] / dg[kl] : b[kl][q];
                });
            }
            sfor<0, NRB>([&](auto q) {
                const T x = qbc<tl>(b[kl][q]);
                /* rows of the pivot's slot at or above it keep their value */
                const T own = fmaT(-a[kl][l], x, b[kl][q]);
                sfor<kl + 1, K>([&](auto k) { b[k][q] = fmaT(-a[k][l], x, b[k][q]); });
                b[kl][q] = (t > tl) ? own : b[kl][q];
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        sfor<0, NRB>([&](auto q) {
            const unsigned col = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB;
            sfor<0, K>([&](auto k) { bstore(b[k][q], rb, b_row[k], col); });
        });
    }
}

/* NRB = 2 right-hand sides per sweep at two wavefronts per SIMD; the next
 * block's loads issued early where the registers allow (FP32; FP64 unit) */
template <class T, int N> hipError_t launch_n(const QuadTrsmArgs<T> &args)
{
    const long waves = (args.nmat + kMatsPerWave - 1) / kMatsPerWave;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    constexpr bool PF_NONUNIT = sizeof(T) == 4;
    if (args.unit)
        hipLaunchKernelGGL((quad_trsm_kernel<T, N, 2, true, true>), grid, block, 0, current_stream(), args);
    else
        hipLaunchKernelGGL((quad_trsm_kernel<T, N, 2, PF_NONUNIT, false>), grid, block, 0, current_stream(), args);
    return hipGetLastError();
}

} /* namespace quadtrsm */
} /* namespace cbk */

#endif /* CBK_QUAD_TRSM_IMPL_H */
