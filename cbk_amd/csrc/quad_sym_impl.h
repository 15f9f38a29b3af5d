/* quad_sym_impl.h -- potrf / posv / potrs and sytrfnp / sysvnp / sytrsnp with
 * one matrix per quad of lanes (quad_common.h, quad_chol.h), 16 matrices per
 * wavefront, the lower triangle of each in registers.
 *
 * A wavefront loads the lower triangles of its 16 matrices (four full
 * 128-byte lines per load instruction at V >= 16 in FP64), factors them
 * (bit-identical to the reference), stores the factor, then solves B in
 * blocks of NRB right-hand sides. Rows past a lane's diagonal and masked
 * stores use an out-of-range offset (zero load / dropped store), so no load
 * or store branches. With info, a matrix whose factorization failed keeps its
 * B (LAPACK dposv skips dpotrs). An order n below the instance's N runs as
 * diag(A, I): rows and columns past n read as zero with 1 on the diagonal and
 * are never stored, B's rows past n as zero; every element of the leading
 * n x n block gets exactly the operations of the order-n factorization, so
 * the factor stays bit-identical to the reference's.
 */
#ifndef CBK_QUAD_SYM_IMPL_H
#define CBK_QUAD_SYM_IMPL_H

#include "cbk_internal.h"
#include "quad_chol.h"

namespace cbk {
namespace quadsym {

using namespace quad;
using team::bload;
using team::bstore;
using team::group_range;
using team::group_rsrc;

template <class T> struct QuadSymArgs {
    T *a; /* lower view: element (i, c) at (i*a_si + c*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long nmat; /* groups * V: every slot, padded ones included */
    int nm;
    int nrhs;
    int *info;
    bool factor; /* false: potrs / sytrsnp on a given factor */
    int q0;      /* first right-hand side of this launch (one-block instances) */
    const int *skip; /* one-block launches after the first: the info the first wrote */
    const int *only; /* if set, only the matrices with only[mat] < 0 are touched */
    int n;           /* the order, n <= N: rows and columns past it read as the identity */
};

/* resident wavefronts per SIMD the kernel is built for: the FP64 order 32
 * holds 144 doubles of A per lane and runs alone on its SIMD */
template <class T, int N> constexpr int waves_per_simd()
{
    return (sizeof(T) == 8 && N > 24) ? 1 : 2;
}

/* Whether one launch loops over all right-hand-side blocks. The FP64 order
 * 32 does not: its triangle spills into the accumulation registers, and with
 * a run-time loop around the sweeps the register allocator spills to scratch;
 * the host launches it once per block of NRB columns instead (the first
 * launch factors, the others solve with the stored factor). */
template <class T, int N> constexpr bool loops_rhs() { return waves_per_simd<T, N>() > 1; }

/* whether the first block of B is loaded with the last panel of A (its loads
 * then overlap the factorization); the FP64 order 24, at two wavefronts per
 * SIMD, has no registers to spare for it */
template <class T, int N> constexpr bool prefetch_b() { return !(sizeof(T) == 8 && N == 24); }

/* MASKED: the instance that the masked passes after the mixed-precision
 * kernel launch (QuadSymArgs::only set), a kernel of its own so that traces
 * keep the two apart; either instance honours only */
/* PAD: the instance for orders below N (run as diag(A, I)); the one for N
 * itself folds the order in, so its loads and stores carry no order masks */
template <class T, int N, int NRB, bool LDL, bool MASKED = false, bool PAD = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_simd<T, N>(),
                                                                     waves_per_simd<T, N>()))) void
quad_sym_kernel(QuadSymArgs<T> p)
{
    constexpr int K = N / 4;
    constexpr unsigned SZ = sizeof(T);
    __shared__ T lds[4 * kMatsPerWave * stage_elems<N>()];

    QuadPlace pl;
    if (!quad_place(p.nmat, p.V, pl)) return; /* whole wavefront */
    if (p.only) {
        /* a masked pass: the other matrices' slots read as zero and their
         * stores drop; a wavefront with none to do leaves at once */
        pl.live = pl.live && pl.mat < p.nm && p.only[pl.mat] < 0;
        if (__builtin_amdgcn_ballot_w64(pl.live) == 0) return;
    }
    const int t = pl.t;
    const unsigned VS = (unsigned)p.V * SZ;
    const __amdgpu_buffer_rsrc_t ra =
        group_rsrc(p.a + pl.w0 * p.a_gs, group_range(pl.groups_left, p.a_gs, SZ));
    const unsigned aslot = pl.slot(p.a_gs, p.V, SZ);
    const unsigned a_col = (unsigned)p.a_sj * VS;
    unsigned a_row[K]; /* byte offset of (row t + 4k, column 0) */
    const int n = PAD ? p.n : N;
    sfor<0, K>([&](auto k) {
        a_row[k] = (aslot == kOOB || t + 4 * k >= n) ? kOOB
                                                     : aslot + (unsigned)((t + 4 * k) * p.a_si) * VS;
    });
    T *L = lds + (threadIdx.x >> 2) * stage_elems<N>();

    const __amdgpu_buffer_rsrc_t rb =
        group_rsrc(p.b + pl.w0 * p.b_gs, group_range(pl.groups_left, p.b_gs, SZ));
    const unsigned bslot = pl.slot(p.b_gs, p.V, SZ);
    const unsigned b_col = (unsigned)p.b_sj * VS;
    unsigned b_row[K];
    sfor<0, K>([&](auto k) {
        b_row[k] = (bslot == kOOB || t + 4 * k >= n) ? kOOB
                                                     : bslot + (unsigned)((t + 4 * k) * p.b_si) * VS;
    });
    auto load_b = [&](T(&b)[K][NRB], int q0) {
        sfor<0, NRB>([&](auto q) {
            const unsigned col = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB;
            sfor<0, K>([&](auto k) { b[k][q] = bload(T(), rb, b_row[k], col); });
        });
    };
    /* the first block of B, loaded with the last panel of A */
    const int qfirst = loops_rhs<T, N>() ? 0 : p.q0;
    T b[K][NRB];

    T a[K][N];
    T dg[K];
    int linfo = 0;
    if (p.factor) {
        /* streamed by 8-wide panels: the lower triangle's columns of panel P
         * (entries right of the diagonal read as 0) are loaded while panel
         * P - 1 is factored, and panel P's factor is stored while panel P + 1
         * is */
        auto load = [&](auto Pc) {
            constexpr int P = Pc;
            sfor<8 * P, 8 * P + 8>([&](auto c) {
                sfor<c / 4, K>([&](auto k) {
                    a[k][c] = bload(T(), ra, (c < n && (c < 4 * k || c <= t + 4 * k)) ? a_row[k] : kOOB,
                                    c * a_col);
                });
                /* the identity past n: 1 on the diagonal (row t + 4k = c) */
                a[c / 4][c] = (c >= n && c == t + 4 * (c / 4)) ? T(1) : a[c / 4][c];
            });
            if constexpr (P == N / 8 - 1 && prefetch_b<T, N>()) {
                if (p.nrhs > 0) load_b(b, qfirst);
            }
        };
        auto store = [&](auto Pc) {
            constexpr int P = Pc;
            sfor<8 * P, 8 * P + 8>([&](auto c) {
                sfor<c / 4, K>([&](auto k) {
                    bstore(a[k][c], ra, (c < n && (c < 4 * k || c < t + 4 * k)) ? a_row[k] : kOOB,
                           c * a_col);
                });
            });
            sfor<2 * P, 2 * P + 2>([&](auto k) {
                const unsigned doff =
                    (a_row[k] == kOOB) ? kOOB : a_row[k] + (unsigned)(t + 4 * k) * a_col;
                bstore(dg[k], ra, doff, 0);
            });
        };
        linfo = quad_factor_panels<T, N, LDL, quad_mfma<T, N>()>(a, dg, t, load, store);
        if (p.info && pl.live && pl.mat < p.nm && t == 0) p.info[pl.mat] = linfo;
        if constexpr (!prefetch_b<T, N>()) {
            if (p.nrhs > 0) load_b(b, qfirst);
        }
    }
    else {
        sfor<0, K>([&](auto k) {
            sfor<0, 4 * k + 4>([&](auto c) {
                a[k][c] = bload(T(), ra, (c < n && (c < 4 * k || c < t + 4 * k)) ? a_row[k] : kOOB,
                                c * a_col);
            });
            const unsigned doff = (a_row[k] == kOOB) ? kOOB : a_row[k] + (unsigned)(t + 4 * k) * a_col;
            dg[k] = (t + 4 * k >= n) ? T(1) : bload(T(), ra, doff, 0);
        });
        if (p.nrhs > 0) load_b(b, qfirst);
    }

    if (p.nrhs == 0) return;

    T dinv[K];
    quad_prepare<T, N>(a, dg, dinv, t);

    /* with info, a failed matrix keeps B: its lanes' stores are dropped */
    const bool keep = (p.info && linfo != 0) || (p.skip && pl.mat < p.nm && p.skip[pl.mat] != 0);
    auto solve_block = [&](T(&bb)[K][NRB], int q0) {
        quad_solve<T, N, NRB, LDL>(a, dinv, bb, L, t);
        sfor<0, NRB>([&](auto q) {
            const unsigned col = (q0 + q < p.nrhs) ? (q0 + q) * b_col : kOOB;
            sfor<0, K>([&](auto k) { bstore(bb[k][q], rb, keep ? kOOB : b_row[k], col); });
        });
    };
    solve_block(b, qfirst);
    if constexpr (loops_rhs<T, N>()) {
        for (int q0 = NRB; q0 < p.nrhs; q0 += NRB) {
            T bn[K][NRB];
            load_b(bn, q0);
            solve_block(bn, q0);
        }
    }
}

template <class T, int N, int NRB, bool LDL, bool MASKED, bool PAD>
hipError_t launch_nrb_mp(const QuadSymArgs<T> &args)
{
    const long waves = (args.nmat + kMatsPerWave - 1) / kMatsPerWave;
    const dim3 grid((unsigned)((waves + 3) / 4)), block(256);
    if (loops_rhs<T, N>() || args.nrhs <= NRB) {
        QuadSymArgs<T> a1 = args;
        a1.q0 = 0;
        a1.skip = nullptr;
        hipLaunchKernelGGL((quad_sym_kernel<T, N, NRB, LDL, MASKED, PAD>), grid, block, 0, current_stream(), a1);
        return hipGetLastError();
    }
    /* one launch per block of NRB right-hand sides; the factor from the first */
    for (int q0 = 0; q0 < args.nrhs; q0 += NRB) {
        QuadSymArgs<T> a1 = args;
        a1.q0 = q0;
        a1.factor = args.factor && q0 == 0;
        a1.skip = (q0 > 0 && args.factor) ? args.info : nullptr;
        if (q0 > 0) a1.info = nullptr;
        hipLaunchKernelGGL((quad_sym_kernel<T, N, NRB, LDL, MASKED, PAD>), grid, block, 0, current_stream(), a1);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <class T, int N, int NRB, bool LDL, bool MASKED>
hipError_t launch_nrb_m(const QuadSymArgs<T> &args)
{
    if (args.n != N) return launch_nrb_mp<T, N, NRB, LDL, MASKED, true>(args);
    return launch_nrb_mp<T, N, NRB, LDL, MASKED, false>(args);
}

template <class T, int N, int NRB, bool LDL> hipError_t launch_nrb(const QuadSymArgs<T> &args)
{
    /* the masked passes follow the mixed-precision kernels, which the FP64
     * order 24 serves */
    if constexpr (sizeof(T) == 8 && N == 24) {
        if (args.only) return launch_nrb_m<T, N, NRB, LDL, true>(args);
    }
    return launch_nrb_m<T, N, NRB, LDL, false>(args);
}

/* one order: NRB = 4 right-hand sides per sweep (2 for the FP64 order-24
 * Cholesky, which spills at 4 with two wavefronts per SIMD) */
template <class T, int N> hipError_t launch_n(bool ldl, const QuadSymArgs<T> &args)
{
    constexpr int NRB_CHOL = (sizeof(T) == 8 && N == 24) ? 2 : 4;
    return ldl ? launch_nrb<T, N, 4, true>(args) : launch_nrb<T, N, NRB_CHOL, false>(args);
}

} /* namespace quadsym */
} /* namespace cbk */

#endif /* CBK_QUAD_SYM_IMPL_H */
