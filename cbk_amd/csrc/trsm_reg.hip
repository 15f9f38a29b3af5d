/* trsm_reg.hip -- trsm_compact, side L, with one lane per matrix and the
 * triangle of op(A) in registers, for m <= 16 (and m <= 24 in FP32).
 * Loops run to the compile-time bound M (8, 16 or 24) with wave-uniform
 * guards on the run-time m, so every register index is a constant; rows past
 * m are neither loaded, computed nor stored, and at m == M (EXACT) the guards
 * fold away.
 *
 * The lane kernel (lane_kernels.hip) reads the triangle from HBM once per
 * right-hand side; the team kernel (trsm_team_impl.h) spreads a matrix over 16
 * lanes and is bound by the latency of its broadcasts. Here a lane loads the
 * triangle E = op(A) once, then takes B a column at a time (load, solve,
 * store), so A and B cross HBM once each; at V = 64 a wavefront's 64 lanes
 * read one 512-byte run per element.
 *
 * The operations and their order are those of the reference's trsm router
 * (src/cbk_trsm_compact.hpp:73-145, 179-228, 239-254), so B is bit-identical
 * to it. With E = op(A) and back = (uplo U) != (trans T), every route reduces
 * on E to:
 *   dot form   x_i = (alpha b_i - sum_l E(i,l) x_l) [/ E(i,i)], i descending
 *              when back (l > i ascending), else ascending (l < i ascending);
 *   axpy form  b := alpha b, then per pivot k (descending when back):
 *              x_k = b_k [/ E(k,k)], b_i -= E(i,k) x_k over the unsolved rows i
 *              (ascending);
 * the axpy form only for the last column of an odd count when B has unit row
 * stride and A's unit stride matches op (A's row stride 1 and no transpose, or
 * its column stride 1 and a transpose), the dot form everywhere else.
 */

#include "cbk_internal.h"
#include "team_common.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace cbk {
namespace {

template <int B, int E, class F> __device__ __forceinline__ void rfor(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        rfor<B + 1, E>(f);
    }
}

constexpr int kBlock = 256;

template <class T> struct RegTrsmArgs {
    const T *a;
    int e_si, e_sj; /* element strides of E = op(A) */
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int m;
    int n;
    T alpha;
    bool unit;
    bool axpy_last;
    int V;
    long nlanes;
};

/* index of the k-th row solved (pivot order) */
template <bool BACK, int M, int K> constexpr int piv() { return BACK ? M - 1 - K : K; }

template <class T, int M, bool BACK, bool EXACT>
__global__ __launch_bounds__(kBlock) void trsm_reg_kernel(RegTrsmArgs<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    const int m = EXACT ? M : p.m;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long w0 = __builtin_amdgcn_readfirstlane(
        (int)(((long)blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) / p.V));
    if (idx >= p.nlanes) return;
    const long g = idx / p.V;
    const int v = (int)(idx % p.V);
    const long groups_left = p.nlanes / p.V - w0;
    const unsigned VS = (unsigned)p.V * SZ;
    auto opaque = [](unsigned x) {
        asm volatile("" : "+s"(x));
        return x;
    };
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + w0 * p.a_gs, team::group_range(groups_left, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + w0 * p.b_gs, team::group_range(groups_left, p.b_gs, SZ));
    const unsigned av = (unsigned)(((g - w0) * p.a_gs + v) * SZ);
    const unsigned bv = (unsigned)(((g - w0) * p.b_gs + v) * SZ);

    /* E's triangle: l > i when BACK, l < i otherwise, and the diagonal */
    T e[M][M];
    {
        const unsigned si = opaque((unsigned)p.e_si * VS), sj = opaque((unsigned)p.e_sj * VS);
        rfor<0, M>([&](auto i) {
            rfor<0, M>([&](auto l) {
                if constexpr (BACK ? (l >= i) : (l <= i))
                    e[i][l] = team::bload(T(), ra, ((l == i && p.unit) || i >= m || l >= m) ? team::kOOB : av,
                                          i * si + l * sj);
            });
        });
    }
    const T alpha = p.alpha;
    const unsigned bsiV = (unsigned)p.b_si * VS, bsjV = (unsigned)p.b_sj * VS;
    for (int j = 0; j < p.n; ++j) {
        const unsigned si = opaque(bsiV), col = (unsigned)j * opaque(bsjV);
        T b[M];
        rfor<0, M>([&](auto i) { b[i] = team::bload(T(), rb, i < m ? bv : team::kOOB, col + i * si); });
        if (p.axpy_last && j == p.n - 1) {
            /* trsm_axpy_col (src/cbk_trsm_compact.hpp:49-67) */
            rfor<0, M>([&](auto i) {
                if (i < m) b[i] = b[i] * alpha;
            });
            rfor<0, M>([&](auto kc) {
                constexpr int k = piv<BACK, M, kc>();
                if (k < m) {
                    const T xk = p.unit ? b[k] : b[k] / e[k][k];
                    b[k] = xk;
                    rfor<BACK ? 0 : k + 1, BACK ? k : M>([&](auto i) {
                        if (i < m) b[i] = team::fmaT(-e[i][k], xk, b[i]);
                    });
                }
            });
        }
        else {
            /* trsm_dot_block / trsm_compact_group_strided, one column
             * (src/cbk_trsm_compact.hpp:73-104, 179-200) */
            rfor<0, M>([&](auto tc) {
                constexpr int i = piv<BACK, M, tc>();
                if (i < m) {
                    T w = b[i] * alpha;
                    rfor<BACK ? i + 1 : 0, BACK ? M : i>([&](auto l) {
                        if (l < m) w = team::fmaT(-e[i][l], b[l], w);
                    });
                    b[i] = p.unit ? w : w / e[i][i];
                }
            });
        }
        rfor<0, M>([&](auto i) { team::bstore(b[i], rb, i < m ? bv : team::kOOB, col + i * si); });
    }
}

template <class T, int M> hipError_t launch(bool back, const RegTrsmArgs<T> &a)
{
    const dim3 grid((unsigned)((a.nlanes + kBlock - 1) / kBlock)), block(kBlock);
    const hipStream_t s = current_stream();
    if (a.m == M) {
        if (back) hipLaunchKernelGGL((trsm_reg_kernel<T, M, true, true>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((trsm_reg_kernel<T, M, false, true>), grid, block, 0, s, a);
    }
    else {
        if (back) hipLaunchKernelGGL((trsm_reg_kernel<T, M, true, false>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((trsm_reg_kernel<T, M, false, false>), grid, block, 0, s, a);
    }
    return hipGetLastError();
}

} /* namespace */

bool trsm_reg_supports(bool left, int m, bool fp64)
{
    return left && m >= 1 && m <= (fp64 ? 16 : 24);
}

template <class T>
hipError_t launch_trsm_reg(bool up, bool tran, bool unit, int m, int n, T alpha, Operand<T> A,
                           Operand<T> B, int V, int nm, bool strided)
{
    RegTrsmArgs<T> a;
    const int asi = A.row ? A.ld : 1, asj = A.row ? 1 : A.ld;
    a.a = A.p;
    a.e_si = tran ? asj : asi;
    a.e_sj = tran ? asi : asj;
    a.a_gs = A.gstride;
    a.b = B.p;
    a.b_si = B.row ? B.ld : 1;
    a.b_sj = B.row ? 1 : B.ld;
    a.b_gs = B.gstride;
    a.m = m;
    a.n = n;
    a.alpha = alpha;
    a.unit = unit;
    /* the router's first branch (A's row stride 1) sees op as given, its
     * second (A's column stride 1) the flipped operand: either takes the axpy
     * form for the last of an odd column count when the operand it sees is
     * not transposed */
    const bool odd = (n % 4 == 1) || (n % 4 == 3);
    a.axpy_last = !strided && odd && a.b_si == 1 && (asi == 1 ? !tran : (asj == 1 && tran));
    a.V = V;
    a.nlanes = (long)((nm + V - 1) / V) * V;
    const bool back = (up != tran);
    if (m <= 8) return launch<T, 8>(back, a);
    if (m <= 16) return launch<T, 16>(back, a);
    if constexpr (sizeof(T) == 4) {
        if (m <= 24) return launch<T, 24>(back, a);
    }
    return hipErrorInvalidValue;
}

template hipError_t launch_trsm_reg<double>(bool, bool, bool, int, int, double, Operand<double>,
                                            Operand<double>, int, int, bool);
template hipError_t launch_trsm_reg<float>(bool, bool, bool, int, int, float, Operand<float>,
                                           Operand<float>, int, int, bool);

} /* namespace cbk */
