/* pack.hip -- dense <-> compact conversion on the device.
 *
 * Pack: nm dense matrices (rows x cols) into the compact layout, matrix
 * g*V+v's element (i, j) at ap[g*gstride + (i*si + j*sj)*V + v]; the slots of
 * the last group beyond nm receive pad_diag on the diagonal and zero
 * elsewhere (the identity for matrix batches, zero for tau batches), as the
 * reference's pack_compact (tests/test_compact_util.hpp:475-506). Unpack is
 * the inverse for the nm real slots (:508-531). Entries of the compact buffer
 * beyond rows/cols in its leading dimension are never touched.
 *
 * Both are transposes of a V x (rows*cols) block per group, staged through
 * LDS: a tile holds kTile elements, kTile / V entries of each of the group's
 * V matrices. Dense reads run along the entry index (contiguous along a
 * column or row of one matrix), compact writes along the slot index (V
 * contiguous values per entry), so both sides are coalesced. Entries are
 * enumerated in the compact layout's order, so consecutive tiles of a group
 * write consecutive memory. HBM-bound: each value is read once and written
 * once.
 */

#include "cbk.h"
#include "cbk_ext.h"
#include "cbk_internal.h"

#include <hip/hip_runtime.h>

#include <climits>

namespace cbk {
namespace {

constexpr int kThreads = 256;
constexpr int kTile = 2048; /* elements per tile */
constexpr long kMaxBlocks = 1L << 16;

template <class T> struct PackArgs {
    const T *a;             /* strided dense batch: matrix idx at a + idx*stride_a */
    const T *const *aptr;   /* or an array of nm matrix pointers (a == null) */
    T *const *wptr;         /* unpack into an array of pointers */
    T *w;                   /* unpack into a strided batch */
    size_t stride_a;
    long dsi, dsj;          /* dense (i, j) strides */
    T *ap;
    const T *cap;
    long csi, csj;          /* compact (i, j) strides, in packs */
    size_t gstride;
    long nfast, nent;       /* entries per compact fast line; entries per matrix */
    bool row;               /* compact entries enumerated row by row */
    T pad;
    int V, vshift, ktshift; /* V = 1 << vshift, entries per tile = 1 << ktshift */
    long nm, ngroups, nchunks;
};

/* entry k (compact order) -> (i, j) */
template <class T> __device__ __forceinline__ void entry_ij(const PackArgs<T> &p, long k, long &i, long &j)
{
    const long s = k / p.nfast, f = k - s * p.nfast;
    i = p.row ? s : f;
    j = p.row ? f : s;
}

template <class T> __device__ __forceinline__ const T *dense_src(const PackArgs<T> &p, long idx)
{
    return p.a ? p.a + (size_t)idx * p.stride_a : p.aptr[idx];
}

template <class T> __device__ __forceinline__ T *dense_dst(const PackArgs<T> &p, long idx)
{
    return p.w ? p.w + (size_t)idx * p.stride_a : p.wptr[idx];
}

template <class T> __global__ __launch_bounds__(kThreads) void pack_kernel(PackArgs<T> p)
{
    __shared__ T tile[kTile + kTile / 2]; /* kt rows of V + 1 values (kt <= kTile / 2) */
    const int V = p.V, kt = 1 << p.ktshift, vs = p.vshift;
    const long ntiles = p.ngroups * p.nchunks;
    for (long tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
        const long g = tl / p.nchunks;
        const long k0 = (tl - g * p.nchunks) << p.ktshift;
        /* dense side: slot v = q / kt, entry kk = q % kt */
        for (int q = threadIdx.x; q < kTile; q += kThreads) {
            const int v = q >> p.ktshift, kk = q & (kt - 1);
            const long k = k0 + kk, idx = g * V + v;
            if (k >= p.nent) continue;
            long i, j;
            entry_ij(p, k, i, j);
            T x;
            if (idx < p.nm)
                x = dense_src(p, idx)[i * p.dsi + j * p.dsj];
            else
                x = (i == j) ? p.pad : T(0);
            tile[kk * (V + 1) + v] = x;
        }
        __syncthreads();
        /* compact side: entry kk = q / V, slot v = q % V */
        T *dst = p.ap + (size_t)g * p.gstride;
        for (int q = threadIdx.x; q < kTile; q += kThreads) {
            const int kk = q >> vs, v = q & (V - 1);
            const long k = k0 + kk;
            if (k >= p.nent) continue;
            long i, j;
            entry_ij(p, k, i, j);
            dst[(size_t)(i * p.csi + j * p.csj) * V + v] = tile[kk * (V + 1) + v];
        }
        __syncthreads();
    }
}

template <class T> __global__ __launch_bounds__(kThreads) void unpack_kernel(PackArgs<T> p)
{
    __shared__ T tile[kTile + kTile / 2];
    const int V = p.V, kt = 1 << p.ktshift, vs = p.vshift;
    const long ntiles = p.ngroups * p.nchunks;
    for (long tl = blockIdx.x; tl < ntiles; tl += gridDim.x) {
        const long g = tl / p.nchunks;
        const long k0 = (tl - g * p.nchunks) << p.ktshift;
        const T *src = p.cap + (size_t)g * p.gstride;
        for (int q = threadIdx.x; q < kTile; q += kThreads) {
            const int kk = q >> vs, v = q & (V - 1);
            const long k = k0 + kk;
            if (k >= p.nent || g * V + v >= p.nm) continue;
            long i, j;
            entry_ij(p, k, i, j);
            tile[kk * (V + 1) + v] = src[(size_t)(i * p.csi + j * p.csj) * V + v];
        }
        __syncthreads();
        for (int q = threadIdx.x; q < kTile; q += kThreads) {
            const int v = q >> p.ktshift, kk = q & (kt - 1);
            const long k = k0 + kk, idx = g * V + v;
            if (k >= p.nent || idx >= p.nm) continue;
            long i, j;
            entry_ij(p, k, i, j);
            dense_dst(p, idx)[i * p.dsi + j * p.dsj] = tile[kk * (V + 1) + v];
        }
        __syncthreads();
    }
}

inline bool opt(char c, char upper) { return c == upper || c == upper + ('a' - 'A'); }
inline long max1(long x) { return x < 1 ? 1 : x; }
inline int log2_vlen(int V)
{
    switch (V) {
    case 2: return 1;
    case 4: return 2;
    case 8: return 3;
    case 16: return 4;
    case 32: return 5;
    case 64: return 6;
    default: return -1;
    }
}

/* the launch shared by every entry point, after its argument checks */
template <class T> int run(bool pack, PackArgs<T> p, bool row, bool arow, int rows, int cols, int lda,
                           int ldap)
{
    p.row = row;
    p.nfast = row ? cols : rows;
    p.nent = (long)rows * cols;
    p.dsi = arow ? lda : 1;
    p.dsj = arow ? 1 : lda;
    p.csi = row ? ldap : 1;
    p.csj = row ? 1 : ldap;
    p.gstride = (size_t)ldap * (row ? rows : cols) * p.V;
    p.ktshift = 11 - p.vshift; /* kTile = 2^11 */
    p.ngroups = (p.nm + p.V - 1) / p.V;
    p.nchunks = (p.nent + (1L << p.ktshift) - 1) >> p.ktshift;
    if (p.ngroups == 0 || p.nent == 0) return 0;
    const long ntiles = p.ngroups * p.nchunks;
    const dim3 grid((unsigned)(ntiles < kMaxBlocks ? ntiles : kMaxBlocks));
    if (pack)
        hipLaunchKernelGGL(pack_kernel<T>, grid, dim3(kThreads), 0, current_stream(), p);
    else
        hipLaunchKernelGGL(unpack_kernel<T>, grid, dim3(kThreads), 0, current_stream(), p);
    return launch_status(hipGetLastError());
}

/* the strided forms; codes in argument order */
template <class T>
int pack_strided(char layout, int rows, int cols, const T *a, char alayout, int lda,
                 size_t stride_a, T *ap, int ldap, T pad, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool acol = opt(alayout, 'C'), arow = opt(alayout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && a == nullptr) return -4;
    if (!acol && !arow) return -5;
    if (lda < max1(arow ? cols : rows)) return -6;
    if (nm > 1 && stride_a < (size_t)lda * (arow ? rows : cols)) return -7;
    if (any && ap == nullptr) return -8;
    if (ldap < max1(row ? cols : rows)) return -9;
    const int vs = log2_vlen(V);
    if (vs < 0) return -11;
    if (nm < 0) return -12;
    PackArgs<T> p{};
    p.a = a;
    p.stride_a = stride_a;
    p.ap = ap;
    p.pad = pad;
    p.V = V;
    p.vshift = vs;
    p.nm = nm;
    return run<T>(true, p, row, arow, rows, cols, lda, ldap);
}

template <class T>
int unpack_strided(char layout, int rows, int cols, const T *ap, int ldap, T *a, char alayout,
                   int lda, size_t stride_a, int V, int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    const bool acol = opt(alayout, 'C'), arow = opt(alayout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && ap == nullptr) return -4;
    if (ldap < max1(row ? cols : rows)) return -5;
    if (any && a == nullptr) return -6;
    if (!acol && !arow) return -7;
    if (lda < max1(arow ? cols : rows)) return -8;
    if (nm > 1 && stride_a < (size_t)lda * (arow ? rows : cols)) return -9;
    const int vs = log2_vlen(V);
    if (vs < 0) return -10;
    if (nm < 0) return -11;
    PackArgs<T> p{};
    p.w = a;
    p.stride_a = stride_a;
    p.cap = ap;
    p.V = V;
    p.vshift = vs;
    p.nm = nm;
    return run<T>(false, p, row, arow, rows, cols, lda, ldap);
}

/* the pointer-array forms of MKL's mkl_?gepack_compact / mkl_?geunpack_compact
 * (dense and compact in the same layout; identity padding) */
template <class T>
int gepack(char layout, int rows, int cols, const T *const *a, int lda, T *ap, int ldap, int V,
           int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && a == nullptr) return -4;
    if (lda < max1(row ? cols : rows)) return -5;
    if (any && ap == nullptr) return -6;
    if (ldap < max1(row ? cols : rows)) return -7;
    const int vs = log2_vlen(V);
    if (vs < 0) return -8;
    if (nm < 0) return -9;
    PackArgs<T> p{};
    p.aptr = a;
    p.ap = ap;
    p.pad = T(1);
    p.V = V;
    p.vshift = vs;
    p.nm = nm;
    return run<T>(true, p, row, row, rows, cols, lda, ldap);
}

template <class T>
int geunpack(char layout, int rows, int cols, T *const *a, int lda, const T *ap, int ldap, int V,
             int nm)
{
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (rows < 0) return -2;
    if (cols < 0) return -3;
    const bool any = nm > 0 && rows > 0 && cols > 0;
    if (any && a == nullptr) return -4;
    if (lda < max1(row ? cols : rows)) return -5;
    if (any && ap == nullptr) return -6;
    if (ldap < max1(row ? cols : rows)) return -7;
    const int vs = log2_vlen(V);
    if (vs < 0) return -8;
    if (nm < 0) return -9;
    PackArgs<T> p{};
    p.wptr = a;
    p.cap = ap;
    p.V = V;
    p.vshift = vs;
    p.nm = nm;
    return run<T>(false, p, row, row, rows, cols, lda, ldap);
}

} /* namespace */
} /* namespace cbk */

extern "C" {

#define CBK_PACK_ENTRY_POINTS(T, p)                                                           \
    int cbk_##p##pack_compact(char layout, int rows, int cols, const T *a, char alayout,      \
                              int lda, size_t stride_a, T *ap, int ldap, T pad_diag, int V,   \
                              int nm)                                                         \
    {                                                                                         \
        return cbk::pack_strided<T>(layout, rows, cols, a, alayout, lda, stride_a, ap, ldap,  \
                                    pad_diag, V, nm);                                         \
    }                                                                                         \
    int cbk_##p##unpack_compact(char layout, int rows, int cols, const T *ap, int ldap, T *a, \
                                char alayout, int lda, size_t stride_a, int V, int nm)        \
    {                                                                                         \
        return cbk::unpack_strided<T>(layout, rows, cols, ap, ldap, a, alayout, lda,          \
                                      stride_a, V, nm);                                       \
    }                                                                                         \
    int cbk_##p##gepack_compact(char layout, int rows, int cols, const T *const *a, int lda,  \
                                T *ap, int ldap, int V, int nm)                               \
    {                                                                                         \
        return cbk::gepack<T>(layout, rows, cols, a, lda, ap, ldap, V, nm);                   \
    }                                                                                         \
    int cbk_##p##geunpack_compact(char layout, int rows, int cols, T *const *a, int lda,      \
                                  const T *ap, int ldap, int V, int nm)                       \
    {                                                                                         \
        return cbk::geunpack<T>(layout, rows, cols, a, lda, ap, ldap, V, nm);                 \
    }

CBK_PACK_ENTRY_POINTS(double, d)
CBK_PACK_ENTRY_POINTS(float, s)
#undef CBK_PACK_ENTRY_POINTS

} /* extern "C" */
