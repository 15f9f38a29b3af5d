/* chol_grid.hip -- potrf / posv / potrs and sytrfnp / sysvnp / sytrsnp at
 * orders 33 to 64, one matrix per wavefront on an 8 x 8 grid of lanes.
 *
 * Lane (r, c) = lane bits [2:0] and [5:3] holds the element (r + 8a, c + 8b)
 * of the lower view for every slot pair a >= b, f[a][b] (36 values; the
 * entries right of the diagonal in the diagonal slots ride along as scratch,
 * never stored). The reference's factorization (src/cbk_potrf_compact.hpp:
 * 79-124, src/cbk_sytrfnp_compact.hpp:96-265) works by 8-wide panels, which
 * are the column slots b here; every element gets the reference's operations
 * in its order, as quad_chol.h states them:
 *   trailing form  at the start of panel P, for each earlier pivot j
 *                  ascending: a(i, c) = fma(-w, L(i, j), a(i, c)) with
 *                  w = L(c, j) (LDL^T: L(c, j) D(j)), L scaled;
 *   panel form     for pivot j of panel P: w = a(c, j) * (1/piv) on the
 *                  unscaled column for the panel's columns c > j,
 *                  a(i, c) = fma(-w, a(i, j), a(i, c)); then column j scaled
 *                  by sqrt(piv) * (1/piv) (LDL^T: 1/piv), the diagonal kept
 *                  as sqrt(piv) (LDL^T: piv);
 * so the factor is bit-identical to the reference's. The values a lane needs
 * from another column lane (L(i, j) for its rows, w for its column) come by
 * ds_bpermute, the pivot by v_readlane. The solves follow quad_chol.h's
 * quad_solve on 8-column blocks of B (lane c holds column q0 + c): forward
 * right-looking in the reference's per-element order, then the reciprocal
 * of the diagonal, backward column-oriented from the bottom (its sums in
 * another order than the reference's row dots: X agrees to rounding).
 * Rows and columns past n are zero and never stored; their pivots are never
 * taken.
 */
#include "cbk_internal.h"
#include "team_common.h"

namespace cbk {
namespace cholgrid {

using team::fmaT;
using team::kOOB;
using team::sfor;
using team::sfor_down;
using team::sqrtT;

constexpr int S = 8;     /* slots per dimension */
constexpr int NMAX = 64;

template <class T> struct CholGridArgs {
    T *a; /* lower view: element (i, c) at (i*a_si + c*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long long count; /* slots, padded ones included */
    int nm, n, nrhs;
    int *info;
    bool factor; /* false: potrs / sytrsnp on a given factor */
    const int *mask; /* if set: only the matrices k < nm with mask[k] < 0 */
};

/* x of the lane whose ds_bpermute byte address is addr */
__device__ __forceinline__ double bperm(double x, int addr)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_ds_bpermute(addr, w.x);
    w.y = __builtin_amdgcn_ds_bpermute(addr, w.y);
    return __builtin_bit_cast(double, w);
}
__device__ __forceinline__ float bperm(float x, int addr)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(addr, __builtin_bit_cast(int, x)));
}

/* x of lane (R, c) of this lane's column lane (ds_swizzle: and 0x18, or R) */
template <int R> __device__ __forceinline__ double rowbc(double x)
{
    constexpr int pat = 0x18 | (R << 5);
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_ds_swizzle(w.x, pat);
    w.y = __builtin_amdgcn_ds_swizzle(w.y, pat);
    return __builtin_bit_cast(double, w);
}
template <int R> __device__ __forceinline__ float rowbc(float x)
{
    constexpr int pat = 0x18 | (R << 5);
    return __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, x), pat));
}

/* x of lane L, wave-uniform */
template <int L> __device__ __forceinline__ double lane_value(double x)
{
    int2 w = __builtin_bit_cast(int2, x);
    w.x = __builtin_amdgcn_readlane(w.x, L);
    w.y = __builtin_amdgcn_readlane(w.y, L);
    return __builtin_bit_cast(double, w);
}
template <int L> __device__ __forceinline__ float lane_value(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), L));
}

template <class T> __device__ __forceinline__ T opaque(T x)
{
    asm volatile("" : "+v"(x));
    return x;
}

template <class T, bool LDL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
chol_grid_kernel(CholGridArgs<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int r = lane & 7, c = lane >> 3;
    /* blocks in XCD order: the 16 matrices of 4 consecutive logical blocks
     * share an XCD (a 128-byte run of each element at V >= 16 in FP64) */
    const long long kq = (long long)team::xcd_block(blockIdx.x) * 4 + team::wave_in_block();
    if (kq >= p.count) return; /* whole wavefront */
    if (p.mask && (kq >= p.nm || p.mask[kq] >= 0)) return;
    const int n = p.n;
    const long long g = kq / p.V, v = kq % p.V;
    const long long groups = p.count / p.V;
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + g * p.a_gs, team::group_range(groups - g, p.a_gs, SZ));
    const unsigned VS = (unsigned)p.V * SZ;
    const unsigned ars = (unsigned)p.a_si * VS, acs = (unsigned)p.a_sj * VS;
    const unsigned aoff = (unsigned)(v * SZ) + r * ars + c * acs;
    /* element (r + 8a, c + 8b) on and below the diagonal, inside the matrix */
    auto a_at = [&](unsigned lo, int rr, int cc, int a, int b, bool diag) {
        const int i = rr + 8 * a, j = cc + 8 * b;
        return (i < n && (diag ? j <= i : j < i)) ? lo : kOOB;
    };

    T f[S][S];
    int linfo = 0;
    const int raddr = 4 * r;     /* ds_bpermute address of lane (r, 0) */
    const int caddr = 4 * c;     /* ... of lane (c, 0) */
    if (p.factor) {
        sfor<0, S>([&](auto a) {
            sfor<0, a + 1>([&](auto b) {
                f[a][b] = team::bload(T(0), ra, a_at(aoff, r, c, a, b, true), a * 8 * ars + b * 8 * acs);
            });
        });
        sfor<0, S>([&](auto Pc) {
            constexpr int P = Pc;
            if (8 * P < n) {
                /* trailing form of the earlier panels' pivots on this panel */
                sfor<0, 8 * P>([&](auto jc) {
                    constexpr int j = jc;
                    constexpr int jl = j % 8, jb = j / 8;
                    /* w = L(c + 8P, j), from lane (c, jl) */
                    T w = bperm(f[P][jb], caddr + 32 * jl);
                    if constexpr (LDL) w = w * lane_value<jl + 8 * jl>(f[jb][jb]);
                    sfor<P, S>([&](auto a) {
                        const T lij = bperm(f[a][jb], raddr + 32 * jl); /* L(r + 8a, j) */
                        f[a][P] = fmaT(-w, lij, f[a][P]);
                    });
                    __builtin_amdgcn_sched_barrier(0);
                });
                /* the panel, right-looking */
                sfor<0, 8>([&](auto jlc) {
                    constexpr int jl = jlc;
                    constexpr int j = 8 * P + jl;
                    if (j < n) {
                        const T piv = lane_value<jl + 8 * jl>(f[P][P]);
                        if (linfo == 0 && (LDL ? (piv == T(0)) : !(piv > T(0)))) linfo = j + 1;
                        const T inv = T(1) / piv;
                        if constexpr (jl + 1 < 8) {
                            /* w = a(c + 8P, j) / piv for the columns c > jl */
                            const T w = bperm(f[P][P], caddr + 32 * jl) * inv;
                            sfor<P, S>([&](auto a) {
                                const T aij = bperm(f[a][P], raddr + 32 * jl); /* a(r + 8a, j) */
                                f[a][P] = (c > jl) ? fmaT(-w, aij, f[a][P]) : f[a][P];
                            });
                        }
                        T d, s;
                        if constexpr (LDL) {
                            d = piv;
                            s = inv;
                        }
                        else {
                            d = sqrtT(piv);
                            s = d * inv;
                        }
                        sfor<P, S>([&](auto a) { f[a][P] = (c == jl) ? f[a][P] * s : f[a][P]; });
                        f[P][P] = (c == jl && r == jl) ? d : f[P][P];
                    }
                    __builtin_amdgcn_sched_barrier(0);
                });
            }
        });
        if (p.info && kq < p.nm && lane == 0) p.info[kq] = linfo;
        const unsigned aoff2 = opaque(aoff);
        const int r2 = opaque(r), c2 = opaque(c);
        sfor<0, S>([&](auto a) {
            sfor<0, a + 1>([&](auto b) {
                team::bstore(f[a][b], ra, a_at(aoff2, r2, c2, a, b, true), a * 8 * ars + b * 8 * acs);
            });
        });
    }
    else {
        sfor<0, S>([&](auto a) {
            sfor<0, a + 1>([&](auto b) {
                f[a][b] = team::bload(T(0), ra, a_at(aoff, r, c, a, b, true), a * 8 * ars + b * 8 * acs);
            });
        });
    }
    if (p.nrhs == 0) return;
    /* with info, a failed matrix keeps B (LAPACK dposv skips dpotrs) */
    const bool keep = p.info && linfo != 0;

    /* 1 / L(i, i) (1 / D(i)) for this lane's rows i = r + 8a, from lane (r, r) */
    T dinv[S];
    sfor<0, S>([&](auto a) {
        const T dv = bperm(f[a][a], 36 * r);
        dinv[a] = (r + 8 * a < n) ? T(1) / dv : T(1);
    });

    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + g * p.b_gs, team::group_range(groups - g, p.b_gs, SZ));
    const unsigned brs = (unsigned)p.b_si * VS, bcs = (unsigned)p.b_sj * VS;
    const unsigned boff = (unsigned)(v * SZ) + r * brs;
    for (int q0 = 0; q0 < p.nrhs; q0 += 8) {
        /* column q0 + c of B: its offset in the lane's voffset (the scalar
         * offset must be wave-uniform) */
        const int q = q0 + c;
        const unsigned boq = boff + (unsigned)q * bcs;
        T bb[S];
        sfor<0, S>([&](auto a) {
            const unsigned o = (r + 8 * a < n && q < p.nrhs) ? boq : kOOB;
            bb[a] = team::bload(T(0), rb, o, a * 8 * brs);
        });
        /* forward: L y = b; step l: y_l = b_l (/ L(l, l)) from its row lane,
         * then the rows below */
        /* by slots of 8 steps: a step past n is an exact no-op (its y, and its
         * L entries, are +0), so only whole slots are guarded and a slot's
         * steps schedule together */
        sfor<0, S>([&](auto lbc) {
            constexpr int lb = lbc;
            if (8 * lb < n) {
                sfor<8 * lb, 8 * lb + 8>([&](auto lc) {
                    constexpr int l = lc;
                    constexpr int ll = l % 8;
                    const T y = rowbc<ll>(LDL ? bb[lb] : bb[lb] * dinv[lb]);
                    sfor<lb, S>([&](auto a) {
                        const T lil = bperm(f[a][lb], raddr + 32 * ll); /* L(r + 8a, l) */
                        if constexpr (a > lb)
                            bb[a] = fmaT(-lil, y, bb[a]);
                        else
                            bb[a] = (r > ll) ? fmaT(-lil, y, bb[a]) : bb[a];
                    });
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        /* Cholesky: y = b / L(i, i); sytrsnp: the reciprocal of D times b */
        sfor<0, S>([&](auto a) { bb[a] = bb[a] * dinv[a]; });
        /* backward: L^T x = y from the bottom; step l: x_l from its row lane,
         * then the rows above with L(l, i) from lane (l % 8, r) */
        sfor_down<0, S>([&](auto lbc) {
            constexpr int lb = lbc;
            if (8 * lb < n) {
                sfor_down<8 * lb, 8 * lb + 8>([&](auto lc) {
                    constexpr int l = lc;
                    constexpr int ll = l % 8;
                    const T x = rowbc<ll>(LDL ? bb[lb] : bb[lb] * dinv[lb]);
                    sfor<0, lb + 1>([&](auto a) {
                        const T cli = bperm(f[lb][a], 4 * ll + 32 * r); /* L(l, r + 8a) */
                        if constexpr (a < lb)
                            bb[a] = fmaT(-cli, x, bb[a]);
                        else
                            bb[a] = (r < ll) ? fmaT(-cli, x, bb[a]) : bb[a];
                    });
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        if constexpr (!LDL) {
            sfor<0, S>([&](auto a) { bb[a] = bb[a] * dinv[a]; });
        }
        const unsigned boq2 = opaque(boq);
        const int r2 = opaque(r);
        sfor<0, S>([&](auto a) {
            const unsigned o = (!keep && r2 + 8 * a < n && q < p.nrhs) ? boq2 : kOOB;
            team::bstore(bb[a], rb, o, a * 8 * brs);
        });
    }
}

/* trsm side L, forward sweep (op(A) lower), orders 33..64: B := alpha
 * op(A)^{-1} B on 8-column blocks of B, per element the reference's sequence
 * (src/cbk_trsm_compact.hpp; quad_trsm_impl.h): the multiply by alpha, the
 * divide on the pivot's own row, then the fma's over l ascending */
template <class T> struct TrsmGridArgs {
    const T *a; /* op(A): element (i, c) at (i*a_si + c*a_sj)*V in the group */
    int a_si, a_sj;
    size_t a_gs;
    T *b;
    int b_si, b_sj;
    size_t b_gs;
    int V;
    long long count;
    int m, nrhs;
    T alpha;
};

template <class T, bool UNIT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void
trsm_grid_kernel(TrsmGridArgs<T> p)
{
    constexpr unsigned SZ = sizeof(T);
    const int lane = threadIdx.x & 63;
    const int r = lane & 7, c = lane >> 3;
    const long long kq = (long long)team::xcd_block(blockIdx.x) * 4 + team::wave_in_block();
    if (kq >= p.count) return; /* whole wavefront */
    const int m = p.m;
    const long long g = kq / p.V, v = kq % p.V;
    const long long groups = p.count / p.V;
    const __amdgpu_buffer_rsrc_t ra =
        team::group_rsrc(p.a + g * p.a_gs, team::group_range(groups - g, p.a_gs, SZ));
    const __amdgpu_buffer_rsrc_t rb =
        team::group_rsrc(p.b + g * p.b_gs, team::group_range(groups - g, p.b_gs, SZ));
    const unsigned VS = (unsigned)p.V * SZ;
    const unsigned ars = (unsigned)p.a_si * VS, acs = (unsigned)p.a_sj * VS;
    const unsigned brs = (unsigned)p.b_si * VS, bcs = (unsigned)p.b_sj * VS;
    const unsigned aoff = (unsigned)(v * SZ) + r * ars + c * acs;
    const unsigned boff = (unsigned)(v * SZ) + r * brs;
    const int raddr = 4 * r;

    /* op(A) strictly below the diagonal, and the diagonal unless unit */
    T f[S][S];
    sfor<0, S>([&](auto a) {
        sfor<0, a + 1>([&](auto b) {
            const int i = r + 8 * a, j = c + 8 * b;
            const bool in = i < m && (j < i || (!UNIT && j == i));
            f[a][b] = team::bload(T(0), ra, in ? aoff : kOOB, a * 8 * ars + b * 8 * acs);
        });
    });
    for (int q0 = 0; q0 < p.nrhs; q0 += 8) {
        const int q = q0 + c;
        const unsigned boq = boff + (unsigned)q * bcs;
        T bb[S];
        sfor<0, S>([&](auto a) {
            const unsigned o = (r + 8 * a < m && q < p.nrhs) ? boq : kOOB;
            bb[a] = team::bload(T(0), rb, o, a * 8 * brs) * p.alpha;
        });
        sfor<0, NMAX>([&](auto lc) {
            constexpr int l = lc;
            constexpr int ll = l % 8, lb = l / 8;
            if (l < m) {
                if constexpr (!UNIT) {
                    const T dl = lane_value<ll + 8 * ll>(f[lb][lb]);
                    bb[lb] = (r == ll) ? bb[lb] / dl : bb[lb];
                }
                const T x = rowbc<ll>(bb[lb]);
                sfor<lb, S>([&](auto a) {
                    const T lil = bperm(f[a][lb], raddr + 32 * ll); /* op(A)(r + 8a, l) */
                    if constexpr (a > lb)
                        bb[a] = fmaT(-lil, x, bb[a]);
                    else
                        bb[a] = (r > ll) ? fmaT(-lil, x, bb[a]) : bb[a];
                });
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        const unsigned boq2 = opaque(boq);
        const int r2 = opaque(r);
        sfor<0, S>([&](auto a) {
            const unsigned o = (r2 + 8 * a < m && q < p.nrhs) ? boq2 : kOOB;
            team::bstore(bb[a], rb, o, a * 8 * brs);
        });
    }
}

} /* namespace cholgrid */

bool chol_grid_supports(int n) { return n > 32 && n <= 64; }

bool trsm_grid_supports(bool left, bool up, bool tran, int m)
{
    return left && up == tran && m > 32 && m <= 64;
}

template <class T>
hipError_t launch_trsm_grid(bool tran, bool unit, int m, int n, T alpha, Operand<T> A, Operand<T> B,
                            int V, int nm)
{
    cholgrid::TrsmGridArgs<T> args;
    args.a = A.p;
    /* op(A): the transposed view when tran (src/cbk_common.hpp:156-177) */
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (tran) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.a_si = si;
    args.a_sj = sj;
    args.a_gs = A.gstride;
    args.b = B.p;
    args.b_si = B.row ? B.ld : 1;
    args.b_sj = B.row ? 1 : B.ld;
    args.b_gs = B.gstride;
    args.V = V;
    args.count = (long long)((nm + V - 1) / V) * V;
    args.m = m;
    args.nrhs = n;
    args.alpha = alpha;
    const unsigned blocks = team::xcd_grid((unsigned)((args.count + 3) / 4));
    if (unit)
        hipLaunchKernelGGL((cholgrid::trsm_grid_kernel<T, true>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((cholgrid::trsm_grid_kernel<T, false>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    return hipGetLastError();
}

template hipError_t launch_trsm_grid<double>(bool, bool, int, int, double, Operand<double>,
                                             Operand<double>, int, int);
template hipError_t launch_trsm_grid<float>(bool, bool, int, int, float, Operand<float>,
                                            Operand<float>, int, int);

template <class T>
hipError_t launch_chol_grid(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                            Operand<T> B, int V, int nm, int *info, const int *mask)
{
    cholgrid::CholGridArgs<T> args;
    args.a = A.p;
    /* the lower view: uplo 'U' reads A transposed (src/cbk_common.hpp:201-206) */
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (up) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.a_si = si;
    args.a_sj = sj;
    args.a_gs = A.gstride;
    args.b = B.p;
    args.b_si = B.row ? B.ld : 1;
    args.b_sj = B.row ? 1 : B.ld;
    args.b_gs = B.gstride;
    args.V = V;
    args.count = (long long)((nm + V - 1) / V) * V;
    args.nm = nm;
    args.n = n;
    args.nrhs = nrhs;
    args.info = info;
    args.factor = factor;
    args.mask = mask;
    const unsigned blocks = team::xcd_grid((unsigned)((args.count + 3) / 4));
    if (ldl)
        hipLaunchKernelGGL((cholgrid::chol_grid_kernel<T, true>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((cholgrid::chol_grid_kernel<T, false>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    return hipGetLastError();
}

template hipError_t launch_chol_grid<double>(bool, bool, bool, int, int, Operand<double>,
                                             Operand<double>, int, int, int *, const int *);
template hipError_t launch_chol_grid<float>(bool, bool, bool, int, int, Operand<float>,
                                            Operand<float>, int, int, int *, const int *);

} /* namespace cbk */
