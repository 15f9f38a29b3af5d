/* batch_gels_impl.h -- least squares (gels, m >= n, trans 'N') over an uneven
 * batch: problems of different m, n, nrhs, each stored column-major on its
 * own, taken from a work queue by persistent grids.
 *
 * Column ownership. A problem is the m x (n + nrhs) matrix [A | B]; lane c of
 * the problem's lane group owns column c in registers, f[i] = row i (rows >= m
 * zero). Householder QR (the reference's geqr2 and larf,
 * src/cbk_geqrf_compact.hpp:74-131, src/cbk_ormqr_compact.hpp:68-105) then
 * runs per reflector kk:
 *   tail   = sum_{i>kk} f_kk[i]^2            (lane kk, its own column)
 *   larfg  : norm = sqrt(x0^2 + tail), beta = x0 >= 0 ? -norm : norm,
 *            tau = (beta - x0)/beta, the body of column kk scaled by
 *            1/(x0 - beta), R(kk,kk) = beta (tau = 0, column unchanged, when
 *            tail = 0)
 *   apply  : every column c > kk, B's included:
 *            w = f_c[kk] + sum_{i>kk} v_i f_c[i], w *= tau,
 *            f_c[kk] -= w, f_c[i] -= v_i w
 * so Q^T B is formed in the B columns as the factorization proceeds (the
 * reference's ormqr, forward). Back substitution R X = (Q^T B)(0:n) then runs
 * in the B columns, column-oriented (x_l = b_l / R(l,l), b_i -= R(i,l) x_l).
 * The reflector element v_i that every column needs lives in lane kk:
 *   - team kernel (n + nrhs <= 16): 16 lanes per problem, 4 problems per
 *     wavefront, v_i reaches the team by a DPP row broadcast fused into the
 *     fma (reflectors and pivots unrolled at compile time);
 *   - wavefront kernel (n + nrhs <= 64): one problem per wavefront, v_i read
 *     with v_readlane into a scalar register; the reflector loop runs at run
 *     time over a rolling register window (below);
 *   - two-array wavefront kernel (n <= 64, nrhs <= 64, n + nrhs > 64): A's
 *     and B's columns in separate register arrays of the same lanes.
 * The team kernel sums in row order as the reference does; the wavefront
 * kernel keeps four partial sums (rows i mod 4). Results agree with the
 * reference to rounding, which the tests bound.
 *
 * Work distribution: one 32-bit counter per class, cleared before the launch;
 * a wavefront takes the next 4 problems (team kernel) or 1 (wavefront kernel)
 * with one atomic add and exits when the queue is empty. The plan orders each
 * class longest first, so the tail of the queue is short problems.
 */
#ifndef CBK_BATCH_GELS_IMPL_H
#define CBK_BATCH_GELS_IMPL_H

#include "cbk_internal.h"
#include "team_common.h"

namespace cbk {
namespace batch {

using namespace team;

template <class T> struct Larfg {
    T tau, inv, rdiag;
};

/* larfg_pack (src/cbk_geqrf_compact.hpp:74-90): no rescaling, direct sqrt */
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

/* ---- team kernel: n + nrhs <= 16, m <= MMAX --------------------------------- */
template <class T, int MMAX>
__global__ __launch_bounds__(256) void gels_team_kernel(BatchArgs<T> p)
{
    const int c = threadIdx.x & (TS - 1);
    const int sub = (threadIdx.x >> 4) & 3; /* team within the wavefront */
    const int lane = threadIdx.x & 63;
    const unsigned count = batch_count(p);
    for (;;) {
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(p.counter, 4u);
        base = __builtin_amdgcn_readfirstlane(base);
        if (base >= count) return; /* the whole wavefront */
        const int k = (int)base + sub;
        const bool live = (unsigned)k < count;
        const BatchProblem d = p.desc[live ? k : (int)base];
        const int m = d.m, n = live ? d.n : 0, w = live ? d.n + d.nrhs : 0;
        const bool isA = c < n, isB = c >= n && c < w;
        const bool has = isA || isB;
        T *col = isA ? p.a + d.a_off + (long long)c * d.lda
                     : p.b + d.b_off + (long long)(isB ? c - n : 0) * d.ldb;

        T f[MMAX];
        sfor<0, MMAX>([&](auto i) { f[i] = T(0); });
        if (has) {
            sfor<0, MMAX>([&](auto i) {
                if (i < m) f[i] = col[i];
            });
        }

        sfor<0, TS>([&](auto kc) {
            constexpr int kk = kc;
            if (kk < n) {
                T tl = T(0);
                sfor<kk + 1, MMAX>([&](auto i) { tl = fmaT(f[i], f[i], tl); });
                const Larfg<T> h = larfg(bcast<kk>(f[kk]), bcast<kk>(tl));
                if (c == kk) {
                    sfor<kk + 1, MMAX>([&](auto i) { f[i] = f[i] * h.inv; });
                    f[kk] = h.rdiag;
                }
                /* the broadcast sources f[kk+1..] of lane kk are final here */
                dpp_ready_from<kk + 1>(f);
                T wv = f[kk];
                sfor<kk + 1, MMAX>([&](auto i) { fma_bcr<kk>(wv, f[i], f[i]); });
                wv = (c > kk) ? wv * h.tau : T(0);
                f[kk] = f[kk] - wv;
                sfor<kk + 1, MMAX>([&](auto i) { fnma_bcr<kk>(f[i], f[i], wv); });
            }
            __builtin_amdgcn_sched_barrier(0);
        });

        /* info: the first exact zero on R's diagonal (LAPACK dtrtrs's test) */
        int bad = 0;
        sfor<0, TS>([&](auto lc) {
            constexpr int l = lc;
            const T dl = bcast<l>(f[l]);
            if (l < n && bad == 0 && dl == T(0)) bad = l + 1;
        });
        if (live && c == 0 && p.info) p.info[d.idx] = bad;
        if (bad == 0 || !p.info) {
            sfor_down<0, TS>([&](auto lc) {
                constexpr int l = lc;
                if (l < n) {
                    const T dl = bcast<l>(f[l]);
                    const T x = isB ? f[l] / dl : T(0);
                    f[l] = isB ? x : f[l];
                    sfor<0, l>([&](auto i) { fnma_bc<l>(f[i], f[i], x); });
                }
            });
        }
        if (has) {
            sfor<0, MMAX>([&](auto i) {
                if (i < m) col[i] = f[i];
            });
        }
    }
}

/* LDS writes of some lanes visible to every lane of the same wavefront (its
 * LDS operations complete in order; the fences keep the compiler from moving
 * accesses across) */
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* ---- wavefront kernel: n + nrhs <= 64, m <= MMAX ----------------------------
 * The reflector loop runs at run time in steps of G = 8: each lane keeps a
 * window of its column that starts at row k0, f[i] = row k0 + i, and the 8
 * reflectors k0 .. k0+7 pivot on window rows 0..7 (compile-time indices);
 * then the window shifts up by 8 (zeros enter at the bottom, where rows >= m
 * already are zero). The reflector's body v_i, which every lane needs, is
 * written once to the wavefront's LDS row by lane kk and read back by all
 * lanes at the same address (a broadcast read), in place of two v_readlane
 * per row for the dot product and two more for the update. The back
 * substitution walks a reversed window up from row n-1 the same way, R's
 * column l broadcast through LDS from lane l. Every index into the windows is
 * a compile-time constant, and rows past the end are zeros that every update
 * leaves zero. */
/* x, as a value the compiler must take to depend on dep. A v_readlane of x
 * then cannot be merged with an earlier one or hoisted above dep; chaining
 * dep to the result four rows back keeps about four broadcast values (eight
 * scalar registers) alive instead of all 2*MMAX. */
template <class T> __device__ __forceinline__ T after(T x, T dep)
{
    asm volatile("" : "+v"(x) : "v"(dep));
    return x;
}

/* the problem the wavefront took: from the table, or computed in compact mode */
template <class T>
__device__ __forceinline__ BatchProblem problem_at(const BatchArgs<T> &p, unsigned k)
{
    if (p.desc) return p.desc[k];
    const long long g = k / p.cV, v = k % p.cV;
    return BatchProblem{p.cm, p.cn, p.cnrhs, p.a_cs, p.b_cs, (int)k,
                        g * p.a_gs + v, g * p.b_gs + v, 0};
}

/* body(i) for the compile-time rows i of [B, E) whose 8-row chunk starts
 * below `live` (wave-uniform): windows shrink as the reflectors advance, and
 * the rows past the live ones are zeros that the skipped updates would leave
 * zero */
template <int B, int E, class F> __device__ __forceinline__ void live_rows(int live, F &&body)
{
    sfor<0, (E + 7) / 8>([&](auto ch) {
        constexpr int lo = ch * 8 > B ? ch * 8 : B;
        constexpr int hi = ch * 8 + 8 < E ? ch * 8 + 8 : E;
        if constexpr (lo < hi) {
            if (ch * 8 < live) sfor<lo, hi>(body);
        }
    });
}

template <class T, int MMAX>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? (MMAX > 32 ? 2 : 4) : (MMAX > 32 ? 4 : 6)))) void
gels_wave_kernel(BatchArgs<T> p)
{
    constexpr int G = 8; /* reflectors per window position */
    __shared__ T vshare[4][MMAX];
    T *vb = vshare[wave_in_block()]; /* this wavefront's broadcast row */
    const int c = threadIdx.x & 63;
    /* masked compact mode: the grid is persistent and wavefront w scans the
     * mask 64 entries at a time (chunks w, w + waves, ...), one coalesced load
     * and a ballot per chunk, then solves the marked matrices of the chunk
     * one after another; a call with no marked matrix costs one pass over
     * the mask */
    const unsigned nwaves = gridDim.x * 4u;
    unsigned chunk = blockIdx.x * 4u + (unsigned)wave_in_block();
    unsigned long long pend = 0;
    bool first_chunk = true;
    const unsigned count = batch_count(p);
    for (int round = 0;; ++round) {
        unsigned kq = 0;
        if (p.counter) {
            if (c == 0) kq = atomicAdd(p.counter, 1u);
            kq = __builtin_amdgcn_readfirstlane(kq);
        }
        else if (p.mask) {
            while (pend == 0) {
                if (!first_chunk) chunk += nwaves;
                first_chunk = false;
                if (chunk >= ((unsigned)p.ninfo + 63u) / 64u) return;
                const unsigned k = chunk * 64u + (unsigned)c;
                pend = __builtin_amdgcn_ballot_w64(k < (unsigned)p.ninfo && p.mask[k] < 0);
            }
            kq = chunk * 64u + (unsigned)__builtin_ctzll(pend);
            pend &= pend - 1ull;
        }
        else {
            /* compact mode: one problem per wavefront; blocks in XCD order, so
             * the 16 consecutive matrices that share the lines of an element
             * (V = 64) are read through one L2 */
            if (round > 0) return;
            kq = xcd_block(blockIdx.x) * 4u + (unsigned)wave_in_block();
        }
        if (kq >= count) return;
        const BatchProblem d = problem_at(p, kq);
        const int m = d.m, n = d.n, w = d.n + d.nrhs;
        const bool isA = c < n, isB = c >= n && c < w;
        const bool has = isA || isB;
        const long long rs = isA ? p.a_rs : p.b_rs;
        T *col = isA ? p.a + d.a_off + (long long)c * d.lda
                     : p.b + d.b_off + (long long)(isB ? c - n : 0) * d.ldb;
        T *tau = p.ctau ? p.ctau + (long long)(kq / p.cV) * p.cn * p.cV + kq % p.cV : nullptr;
        /* where the B lanes' results go: B itself, or X (rows < n) */
        const bool xout = isB && p.x;
        T *ocol = xout ? p.x + (long long)(kq / p.cV) * p.x_gs + kq % p.cV + (long long)(c - n) * p.x_cs
                       : col;
        const long long ors = xout ? p.x_rs : rs;

        T f[MMAX]; /* window: f[i] = row k0 + i of the lane's column */
        sfor<0, MMAX>([&](auto i) { f[i] = T(0); });
        if (has) {
            sfor<0, MMAX>([&](auto i) {
                if (i < m) f[i] = col[i * rs];
            });
        }

        T own = T(0); /* R(c,c) of an A lane */
        for (int k0 = 0; k0 < n; k0 += G) {
            const int live = m - k0; /* window rows that can be nonzero */
            sfor<0, G>([&](auto jc) {
                constexpr int j = jc; /* reflector kk = k0 + j pivots on window row j */
                const int kk = k0 + j;
                if (kk < n) {
                    T t4[4] = {T(0), T(0), T(0), T(0)};
                    live_rows<j + 1, MMAX>(live, [&](auto i) { t4[i % 4] = fmaT(f[i], f[i], t4[i % 4]); });
                    const T tl = (t4[0] + t4[1]) + (t4[2] + t4[3]);
                    const Larfg<T> h = larfg(rdlane(f[j], kk), rdlane(tl, kk));
                    if (tau && c == 0) tau[(long long)kk * p.cV] = h.tau;
                    if (c == kk) {
                        live_rows<j + 1, MMAX>(live, [&](auto i) {
                            f[i] = f[i] * h.inv;
                            vb[i] = f[i];
                        });
                        f[j] = h.rdiag;
                    }
                    wave_lds_sync();
                    /* w = f[j] + sum v_i f[i], four partial sums; v_i from the
                     * wavefront's LDS row, read by every lane at once */
                    T acc[4] = {f[j], T(0), T(0), T(0)};
                    live_rows<j + 1, MMAX>(live, [&](auto i) { acc[i % 4] = fmaT(vb[i], f[i], acc[i % 4]); });
                    T wv = (acc[0] + acc[1]) + (acc[2] + acc[3]);
                    wv = (c > kk) ? wv * h.tau : T(0);
                    f[j] = f[j] - wv;
                    live_rows<j + 1, MMAX>(live, [&](auto i) { f[i] = fmaT(-vb[i], wv, f[i]); });
                    own = (c == kk) ? f[j] : own;
                    if (has) ocol[(long long)kk * ors] = f[j];
                    wave_lds_sync(); /* every read of vb before the next write */
                }
                __builtin_amdgcn_sched_barrier(0);
            });
            if (k0 + G < n) {
                live_rows<0, MMAX - G>(live, [&](auto i) { f[i] = f[i + G]; });
                sfor<MMAX - G, MMAX>([&](auto i) { f[i] = T(0); });
            }
        }
        if (has && !xout) {
            /* rows n..m-1: Q^T B below the solution, reflector bodies of A */
            const int kb = n > 0 ? ((n - 1) / G) * G : 0; /* the last window's first row */
            sfor<0, MMAX>([&](auto i) {
                const int row = kb + i;
                if (row >= n && row < m) col[(long long)row * rs] = f[i];
            });
        }

        /* info: the first exact zero on R's diagonal (LAPACK dtrtrs's test) */
        int bad = 0;
        for (int l = 0; l < n; ++l) {
            if (rdlane(own, l) == T(0)) {
                bad = l + 1;
                break;
            }
        }
        if (c == 0 && p.info && d.idx < p.ninfo) p.info[d.idx] = bad;
        if (bad != 0 && p.info) continue; /* LAPACK dgels stops; without info, solve as gels_compact */

        /* R X = (Q^T B)(0:n), column-oriented from row n-1 up, G steps per
         * window: g[i] = row lb - i of the lane's column (R's column for an A
         * lane, the right-hand side for a B lane); step l = lb - j takes
         * x_l = b_l / R(l,l), then b_r -= R(r,l) x_l for rows r < l, with
         * R(r,l) broadcast from lane l's window through LDS */
        T g[MMAX];
        sfor<0, MMAX>([&](auto i) { g[i] = T(0); });
        if (has) {
            sfor<0, MMAX>([&](auto i) {
                if (i < n) g[i] = ocol[(long long)(n - 1 - i) * ors];
            });
        }
        for (int lb = n - 1; lb >= 0; lb -= G) {
            sfor<0, G>([&](auto jc) {
                constexpr int j = jc;
                const int l = lb - j;
                if (l >= 0) {
                    const T dl = rdlane(own, l);
                    const T x = isB ? g[j] / dl : T(0);
                    if (isB) ocol[(long long)l * ors] = x;
                    if (c == l) live_rows<j + 1, MMAX>(lb + 1, [&](auto i) { vb[i] = g[i]; });
                    wave_lds_sync();
                    live_rows<j + 1, MMAX>(lb + 1, [&](auto i) { g[i] = fmaT(-vb[i], x, g[i]); });
                    g[j] = isB ? x : g[j];
                    wave_lds_sync();
                }
                __builtin_amdgcn_sched_barrier(0);
            });
            if (lb - G >= 0) {
                live_rows<0, MMAX - G>(lb + 1, [&](auto i) { g[i] = g[i + G]; });
                sfor<MMAX - G, MMAX>([&](auto i) { g[i] = T(0); });
            }
        }
    }
}

/* ---- wavefront kernel, two problems per wavefront: n + nrhs <= 32 ----------
 * Lanes 0-31 take one problem of the queue and lanes 32-63 the next, each
 * half with the windows and LDS broadcasts of gels_wave_kernel (an LDS row
 * per half). The loops run to the larger of the two bounds; a half whose
 * problem is shorter idles through the extra steps with zero updates (its
 * LDS row keeps finite values from its own earlier steps, or the zeros it
 * starts with, so a zero weight leaves its columns unchanged). Rows that
 * leave a window below a half's last reflector are final and stored then. */
template <class T, int MMAX>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? (MMAX > 32 ? 2 : 4) : (MMAX > 32 ? 4 : 6)))) void
gels_wave_pair_kernel(BatchArgs<T> p)
{
    constexpr int G = 8;
    __shared__ T vshare[4][2][MMAX];
    const int lane = threadIdx.x & 63;
    const int hf = lane >> 5, c = lane & 31;
    T *vb = vshare[wave_in_block()][hf];
    sfor<0, (MMAX + 31) / 32>([&](auto r) {
        if (c + 32 * r < MMAX) vb[c + 32 * r] = T(0);
    });
    wave_lds_sync();
    auto both = [&](int x, auto op) {
        return op(__builtin_amdgcn_readlane(x, 0), __builtin_amdgcn_readlane(x, 32));
    };
    auto imax = [](int a, int b) { return a > b ? a : b; };
    const unsigned count = batch_count(p);
    for (;;) {
        unsigned k0q = 0;
        if (lane == 0) k0q = atomicAdd(p.counter, 2u);
        k0q = __builtin_amdgcn_readfirstlane(k0q);
        if (k0q >= count) return;
        const unsigned kq = k0q + (unsigned)hf;
        const bool real = kq < count;
        const BatchProblem d = p.desc[real ? kq : k0q];
        const int m = real ? d.m : 0, n = real ? d.n : 0, w = real ? d.n + d.nrhs : 0;
        const int nmax = both(n, imax), mmax = both(m, imax);
        const bool isA = c < n, isB = c >= n && c < w;
        const bool has = isA || isB;
        T *col = isA ? p.a + d.a_off + (long long)c * d.lda
                     : p.b + d.b_off + (long long)(isB ? c - n : 0) * d.ldb;

        T f[MMAX]; /* window: f[i] = row k0 + i */
        sfor<0, MMAX>([&](auto i) { f[i] = T(0); });
        if (has) {
            sfor<0, MMAX>([&](auto i) {
                if (i < m) f[i] = col[i];
            });
        }
        T own = T(0);
        for (int k0 = 0; k0 < nmax; k0 += G) {
            const int live = mmax - k0;
            sfor<0, G>([&](auto jc) {
                constexpr int j = jc;
                const int kk = k0 + j;
                if (kk < nmax) {
                    const bool act = kk < n;
                    T t4[4] = {T(0), T(0), T(0), T(0)};
                    live_rows<j + 1, MMAX>(live, [&](auto i) { t4[i % 4] = fmaT(f[i], f[i], t4[i % 4]); });
                    const T tl = (t4[0] + t4[1]) + (t4[2] + t4[3]);
                    const T x0 = hf ? rdlane(f[j], 32 + kk) : rdlane(f[j], kk);
                    const T tt = hf ? rdlane(tl, 32 + kk) : rdlane(tl, kk);
                    const Larfg<T> h = larfg(x0, tt);
                    if (act && c == kk) {
                        live_rows<j + 1, MMAX>(live, [&](auto i) {
                            f[i] = f[i] * h.inv;
                            vb[i] = f[i];
                        });
                        f[j] = h.rdiag;
                    }
                    wave_lds_sync();
                    T acc[4] = {f[j], T(0), T(0), T(0)};
                    live_rows<j + 1, MMAX>(live, [&](auto i) { acc[i % 4] = fmaT(vb[i], f[i], acc[i % 4]); });
                    T wv = (acc[0] + acc[1]) + (acc[2] + acc[3]);
                    wv = (act && c > kk) ? wv * h.tau : T(0);
                    f[j] = f[j] - wv;
                    live_rows<j + 1, MMAX>(live, [&](auto i) { f[i] = fmaT(-vb[i], wv, f[i]); });
                    own = (act && c == kk) ? f[j] : own;
                    if (has && act) col[kk] = f[j];
                    wave_lds_sync();
                }
                __builtin_amdgcn_sched_barrier(0);
            });
            if (k0 + G < nmax) {
                if (has) {
                    sfor<0, G>([&](auto i) {
                        const int row = k0 + i;
                        if (row >= n && row < m) col[row] = f[i];
                    });
                }
                live_rows<0, MMAX - G>(live, [&](auto i) { f[i] = f[i + G]; });
                sfor<MMAX - G, MMAX>([&](auto i) { f[i] = T(0); });
            }
        }
        if (has) {
            const int kb = nmax > 0 ? ((nmax - 1) / G) * G : 0;
            sfor<0, MMAX>([&](auto i) {
                const int row = kb + i;
                if (row >= n && row < m) col[row] = f[i];
            });
        }

        int bad = 0;
        for (int l = 0; l < nmax; ++l) {
            const T o = hf ? rdlane(own, 32 + l) : rdlane(own, l);
            if (bad == 0 && l < n && o == T(0)) bad = l + 1;
        }
        if (c == 0 && real && p.info) p.info[d.idx] = bad;
        /* LAPACK dgels stops at a zero diagonal when info is wanted */
        const int nb = (bad != 0 && p.info) ? 0 : n;
        const int nbmax = both(nb, imax);

        T g[MMAX]; /* window: g[i] = row lb - i */
        sfor<0, MMAX>([&](auto i) { g[i] = T(0); });
        if (has) {
            sfor<0, MMAX>([&](auto i) {
                const int row = nbmax - 1 - i;
                if (row >= 0 && row < nb) g[i] = col[row];
            });
        }
        for (int lb = nbmax - 1; lb >= 0; lb -= G) {
            sfor<0, G>([&](auto jc) {
                constexpr int j = jc;
                const int l = lb - j;
                if (l >= 0) {
                    const bool act = l < nb;
                    const T dl = hf ? rdlane(own, 32 + l) : rdlane(own, l);
                    const T x = (isB && act) ? g[j] / dl : T(0);
                    if (isB && act) col[l] = x;
                    if (act && c == l) live_rows<j + 1, MMAX>(lb + 1, [&](auto i) { vb[i] = g[i]; });
                    wave_lds_sync();
                    live_rows<j + 1, MMAX>(lb + 1, [&](auto i) { g[i] = fmaT(-vb[i], x, g[i]); });
                    g[j] = (isB && act) ? x : g[j];
                    wave_lds_sync();
                }
                __builtin_amdgcn_sched_barrier(0);
            });
            if (lb - G >= 0) {
                live_rows<0, MMAX - G>(lb + 1, [&](auto i) { g[i] = g[i + G]; });
                sfor<MMAX - G, MMAX>([&](auto i) { g[i] = T(0); });
            }
        }
    }
}

/* ---- wavefront kernel, two arrays: n <= 64, nrhs <= 64, m <= 64 ------------
 * For n + nrhs > 64 the columns of [A | B] outnumber the lanes: lane c holds
 * A's column c (c < n) in f and B's column c (c < nrhs) in fb. Both take the
 * same reflectors (one pair of v_readlane per row serves both), with the
 * rolling windows of gels_wave_kernel. */
template <class T>
__global__ __launch_bounds__(256) void gels_wave2_kernel(BatchArgs<T> p)
{
    constexpr int MMAX = 64;
    const int c = threadIdx.x & 63;
    const unsigned count = batch_count(p);
    for (;;) {
        unsigned kq = 0;
        if (c == 0) kq = atomicAdd(p.counter, 1u);
        kq = __builtin_amdgcn_readfirstlane(kq);
        if (kq >= count) return;
        const BatchProblem d = p.desc[kq];
        const int m = d.m, n = d.n;
        const bool isA = c < n, isB = c < d.nrhs;
        T *acol = p.a + d.a_off + (long long)(isA ? c : 0) * d.lda;
        T *bcol = p.b + d.b_off + (long long)(isB ? c : 0) * d.ldb;

        T f[MMAX], fb[MMAX];
        sfor<0, MMAX>([&](auto i) {
            f[i] = T(0);
            fb[i] = T(0);
        });
        if (isA) {
            sfor<0, MMAX>([&](auto i) {
                if (i < m) f[i] = acol[i];
            });
        }
        if (isB) {
            sfor<0, MMAX>([&](auto i) {
                if (i < m) fb[i] = bcol[i];
            });
        }

        T own = T(0);
        for (int kk = 0; kk < n; ++kk) {
            const int live = m - kk;
            T tl = T(0);
            live_rows<1, MMAX>(live, [&](auto i) { tl = fmaT(f[i], f[i], tl); });
            const Larfg<T> h = larfg(rdlane(f[0], kk), rdlane(tl, kk));
            if (c == kk) {
                live_rows<1, MMAX>(live, [&](auto i) { f[i] = f[i] * h.inv; });
                f[0] = h.rdiag;
            }
            T acc[4] = {f[0], T(0), T(0), T(0)};
            T accb[4] = {fb[0], T(0), T(0), T(0)};
            live_rows<1, MMAX>(live, [&](auto i) {
                const T fi = (i > 4) ? after(f[i], acc[i % 4]) : f[i];
                const T v = rdlane(fi, kk);
                acc[i % 4] = fmaT(v, fi, acc[i % 4]);
                accb[i % 4] = fmaT(v, fb[i], accb[i % 4]);
            });
            T wv = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            T wb = (accb[0] + accb[1]) + (accb[2] + accb[3]);
            wv = (c > kk) ? wv * h.tau : T(0);
            wb = isB ? wb * h.tau : T(0);
            f[0] = f[0] - wv;
            fb[0] = fb[0] - wb;
            live_rows<1, MMAX>(live, [&](auto i) {
                const T fi = after(f[i], (i > 4) ? f[i - 4] : wv);
                const T v = rdlane(fi, kk);
                f[i] = fmaT(-v, wv, fi);
                fb[i] = fmaT(-v, wb, fb[i]);
            });
            own = (c == kk) ? f[0] : own;
            if (isA) acol[kk] = f[0];
            if (isB) bcol[kk] = fb[0];
            live_rows<0, MMAX - 1>(live, [&](auto i) {
                f[i] = f[i + 1];
                fb[i] = fb[i + 1];
            });
            f[MMAX - 1] = T(0);
            fb[MMAX - 1] = T(0);
        }
        sfor<0, MMAX>([&](auto i) {
            if (i < m - n) {
                if (isA) acol[n + i] = f[i];
                if (isB) bcol[n + i] = fb[i];
            }
        });

        int bad = 0;
        for (int l = 0; l < n; ++l) {
            if (rdlane(own, l) == T(0)) {
                bad = l + 1;
                break;
            }
        }
        if (c == 0 && p.info) p.info[d.idx] = bad;
        if (bad != 0 && p.info) continue; /* LAPACK dgels stops; without info, solve as gels_compact */

        /* R X = (Q^T B)(0:n): g (R's columns) and gb (the right-hand sides),
         * element j = row l - j at step l */
        T g[MMAX], gb[MMAX];
        sfor<0, MMAX>([&](auto j) {
            g[j] = T(0);
            gb[j] = T(0);
        });
        if (isA) {
            sfor<0, MMAX>([&](auto j) {
                if (j < n) g[j] = acol[n - 1 - j];
            });
        }
        if (isB) {
            sfor<0, MMAX>([&](auto j) {
                if (j < n) gb[j] = bcol[n - 1 - j];
            });
        }
        for (int l = n - 1; l >= 0; --l) {
            const T dl = rdlane(own, l);
            const T x = isB ? gb[0] / dl : T(0);
            if (isB) bcol[l] = x;
            live_rows<1, MMAX>(l + 1, [&](auto j) {
                const T gj = after(g[j], (j > 4) ? gb[j - 4] : x);
                gb[j] = fmaT(-rdlane(gj, l), x, gb[j]);
            });
            live_rows<0, MMAX - 1>(l + 1, [&](auto j) {
                g[j] = g[j + 1];
                gb[j] = gb[j + 1];
            });
            g[MMAX - 1] = T(0);
            gb[MMAX - 1] = T(0);
        }
    }
}

template <class T> hipError_t launch_class(int cls, const BatchArgs<T> &args, unsigned blocks,
                                           hipStream_t s)
{
    const dim3 grid(blocks), block(256);
    switch (cls) {
    case 0: hipLaunchKernelGGL((gels_team_kernel<T, 16>), grid, block, 0, s, args); break;
    case 1: hipLaunchKernelGGL((gels_team_kernel<T, 32>), grid, block, 0, s, args); break;
    case 2: hipLaunchKernelGGL((gels_team_kernel<T, 64>), grid, block, 0, s, args); break;
    case 3: hipLaunchKernelGGL((gels_wave_kernel<T, 32>), grid, block, 0, s, args); break;
    case 4: hipLaunchKernelGGL((gels_wave_kernel<T, 64>), grid, block, 0, s, args); break;
    case 5: hipLaunchKernelGGL((gels_wave2_kernel<T>), grid, block, 0, s, args); break;
    case 7: hipLaunchKernelGGL((gels_wave_pair_kernel<T, 32>), grid, block, 0, s, args); break;
    case 8: hipLaunchKernelGGL((gels_wave_pair_kernel<T, 64>), grid, block, 0, s, args); break;
    case 9:
    case 10: return launch_qr_grid_batch<T>(cls, args, blocks, s);
    case 11: return launch_qr_grid64_batch<T>(args, blocks, s);
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

} /* namespace batch */

template <class T>
hipError_t launch_gels_batch(int cls, const BatchArgs<T> &args, unsigned blocks, hipStream_t s)
{
    return batch::launch_class<T>(cls, args, blocks, s);
}

template <class T>
hipError_t launch_qr_wave(int p, int q, int nrhs, bool transposed, Operand<T> A, T *tau,
                          Operand<T> B, int V, int nm, int *info)
{
    BatchArgs<T> args;
    int si = A.row ? A.ld : 1, sj = A.row ? 1 : A.ld;
    if (transposed) {
        const int tmp = si;
        si = sj;
        sj = tmp;
    }
    args.a = A.p;
    args.b = B.p;
    args.info = info;
    args.desc = nullptr;
    args.counter = nullptr;
    args.count = ((nm + V - 1) / V) * V; /* padded slots too, as the reference */
    args.tau = nullptr;
    args.tran = false;
    args.a_rs = si * V;
    args.a_cs = sj * V;
    args.b_rs = (B.row ? B.ld : 1) * V;
    args.b_cs = (B.row ? 1 : B.ld) * V;
    args.ninfo = nm;
    args.cm = p;
    args.cn = q;
    args.cnrhs = nrhs;
    args.cV = V;
    args.a_gs = (long long)A.gstride;
    args.b_gs = (long long)B.gstride;
    args.ctau = tau;
    args.mask = nullptr;
    args.x = nullptr;
    args.x_rs = args.x_cs = 0;
    args.x_gs = 0;
    args.dcount = nullptr;
    const unsigned blocks = team::xcd_grid((unsigned)((args.count + 3) / 4));
    if (p <= 32)
        hipLaunchKernelGGL((batch::gels_wave_kernel<T, 32>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    else
        hipLaunchKernelGGL((batch::gels_wave_kernel<T, 64>), dim3(blocks), dim3(256), 0,
                           current_stream(), args);
    return hipGetLastError();
}

template <class T> hipError_t gels_batch_occupancy(int cls, int *blocks)
{
    using namespace batch;
    switch (cls) {
    case 0: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_team_kernel<T, 16>, 256, 0);
    case 1: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_team_kernel<T, 32>, 256, 0);
    case 2: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_team_kernel<T, 64>, 256, 0);
    case 3: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_wave_kernel<T, 32>, 256, 0);
    case 4: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_wave_kernel<T, 64>, 256, 0);
    case 5: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_wave2_kernel<T>, 256, 0);
    case 7: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_wave_pair_kernel<T, 32>, 256, 0);
    case 8: return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, gels_wave_pair_kernel<T, 64>, 256, 0);
    case 9:
    case 10: return qr_grid_batch_occupancy<T>(cls, blocks);
    case 11: return qr_grid64_batch_occupancy<T>(blocks);
    default: *blocks = 0; return hipErrorInvalidValue;
    }
}

} /* namespace cbk */

#endif /* CBK_BATCH_GELS_IMPL_H */
