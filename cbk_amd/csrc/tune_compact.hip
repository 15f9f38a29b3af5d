/* tune_compact.hip -- the per-shape kernel-family table of the compact
 * routines and the tuner that fills it (include/cbk_ext.h,
 * cbk_autotune_compact). No reference counterpart: the reference's kernels
 * have one code path per routine, with compile-time block sizes
 * (src/cbk_potrf_compact.hpp:71-74, src/cbk_common.hpp:272-275). */

#include "cbk.h"
#include "cbk_ext.h"
#include "cbk_internal.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

namespace cbk {
namespace {

inline bool opt(char c, char upper) { return c == upper || c == upper + ('a' - 'A'); }
inline bool vlen_ok(int V) { return V == 2 || V == 4 || V == 8 || V == 16 || V == 32 || V == 64; }

using Key = std::tuple<int, int, int, int, int, int, int>; /* device, class, fp64, p, q, nrhs, V */

std::mutex g_mu;
std::map<Key, int> g_table;
std::atomic<int> g_entries{0};

int current_device()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    return dev;
}

void put_family(int cls, bool fp64, int p, int q, int nrhs, int V, int family)
{
    const Key k{current_device(), cls, fp64, p, q, nrhs, V};
    std::lock_guard<std::mutex> lk(g_mu);
    if (family == CBK_FAMILY_DEFAULT)
        g_table.erase(k);
    else
        g_table[k] = family;
    g_entries.store((int)g_table.size());
}

/* the table key of a tuner call, or a negative code */
struct Shape {
    int cls;
    bool fp64;
    int p, q, nrhs;
};

int shape_of(char prec, const char *routine, char uplo, char trans, int m, int n, int nrhs,
             Shape &s)
{
    if (!opt(prec, 'D') && !opt(prec, 'S')) return -1;
    if (!routine) return -2;
    s.fp64 = opt(prec, 'D');
    const bool up = opt(uplo, 'U'), lo = opt(uplo, 'L');
    const bool tr = opt(trans, 'T') || opt(trans, 'C'), nt = opt(trans, 'N');
    if (!strcmp(routine, "posv") || !strcmp(routine, "sysvnp")) {
        s.cls = routine[1] == 'o' ? kTuneChol : kTuneLdl;
        if (n < 1 || n > 64) return -6;
        if (nrhs < 0 || nrhs > 64) return -7;
        s.p = n;
        s.q = 0;
        s.nrhs = nrhs;
        return 0;
    }
    if (!strcmp(routine, "trsm")) {
        if (!(up || lo) || !(tr || nt)) return -3;
        if (m < 1 || m > 64) return -5;
        if (n < 1 || n > 64) return -6;
        s.cls = (up == tr) ? kTuneTrsmFwd : kTuneTrsmBwd;
        s.p = m;
        s.q = n;
        s.nrhs = 0;
        return 0;
    }
    if (!strcmp(routine, "gels") || !strcmp(routine, "geqrf")) {
        const bool gels = routine[1] == 'e' && routine[2] == 'l';
        if (m < 1 || m > 64) return -5;
        if (n < 1 || n > m) return -6;
        if (gels && (nrhs < 1 || n + nrhs > 64)) return -7;
        s.cls = gels ? kTuneGels : kTuneGeqrf;
        s.p = m;
        s.q = n;
        s.nrhs = gels ? nrhs : 0;
        return 0;
    }
    return -2;
}

/* the families that can serve the shape at V (the dispatcher's predicates) */
std::vector<int> candidates(const Shape &s, int V)
{
    std::vector<int> f;
    switch (s.cls) {
    case kTuneChol:
    case kTuneLdl:
        if (reg_chol_supports(s.p, s.fp64)) f.push_back(CBK_FAMILY_REGISTER);
        if (quad_sym_supports(s.p, s.fp64)) f.push_back(CBK_FAMILY_QUAD);
        if (sym_team_supports(s.p) && V >= 4) f.push_back(CBK_FAMILY_TEAM);
        if (chol_grid_supports(s.p)) f.push_back(CBK_FAMILY_GRID);
        break;
    case kTuneTrsmFwd:
    case kTuneTrsmBwd: {
        const bool up = false, tran = s.cls == kTuneTrsmBwd; /* a representative of each sweep */
        if (trsm_reg_supports(true, s.p, s.fp64)) f.push_back(CBK_FAMILY_REGISTER);
        if (quad_trsm_supports(true, up, tran, s.p, s.fp64)) f.push_back(CBK_FAMILY_QUAD);
        if (V >= 4 && trsm_team_supports(true, up, tran, s.p)) f.push_back(CBK_FAMILY_TEAM);
        if (trsm_grid_supports(true, up, tran, s.p)) f.push_back(CBK_FAMILY_GRID);
        break;
    }
    case kTuneGels:
    case kTuneGeqrf: {
        const int r = s.nrhs;
        if (qr_reg_supports(s.p, s.q, s.fp64)) f.push_back(CBK_FAMILY_REGISTER);
        if (V >= 4 && qr_team_supports(s.p, s.q, r, s.fp64)) f.push_back(CBK_FAMILY_TEAM);
        if (V >= 4 && qr_cols_supports(s.p, s.q, r)) f.push_back(CBK_FAMILY_COLUMNS);
        if (s.cls == kTuneGels && qr_wave_supports(s.p, s.q, r)) f.push_back(CBK_FAMILY_WAVE);
        if (qr_grid_supports(s.p, s.q, r)) f.push_back(CBK_FAMILY_GRID);
        break;
    }
    }
    f.push_back(CBK_FAMILY_LANE);
    return f;
}

/* ---- synthetic operands, filled on the device ---------------------------- */

__device__ __forceinline__ double unit_hash(unsigned long long x)
{
    /* SplitMix64 finalizer, then 53 bits to [-1, 1) */
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * 0x1p-52 - 1.0;
}

/* kind 0: symmetric, diagonal rows + 1 (diagonally dominant: SPD, and an
 * unpivoted LDL^T without growth); 1: lower triangle, diagonal 2 + U/2,
 * off-diagonal U/(2 rows); 2: uniform (-1, 1). Column-major compact, ld =
 * rows, every slot (padding included). */
template <class T>
__global__ void fill_compact(T *a, int rows, int cols, int V, long total, int kind, unsigned seed)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const long v = idx % V, rest = idx / V;
    const long i = rest % rows, j = (rest / rows) % cols, g = rest / ((long)rows * cols);
    const long mat = g * V + v;
    const long lo = i < j ? i : j, hi = i < j ? j : i;
    double x;
    if (kind == 0) {
        x = unit_hash(((unsigned long long)mat << 16 | (unsigned long long)(lo * 64 + hi)) ^
                      ((unsigned long long)seed << 48));
        if (i == j) x = (double)rows + 1.0 + 0.5 * x;
    }
    else if (kind == 1) {
        x = unit_hash(((unsigned long long)mat << 16 | (unsigned long long)(i * 64 + j)) ^
                      ((unsigned long long)seed << 48));
        x = (i == j) ? 2.0 + 0.5 * x : (i > j ? x / (2.0 * rows) : 0.0);
    }
    else {
        x = unit_hash(((unsigned long long)mat << 16 | (unsigned long long)(i * 64 + j)) ^
                      ((unsigned long long)seed << 48));
    }
    a[idx] = (T)x;
}

template <class T>
hipError_t fill(T *a, int rows, int cols, int V, long groups, int kind, unsigned seed,
                hipStream_t s)
{
    const long total = groups * V * rows * cols;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_compact<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a,
                       rows, cols, V, total, kind, seed);
    return hipGetLastError();
}

/* one timed call of the routine on (A, B, tau) at V; 0 or the routine's code */
template <class T>
int run_once(const Shape &s, char uplo, char trans, T *A, T *B, T *tau, int V, int nm)
{
    const bool d = sizeof(T) == 8;
    const int p = s.p, q = s.q, r = s.nrhs;
    switch (s.cls) {
    case kTuneChol:
        return d ? dposv_compact('C', uplo, p, r, (double *)A, p, (double *)B, p, V, nm)
                 : sposv_compact('C', uplo, p, r, (float *)A, p, (float *)B, p, V, nm);
    case kTuneLdl:
        return d ? dsysvnp_compact('C', uplo, p, r, (double *)A, p, (double *)B, p, V, nm)
                 : ssysvnp_compact('C', uplo, p, r, (float *)A, p, (float *)B, p, V, nm);
    case kTuneTrsmFwd:
    case kTuneTrsmBwd:
        return d ? dtrsm_compact('C', 'L', uplo, trans, 'N', p, q, 1.0, (double *)A, p,
                                 (double *)B, p, V, nm)
                 : strsm_compact('C', 'L', uplo, trans, 'N', p, q, 1.0f, (float *)A, p, (float *)B,
                                 p, V, nm);
    case kTuneGels:
        return d ? dgels_compact('C', 'N', p, q, r, (double *)A, p, (double *)B, p, (double *)tau,
                                 V, nm)
                 : sgels_compact('C', 'N', p, q, r, (float *)A, p, (float *)B, p, (float *)tau, V,
                                 nm);
    default:
        return d ? dgeqrf_compact('C', p, q, (double *)A, p, (double *)tau, V, nm)
                 : sgeqrf_compact('C', p, q, (float *)A, p, (float *)tau, V, nm);
    }
}

/* time every candidate at V (best of 3 after a warm-up), install the fastest;
 * *best_ms the winner's time */
template <class T>
int tune_at(const Shape &s, char uplo, char trans, int V, int nm, int *family, float *best_ms)
{
    const long groups = (nm + V - 1) / V;
    int arows = s.p, acols = s.cls == kTuneGels || s.cls == kTuneGeqrf ? s.q : s.p;
    int bcols = (s.cls == kTuneTrsmFwd || s.cls == kTuneTrsmBwd) ? s.q : s.nrhs;
    const int akind = (s.cls == kTuneChol || s.cls == kTuneLdl) ? 0
                      : (s.cls == kTuneTrsmFwd || s.cls == kTuneTrsmBwd) ? 1
                                                                         : 2;
    const size_t na = (size_t)groups * V * arows * acols, nb = (size_t)groups * V * s.p * bcols,
                 nt = (size_t)groups * V * std::max(1, s.q);
    T *A = nullptr, *A0 = nullptr, *B = nullptr, *B0 = nullptr, *tau = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    void *prev = cbk_get_stream();
    hipError_t e = hipMalloc(&A, na * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&A0, na * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&B, std::max<size_t>(nb, 1) * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&B0, std::max<size_t>(nb, 1) * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&tau, nt * sizeof(T));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = fill<T>(A0, arows, acols, V, groups, akind, 2026u, st);
    if (e == hipSuccess && nb) e = fill<T>(B0, s.p, bcols, V, groups, 2, 2027u, st);
    int code = 0;
    float best = 1e30f;
    int winner = CBK_FAMILY_DEFAULT;
    if (e == hipSuccess) {
        cbk_set_stream(st);
        for (int f : candidates(s, V)) {
            put_family(s.cls, s.fp64, s.p, s.q, s.nrhs, V, f);
            float t_f = 1e30f;
            for (int rep = 0; rep < 4 && e == hipSuccess && code == 0; ++rep) {
                e = hipMemcpyAsync(A, A0, na * sizeof(T), hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess && nb)
                    e = hipMemcpyAsync(B, B0, nb * sizeof(T), hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess) e = hipEventRecord(e0, st);
                if (e == hipSuccess) code = run_once<T>(s, uplo, trans, A, B, tau, V, nm);
                if (e == hipSuccess) e = hipEventRecord(e1, st);
                if (e == hipSuccess) e = hipEventSynchronize(e1);
                float ms = 0.f;
                if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
                if (e == hipSuccess && rep > 0) t_f = std::min(t_f, ms);
            }
            if (e != hipSuccess || code != 0) break;
            if (t_f < best) {
                best = t_f;
                winner = f;
            }
        }
        cbk_set_stream(prev);
    }
    if (e == hipSuccess && code == 0) put_family(s.cls, s.fp64, s.p, s.q, s.nrhs, V, winner);
    else put_family(s.cls, s.fp64, s.p, s.q, s.nrhs, V, CBK_FAMILY_DEFAULT);
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    for (T *x : {A, A0, B, B0, tau})
        if (x) (void)hipFree(x);
    if (e != hipSuccess) return launch_status(e);
    if (code != 0) return code;
    *family = winner;
    *best_ms = best;
    return 0;
}

} /* namespace */

int tuned_family(int cls, bool fp64, int p, int q, int nrhs, int V)
{
    if (g_entries.load(std::memory_order_relaxed) == 0) return CBK_FAMILY_DEFAULT;
    const Key k{current_device(), cls, fp64, p, q, nrhs, V};
    std::lock_guard<std::mutex> lk(g_mu);
    const auto it = g_table.find(k);
    return it == g_table.end() ? CBK_FAMILY_DEFAULT : it->second;
}

} /* namespace cbk */

using namespace cbk;

extern "C" {

int cbk_autotune_compact(char prec, const char *routine, char uplo, char trans, int m, int n,
                         int nrhs, int V, int nm, int *family, int *vbest)
{
    Shape s;
    const int code = shape_of(prec, routine, uplo, trans, m, n, nrhs, s);
    if (code != 0) return code;
    if (V != 0 && !vlen_ok(V)) return -8;
    if (nm < 0) return -9;
    if (nm == 0) nm = 1 << 17;
    const char u = opt(uplo, 'U') ? 'U' : 'L', t = (opt(trans, 'T') || opt(trans, 'C')) ? 'T' : 'N';
    const int vs[3] = {16, 32, 64};
    float best_rate = 0.f;
    int best_f = CBK_FAMILY_DEFAULT, best_v = V;
    for (int i = 0; i < (V == 0 ? 3 : 1); ++i) {
        const int v = V == 0 ? vs[i] : V;
        int f = 0;
        float ms = 0.f;
        const int r = s.fp64 ? tune_at<double>(s, u, t, v, nm, &f, &ms)
                             : tune_at<float>(s, u, t, v, nm, &f, &ms);
        if (r != 0) return r;
        const float rate = (float)nm / ms;
        if (rate > best_rate) {
            best_rate = rate;
            best_f = f;
            best_v = v;
        }
    }
    if (family) *family = best_f;
    if (vbest) *vbest = best_v;
    return 0;
}

int cbk_set_kernel_family(char prec, const char *routine, char uplo, char trans, int m, int n,
                          int nrhs, int V, int family)
{
    Shape s;
    const int code = shape_of(prec, routine, uplo, trans, m, n, nrhs, s);
    if (code != 0) return code;
    if (!vlen_ok(V)) return -8;
    if (family < CBK_FAMILY_DEFAULT || family > CBK_FAMILY_GRID) return -10;
    put_family(s.cls, s.fp64, s.p, s.q, s.nrhs, V, family);
    return 0;
}

int cbk_get_kernel_family(char prec, const char *routine, char uplo, char trans, int m, int n,
                          int nrhs, int V)
{
    Shape s;
    const int code = shape_of(prec, routine, uplo, trans, m, n, nrhs, s);
    if (code != 0) return code;
    if (!vlen_ok(V)) return -8;
    return tuned_family(s.cls, s.fp64, s.p, s.q, s.nrhs, V);
}

} /* extern "C" */
