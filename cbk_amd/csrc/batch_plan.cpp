/* batch_plan.cpp -- uneven batches: the plan (size classes, queue order,
 * workspace layout, the recorded hipGraph), the dgels_batch / sgels_batch /
 * dsgels_batch entry points and the tuner that picks each class's kernel
 * (include/cbk_ext.h declares the ABI; no reference counterpart).
 *
 * Workspace layout of a plain plan (device, caller-owned):
 *   [0, 256)                 one 32-bit work counter per class
 *   [256, 256 + 48 nprob)    the problem table, grouped by class, each class
 *                            ordered by estimated work, largest first
 *   then, 256-aligned        tau scratch of the general kernel's problems
 *
 * A mixed plan (dsgels_batch) keeps, each 256-aligned after the counters
 * (FP32 class queues, the FP64 queues of the problems without an FP32 path
 * and of the fallback lists, the fallback lists' fills): the FP32 table of
 * the problems the FP32 path serves, the FP64 lists (one segment per kernel:
 * the problems without an FP32 path, fixed at plan time, then room for that
 * kernel's refined problems that fall back), the refinement table grouped
 * by shape bucket, the list of the problems without an FP32 path, the FP32
 * info scratch, the FP64 tau scratch and the FP32 copies of A and B (at the
 * caller's offsets).
 */

#include "cbk.h"
#include "cbk_ext.h"
#include "cbk_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <new>
#include <random>
#include <vector>

using cbk::BatchArgs;
using cbk::BatchProblem;
using cbk::kBatchClasses;
using cbk::kBatchGeneral;
using cbk::kBatchKernels;

static_assert(CBK_BATCH_NCLASS == kBatchClasses, "class count");

namespace {

constexpr size_t kCounterBytes = 256;
constexpr int kMaxDevices = 64;

/* one tuning table per device; defaults: the team kernel for every problem
 * it holds, grids from the occupancy of each kernel */
struct TuningSlot {
    cbk_tuning t;
    bool set;
};
std::mutex g_tune_mu;
TuningSlot g_tune[kMaxDevices];

cbk_tuning default_tuning()
{
    cbk_tuning t;
    for (int b = 0; b < 3; ++b)
        t.narrow_max_cols[b] = 16;
    for (int c = 0; c < kBatchClasses; ++c)
        t.blocks_per_cu[c] = 0; /* 0: the kernel's occupancy */
    return t;
}

cbk_tuning tuning_for(int dev)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (dev >= 0 && dev < kMaxDevices && g_tune[dev].set) return g_tune[dev].t;
    return default_tuning();
}

inline bool opt(char c, char upper) { return c == upper || c == upper + ('a' - 'A'); }
inline int max1(int x) { return x < 1 ? 1 : x; }

int classify(const cbk_gels_problem &p, bool tran, const cbk_tuning &t)
{
    if (tran || p.m < p.n || p.n == 0 || p.nrhs == 0 || p.m > 64) return kBatchGeneral;
    const int w = p.n + p.nrhs;
    const int mb = p.m <= 16 ? 0 : p.m <= 32 ? 1 : 2;
    if (w <= 16 && w <= t.narrow_max_cols[mb]) return mb;
    if (w <= 64) return p.m <= 32 ? 3 : 4;
    return p.nrhs <= 64 ? 5 : kBatchGeneral;
}

/* Householder QR plus Q^T B plus the solve, in multiply-adds */
double work_of(const cbk_gels_problem &p)
{
    const double m = p.m, n = p.n, r = p.nrhs;
    return m * n * n - n * n * n / 3.0 + 2.0 * m * n * r + n * n * r;
}

/* the kernel a problem of class c runs on: the wavefront classes' problems
 * with n <= 32 and nrhs <= 8 on the lane-grid kernel (9, 10), two per
 * wavefront; the wider ones of classes 4 and 5 with nrhs <= 8 on the 8 x 8
 * lane grid (11), one per wavefront; the other problems of classes 3 and 4
 * with n + nrhs <= 32 two per wavefront (7, 8) */
int kernel_of(int c, const cbk_gels_problem &p)
{
    if ((c == 3 || c == 4) && p.n <= 32 && p.nrhs <= 8) return c + 6;
    if ((c == 4 || c == 5) && p.nrhs <= 8) return 11;
    return (c == 3 || c == 4) && p.n + p.nrhs <= 32 ? c + 4 : c;
}
/* the public class of kernel k (kernel 11 is tuned as class 5) */
int class_of_kernel(int k)
{
    return k == 11 ? 5 : k >= 9 ? k - 6 : k >= kBatchClasses ? k - 4 : k;
}

int problems_per_block(int k)
{
    return k <= 2 ? 16 : k <= 5 || k == 11 ? 4 : k == kBatchGeneral ? 256 : 8;
}

/* a problem the mixed path refines: trans 'N', 1 <= n <= m <= 64, n <= 32,
 * 1 <= nrhs <= 8 */
bool mixed_eligible(const cbk_gels_problem &p, bool tran)
{
    return !tran && p.n >= 1 && p.n <= p.m && p.m <= 64 && p.n <= 32 && p.nrhs >= 1 &&
           p.nrhs <= 8;
}

/* FP32 queues, FP64 queues of the problems without an FP32 path, FP64
 * queues of the fallback lists, the fallback lists' fills */
constexpr int kMixedCounters = 4 * kBatchKernels;

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

/* fbtable, per FP64 kernel c: the nstat[c] problems without an FP32 path
 * (fixed at plan time) from fb_begin[c], then room for the ndyn[c] refined
 * problems of kernel c that may fall back (filled on the device) */
struct MixedLayout {
    std::vector<BatchProblem> table32, fbtable, inel;
    std::vector<cbk::MixedProblem> reftable;
    int begin32[kBatchKernels + 1] = {};
    int fb_begin[kBatchKernels] = {}, nstat[kBatchKernels] = {}, ndyn[kBatchKernels] = {};
    int ref_begin[cbk::kRefineBuckets + 1] = {};
    long long ext_a = 0, ext_b = 0, tau64 = 0;
    size_t off32 = 0, offfb = 0, offref = 0, offinel = 0, offinfo = 0, offtau = 0, offa = 0,
           offb = 0;
};

} /* namespace */

struct cbk_batch_plan {
    char prec = 'D';
    bool tran = false;
    int nprob = 0;
    std::vector<BatchProblem> table;
    int begin[kBatchKernels + 1] = {}; /* per kernel */
    long long tau_elems = 0;
    size_t table_off = kCounterBytes, tau_off = 0, bytes = 0;
    cbk_tuning tune{};
    int device = 0;
    /* device state: the table upload and the recorded graph */
    void *uploaded = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *key[4] = {nullptr, nullptr, nullptr, nullptr};
    hipStream_t cap = nullptr, side[kBatchKernels] = {};
    hipEvent_t fork = nullptr, join[kBatchKernels] = {};
    /* mixed plans (prec 'M') */
    MixedLayout *mx = nullptr;
    void *mkey[6] = {};
    hipEvent_t fork2 = nullptr, join2[kBatchKernels] = {};
    hipEvent_t fork3 = nullptr, mjoin[kBatchKernels] = {}, rjoin[cbk::kRefineBuckets] = {};
    hipStream_t mside[kBatchKernels] = {}, rside[cbk::kRefineBuckets] = {};

    ~cbk_batch_plan()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (int c = 0; c < kBatchKernels; ++c) {
            if (side[c]) (void)hipStreamDestroy(side[c]);
            if (join[c]) (void)hipEventDestroy(join[c]);
        }
        if (cap) (void)hipStreamDestroy(cap);
        if (fork) (void)hipEventDestroy(fork);
        for (int c = 0; c < kBatchKernels; ++c) {
            if (join2[c]) (void)hipEventDestroy(join2[c]);
            if (mjoin[c]) (void)hipEventDestroy(mjoin[c]);
            if (mside[c]) (void)hipStreamDestroy(mside[c]);
        }
        for (int b = 0; b < cbk::kRefineBuckets; ++b) {
            if (rjoin[b]) (void)hipEventDestroy(rjoin[b]);
            if (rside[b]) (void)hipStreamDestroy(rside[b]);
        }
        if (fork2) (void)hipEventDestroy(fork2);
        if (fork3) (void)hipEventDestroy(fork3);
        delete mx;
    }
};

namespace {

hipError_t ensure_streams(cbk_batch_plan *pl)
{
    if (pl->cap) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->fork2, hipEventDisableTiming);
    for (int c = 0; c < kBatchKernels && e == hipSuccess; ++c) {
        e = hipStreamCreateWithFlags(&pl->side[c], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->join[c], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->join2[c], hipEventDisableTiming);
    }
    return e;
}

/* the arguments of one class launch over count problems at desc */
/* the extra branches of a mixed plan's graph */
hipError_t ensure_mixed_streams(cbk_batch_plan *pl)
{
    if (pl->fork3) return hipSuccess;
    hipError_t e = hipEventCreateWithFlags(&pl->fork3, hipEventDisableTiming);
    for (int c = 0; c < kBatchKernels && e == hipSuccess; ++c) {
        e = hipStreamCreateWithFlags(&pl->mside[c], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->mjoin[c], hipEventDisableTiming);
    }
    for (int b = 0; b < cbk::kRefineBuckets && e == hipSuccess; ++b) {
        e = hipStreamCreateWithFlags(&pl->rside[b], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->rjoin[b], hipEventDisableTiming);
    }
    return e;
}

template <class T>
BatchArgs<T> list_args(T *ap, T *bp, int *info, const BatchProblem *desc, int count,
                       unsigned *counter, T *tau, bool tran, const unsigned *dcount)
{
    BatchArgs<T> a;
    a.a = ap;
    a.b = bp;
    a.info = info;
    a.desc = desc;
    a.counter = counter;
    a.count = count;
    a.tau = tau;
    a.tran = tran;
    a.a_rs = a.b_rs = 1;
    a.ninfo = 0x7fffffff;
    a.cm = a.cn = a.cnrhs = a.cV = a.a_cs = a.b_cs = 0;
    a.a_gs = a.b_gs = 0;
    a.ctau = nullptr;
    a.mask = nullptr;
    a.x = nullptr;
    a.x_rs = a.x_cs = 0;
    a.x_gs = 0;
    a.dcount = dcount;
    return a;
}

template <class T>
BatchArgs<T> class_args(const cbk_batch_plan *pl, int c, T *ap, T *bp, int *info, char *work)
{
    return list_args<T>(ap, bp, info,
                        reinterpret_cast<const BatchProblem *>(work + pl->table_off) + pl->begin[c],
                        pl->begin[c + 1] - pl->begin[c], reinterpret_cast<unsigned *>(work) + c,
                        reinterpret_cast<T *>(work + pl->tau_off), pl->tran, nullptr);
}

/* blocks of class c's persistent grid: enough resident blocks to fill the
 * device (the tuned count per compute unit, else the kernel's occupancy), no
 * more than the class has problems for */
template <class T>
hipError_t class_blocks_n(const cbk_batch_plan *pl, int c, long count, unsigned *blocks)
{
    const long need = (count + problems_per_block(c) - 1) / problems_per_block(c);
    if (c == kBatchGeneral) {
        *blocks = (unsigned)need;
        return hipSuccess;
    }
    int per_cu = pl->tune.blocks_per_cu[class_of_kernel(c)];
    hipError_t e = hipSuccess;
    if (per_cu <= 0) e = cbk::gels_batch_occupancy<T>(c, &per_cu);
    if (e != hipSuccess) return e;
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, pl->device);
    if (e != hipSuccess) return e;
    const long have = (long)std::max(1, per_cu) * cus;
    *blocks = (unsigned)std::min(need, have);
    return hipSuccess;
}

template <class T> hipError_t class_blocks(const cbk_batch_plan *pl, int c, unsigned *blocks)
{
    return class_blocks_n<T>(pl, c, pl->begin[c + 1] - pl->begin[c], blocks);
}

template <class T>
hipError_t launch_args(int c, const BatchArgs<T> &args, unsigned blocks, hipStream_t s)
{
    if (c == kBatchGeneral) return cbk::launch_gels_batch_lane<T>(args, s);
    return cbk::launch_gels_batch<T>(c, args, blocks, s);
}

template <class T>
hipError_t launch_class(const cbk_batch_plan *pl, int c, unsigned blocks, T *ap, T *bp,
                        int *info, char *work, hipStream_t s)
{
    return launch_args<T>(c, class_args<T>(pl, c, ap, bp, info, work), blocks, s);
}

/* counters cleared by a kernel (no memset node), then every non-empty class
 * on its own branch of the graph */
template <class T>
hipError_t record_graph(cbk_batch_plan *pl, T *ap, T *bp, int *info, char *work)
{
    hipError_t e = ensure_streams(pl);
    if (e != hipSuccess) return e;
    unsigned blocks[kBatchKernels] = {};
    for (int c = 0; c < kBatchKernels && e == hipSuccess; ++c)
        if (pl->begin[c + 1] > pl->begin[c]) e = class_blocks<T>(pl, c, &blocks[c]);
    if (e != hipSuccess) return e;
    if (pl->exec) {
        (void)hipGraphExecDestroy(pl->exec);
        pl->exec = nullptr;
    }
    if (pl->graph) {
        (void)hipGraphDestroy(pl->graph);
        pl->graph = nullptr;
    }
    e = hipStreamBeginCapture(pl->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = cbk::launch_clear_u32(reinterpret_cast<unsigned *>(work), kBatchKernels, pl->cap);
    if (le == hipSuccess) le = hipEventRecord(pl->fork, pl->cap);
    for (int c = 0; c < kBatchKernels && le == hipSuccess; ++c) {
        if (pl->begin[c + 1] == pl->begin[c]) continue;
        le = hipStreamWaitEvent(pl->side[c], pl->fork, 0);
        if (le == hipSuccess) le = launch_class<T>(pl, c, blocks[c], ap, bp, info, work, pl->side[c]);
        if (le == hipSuccess) le = hipEventRecord(pl->join[c], pl->side[c]);
        if (le == hipSuccess) le = hipStreamWaitEvent(pl->cap, pl->join[c], 0);
    }
    e = hipStreamEndCapture(pl->cap, &pl->graph);
    if (le != hipSuccess) return le;
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return e;
    pl->key[0] = ap;
    pl->key[1] = bp;
    pl->key[2] = info;
    pl->key[3] = work;
    return hipSuccess;
}

template <class T>
int gels_batch(cbk_batch_plan *pl, char prec, T *ap, T *bp, int *info, void *work, size_t lwork)
{
    if (!pl || pl->prec != prec) return -1;
    if (pl->nprob == 0) return 0;
    if (!ap) return -2;
    if (!bp) return -3;
    if (!work && pl->bytes) return -5;
    if (lwork < pl->bytes) return -6;
    char *w = static_cast<char *>(work);
    const hipStream_t st = cbk::current_stream();
    hipError_t e = hipSuccess;
    int dev = 0;
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return cbk::launch_status(e);
    if (dev != pl->device) return cbk::launch_status(hipErrorInvalidDevice);
    if (pl->uploaded != work) {
        e = hipMemcpyAsync(w + pl->table_off, pl->table.data(),
                           pl->table.size() * sizeof(BatchProblem), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return cbk::launch_status(e);
        pl->uploaded = work;
    }
    if (!pl->exec || pl->key[0] != (void *)ap || pl->key[1] != (void *)bp ||
        pl->key[2] != (void *)info || pl->key[3] != work) {
        e = record_graph<T>(pl, ap, bp, info, w);
        if (e != hipSuccess) return cbk::launch_status(e);
    }
    return cbk::launch_status(hipGraphLaunch(pl->exec, st));
}

/* ---- the mixed-precision batch ---------------------------------------------- */

/* sort the problems into the mixed plan's tables (host only) */
void build_mixed(cbk_batch_plan *pl, const cbk_gels_problem *desc)
{
    MixedLayout &L = *pl->mx;
    const int np = pl->nprob;
    std::vector<int> k64(np), el(np);
    for (int k = 0; k < np; ++k) {
        k64[k] = kernel_of(classify(desc[k], pl->tran, pl->tune), desc[k]);
        el[k] = mixed_eligible(desc[k], pl->tran);
        (el[k] ? L.ndyn : L.nstat)[k64[k]]++;
    }
    for (int c = 1; c < kBatchKernels; ++c)
        L.fb_begin[c] = L.fb_begin[c - 1] + L.nstat[c - 1] + L.ndyn[c - 1];
    L.fbtable.assign(np, BatchProblem{0, 0, 0, 1, 1, 0, 0, 0, 0});
    std::vector<int> order, placed(kBatchKernels, 0);
    /* the problems without an FP32 path, each kernel's longest first */
    std::vector<int> stat;
    for (int k = 0; k < np; ++k)
        if (!el[k]) stat.push_back(k);
    std::stable_sort(stat.begin(), stat.end(),
                     [&](int x, int y) { return work_of(desc[x]) > work_of(desc[y]); });
    for (int k : stat) {
        const cbk_gels_problem &p = desc[k];
        BatchProblem b{p.m, p.n, p.nrhs, p.lda, p.ldb, k, p.a_off, p.b_off, 0};
        if (k64[k] == kBatchGeneral) {
            b.tau_off = L.tau64;
            L.tau64 += std::min(p.m, p.n);
        }
        L.fbtable[L.fb_begin[k64[k]] + placed[k64[k]]++] = b;
        L.inel.push_back(b);
    }
    for (int k = 0; k < np; ++k) {
        const cbk_gels_problem &p = desc[k];
        if (!el[k]) continue;
        order.push_back(k);
        L.ext_a = std::max(L.ext_a, p.a_off + (long long)p.lda * (p.n - 1) + p.m);
        L.ext_b = std::max(L.ext_b, p.b_off + (long long)p.ldb * (p.nrhs - 1) + p.m);
    }
    /* FP32 classes as the FP64 ones (the classes do not depend on precision),
     * each ordered longest first */
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (k64[x] != k64[y]) return k64[x] < k64[y];
        return work_of(desc[x]) > work_of(desc[y]);
    });
    for (int k : order) {
        const cbk_gels_problem &p = desc[k];
        L.table32.push_back(BatchProblem{p.m, p.n, p.nrhs, p.lda, p.ldb, k, p.a_off, p.b_off, 0});
        L.begin32[k64[k] + 1]++;
    }
    for (int c = 0; c < kBatchKernels; ++c)
        L.begin32[c + 1] += L.begin32[c];
    std::vector<int> bucket(np, 0);
    for (int k : order)
        bucket[k] = cbk::mixed_refine_bucket(desc[k].m, desc[k].n, desc[k].nrhs);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (bucket[x] != bucket[y]) return bucket[x] < bucket[y];
        return work_of(desc[x]) > work_of(desc[y]);
    });
    for (int k : order) {
        const cbk_gels_problem &p = desc[k];
        L.reftable.push_back(
            cbk::MixedProblem{p.m, p.n, p.nrhs, p.lda, p.ldb, k, p.a_off, p.b_off, k64[k], 0});
        L.ref_begin[bucket[k] + 1]++;
    }
    for (int b = 0; b < cbk::kRefineBuckets; ++b)
        L.ref_begin[b + 1] += L.ref_begin[b];
    size_t off = kCounterBytes;
    L.off32 = off;
    off = align256(off + L.table32.size() * sizeof(BatchProblem));
    L.offfb = off;
    off = align256(off + L.fbtable.size() * sizeof(BatchProblem));
    L.offref = off;
    off = align256(off + L.reftable.size() * sizeof(cbk::MixedProblem));
    L.offinel = off;
    off = align256(off + L.inel.size() * sizeof(BatchProblem));
    L.offinfo = off;
    off = align256(off + (size_t)np * sizeof(int));
    L.offtau = off;
    off = align256(off + (size_t)L.tau64 * sizeof(double));
    L.offa = off;
    off = align256(off + (size_t)L.ext_a * sizeof(float));
    L.offb = off;
    off = align256(off + (size_t)L.ext_b * sizeof(float));
    pl->bytes = off;
}

template <class E>
hipError_t upload(char *w, size_t off, const std::vector<E> &v, hipStream_t s)
{
    if (v.empty()) return hipSuccess;
    return hipMemcpyAsync(w + off, v.data(), v.size() * sizeof(E), hipMemcpyHostToDevice, s);
}

/* Counters cleared; the conversion and copies; then on their own branches,
 * concurrently: every FP32 class followed (after all of them) by the
 * refinement buckets, and the FP64 classes of the problems without an FP32
 * path; after the refinement, the FP64 classes over the fallback lists. */
hipError_t record_graph_mixed(cbk_batch_plan *pl, double *ap, const double *bp, double *xp,
                              int *iter, int *info, char *work)
{
    MixedLayout &L = *pl->mx;
    constexpr int K = kBatchKernels;
    hipError_t e = ensure_streams(pl);
    if (e == hipSuccess) e = ensure_mixed_streams(pl);
    if (e != hipSuccess) return e;
    unsigned b32[K] = {}, bst[K] = {}, bdyn[K] = {};
    for (int c = 0; c < K && e == hipSuccess; ++c) {
        if (L.begin32[c + 1] > L.begin32[c])
            e = class_blocks_n<float>(pl, c, L.begin32[c + 1] - L.begin32[c], &b32[c]);
        if (e == hipSuccess && L.nstat[c] > 0) e = class_blocks_n<double>(pl, c, L.nstat[c], &bst[c]);
        if (e == hipSuccess && L.ndyn[c] > 0) e = class_blocks_n<double>(pl, c, L.ndyn[c], &bdyn[c]);
    }
    if (e != hipSuccess) return e;
    if (pl->exec) {
        (void)hipGraphExecDestroy(pl->exec);
        pl->exec = nullptr;
    }
    if (pl->graph) {
        (void)hipGraphDestroy(pl->graph);
        pl->graph = nullptr;
    }
    unsigned *ctr = reinterpret_cast<unsigned *>(work);
    unsigned *q32 = ctr, *qst = ctr + K, *qdyn = ctr + 2 * K, *fill = ctr + 3 * K;
    float *a32 = reinterpret_cast<float *>(work + L.offa);
    float *b32p = reinterpret_cast<float *>(work + L.offb);
    int *info32 = reinterpret_cast<int *>(work + L.offinfo);
    double *tau = reinterpret_cast<double *>(work + L.offtau);
    BatchProblem *fb = reinterpret_cast<BatchProblem *>(work + L.offfb);
    const BatchProblem *t32 = reinterpret_cast<const BatchProblem *>(work + L.off32);
    const cbk::MixedProblem *ref = reinterpret_cast<const cbk::MixedProblem *>(work + L.offref);
    hipStream_t cap = pl->cap;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    hipError_t le = cbk::launch_clear_u32(ctr, kMixedCounters, cap);
    if (le == hipSuccess) {
        const cbk::MixedPrepArgs pa{ap, bp, a32, b32p, xp, ref, nullptr, (int)L.reftable.size(), iter};
        le = cbk::launch_mixed_prep(pa, cap);
    }
    if (le == hipSuccess) {
        const cbk::MixedPrepArgs pc{ap, bp, a32, b32p, xp, nullptr,
                                    reinterpret_cast<const BatchProblem *>(work + L.offinel),
                                    (int)L.inel.size(), iter};
        le = cbk::launch_mixed_prep(pc, cap);
    }
    if (le == hipSuccess) le = hipEventRecord(pl->fork, cap);
    /* the FP64 problems without an FP32 path: independent of everything after */
    for (int c = 0; c < K && le == hipSuccess; ++c) {
        if (L.nstat[c] == 0) continue;
        hipStream_t s = pl->mside[c];
        le = hipStreamWaitEvent(s, pl->fork, 0);
        const BatchArgs<double> a = list_args<double>(ap, xp, info, fb + L.fb_begin[c], L.nstat[c],
                                                      qst + c, tau, pl->tran, nullptr);
        if (le == hipSuccess) le = launch_args<double>(c, a, bst[c], s);
        if (le == hipSuccess) le = hipEventRecord(pl->mjoin[c], s);
    }
    /* the FP32 classes */
    for (int c = 0; c < K && le == hipSuccess; ++c) {
        const int cnt = L.begin32[c + 1] - L.begin32[c];
        if (cnt == 0) continue;
        le = hipStreamWaitEvent(pl->side[c], pl->fork, 0);
        const BatchArgs<float> a = list_args<float>(a32, b32p, info32, t32 + L.begin32[c], cnt,
                                                    q32 + c, nullptr, false, nullptr);
        if (le == hipSuccess) le = launch_args<float>(c, a, b32[c], pl->side[c]);
        if (le == hipSuccess) le = hipEventRecord(pl->join[c], pl->side[c]);
        if (le == hipSuccess) le = hipStreamWaitEvent(cap, pl->join[c], 0);
    }
    /* the refinement buckets, on branches of their own */
    if (le == hipSuccess) le = hipEventRecord(pl->fork2, cap);
    for (int b = 0; b < cbk::kRefineBuckets && le == hipSuccess; ++b) {
        const int cnt = L.ref_begin[b + 1] - L.ref_begin[b];
        if (cnt == 0) continue;
        hipStream_t s = pl->rside[b];
        le = hipStreamWaitEvent(s, pl->fork2, 0);
        cbk::MixedRefineArgs ra;
        ra.a = ap;
        ra.b = bp;
        ra.x = xp;
        ra.a32 = a32;
        ra.b32 = b32p;
        ra.iter = iter;
        ra.info = info;
        ra.desc = ref + L.ref_begin[b];
        ra.count = cnt;
        ra.fb = fb;
        ra.fb_fill = fill;
        for (int c = 0; c < K; ++c)
            ra.fb_begin[c] = L.fb_begin[c] + L.nstat[c];
        ra.itermax = 30;
        if (le == hipSuccess) le = cbk::launch_mixed_refine(b, ra, s);
        if (le == hipSuccess) le = hipEventRecord(pl->rjoin[b], s);
        if (le == hipSuccess) le = hipStreamWaitEvent(cap, pl->rjoin[b], 0);
    }
    /* FP64 gels, in place on A and X, over the fallback lists */
    if (le == hipSuccess) le = hipEventRecord(pl->fork3, cap);
    for (int c = 0; c < K && le == hipSuccess; ++c) {
        if (L.ndyn[c] == 0) continue;
        le = hipStreamWaitEvent(pl->side[c], pl->fork3, 0);
        const BatchArgs<double> a =
            list_args<double>(ap, xp, info, fb + L.fb_begin[c] + L.nstat[c], L.ndyn[c], qdyn + c,
                              tau, pl->tran, fill + c);
        if (le == hipSuccess) le = launch_args<double>(c, a, bdyn[c], pl->side[c]);
        if (le == hipSuccess) le = hipEventRecord(pl->join2[c], pl->side[c]);
        if (le == hipSuccess) le = hipStreamWaitEvent(cap, pl->join2[c], 0);
    }
    for (int c = 0; c < K && le == hipSuccess; ++c)
        if (L.nstat[c] > 0) le = hipStreamWaitEvent(cap, pl->mjoin[c], 0);
    e = hipStreamEndCapture(cap, &pl->graph);
    if (le != hipSuccess) return le;
    if (e != hipSuccess) return e;
    e = hipGraphInstantiate(&pl->exec, pl->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return e;
    void *key[6] = {ap, const_cast<double *>(bp), xp, iter, info, work};
    for (int i = 0; i < 6; ++i)
        pl->mkey[i] = key[i];
    return hipSuccess;
}

int dsgels_batch_impl(cbk_batch_plan *pl, double *ap, const double *bp, double *xp, int *iter,
                      int *info, void *work, size_t lwork)
{
    if (!pl || pl->prec != 'M') return -1;
    if (pl->nprob == 0) return 0;
    if (!ap) return -2;
    if (!bp) return -3;
    if (!xp) return -4;
    if (!iter) return -5;
    if (!info) return -6;
    if (!work && pl->bytes) return -7;
    if (lwork < pl->bytes) return -8;
    char *w = static_cast<char *>(work);
    const hipStream_t st = cbk::current_stream();
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return cbk::launch_status(e);
    if (dev != pl->device) return cbk::launch_status(hipErrorInvalidDevice);
    MixedLayout &L = *pl->mx;
    if (pl->uploaded != work) {
        e = upload(w, L.off32, L.table32, st);
        if (e == hipSuccess) e = upload(w, L.offfb, L.fbtable, st);
        if (e == hipSuccess) e = upload(w, L.offref, L.reftable, st);
        if (e == hipSuccess) e = upload(w, L.offinel, L.inel, st);
        if (e != hipSuccess) return cbk::launch_status(e);
        pl->uploaded = work;
    }
    void *key[6] = {ap, const_cast<double *>(bp), xp, iter, info, work};
    bool same = pl->exec != nullptr;
    for (int i = 0; i < 6; ++i)
        same = same && pl->mkey[i] == key[i];
    if (!same) {
        e = record_graph_mixed(pl, ap, bp, xp, iter, info, w);
        if (e != hipSuccess) return cbk::launch_status(e);
    }
    return cbk::launch_status(hipGraphLaunch(pl->exec, st));
}

/* ---- the tuner ------------------------------------------------------------- */

/* time one class kernel over `count` copies of an m x n problem with nrhs
 * right-hand sides (uniform(-1,1) data), best of 3, in milliseconds; per_cu:
 * the persistent grid's blocks per compute unit (0: the kernel's occupancy) */
template <class T>
hipError_t time_class(int cls, int m, int n, int nrhs, int count, int dev, float *ms,
                      int per_cu = 0)
{
    cbk_batch_plan pl;
    pl.prec = sizeof(T) == 8 ? 'D' : 'S';
    pl.device = dev;
    pl.tune = default_tuning();
    pl.tune.blocks_per_cu[class_of_kernel(cls)] = per_cu;
    pl.nprob = count;
    pl.table.resize(count);
    const long long sa = (long long)m * n, sb = (long long)m * nrhs;
    for (int k = 0; k < count; ++k)
        pl.table[k] = BatchProblem{m, n, nrhs, m, m, k, k * sa, k * sb, 0};
    for (int c = 0; c <= kBatchKernels; ++c)
        pl.begin[c] = c <= cls ? 0 : count;
    pl.tau_off = pl.table_off + ((count * sizeof(BatchProblem) + 255) / 256) * 256;
    pl.bytes = pl.tau_off;
    std::vector<T> ha(count * sa), hb(count * sb);
    std::mt19937_64 g(2026);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    for (auto &x : ha) x = (T)u(g);
    for (auto &x : hb) x = (T)u(g);
    T *da = nullptr, *db = nullptr;
    char *dw = nullptr;
    int *di = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipMalloc(&da, ha.size() * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&db, hb.size() * sizeof(T));
    if (e == hipSuccess) e = hipMalloc(&dw, pl.bytes);
    if (e == hipSuccess) e = hipMalloc(&di, count * sizeof(int));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dw + pl.table_off, pl.table.data(), count * sizeof(BatchProblem),
                           hipMemcpyHostToDevice, s);
    unsigned blocks = 0;
    if (e == hipSuccess) e = class_blocks<T>(&pl, cls, &blocks);
    float best = 1e30f;
    for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {
        e = hipMemcpyAsync(da, ha.data(), ha.size() * sizeof(T), hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(db, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = cbk::launch_clear_u32(reinterpret_cast<unsigned *>(dw), kBatchKernels, s);
        if (e == hipSuccess) e = hipEventRecord(e0, s);
        if (e == hipSuccess) e = launch_class<T>(&pl, cls, blocks, da, db, di, dw, s);
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess && rep > 0) best = std::min(best, t);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dw) (void)hipFree(dw);
    if (di) (void)hipFree(di);
    *ms = best;
    return e;
}

} /* namespace */

extern "C" {

int cbk_gels_batch_plan(char prec, char trans, int nprob, const cbk_gels_problem *desc,
                        cbk_batch_plan_t *plan, int *bad)
{
    if (!opt(prec, 'D') && !opt(prec, 'S') && !opt(prec, 'M')) return -1;
    if (!opt(trans, 'N') && !opt(trans, 'T') && !opt(trans, 'C')) return -2;
    if (nprob < 0) return -3;
    if (!desc && nprob > 0) return -4;
    if (!plan) return -6;
    for (int k = 0; k < nprob; ++k) {
        const cbk_gels_problem &p = desc[k];
        const int mx = p.m > p.n ? p.m : p.n;
        if (p.m < 0 || p.n < 0 || p.nrhs < 0 || p.lda < max1(p.m) || p.ldb < max1(mx) ||
            p.a_off < 0 || p.b_off < 0) {
            if (bad) *bad = k;
            return -5;
        }
    }
    cbk_batch_plan *pl = new (std::nothrow) cbk_batch_plan;
    if (!pl) return cbk::launch_status(hipErrorOutOfMemory);
    pl->prec = opt(prec, 'D') ? 'D' : opt(prec, 'S') ? 'S' : 'M';
    pl->tran = !opt(trans, 'N');
    pl->nprob = nprob;
    if (hipGetDevice(&pl->device) != hipSuccess) pl->device = 0;
    pl->tune = tuning_for(pl->device);
    if (pl->prec == 'M') {
        pl->mx = new (std::nothrow) MixedLayout;
        if (!pl->mx) {
            delete pl;
            return cbk::launch_status(hipErrorOutOfMemory);
        }
        build_mixed(pl, desc);
        *plan = pl;
        return 0;
    }
    std::vector<int> cls(nprob);
    std::vector<int> order(nprob);
    for (int k = 0; k < nprob; ++k) {
        cls[k] = kernel_of(classify(desc[k], pl->tran, pl->tune), desc[k]);
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (cls[x] != cls[y]) return cls[x] < cls[y];
        return work_of(desc[x]) > work_of(desc[y]);
    });
    pl->table.resize(nprob);
    const size_t esz = pl->prec == 'D' ? 8 : 4;
    for (int i = 0; i < nprob; ++i) {
        const cbk_gels_problem &p = desc[order[i]];
        BatchProblem &b = pl->table[i];
        b = BatchProblem{p.m, p.n, p.nrhs, p.lda, p.ldb, order[i], p.a_off, p.b_off, 0};
        if (cls[order[i]] == kBatchGeneral) {
            b.tau_off = pl->tau_elems;
            pl->tau_elems += std::min(p.m, p.n);
        }
        pl->begin[cls[order[i]] + 1]++;
    }
    for (int c = 0; c < kBatchKernels; ++c)
        pl->begin[c + 1] += pl->begin[c];
    pl->tau_off = pl->table_off + ((nprob * sizeof(BatchProblem) + 255) / 256) * 256;
    pl->bytes = pl->tau_off + (size_t)pl->tau_elems * esz;
    *plan = pl;
    return 0;
}

size_t cbk_batch_plan_worksize(cbk_batch_plan_t plan) { return plan ? plan->bytes : 0; }

int cbk_batch_plan_classes(cbk_batch_plan_t plan, int *count)
{
    if (!plan || !count) return -1;
    for (int c = 0; c < kBatchClasses; ++c) count[c] = 0;
    for (int k = 0; k < kBatchKernels; ++k)
        count[class_of_kernel(k)] +=
            plan->mx ? plan->mx->nstat[k] + plan->mx->ndyn[k] : plan->begin[k + 1] - plan->begin[k];
    return kBatchClasses;
}

int dgels_batch(cbk_batch_plan_t plan, double *ap, double *bp, int *info, void *work,
                size_t lwork)
{
    return gels_batch<double>(plan, 'D', ap, bp, info, work, lwork);
}

int sgels_batch(cbk_batch_plan_t plan, float *ap, float *bp, int *info, void *work, size_t lwork)
{
    return gels_batch<float>(plan, 'S', ap, bp, info, work, lwork);
}

int dsgels_batch(cbk_batch_plan_t plan, double *ap, const double *bp, double *xp, int *iter,
                 int *info, void *work, size_t lwork)
{
    return dsgels_batch_impl(plan, ap, bp, xp, iter, info, work, lwork);
}

int cbk_batch_plan_destroy(cbk_batch_plan_t plan)
{
    delete plan;
    return 0;
}

int cbk_get_tuning(cbk_tuning *t)
{
    if (!t) return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    *t = tuning_for(dev);
    return 0;
}

int cbk_set_tuning(const cbk_tuning *t)
{
    if (!t) return -1;
    for (int b = 0; b < 3; ++b)
        if (t->narrow_max_cols[b] < 0 || t->narrow_max_cols[b] > 16) return -1;
    for (int c = 0; c < kBatchClasses; ++c)
        if (t->blocks_per_cu[c] < 0 || t->blocks_per_cu[c] > 64) return -1;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    if (dev < 0 || dev >= kMaxDevices) return -1;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[dev].t = *t;
    g_tune[dev].set = true;
    return 0;
}

/* Per row bucket (16, 32, 64 rows) and column count w = n + nrhs in {4, 8,
 * 12, 16} (n = w - 2, nrhs = 2; m = the bucket), time the team kernel and the
 * wavefront kernel on 16384 identical problems; narrow_max_cols[b] is the
 * largest w up to which the team kernel is not slower at every w tried.
 * Then per class, at a shape typical of it, the persistent grid at half,
 * once and twice the kernel's occupancy per compute unit; blocks_per_cu[c]
 * keeps the fastest (0 when the occupancy itself is). Runs in FP64. */
int cbk_autotune(void)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return cbk::launch_status(e);
    cbk_tuning t = default_tuning();
    const int rows[3] = {16, 32, 64};
    for (int b = 0; b < 3; ++b) {
        int maxw = 0;
        for (int w = 4; w <= 16; w += 4) {
            const int nrhs = 2, n = w - nrhs, m = rows[b];
            float tn = 0.f, tw = 0.f;
            e = time_class<double>(b, m, n, nrhs, 16384, dev, &tn);
            /* the kernel such a problem would take otherwise: the lane grid */
            if (e == hipSuccess) e = time_class<double>(m <= 32 ? 9 : 10, m, n, nrhs, 16384, dev, &tw);
            if (e != hipSuccess) return cbk::launch_status(e);
            if (tn > tw) break;
            maxw = w;
        }
        t.narrow_max_cols[b] = maxw;
    }
    /* grid sizes: class, the kernel behind it at the shape, m, n, nrhs, count */
    const int shapes[6][6] = {{0, 0, 16, 6, 2, 16384},  {1, 1, 32, 6, 2, 16384},
                              {2, 2, 64, 6, 2, 16384},  {3, 9, 32, 20, 4, 16384},
                              {4, 10, 64, 24, 4, 16384}, {5, 11, 64, 60, 6, 4096}};
    for (const auto &sh : shapes) {
        int occ = 0;
        e = cbk::gels_batch_occupancy<double>(sh[1], &occ);
        if (e != hipSuccess) return cbk::launch_status(e);
        occ = occ < 1 ? 1 : occ;
        const int tries[3] = {occ, occ > 1 ? occ / 2 : 1, 2 * occ};
        float best = 0.f;
        int pick = 0;
        for (int i = 0; i < 3; ++i) {
            if (i == 1 && tries[1] == occ) continue;
            float ms = 0.f;
            e = time_class<double>(sh[1], sh[2], sh[3], sh[4], sh[5], dev, &ms, tries[i]);
            if (e != hipSuccess) return cbk::launch_status(e);
            /* another grid must win by more than 2 % to replace the occupancy */
            if (i == 0 || ms < 0.98f * best) {
                best = ms;
                pick = i == 0 ? 0 : tries[i];
            }
        }
        t.blocks_per_cu[sh[0]] = pick > 64 ? 64 : pick;
    }
    return cbk_set_tuning(&t) == 0 ? 0 : -1;
}

} /* extern "C" */
