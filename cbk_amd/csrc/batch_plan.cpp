/* batch_plan.cpp -- uneven batches: the plan (size classes, queue order,
 * workspace layout, the recorded hipGraph), the dgels_batch / sgels_batch
 * entry points and the tuner that picks each class's kernel
 * (include/cbk_ext.h declares the ABI; no reference counterpart).
 *
 * Workspace layout (device, caller-owned):
 *   [0, 256)                 one 32-bit work counter per class
 *   [256, 256 + 48 nprob)    the problem table, grouped by class, each class
 *                            ordered by estimated work, largest first
 *   then, 256-aligned        tau scratch of the general kernel's problems
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
 * with n + nrhs <= 32 go two per wavefront (kernels 7 and 8) */
int kernel_of(int c, const cbk_gels_problem &p)
{
    return (c == 3 || c == 4) && p.n + p.nrhs <= 32 ? c + 4 : c;
}
/* the public class of kernel k */
int class_of_kernel(int k) { return k >= kBatchClasses ? k - 4 : k; }

int problems_per_block(int k) { return k <= 2 ? 16 : k <= 5 ? 4 : k == kBatchGeneral ? 256 : 8; }

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
    }
};

namespace {

hipError_t ensure_streams(cbk_batch_plan *pl)
{
    if (pl->cap) return hipSuccess;
    hipError_t e = hipStreamCreateWithFlags(&pl->cap, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->fork, hipEventDisableTiming);
    for (int c = 0; c < kBatchKernels && e == hipSuccess; ++c) {
        e = hipStreamCreateWithFlags(&pl->side[c], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&pl->join[c], hipEventDisableTiming);
    }
    return e;
}

template <class T>
BatchArgs<T> class_args(const cbk_batch_plan *pl, int c, T *ap, T *bp, int *info, char *work)
{
    BatchArgs<T> a;
    a.a = ap;
    a.b = bp;
    a.info = info;
    a.desc = reinterpret_cast<const BatchProblem *>(work + pl->table_off) + pl->begin[c];
    a.counter = reinterpret_cast<unsigned *>(work) + c;
    a.count = pl->begin[c + 1] - pl->begin[c];
    a.tau = reinterpret_cast<T *>(work + pl->tau_off);
    a.tran = pl->tran;
    a.a_rs = a.b_rs = 1;
    a.ninfo = 0x7fffffff;
    a.cm = a.cn = a.cnrhs = a.cV = a.a_cs = a.b_cs = 0;
    a.a_gs = a.b_gs = 0;
    a.ctau = nullptr;
    a.mask = nullptr;
    a.x = nullptr;
    a.x_rs = a.x_cs = 0;
    a.x_gs = 0;
    return a;
}

/* blocks of class c's persistent grid: enough resident blocks to fill the
 * device (the tuned count per compute unit, else the kernel's occupancy), no
 * more than the class has problems for */
template <class T> hipError_t class_blocks(const cbk_batch_plan *pl, int c, unsigned *blocks)
{
    const long count = pl->begin[c + 1] - pl->begin[c];
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

template <class T>
hipError_t launch_class(const cbk_batch_plan *pl, int c, unsigned blocks, T *ap, T *bp,
                        int *info, char *work, hipStream_t s)
{
    const BatchArgs<T> args = class_args<T>(pl, c, ap, bp, info, work);
    if (c == kBatchGeneral) return cbk::launch_gels_batch_lane<T>(args, s);
    return cbk::launch_gels_batch<T>(c, args, blocks, s);
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

/* ---- the tuner ------------------------------------------------------------- */

/* time one class kernel over `count` copies of an m x n problem with nrhs
 * right-hand sides (uniform(-1,1) data), best of 3, in milliseconds */
template <class T>
hipError_t time_class(int cls, int m, int n, int nrhs, int count, int dev, float *ms)
{
    cbk_batch_plan pl;
    pl.prec = sizeof(T) == 8 ? 'D' : 'S';
    pl.device = dev;
    pl.tune = default_tuning();
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
    if (!opt(prec, 'D') && !opt(prec, 'S')) return -1;
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
    pl->prec = opt(prec, 'D') ? 'D' : 'S';
    pl->tran = !opt(trans, 'N');
    pl->nprob = nprob;
    if (hipGetDevice(&pl->device) != hipSuccess) pl->device = 0;
    pl->tune = tuning_for(pl->device);
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
        count[class_of_kernel(k)] += plan->begin[k + 1] - plan->begin[k];
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
 * Grid sizes stay at each kernel's occupancy. Runs in FP64. */
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
            /* the wavefront kernel such a problem would take: two per wavefront */
            if (e == hipSuccess) e = time_class<double>(m <= 32 ? 7 : 8, m, n, nrhs, 16384, dev, &tw);
            if (e != hipSuccess) return cbk::launch_status(e);
            if (tn > tw) break;
            maxw = w;
        }
        t.narrow_max_cols[b] = maxw;
    }
    return cbk_set_tuning(&t) == 0 ? 0 : -1;
}

} /* extern "C" */
