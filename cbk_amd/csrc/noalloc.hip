/* noalloc.hip -- dsposv_compact_noalloc: dsposv_compact for callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). The workspace the lane path needs
 * comes from a static device array of the library, one per GPU, whose
 * address is looked up once per device; the call itself allocates nothing.
 */

#include "cbk.h"
#include "cbk_ext.h"
#include "cbk_internal.h"

#include <hip/hip_runtime.h>

#include <mutex>

namespace cbk {
namespace {

/* the largest workspace a no-allocation call can use */
constexpr size_t kScratchBytes = CBK_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;

} /* namespace */
} /* namespace cbk */

/* Every GPU the library runs on holds its own copy until the process ends.
 * In host code the name is only a handle for hipGetSymbolAddress: the device
 * address of the current GPU's copy is what the kernels receive. */
static __device__ __align__(256) unsigned char scratch[cbk::kScratchBytes];

namespace cbk {
namespace {

std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices];

/* the current device's copy of scratch (looked up on the first call on that
 * device, with it current), or null with the HIP error recorded */
void *scratch_address(hipError_t *err)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess && (dev < 0 || dev >= kMaxDevices)) e = hipErrorInvalidDevice;
    if (e != hipSuccess) {
        *err = e;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            *err = e != hipSuccess ? e : hipErrorInvalidSymbol;
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

inline bool opt(char c, char upper) { return c == upper || c == upper + ('a' - 'A'); }
inline int max1(int x) { return x < 1 ? 1 : x; }
inline bool vlen_ok(int V)
{
    return V == 2 || V == 4 || V == 8 || V == 16 || V == 32 || V == 64;
}

} /* namespace */
} /* namespace cbk */

extern "C" size_t cbk_noalloc_scratch_bytes(void) { return cbk::kScratchBytes; }

extern "C" int dsposv_compact_noalloc(char layout, char uplo, int n, int nrhs, double *ap,
                                      int ldap, const double *bp, int ldbp, double *xp,
                                      int ldxp, int *iter, int *info, int V, int nm)
{
    using namespace cbk;
    /* dsposv_compact's checks in argument order, without work and lwork */
    const bool col = opt(layout, 'C'), row = opt(layout, 'R');
    if (!col && !row) return -1;
    if (!opt(uplo, 'L') && !opt(uplo, 'U')) return -2;
    if (n < 0) return -3;
    if (nrhs < 0) return -4;
    if (ldap < max1(n)) return -6;
    if (ldbp < max1(row ? nrhs : n)) return -8;
    if (ldxp < max1(row ? nrhs : n)) return -10;
    if (nm > 0 && iter == nullptr) return -11;
    if (nm > 0 && info == nullptr) return -12;
    if (!vlen_ok(V)) return -13;
    if (nm < 0) return -14;
    const size_t need = dsposv_compact_worksize(layout, uplo, n, nrhs, ldap, ldbp, ldxp, V, nm);
    if (need > kScratchBytes) return CBK_ERR_SCRATCH;
    void *work = nullptr;
    if (need > 0) {
        hipError_t e = hipSuccess;
        work = scratch_address(&e);
        if (work == nullptr) return launch_status(e);
    }
    return dsposv_compact(layout, uplo, n, nrhs, ap, ldap, bp, ldbp, xp, ldxp, work, need, iter,
                          info, V, nm);
}
