/* offsets.h -- the 32-bit offset bound of the buffer-resource kernels
 * (host code only; no HIP dependency, so a CPU test can compile it). */
#ifndef CBK_OFFSETS_H
#define CBK_OFFSETS_H

#include <cstddef>

namespace cbk {

/* The register, team and quad kernels address a wavefront's operands through
 * buffer resources with 32-bit byte offsets: a lane's slot ((g - w0) *
 * gstride + v, where a wavefront spans up to 64 / V groups) plus an element's
 * place ((i * si + j * sj) * V), times the element size, with 2^31 marking
 * "out of range". An operand whose largest such offset over its rows x cols
 * referenced elements does not stay below 2^31 - 1 (a leading dimension in
 * the hundreds of thousands) goes to the lane kernels, which use 64-bit
 * addresses. The dispatcher asks this for every operand of a call. */
inline bool offsets_fit_bytes(size_t gstride, int ld, bool row, int rows, int cols, int V,
                              size_t elem)
{
    const unsigned long long gpw = V < 64 ? (unsigned long long)(64 / V) : 1ull;
    const unsigned long long si = row ? (unsigned long long)ld : 1ull;
    const unsigned long long sj = row ? 1ull : (unsigned long long)ld;
    const unsigned long long r = rows > 0 ? (unsigned long long)(rows - 1) : 0ull;
    const unsigned long long c = cols > 0 ? (unsigned long long)(cols - 1) : 0ull;
    const unsigned long long last =
        (gpw - 1) * (unsigned long long)gstride + (r * si + c * sj) * (unsigned long long)V +
        (unsigned long long)(V - 1);
    return (last + 1) * elem <= 0x7fffffffull;
}
} /* namespace cbk */

#endif /* CBK_OFFSETS_H */
