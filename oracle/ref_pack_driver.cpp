/* ref_pack_driver.cpp -- C entry points over the reference's own dense <->
 * compact utilities (tests/test_compact_util.hpp:475-531 of the reference,
 * included from where it lies), built by `make ref` into
 * oracle/_ref/libcbkref_pack.so. Test infrastructure only: the checker of
 * the library's pack/unpack and of the oracle's restatement of them.
 *
 * The reference's utilities take a column-major dense batch with ld = rows
 * (MatrixBatch) and V in {2, 4, 8, 16}; these wrappers copy a caller's
 * contiguous batch in and out of one. */
#include "test_compact_util.hpp"

#include <cstring>

namespace {

using cbk::test::MatrixBatch;
using cbk::test::pack_compact;
using cbk::test::unpack_compact;

template <class T>
int pack(int rowmajor, int rows, int cols, const T *a, T *ap, int ldap, T pad, int V, int nm)
{
    MatrixBatch<T> M(nm, rows, cols);
    std::memcpy(M.data(), a, sizeof(T) * (size_t)nm * rows * cols);
    pack_compact(M, ap, ldap, V, rowmajor != 0, pad);
    return 0;
}

template <class T>
int unpack(int rowmajor, int rows, int cols, const T *ap, int ldap, T *a, int V, int nm)
{
    MatrixBatch<T> M(nm, rows, cols);
    unpack_compact(M, ap, ldap, V, rowmajor != 0);
    std::memcpy(a, M.data(), sizeof(T) * (size_t)nm * rows * cols);
    return 0;
}

} /* namespace */

extern "C" {
int refpack_d(int rowmajor, int rows, int cols, const double *a, double *ap, int ldap,
              double pad, int V, int nm)
{
    return pack<double>(rowmajor, rows, cols, a, ap, ldap, pad, V, nm);
}
int refunpack_d(int rowmajor, int rows, int cols, const double *ap, int ldap, double *a, int V,
                int nm)
{
    return unpack<double>(rowmajor, rows, cols, ap, ldap, a, V, nm);
}
int refpack_s(int rowmajor, int rows, int cols, const float *a, float *ap, int ldap, float pad,
              int V, int nm)
{
    return pack<float>(rowmajor, rows, cols, a, ap, ldap, pad, V, nm);
}
}
