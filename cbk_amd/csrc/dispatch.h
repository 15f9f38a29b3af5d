/* dispatch.h -- pick the kernel that serves a call: a tuned team kernel when
 * one covers the shape, layout and V, the general lane kernel otherwise. */
#ifndef CBK_DISPATCH_H
#define CBK_DISPATCH_H

#include "cbk_internal.h"

namespace cbk {

/* potrf / sytrfnp (factor, nrhs == 0), potrs / sytrsnp (!factor), posv /
 * sysvnp (factor, nrhs > 0) */
template <class T>
hipError_t dispatch_sym(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                        Operand<T> B, int V, int nm, int *info);

template <class T>
hipError_t dispatch_trsm(bool left, bool up, bool tran, bool unit, int m, int n, T alpha,
                         Operand<T> A, Operand<T> B, int V, int nm);

template <class T>
hipError_t dispatch_gels(bool tran, int m, int n, int nrhs, Operand<T> A, Operand<T> B,
                         T *tau, int V, int nm, int *info);

template <class T> hipError_t dispatch_geqrf(int m, int n, Operand<T> A, T *tau, int V, int nm);

} /* namespace cbk */

#endif /* CBK_DISPATCH_H */
