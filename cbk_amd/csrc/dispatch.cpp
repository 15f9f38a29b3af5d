/* dispatch.cpp -- kernel selection (see dispatch.h). */

#include "dispatch.h"

#include "cbk_ext.h"

namespace cbk {

template <class T>
hipError_t dispatch_sym(bool ldl, bool factor, bool up, int n, int nrhs, Operand<T> A,
                        Operand<T> B, int V, int nm, int *info)
{
    /* 32-bit offsets of the register and team kernels (offsets_fit) */
    const bool fit = offsets_fit(A, n, n, V) && offsets_fit(B, n, nrhs, V);
    const bool fp64 = sizeof(T) == 8;
    /* a family the tuner chose for this shape, where it serves the call */
    if (fit && current_policy() == CBK_POLICY_AUTO) {
        switch (tuned_family(ldl ? kTuneLdl : kTuneChol, fp64, n, 0, nrhs, V)) {
        case CBK_FAMILY_REGISTER:
            if (reg_chol_supports(n, fp64))
                return launch_reg_chol<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
            break;
        case CBK_FAMILY_QUAD:
            if (quad_sym_supports(n, fp64))
                return launch_quad_sym<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
            break;
        case CBK_FAMILY_TEAM:
            if (sym_team_supports(n) && V >= 4)
                return launch_sym_team<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
            break;
        case CBK_FAMILY_GRID:
            if (chol_grid_supports(n))
                return launch_chol_grid<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
            break;
        case CBK_FAMILY_LANE: return launch_chol_lane<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
        default: break;
        }
    }
    /* a matrix per 4 lanes at its orders (at FP32 24 measured faster than the
     * whole triangle in one lane: ssysvnp nrhs = 24 450 vs 523 us), else the
     * whole triangle in one lane's registers where it fits (faster than the
     * team kernel at its orders 16 and 24) */
    if (fit && current_policy() == CBK_POLICY_AUTO && (n == 24 || n == 32) && quad_sym_supports(n, fp64))
        return launch_quad_sym<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
    if (fit && current_policy() != CBK_POLICY_LANE && reg_chol_supports(n, fp64))
        return launch_reg_chol<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
    /* the other orders up to 32 on the next quad instance, as diag(A, I)
     * (dposv n = 20 / 28: 1049 / 2509 us on the lane kernel) */
    if (fit && current_policy() == CBK_POLICY_AUTO && quad_sym_supports(n, fp64))
        return launch_quad_sym<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
    if (fit && current_policy() == CBK_POLICY_AUTO && sym_team_supports(n) && V >= 4)
        return launch_sym_team<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
    /* orders 33..64: a matrix per wavefront on the 8 x 8 lane grid (dposv
     * n = 48: 13.9 ms on the lane kernel) */
    if (fit && current_policy() != CBK_POLICY_LANE && chol_grid_supports(n))
        return launch_chol_grid<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
    return launch_chol_lane<T>(ldl, factor, up, n, nrhs, A, B, V, nm, info);
}

template <class T>
hipError_t dispatch_trsm(bool left, bool up, bool tran, bool unit, int m, int n, T alpha,
                         Operand<T> A, Operand<T> B, int V, int nm)
{
    const int s = left ? m : n;
    const bool fit = offsets_fit(A, s, s, V) && offsets_fit(B, m, n, V);
    const bool fp64 = sizeof(T) == 8;
    if (left && fit && current_policy() == CBK_POLICY_AUTO) {
        /* op(A) lower (up == tran) is the forward sweep */
        switch (tuned_family(up == tran ? kTuneTrsmFwd : kTuneTrsmBwd, fp64, m, n, 0, V)) {
        case CBK_FAMILY_REGISTER:
            if (trsm_reg_supports(left, m, fp64))
                return launch_trsm_reg<T>(up, tran, unit, m, n, alpha, A, B, V, nm);
            break;
        case CBK_FAMILY_QUAD:
            if (quad_trsm_supports(left, up, tran, m, fp64))
                return launch_quad_trsm<T>(tran, unit, m, n, alpha, A, B, V, nm);
            break;
        case CBK_FAMILY_TEAM:
            if (V >= 4 && trsm_team_supports(left, up, tran, m))
                return launch_trsm_team<T>(up, tran, unit, m, n, alpha, A, B, V, nm);
            break;
        case CBK_FAMILY_GRID:
            if (trsm_grid_supports(left, up, tran, m))
                return launch_trsm_grid<T>(tran, unit, m, n, alpha, A, B, V, nm);
            break;
        case CBK_FAMILY_LANE:
            return launch_trsm_lane<T>(left, up, tran, unit, m, n, alpha, A, B, V, nm);
        default: break;
        }
    }
    if (!left && fit && current_policy() != CBK_POLICY_LANE) {
        /* X op(A) = alpha B is op(A)^T X^T = alpha B^T: the side-L kernels on
         * B's transposed view (the same buffer with its strides swapped) with
         * trans flipped. Per element of X this is the reference's side-R
         * sequence (src/cbk_trsm_compact.hpp:206-226): the multiply by alpha,
         * the fma's over l ascending, then the divide -- the side-L kernels'
         * row-dot sequence (never the axpy form of the side-L router, whose
         * back sweep subtracts in descending order) -- so the result is
         * bit-identical. */
        const Operand<T> Bt{B.p, B.ld, !B.row, B.gstride};
        if (trsm_reg_supports(true, n, fp64))
            return launch_trsm_reg<T>(up, !tran, unit, n, m, alpha, A, Bt, V, nm, true);
        if (quad_trsm_supports(true, up, !tran, n, fp64))
            return launch_quad_trsm<T>(!tran, unit, n, m, alpha, A, Bt, V, nm);
        if (V >= 4 && trsm_team_supports(true, up, !tran, n))
            return launch_trsm_team<T>(up, !tran, unit, n, m, alpha, A, Bt, V, nm);
        if (trsm_grid_supports(true, up, !tran, n))
            return launch_trsm_grid<T>(!tran, unit, n, m, alpha, A, Bt, V, nm);
    }
    if (fit && current_policy() != CBK_POLICY_LANE && trsm_reg_supports(left, m, sizeof(T) == 8))
        return launch_trsm_reg<T>(up, tran, unit, m, n, alpha, A, B, V, nm);
    if (fit && current_policy() != CBK_POLICY_LANE && quad_trsm_supports(left, up, tran, m, sizeof(T) == 8))
        return launch_quad_trsm<T>(tran, unit, m, n, alpha, A, B, V, nm);
    if (fit && current_policy() != CBK_POLICY_LANE && V >= 4 && trsm_team_supports(left, up, tran, m))
        return launch_trsm_team<T>(up, tran, unit, m, n, alpha, A, B, V, nm);
    if (fit && current_policy() != CBK_POLICY_LANE && trsm_grid_supports(left, up, tran, m))
        return launch_trsm_grid<T>(tran, unit, m, n, alpha, A, B, V, nm);
    return launch_trsm_lane<T>(left, up, tran, unit, m, n, alpha, A, B, V, nm);
}

template <class T>
hipError_t dispatch_gels(bool tran, int m, int n, int nrhs, Operand<T> A, Operand<T> B,
                         T *tau, int V, int nm, int *info)
{
    const bool tall = m >= n;
    const int p = tall ? m : n, q = tall ? n : m;
    const bool overdet = (tall != tran) && offsets_fit(A, m, n, V) &&
                         offsets_fit(B, p, nrhs, V); /* src/cbk_gels_compact.hpp:124 */
    const bool fp64 = sizeof(T) == 8;
    if (overdet && current_policy() == CBK_POLICY_AUTO) {
        switch (tuned_family(kTuneGels, fp64, p, q, nrhs, V)) {
        case CBK_FAMILY_REGISTER:
            if (qr_reg_supports(p, q, fp64))
                return launch_qr_reg<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
            break;
        case CBK_FAMILY_TEAM:
            if (V >= 4 && qr_team_supports(p, q, nrhs, fp64))
                return launch_qr_team<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
            break;
        case CBK_FAMILY_COLUMNS:
            if (V >= 4 && qr_cols_supports(p, q, nrhs))
                return launch_qr_cols<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
            break;
        case CBK_FAMILY_WAVE:
            if (qr_wave_supports(p, q, nrhs))
                return launch_qr_wave<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
            break;
        case CBK_FAMILY_GRID:
            if (qr_grid_supports(p, q, nrhs))
                return launch_qr_grid<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
            break;
        case CBK_FAMILY_LANE: return launch_gels_lane<T>(tran, m, n, nrhs, A, B, tau, V, nm, info);
        default: break;
        }
    }
    /* F in one lane's registers where it fits; then the row team at the
     * orders it is tuned for (its butterfly sums measured faster there), the
     * column team for the other small shapes */
    if (current_policy() != CBK_POLICY_LANE && overdet && qr_reg_supports(p, q, fp64))
        return launch_qr_reg<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
    /* (the lane grid first at more than 32 rows: sgels 64x32 nrhs 2 1518 us
     * against the row team's 2470 us) */
    if (current_policy() == CBK_POLICY_AUTO && overdet && V >= 4 &&
        qr_team_supports(p, q, nrhs, fp64) && !(p > 32 && qr_grid_supports(p, q, nrhs)))
        return launch_qr_team<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
    if (current_policy() == CBK_POLICY_AUTO && overdet && V >= 4 && qr_cols_supports(p, q, nrhs))
        return launch_qr_cols<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
    if (current_policy() != CBK_POLICY_LANE && overdet && qr_grid_supports(p, q, nrhs))
        return launch_qr_grid<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
    if (current_policy() != CBK_POLICY_LANE && overdet && qr_wave_supports(p, q, nrhs))
        return launch_qr_wave<T>(p, q, nrhs, !tall, A, tau, B, V, nm, info);
    return launch_gels_lane<T>(tran, m, n, nrhs, A, B, tau, V, nm, info);
}

template <class T> hipError_t dispatch_geqrf(int m, int n, Operand<T> A, T *tau, int V, int nm)
{
    const Operand<T> none = {nullptr, 1, A.row, 0};
    const bool tall = m >= n && offsets_fit(A, m, n, V);
    const bool fp64 = sizeof(T) == 8;
    if (tall && current_policy() == CBK_POLICY_AUTO) {
        switch (tuned_family(kTuneGeqrf, fp64, m, n, 0, V)) {
        case CBK_FAMILY_REGISTER:
            if (qr_reg_supports(m, n, fp64))
                return launch_qr_reg<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
            break;
        case CBK_FAMILY_TEAM:
            if (V >= 4 && qr_team_supports(m, n, 0, fp64))
                return launch_qr_team<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
            break;
        case CBK_FAMILY_COLUMNS:
            if (V >= 4 && qr_cols_supports(m, n, 0))
                return launch_qr_cols<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
            break;
        case CBK_FAMILY_GRID:
            if (qr_grid_supports(m, n, 0))
                return launch_qr_grid<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
            break;
        case CBK_FAMILY_LANE: return launch_geqrf_lane<T>(m, n, A, tau, V, nm);
        default: break;
        }
    }
    if (current_policy() != CBK_POLICY_LANE && tall && qr_reg_supports(m, n, fp64))
        return launch_qr_reg<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
    if (current_policy() == CBK_POLICY_AUTO && tall && V >= 4 && qr_team_supports(m, n, 0, fp64) &&
        !(m > 32 && qr_grid_supports(m, n, 0)))
        return launch_qr_team<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
    if (current_policy() == CBK_POLICY_AUTO && tall && V >= 4 && qr_cols_supports(m, n, 0))
        return launch_qr_cols<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
    if (current_policy() != CBK_POLICY_LANE && tall && qr_grid_supports(m, n, 0))
        return launch_qr_grid<T>(m, n, 0, false, A, tau, none, V, nm, nullptr);
    return launch_geqrf_lane<T>(m, n, A, tau, V, nm);
}

#define CBK_INSTANTIATE_DISPATCH(T)                                                      \
    template hipError_t dispatch_sym<T>(bool, bool, bool, int, int, Operand<T>,           \
                                        Operand<T>, int, int, int *);                     \
    template hipError_t dispatch_trsm<T>(bool, bool, bool, bool, int, int, T, Operand<T>, \
                                         Operand<T>, int, int);                           \
    template hipError_t dispatch_gels<T>(bool, int, int, int, Operand<T>, Operand<T>,     \
                                         T *, int, int, int *);                           \
    template hipError_t dispatch_geqrf<T>(int, int, Operand<T>, T *, int, int);

CBK_INSTANTIATE_DISPATCH(double)
CBK_INSTANTIATE_DISPATCH(float)

} /* namespace cbk */
