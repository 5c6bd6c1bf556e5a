// Coefficients of the Chebyshev tau system for u'' - lam u = f with the two
// boundary rows sum_even u_k = bc_e, sum_odd u_k = bc_o (the system that
// bla/setmacrhs.f:23-74 builds and bla/trid.f:120-199 solves).
//
// For k = 2..n-1 the row relates a_k to g = f + lam a, the u'' coefficients
// truncated to degree n-3:
//   L_k a_{k-2} + D_k a_k + U_k a_{k+2} = R_k
//   L_k = -c_{k-2} lam / (4k(k-1))            (c_0 = 2, c_j = 1)
//   D_k = 1 + [k <= n-3] lam / (2(k^2-1))
//   U_k = -[k <= n-5] lam / (4k(k+1))
//   R_k = c_{k-2} f_{k-2}/(4k(k-1)) - [k<=n-3] f_k/(2(k^2-1))
//         + [k<=n-5] f_{k+2}/(4k(k+1))
// Each parity class is a tridiagonal chain closed by its boundary row.
// Bottom-up elimination gives gamma_k = L_k / piv_k with
// piv_k = D_k - U_k gamma_{k+2}, and for a right-hand side
// rho_k = (R_k - U_k rho_{k+2}) / piv_k; then a_k = rho_k - gamma_k a_{k-2}
// is affine in the first coefficient of the class, which the boundary row
// fixes.
#pragma once

#include <hip/hip_runtime.h>

namespace sg {

__device__ __forceinline__ double tau_low(int k, double lam)
{
    const double kk = (double)k;
    const double c = (k == 2) ? 2.0 : 1.0;
    return -c * lam / (4.0 * kk * (kk - 1.0));
}

__device__ __forceinline__ double tau_diag(int n, int k, double lam)
{
    const double kk = (double)k;
    return 1.0 + (k <= n - 3 ? lam / (2.0 * (kk * kk - 1.0)) : 0.0);
}

__device__ __forceinline__ double tau_up(int n, int k, double lam)
{
    const double kk = (double)k;
    return (k <= n - 5) ? -lam / (4.0 * kk * (kk + 1.0)) : 0.0;
}

__device__ __forceinline__ double tau_rhs(int n, int k, double fm2, double f0, double fp2)
{
    const double kk = (double)k;
    const double c = (k == 2) ? 2.0 : 1.0;
    double R = c * fm2 / (4.0 * kk * (kk - 1.0));
    if (k <= n - 3)
        R -= f0 / (2.0 * (kk * kk - 1.0));
    if (k <= n - 5)
        R += fp2 / (4.0 * kk * (kk + 1.0));
    return R;
}

}  // namespace sg
