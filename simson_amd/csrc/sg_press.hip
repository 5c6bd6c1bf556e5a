// Pressure of the current state (bla.f:2486-2527): nonlinp.f over all y
// planes (the step-2 kernels in their pressure variant write H_1..3 and
// W = u.u), the Chebyshev analysis of those and of v (the transform
// products), then linearp.f per (kx,kz) column in k_pressure below, and the
// Chebyshev synthesis of p.  Computed only when a field is saved, so the
// column work is one thread per column.
#include "sg_internal.h"
#include "sg_tau.h"

namespace sg {

namespace {

// Neumann tau solve of u'' - lam u = f (setmac.f neumann = 1, crhsc.f,
// tridc.f): interior rows as the Dirichlet system (sg_tau.h), boundary rows
// sum_{k even} k^2 u_k = bce, sum_{k odd} k^2 u_k = bco; a boundary row that
// vanishes after elimination (lam = 0, even class) gives u_first = 0
// (tridc.f:127-131).  gp holds (gamma_k, 1/pivot_k) for lam.
__device__ void solve_neumann(int n, double lam, const double *f, double bce, double bco,
                              const double2 *gp, double *a, double *rho)
{
    for (int c = 0; c < 2; c++) {
        const int klast = ((n - 1 - c) & 1) == 0 ? n - 1 : n - 2;
        double r_up = 0.0;
        for (int k = klast; k >= c + 2; k -= 2) {
            const double R = tau_rhs(n, k, f[k - 2], f[k], k + 2 < n ? f[k + 2] : 0.0);
            const double rk = (R - tau_up(n, k, lam) * r_up) * gp[k].y;
            rho[k] = rk;
            r_up = rk;
        }
        double P = 0.0, Q = 1.0, sP = 0.0, sQ = (double)(c * c);
        for (int k = c + 2; k <= klast; k += 2) {
            const double g = gp[k].x, kk = (double)k * (double)k;
            P = rho[k] - g * P;
            Q = -g * Q;
            sP += kk * P;
            sQ += kk * Q;
        }
        double prev = sQ == 0.0 ? 0.0 : ((c == 0 ? bce : bco) - sP) / sQ;
        a[c] = prev;
        for (int k = c + 2; k <= klast; k += 2) {
            prev = rho[k] - gp[k].x * prev;
            a[k] = prev;
        }
    }
}

constexpr int kMaxNy = 384;

// Qc: Chebyshev columns [5][nz][nxh][ny] of H1, H2, H3, W (each the
// coefficients of field/(nxp nzp)) and v (the same scaling, so times
// nxp nzp here); out: p as Chebyshev columns [nz][nxh][ny].
__global__ void __launch_bounds__(64) k_pressure(const double2 *__restrict__ Qc,
                                                 const double2 *__restrict__ gk,
                                                 const double *__restrict__ alfa,
                                                 const double *__restrict__ beta, int nz, int nxh,
                                                 int n, double re, double nxpnzp,
                                                 double *__restrict__ scratch, double2 *out)
{
    const int ncol = nz * nxh;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol)
        return;
    const int z = t / nxh, x = t - z * nxh;
    if (z == nz / 2)                             // linearp.f:77, z oddball
        return;
    const size_t fs = (size_t)ncol * n, o = (size_t)t * n;
    double *fr = scratch + (size_t)t * 6 * n, *fi = fr + n, *ar = fi + n, *ai = ar + n;
    double *rho = ai + n, *h2r = rho + n;
    const double al = alfa[x], be = beta[z], k2 = al * al + be * be;
    // f = d(H2)/dy + i (alpha H1 + beta H3) (linearp.f:185-196, nonlinp.f:229-247)
    double b2r = 0.0, b2i = 0.0, b1r = 0.0, b1i = 0.0;
    for (int k = n - 1; k >= 0; k--) {
        const double2 h2 = Qc[fs + o + k];
        h2r[k] = h2.x;
        // dcheb.f:24-36 run downwards: b_k = b_{k+2} + 2(k+1) a_{k+1}
        double dr = 0.0, di = 0.0;
        if (k <= n - 2) {
            const double2 hn = Qc[fs + o + k + 1];
            dr = b2r + 2.0 * (double)(k + 1) * hn.x;
            di = b2i + 2.0 * (double)(k + 1) * hn.y;
        }
        b2r = b1r;
        b2i = b1i;
        b1r = dr;
        b1i = di;
        if (k == 0) {
            dr *= 0.5;
            di *= 0.5;
        }
        const double2 h1 = Qc[o + k], h3 = Qc[2 * fs + o + k];
        fr[k] = dr - (al * h1.y + be * h3.y);
        fi[k] = di + (al * h1.x + be * h3.x);
    }
    // Neumann data dp/dy(+-1) = H2 + (v'' - k^2 v)/Re (linearp.f:117-162;
    // the time-derivative term there vanishes with no-slip walls)
    double h2pr = 0.0, h2pi = 0.0, h2mr = 0.0, h2mi = 0.0;
    double vpr = 0.0, vpi = 0.0, vmr = 0.0, vmi = 0.0, dpr = 0.0, dpi = 0.0, dmr = 0.0, dmi = 0.0;
    for (int k = 0; k < n; k++) {
        const double sg = (k & 1) ? -1.0 : 1.0, kk = (double)k * (double)k;
        const double w2 = kk * (kk - 1.0) / 3.0;
        const double2 h2 = Qc[fs + o + k], vk = Qc[4 * fs + o + k];
        const double vr = vk.x * nxpnzp, vi = vk.y * nxpnzp;
        h2pr += h2.x;
        h2pi += h2.y;
        h2mr += sg * h2.x;
        h2mi += sg * h2.y;
        vpr += vr;
        vpi += vi;
        vmr += sg * vr;
        vmi += sg * vi;
        dpr += w2 * vr;
        dpi += w2 * vi;
        dmr += sg * w2 * vr;
        dmi += sg * w2 * vi;
    }
    const double v1r = h2pr + (dpr - k2 * vpr) / re, v1i = h2pi + (dpi - k2 * vpi) / re;
    const double v2r = h2mr + (dmr - k2 * vmr) / re, v2i = h2mi + (dmi - k2 * vmi) / re;
    const double2 *g = gk + o;
    solve_neumann(n, k2, fr, 0.5 * (v1r - v2r), 0.5 * (v1r + v2r), g, ar, rho);
    solve_neumann(n, k2, fi, 0.5 * (v1i - v2i), 0.5 * (v1i + v2i), g, ai, rho);
    if (t == 0) {
        // mean pressure: the integral of H2 (icheb.f:24-36) with the value
        // E(-1) at the lower wall (linearp.f:234-244)
        double e0 = 0.0;
        for (int k = 0; k < n; k++)
            e0 += ((k & 1) ? -0.5 : 0.5) * Qc[3 * fs + o + k].x;
        h2r[n - 1] = 0.0;
        for (int k = 2; k < n - 1; k++)
            ar[k] = (h2r[k - 1] - h2r[k + 1]) * (0.5 / (double)k);
        ar[n - 1] = h2r[n - 2] * (0.5 / (double)(n - 1));
        ar[1] = h2r[0] - 0.5 * h2r[2];
        ar[0] = e0;
        for (int k = 1; k + 1 < n; k += 2)
            ar[0] += ar[k] - ar[k + 1];
        for (int k = 0; k < n; k++)
            ai[k] = 0.0;
    }
    // p = phi - E, E = W / 2 (linearp.f:246-252)
    for (int k = 0; k < n; k++) {
        const double2 w = Qc[3 * fs + o + k];
        out[o + k] = make_double2(ar[k] - 0.5 * w.x, ai[k] - 0.5 * w.y);
    }
}

}  // namespace

int pressure_max_ny() { return kMaxNy; }

void launch_pressure_cols(const double2 *Qc, const double2 *gk, const double *alfa,
                          const double *beta, int nz, int nxh, int ny, double re,
                          double nxpnzp, double *scratch, double2 *out, hipStream_t s)
{
    const int ncol = nz * nxh;
    hipLaunchKernelGGL(k_pressure, dim3((ncol + 63) / 64), dim3(64), 0, s, Qc, gk, alfa, beta, nz,
                       nxh, ny, re, nxpnzp, scratch, out);
}

}  // namespace sg
