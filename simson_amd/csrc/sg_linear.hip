// Step 3 of the bla substep on the GPU (bla/linearbl.f, tau method, called
// per z plane from bla/bla.f:2433-2475) and the start-up partial RHS
// (bla/prhs.f:93-709).
//
//   k_gemm  batched y Chebyshev transforms as dense products with the
//           (ny x ny) transform matrix: analysis of H (vchbf + scaling,
//           linearbl.f:262-273) and synthesis of u,v,w,om_x,om_z (vchbb,
//           linearbl.f:1830-1846);
//   k_tau   one thread per (kx,kz) column: Helmholtz tau solves for phi,
//           omega_y and v (setmacrhs.f + trid.f), the homogeneous
//           superposition for dv/dy = 0 at the walls, u, w, om_x, om_z,
//           the k = 0 mean flow with the constant-flux px update, and the
//           next partial right-hand sides;
//   k_prhs  one thread per column: om_x, om_z, p_v, p_omega and puw from the
//           velocities at start-up.
#include "sg_internal.h"

namespace sg {

namespace {

// a column of a per-(kx,kz) array: element k at p[k * s]
struct ColView {
    double *p;
    size_t s;
    __device__ __forceinline__ double &operator[](int k) const { return p[(size_t)k * s]; }
};

// dcheb.f:24-36: b = d/dy a in Chebyshev space (b_{n-1} = 0, b_0 halved)
__device__ void dcheb(int n, ColView a, ColView b)
{
    double bk2 = 0.0, bk1 = 0.0;
    b[n - 1] = 0.0;
    for (int k = n - 2; k >= 1; k--) {
        const double bk = bk2 + 2.0 * (double)(k + 1) * a[k + 1];
        b[k] = bk;
        bk2 = bk1;
        bk1 = bk;
    }
    b[0] = 0.5 * (bk2 + 2.0 * a[1]);
}

// Chebyshev tau solve of u'' - lam u = f (setmacrhs.f:23-74, trid.f:120-199)
// with sum_even u_k = bce, sum_odd u_k = bco.  Tau rows k = 2..n-1 relate
// a_k to g = f + lam a truncated to degree n-3; each parity class is a
// tridiagonal chain plus one full row, reduced bottom-up to
// a_k = rho_k - gam_k a_{k-2}, closed by the boundary row, then substituted.
// f may alias nothing else; fzero treats f as identically zero.
__device__ void tau_solve(int n, double lam, ColView f, bool fzero, double bce,
                          double bco, ColView a, ColView gam, ColView rho)
{
    for (int par = 0; par < 2; par++) {
        const int k0 = par;
        const int klast = ((n - 1 - par) & 1) == 0 ? n - 1 : n - 2;
        double g_up = 0.0, r_up = 0.0;
        for (int k = klast; k >= k0 + 2; k -= 2) {
            const double kk = (double)k;
            const double ckm2 = (k == 2) ? 2.0 : 1.0;
            const double L = -ckm2 * lam / (4.0 * kk * (kk - 1.0));
            const double D = 1.0 + (k <= n - 3 ? lam / (2.0 * (kk * kk - 1.0)) : 0.0);
            const double U = (k <= n - 5) ? -lam / (4.0 * kk * (kk + 1.0)) : 0.0;
            double R = 0.0;
            if (!fzero) {
                R = ckm2 * f[k - 2] / (4.0 * kk * (kk - 1.0));
                if (k <= n - 3)
                    R -= f[k] / (2.0 * (kk * kk - 1.0));
                if (k <= n - 5)
                    R += f[k + 2] / (4.0 * kk * (kk + 1.0));
            }
            const double piv = D - U * g_up;
            const double gk = L / piv;
            const double rk = (R - U * r_up) / piv;
            gam[k] = gk;
            rho[k] = rk;
            g_up = gk;
            r_up = rk;
        }
        double P = 0.0, Q = 1.0, sP = 0.0, sQ = 1.0;
        for (int k = k0 + 2; k <= klast; k += 2) {
            const double gk = gam[k];
            P = rho[k] - gk * P;
            Q = -gk * Q;
            sP += P;
            sQ += Q;
        }
        const double bc = par == 0 ? bce : bco;
        double prev = (bc - sP) / sQ;
        a[k0] = prev;
        for (int k = k0 + 2; k <= klast; k += 2) {
            prev = rho[k] - gam[k] * prev;
            a[k] = prev;
        }
    }
}

__device__ __forceinline__ double mass_flux(int n, ColView u)
{
    double m = 0.0;
    for (int k = 0; k < n; k += 2)
        m -= u[k] * 2.0 / (double)(k * k - 1);
    return m;
}

// ---------------------------------------------------------------------
// batched Chebyshev transform: Y[b] = A (ny x ny) * X[b] (ny x ncols)
// ---------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 16;

__global__ void __launch_bounds__(256) k_gemm(const double *__restrict__ A, GemmIO io)
{
    __shared__ double As[GBK][GBM + 1];
    __shared__ double Xs[GBK][GBN + 1];
    const int b = blockIdx.z;
    const int f = b / io.nz, z = b - f * io.nz;
    if (z == io.skip_z)
        return;
    const int ny = io.ny, nc = io.ncols;
    const double *X = io.in + ((size_t)io.in_slot[f] * io.nz + z) * ny * nc;
    double *Y = io.out + ((size_t)io.out_slot[f] * io.nz + z) * ny * nc;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            acc[i][j] = 0.0;
    for (int k0 = 0; k0 < ny; k0 += GBK) {
        for (int i = 0; i < 4; i++) {
            const int e = tid + 256 * i;
            const int mm = e / GBK, kk = e - mm * GBK;
            const int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < ny && gk < ny) ? A[(size_t)gm * ny + gk] : 0.0;
            const int k2 = e / GBN, nn = e - k2 * GBN;
            const int gk2 = k0 + k2, gn = n0 + nn;
            Xs[k2][nn] = (gk2 < ny && gn < nc) ? X[(size_t)gk2 * nc + gn] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk++) {
            double av[4], xv[4];
            for (int i = 0; i < 4; i++)
                av[i] = As[kk][ty + 16 * i];
            for (int j = 0; j < 4; j++)
                xv[j] = Xs[kk][tx + 16 * j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    acc[i][j] = fma(av[i], xv[j], acc[i][j]);
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++) {
        const int gm = m0 + ty + 16 * i;
        if (gm >= ny)
            continue;
        for (int j = 0; j < 4; j++) {
            const int gn = n0 + tx + 16 * j;
            if (gn < nc)
                Y[(size_t)gm * nc + gn] = acc[i][j];
        }
    }
}

// views into the per-column data
struct Cols {
    const ColArgs &a;
    int t, z, x, ncol;
    __device__ ColView W(int i) const
    {
        return ColView{a.W + (size_t)i * a.ny * ncol + t, (size_t)ncol};
    }
    // component c (0 re, 1 im) of slot s of a [slot][nz][ny][nxh] complex array
    __device__ ColView cplx(double2 *base, int s, int c) const
    {
        double *p = (double *)(base + (((size_t)s * a.nz + z) * a.ny) * a.nxh + x) + c;
        return ColView{p, (size_t)2 * a.nxh};
    }
};

__global__ void __launch_bounds__(256) k_tau(ColArgs a, double an, double bn, double anp1,
                                             double bnp1)
{
    const int ncol = a.nz * a.nxh;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol)
        return;
    const int z = t / a.nxh, x = t - z * a.nxh;
    if (z == a.nz / 2)                    // linearbl.f:227, z oddball
        return;
    const int n = a.ny;
    Cols c{a, t, z, x, ncol};
    const double re = a.re;
    const double alfa = a.alfa[x], beta = a.beta[z];
    const double k2 = alfa * alfa + beta * beta;
    const double lam2 = k2 + 2. * re / (an + bn);
    const bool k0 = (z == 0 && x == 0);
    const double k2i = k0 ? 0. : 1. / k2;
    const double c1 = 2. * re / (an + bn) * an;
    const double c1n = 2. * re / (anp1 + bnp1);
    const double c2 = 2. * re * bnp1 / (anp1 + bnp1);

    ColView H1r = c.cplx(a.C, 0, 0), H1i = c.cplx(a.C, 0, 1);
    ColView H2r = c.cplx(a.C, 1, 0), H2i = c.cplx(a.C, 1, 1);
    ColView H3r = c.cplx(a.C, 2, 0), H3i = c.cplx(a.C, 2, 1);
    ColView pvr_in = c.cplx(a.U, 5, 0), pvi_in = c.cplx(a.U, 5, 1);
    ColView pwr_in = c.cplx(a.U, 6, 0), pwi_in = c.cplx(a.U, 6, 1);

    ColView hvr = c.W(0), hvi = c.W(1), hwr = c.W(2), hwi = c.W(3);
    ColView fvr = c.W(4), fvi = c.W(5), fwr = c.W(6), fwi = c.W(7);
    ColView phr = c.W(8), phi = c.W(9), phh = c.W(10);
    ColView omr = c.W(11), omi = c.W(12);
    ColView vr = c.W(13), vi = c.W(14), vh = c.W(15);
    ColView dvr = c.W(16), dvi = c.W(17), dvh = c.W(18);
    ColView domr = c.W(19), domi = c.W(20);
    ColView t1 = c.W(21), t2 = c.W(22), gam = c.W(23), rho = c.W(24);

    // h_omega, f_omega and g = i alpha H1 + i beta H3 (linearbl.f:325-343)
    for (int k = 0; k < n; k++) {
        const double h1r = H1r[k], h1i = H1i[k], h3r = H3r[k], h3i = H3i[k];
        const double wr = -beta * h1i + alfa * h3i;
        const double wi = beta * h1r - alfa * h3r;
        hwr[k] = wr;
        hwi[k] = wi;
        fwr[k] = pwr_in[k] - c1 * wr;
        fwi[k] = pwi_in[k] - c1 * wi;
        t1[k] = -alfa * h1i - beta * h3i;
        t2[k] = alfa * h1r + beta * h3r;
    }
    // h_v = -(dg/dy + k^2 H2), f_v = p_v - c1 h_v (linearbl.f:345-361)
    dcheb(n, t1, phr);
    dcheb(n, t2, phi);
    for (int k = 0; k < n; k++) {
        const double vr_ = -(phr[k] + k2 * H2r[k]);
        const double vi_ = -(phi[k] + k2 * H2i[k]);
        hvr[k] = vr_;
        hvi[k] = vi_;
        fvr[k] = pvr_in[k] - c1 * vr_;
        fvi[k] = pvi_in[k] - c1 * vi_;
    }
    ColView fuw1{a.mf + 0 * n, 1}, fuw2{a.mf + 1 * n, 1}, huw1{a.mf + 2 * n, 1},
        huw2{a.mf + 3 * n, 1}, u0{a.mf + 4 * n, 1}, w0{a.mf + 5 * n, 1},
        du0{a.mf + 6 * n, 1}, dw0{a.mf + 7 * n, 1};
    double px = 0.;
    if (k0) {                              // linearbl.f:303-319, 402-411
        px = *a.px;
        for (int k = 0; k < n; k++) {
            fuw1[k] = a.puw[k] - c1 * H1r[k];
            fuw2[k] = a.puw[n + k] - c1 * H3r[k];
            huw1[k] = H1r[k];
            huw2[k] = H3r[k];
            fvr[k] = fvi[k] = fwr[k] = fwi[k] = 0.;
        }
        fuw1[0] = fuw1[0] + 2. * re * px;
    }
    // Helmholtz solves with lam^2 (linearbl.f:625-626)
    tau_solve(n, lam2, fvr, false, 0., 0., phr, gam, rho);
    tau_solve(n, lam2, fvi, false, 0., 0., phi, gam, rho);
    tau_solve(n, lam2, fvr, true, 1., 1., phh, gam, rho);
    tau_solve(n, lam2, fwr, false, 0., 0., omr, gam, rho);
    tau_solve(n, lam2, fwi, false, 0., 0., omi, gam, rho);
    // v from phi with k^2 (linearbl.f:1000-1001)
    tau_solve(n, k2, phr, false, 0., 0., vr, gam, rho);
    tau_solve(n, k2, phi, false, 0., 0., vi, gam, rho);
    tau_solve(n, k2, phh, false, 0., 0., vh, gam, rho);
    // homogeneous superposition per parity: dv/dy = 0 at both walls
    // (linearbl.f:1023-1070)
    dcheb(n, vr, dvr);
    dcheb(n, vi, dvi);
    dcheb(n, vh, dvh);
    double Sr[2] = {0., 0.}, Si[2] = {0., 0.}, Sh[2] = {0., 0.};
    for (int k = 0; k < n; k++) {
        Sr[k & 1] += dvr[k];
        Si[k & 1] += dvi[k];
        Sh[k & 1] += dvh[k];
    }
    double cr[2], ci[2];
    for (int s = 0; s < 2; s++) {
        cr[s] = (0. - Sr[1 - s]) * (1. / Sh[1 - s]);
        ci[s] = (0. - Si[1 - s]) * (1. / Sh[1 - s]);
    }
    for (int k = 0; k < n; k++) {
        const double h = vh[k];
        vr[k] = vr[k] + cr[k & 1] * h;
        vi[k] = vi[k] + ci[k & 1] * h;
    }
    // with ibc = 0 the 3x3 wall system of evald/evalbc/solvbc has a zero
    // right-hand side: nothing to add.
    // dv, phi = D^2 v - k^2 v, d omega (linearbl.f:1151-1166)
    dcheb(n, vr, dvr);
    dcheb(n, dvr, t1);
    dcheb(n, vi, dvi);
    dcheb(n, dvi, t2);
    for (int k = 0; k < n; k++) {
        phr[k] = t1[k] - k2 * vr[k];
        phi[k] = t2[k] - k2 * vi[k];
    }
    dcheb(n, omr, domr);
    dcheb(n, omi, domi);
    if (k0) {
        // mean flow and constant-flux px (linearbl.f:1216-1417, tau, ibc 0)
        const double bue = .5 * (a.u0upp + a.u0low), buo = .5 * (a.u0upp - a.u0low);
        const double bwe = .5 * (a.w0upp + a.w0low), bwo = .5 * (a.w0upp - a.w0low);
        tau_solve(n, lam2, fuw1, false, bue, buo, u0, gam, rho);
        tau_solve(n, lam2, fuw2, false, bwe, bwo, w0, gam, rho);
        if (a.cflux) {
            const double m1 = mass_flux(n, u0);
            fuw1[0] = fuw1[0] + 2. * re * px;
            tau_solve(n, lam2, fuw1, false, bue, buo, u0, gam, rho);
            const double m2 = mass_flux(n, u0);
            fuw1[0] = fuw1[0] - 4. * re * px;
            px = px * (1. + (a.mflux - m1) / (m2 - m1));
            fuw1[0] = fuw1[0] + 2. * re * px;
            tau_solve(n, lam2, fuw1, false, bue, buo, u0, gam, rho);
            *a.px = px;
        }
        dcheb(n, u0, du0);
        dcheb(n, w0, dw0);
        dcheb(n, du0, t1);
        dcheb(n, dw0, t2);
        for (int k = 0; k < n; k++) {
            a.puw[k] = -t1[k] - c1n * u0[k] - c2 * huw1[k];
            a.puw[n + k] = -t2[k] - c1n * w0[k] - c2 * huw2[k];
        }
    }
    // next partial right-hand sides (linearbl.f:1761-1784, tau branch);
    // f arrays are free now and receive p_v, p_omega
    const double k2m = k2 - c1n;
    dcheb(n, phr, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        fvr[k] = -t2[k] + k2m * phr[k] - c2 * hvr[k];
    dcheb(n, phi, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        fvi[k] = -t2[k] + k2m * phi[k] - c2 * hvi[k];
    dcheb(n, domr, t1);
    for (int k = 0; k < n; k++)
        fwr[k] = -t1[k] + k2m * omr[k] - c2 * hwr[k];
    dcheb(n, domi, t1);
    for (int k = 0; k < n; k++)
        fwi[k] = -t1[k] + k2m * omi[k] - c2 * hwi[k];
    // u, w, om_x, om_z (linearbl.f:1190-1212) and the outputs in
    // Chebyshev space; the k = 0 column takes the mean flow (1640-1652)
    double2 *Cb = a.C + (((size_t)z * a.ny) * a.nxh + x);
    const size_t fs = (size_t)a.nz * a.ny * a.nxh;   // slot stride
    double2 *Ub = a.U + (((size_t)z * a.ny) * a.nxh + x);
    for (int k = 0; k < n; k++) {
        const size_t o = (size_t)k * a.nxh;
        double2 uu, vv, ww, ox, oz;
        if (!k0) {
            uu = make_double2((-alfa * dvi[k] + beta * omi[k]) * k2i,
                              (alfa * dvr[k] - beta * omr[k]) * k2i);
            ww = make_double2((-alfa * omi[k] - beta * dvi[k]) * k2i,
                              (alfa * omr[k] + beta * dvr[k]) * k2i);
            oz = make_double2((-beta * domi[k] + alfa * phi[k]) * k2i,
                              (beta * domr[k] - alfa * phr[k]) * k2i);
            ox = make_double2((-alfa * domi[k] - beta * phi[k]) * k2i,
                              (alfa * domr[k] + beta * phr[k]) * k2i);
            vv = make_double2(vr[k], vi[k]);
        } else {
            uu = make_double2(u0[k], 0.);
            ww = make_double2(w0[k], 0.);
            ox = make_double2(dw0[k], 0.);
            oz = make_double2(-du0[k], 0.);
            vv = make_double2(0., 0.);
        }
        Cb[0 * fs + o] = uu;
        Cb[1 * fs + o] = vv;
        Cb[2 * fs + o] = ww;
        Cb[3 * fs + o] = ox;
        Cb[4 * fs + o] = oz;
        Ub[5 * fs + o] = make_double2(fvr[k], fvi[k]);
        Ub[6 * fs + o] = make_double2(fwr[k], fwi[k]);
    }
}

__global__ void __launch_bounds__(256) k_prhs(ColArgs a, double an)
{
    const int ncol = a.nz * a.nxh;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol)
        return;
    const int z = t / a.nxh, x = t - z * a.nxh;
    const int n = a.ny;
    Cols c{a, t, z, x, ncol};
    const double re = a.re;
    const double alfa = a.alfa[x], beta = a.beta[z];
    const double k2 = alfa * alfa + beta * beta;
    const double lam = k2 - 2. * re / an;
    ColView uR = c.cplx(a.C, 0, 0), uI = c.cplx(a.C, 0, 1);
    ColView vR = c.cplx(a.C, 1, 0), vI = c.cplx(a.C, 1, 1);
    ColView wR = c.cplx(a.C, 2, 0), wI = c.cplx(a.C, 2, 1);
    ColView oyr = c.W(0), oyi = c.W(1), dwr = c.W(2), dwi = c.W(3);
    ColView dur = c.W(4), dui = c.W(5), t1 = c.W(6), t2 = c.W(7);
    ColView phr = c.W(8), phi = c.W(9), pwr = c.W(10), pwi = c.W(11);
    ColView pvr = c.W(12), pvi = c.W(13);
    if (z == 0 && x == 0) {              // prhs.f:534-546
        dcheb(n, uR, t1);
        dcheb(n, t1, t2);
        for (int k = 0; k < n; k++)
            a.puw[k] = -2. * re / an * uR[k] - t2[k];
        dcheb(n, wR, t1);
        dcheb(n, t1, t2);
        for (int k = 0; k < n; k++)
            a.puw[n + k] = -2. * re / an * wR[k] - t2[k];
    }
    for (int k = 0; k < n; k++) {       // prhs.f:563-573
        oyr[k] = -uI[k] * beta + wI[k] * alfa;
        oyi[k] = uR[k] * beta - wR[k] * alfa;
    }
    dcheb(n, wR, dwr);
    dcheb(n, wI, dwi);
    dcheb(n, uR, dur);
    dcheb(n, uI, dui);
    // phi = D^2 v - k^2 v ; p_omega = -D^2 om_y + lam om_y ; p_v = -D^2 phi + lam phi
    dcheb(n, vR, t1);
    dcheb(n, t1, phr);
    dcheb(n, vI, t1);
    dcheb(n, t1, phi);
    for (int k = 0; k < n; k++) {
        phr[k] = phr[k] - k2 * vR[k];
        phi[k] = phi[k] - k2 * vI[k];
    }
    dcheb(n, oyr, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        pwr[k] = -t2[k] + lam * oyr[k];
    dcheb(n, oyi, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        pwi[k] = -t2[k] + lam * oyi[k];
    dcheb(n, phr, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        pvr[k] = -t2[k] + lam * phr[k];
    dcheb(n, phi, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        pvi[k] = -t2[k] + lam * phi[k];
    double2 *Cb = a.C + (((size_t)z * a.ny) * a.nxh + x);
    double2 *Ub = a.U + (((size_t)z * a.ny) * a.nxh + x);
    const size_t fs = (size_t)a.nz * a.ny * a.nxh;
    for (int k = 0; k < n; k++) {       // prhs.f:578-590 then storage
        const size_t o = (size_t)k * a.nxh;
        const double vr = vR[k], vi = vI[k];
        Cb[3 * fs + o] = make_double2(dwr[k] + vi * beta, dwi[k] - vr * beta);
        Cb[4 * fs + o] = make_double2(-dur[k] - vi * alfa, -dui[k] + vr * alfa);
        Ub[5 * fs + o] = make_double2(pvr[k], pvi[k]);
        Ub[6 * fs + o] = make_double2(pwr[k], pwi[k]);
    }
}

}  // namespace

void launch_cheb_gemm(const double *A, const GemmIO &io, hipStream_t s)
{
    dim3 grid((io.ncols + GBN - 1) / GBN, (io.ny + GBM - 1) / GBM, io.nf * io.nz);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, A, io);
}

void launch_tau(const ColArgs &a, double an, double bn, double anp1, double bnp1, hipStream_t s)
{
    const int ncol = a.nz * a.nxh;
    hipLaunchKernelGGL(k_tau, dim3((ncol + 255) / 256), dim3(256), 0, s, a, an, bn, anp1, bnp1);
}

void launch_prhs(const ColArgs &a, double an, hipStream_t s)
{
    const int ncol = a.nz * a.nxh;
    hipLaunchKernelGGL(k_prhs, dim3((ncol + 255) / 256), dim3(256), 0, s, a, an);
}

}  // namespace sg
