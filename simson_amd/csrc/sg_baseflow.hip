// Boundary-layer base flows of spatial simulations (host side).
//
// * The Falkner-Skan-Cooke similarity table that bls/fsc.f writes to
//   fsc.dat and bla reads (rbla.f): f''' = -f f'' - beta (1 - f'^2),
//   f(0) = f'(0) = 0, f'(inf) = 1, beta = 2m/(m+1), and Cooke's
//   g'' = -f g', g(0) = 0, g(inf) = 1, on eta_i = ymax i/(n-1).  Solved by
//   shooting on s = f''(0) with Newton's method, the classical fourth-order
//   Runge-Kutta scheme with one step per table interval for the profile and
//   its derivative with respect to s.  The table's own truncation error
//   (~1e-14 for Blasius at n = 20001, ymax = 30) is part of the base flow
//   the reference integrates, so the same discretisation is kept here.
// * The temporal boundary layers of cbflow.f:58-84 (fltype 3, -1, -2): the
//   similarity profile at the leading-edge distance x0 at every x, and the
//   forcing of nonlinbl.f:771-795 that keeps such a parallel profile from
//   diffusing.
// * The spatial boundary layer of cbflow.f:161-300 (fltype 6, 7, 8): at
//   every point of the 3/2 x grid the similarity profile at the local
//   distance from the leading edge, blended with the profile one box
//   length downstream over the blending zone (bstart, blength), and the
//   wall-normal velocity from continuity, V = -int_{-1}^{y} dU/dx.
// * The derivatives of the base flow at y = +1 that the free-stream
//   condition of ibc = 101 imposes (bflow.f:80-148), as x Fourier
//   coefficients normalised by 1/nxp.
#include <cmath>
#include <vector>

#include "sg_internal.h"

namespace sg {

namespace {

constexpr double kPi = 3.14159265358979323846;

// one classical Runge-Kutta step of size h, with the increments k_i = h f
// formed first (the evaluation order of fsc.f's rk4, so that the table
// rounds as the reference's does)
template <int N, typename F>
void rk4(const double (&y0)[N], double (&y1)[N], double h, F &&rhs)
{
    double k1[N], k2[N], k3[N], k4[N], y[N];
    rhs(y0, k1);
    for (int i = 0; i < N; i++) {
        k1[i] = h * k1[i];
        y[i] = y0[i] + .5 * k1[i];
    }
    rhs(y, k2);
    for (int i = 0; i < N; i++) {
        k2[i] = h * k2[i];
        y[i] = y0[i] + .5 * k2[i];
    }
    rhs(y, k3);
    for (int i = 0; i < N; i++) {
        k3[i] = h * k3[i];
        y[i] = y0[i] + k3[i];
    }
    rhs(y, k4);
    for (int i = 0; i < N; i++) {
        k4[i] = h * k4[i];
        y1[i] = y0[i] + (k1[i] + 2. * k2[i] + 2. * k3[i] + k4[i]) / 6.;
    }
}

// Fourth-order interpolation in a uniform table (cubip.f): the cubic
// through the four entries around x, so that the interval holding x is the
// middle one (the first and last intervals use the end four), and the last
// entry beyond the table.  With p = 2 x/dx - (2m + 1) the nodes sit at
// p = -3, -1, 1, 3; the cubic is evaluated in monomial form, the form whose
// rounding the base flow's free-stream derivatives inherit (they amplify
// it by ~ny^2, see spatial_base_flow).
double interp4(double x, const double *tab, int stride, double dx, int n)
{
    if (x > dx * (double)(n - 1))
        return tab[(size_t)(n - 1) * stride];
    int m = (int)(x / dx);                  // 1-based start of the stencil
    m = m < 1 ? 1 : (m > n - 3 ? n - 3 : m);
    const double f0 = tab[(size_t)(m - 1) * stride], f1 = tab[(size_t)m * stride];
    const double f2 = tab[(size_t)(m + 1) * stride], f3 = tab[(size_t)(m + 2) * stride];
    const double p = 2. * x / dx - (double)(2 * m + 1);
    // Lagrange basis on the nodes expanded in powers of p
    const double c0 = (-3. * f0 + 27. * f1 + 27. * f2 - 3. * f3) / 48.;
    const double c1 = (f0 - 27. * f1 + 27. * f2 - f3) / 48.;
    const double c2 = (3. * f0 - 3. * f1 - 3. * f2 + 3. * f3) / 48.;
    const double c3 = (-f0 + 3. * f1 - 3. * f2 + f3) / 48.;
    return c0 + c1 * p + c2 * p * p + c3 * p * p * p;
}

// xlim.f: x itself above xlow, else a smooth floor that reaches xlow/2
double xfloor(double x, double xlow)
{
    if (x >= xlow)
        return x;
    if (x <= 0.)
        return .5 * xlow;
    const double r = x / xlow;
    return xlow * (.5 + r * r * r - .5 * r * r * r * r);
}

// Chebyshev coefficients of the values at eta_j = cos(pi j/(n-1))
// (DCT-I with half weights at both ends, scaled by 2/(n-1)), and back
void cheb_coeffs(int n, const double *u, double *a, const double *cs)
{
    for (int k = 0; k < n; k++) {
        double s = 0.;
        for (int j = 0; j < n; j++) {
            const double w = (j == 0 || j == n - 1) ? .5 : 1.;
            s += w * u[j] * cs[(size_t)((long)j * k % (2 * (n - 1)))];
        }
        const double ck = (k == 0 || k == n - 1) ? .5 : 1.;
        a[k] = s * ck * 2. / (double)(n - 1);
    }
}

void cheb_values(int n, const double *a, double *u, const double *cs)
{
    for (int j = 0; j < n; j++) {
        double s = 0.;
        for (int k = 0; k < n; k++)
            s += a[k] * cs[(size_t)((long)j * k % (2 * (n - 1)))];
        u[j] = s;
    }
}

}  // namespace

int fsc_table(double m, int n, double ymax, double eps, double *tab, double *dstar2)
{
    if (n < 4 || !(ymax > 0.) || !(eps > 0.) || m < -0.0904 || !tab)
        return SIMSON_E_ARG;
    const double beta = 2. * m / (m + 1.);
    // steps between the rounded grid points eta_i = ymax i/(n-1)
    auto step = [&](int i) {
        return ymax * (double)i / (double)(n - 1) - ymax * (double)(i - 1) / (double)(n - 1);
    };
    // profile (f, f', f'') and its derivative with respect to s = f''(0)
    auto frhs = [beta](const double (&y)[6], double (&d)[6]) {
        d[0] = y[1];
        d[1] = y[2];
        d[2] = -y[0] * y[2] - beta * (1. - y[1] * y[1]);
        d[3] = y[4];
        d[4] = y[5];
        d[5] = -(y[0] * y[5] + y[2] * y[3]) + 2. * beta * y[4] * y[1];
    };
    std::vector<double> F((size_t)n * 3);
    double s = 1.1, res = 1.;
    int iter = 0;
    while (res > eps) {
        if (++iter > 200)
            return SIMSON_E_ARG;
        double y[6] = {0., 0., s, 0., 0., 1.};
        F[0] = 0.;
        F[1] = 0.;
        F[2] = s;
        for (int i = 1; i < n; i++) {
            double y1[6];
            rk4(y, y1, step(i), frhs);
            for (int q = 0; q < 6; q++)
                y[q] = y1[q];
            F[3 * (size_t)i] = y[0];
            F[3 * (size_t)i + 1] = y[1];
            F[3 * (size_t)i + 2] = y[2];
        }
        if (!std::isfinite(y[1]) || !std::isfinite(y[4]) || y[4] == 0.)
            return SIMSON_E_ARG;
        res = std::fabs(y[1] - 1.);
        s -= (y[1] - 1.) / y[4];
    }
    // Cooke's g with g'(0) = f''(0), integrated with f, then scaled to
    // g(ymax) = 1
    auto grhs = [beta](const double (&y)[5], double (&d)[5]) {
        d[0] = y[1];
        d[1] = -y[1] * y[2];
        d[2] = y[3];
        d[3] = y[4];
        d[4] = -y[2] * y[4] - beta * (1. - y[3] * y[3]);
    };
    std::vector<double> G((size_t)n * 2);
    {
        double y[5] = {0., F[2], 0., 0., F[2]};
        G[0] = 0.;
        G[1] = F[2];
        for (int i = 1; i < n; i++) {
            double y1[5];
            rk4(y, y1, step(i), grhs);
            for (int q = 0; q < 5; q++)
                y[q] = y1[q];
            G[2 * (size_t)i] = y[0];
            G[2 * (size_t)i + 1] = y[1];
        }
    }
    const double gs = 1. / G[2 * (size_t)(n - 1)];
    for (int i = 0; i < n; i++) {
        const double f = F[3 * (size_t)i], f1 = F[3 * (size_t)i + 1], f2 = F[3 * (size_t)i + 2];
        const double g = G[2 * (size_t)i] * gs, g1 = G[2 * (size_t)i + 1] * gs;
        double *r = tab + (size_t)i * 8;
        r[0] = ymax * (double)i / (double)(n - 1);
        r[1] = f;
        r[2] = f1;
        r[3] = f2;
        r[4] = -f * f2 - beta * (1. - f1 * f1);
        r[5] = g;
        r[6] = g1;
        r[7] = -g1 * f;
    }
    if (dstar2)
        *dstar2 = ymax - F[3 * (size_t)(n - 1)];
    return 0;
}

namespace {

bool parallel_type(int fltype)
{
    return fltype == 3 || fltype == -1 || fltype == -2;
}

}  // namespace

int spatial_base_flow(const simson_baseflow_t *b, const double *tab, int ntab, double *base,
                      double *bjr, double *bji, double *du0upp)
{
    if (!b || !tab || !base || ntab < 4 || b->nx < 4 || b->nx % 4 != 0 || b->ny < 5 ||
        !(b->xl > 0.) || !(b->re > 0.) || !(b->dstar > 0.) || !(b->dstar2 > 0.) ||
        !(b->fltype == 6 || b->fltype == 7 || b->fltype == 8 || parallel_type(b->fltype)) ||
        (!parallel_type(b->fltype) && !(b->blength > 0.)))
        return SIMSON_E_ARG;
    const int nxp = 3 * b->nx / 2, hp = nxp / 2, ny = b->ny, nxh = b->nx / 2;
    const double xl = b->xl, re = b->re, rlam = b->rlam;
    const double ymax = tab[(size_t)(ntab - 1) * 8];
    const double dyb = ymax / (double)(ntab - 1);
    // leading-edge distance of the box centre (rbla.f)
    const double x0 = b->dstar * b->dstar * re * (rlam + 1.) / (b->dstar2 * b->dstar2 * 2.);
    // Cooke's g enters only the Falkner-Skan-Cooke flows (fltype 8, -2);
    // the other types have no spanwise base velocity (rbla.f zeroes g', g''
    // and rdiscbl.f sets spanv = 0 for them)
    const double spanv = (b->fltype == 8 || b->fltype == -2) ? b->spanv : 0.;
    std::vector<double> U((size_t)ny * nxp), V((size_t)ny * nxp), W((size_t)ny * nxp);
    std::vector<double> eta(ny);
    for (int y = 0; y < ny; y++)
        eta[y] = std::cos(kPi * (double)y / (double)(ny - 1));
    if (parallel_type(b->fltype)) {
        // temporal boundary layers (cbflow.f:58-84): the profile at x0 at
        // every x, no wall-normal velocity
        const double s = std::sqrt(re / (2. * x0));
        for (int y = 0; y < ny; y++) {
            const double e = (1. + eta[y]) * s;
            const double u = interp4(e, tab + 2, 8, dyb, ntab) + b->u0low;
            const double w = spanv * interp4(e, tab + 5, 8, dyb, ntab) + b->w0low;
            for (int j = 0; j < nxp; j++) {
                U[(size_t)y * nxp + j] = u;
                V[(size_t)y * nxp + j] = 0.;
                W[(size_t)y * nxp + j] = w;
            }
        }
    }
    const double xmin = .5 * x0;
    for (int j = 0; j < nxp && !parallel_type(b->fltype); j++) {
        // x_j = -xl/2 + j xl/nxp + xsc, wrapped into [bstart, bstart + xl)
        const double xc = (double)(j - hp) / (double)nxp * xl + b->xsc;
        const double xb = xc - ((double)(int)((xc - b->bstart) / xl + 1.) - 1.) * xl;
        const double bs = smooth_step((xb - b->bstart) / b->blength);
        const double xa = xfloor(xb + x0, xmin), xd = xfloor(xb + x0 + xl, xmin);
        const double ua = std::pow(xa / x0, rlam), ud = std::pow(xd / x0, rlam);
        const double sa = std::sqrt((rlam + 1.) * ua * re / (2. * xa));
        const double sd = std::sqrt((rlam + 1.) * ud * re / (2. * xd));
        for (int y = 0; y < ny; y++) {
            const double yb = 1. + eta[y];
            const double fa = interp4(yb * sa, tab + 2, 8, dyb, ntab);
            const double fd = interp4(yb * sd, tab + 2, 8, dyb, ntab);
            U[(size_t)y * nxp + j] = ua * fa * bs + ud * fd * (1. - bs) + b->u0low;
            double w = b->w0low;
            if (spanv != 0.) {
                const double ga = interp4(yb * sa, tab + 5, 8, dyb, ntab);
                const double gd = interp4(yb * sd, tab + 5, 8, dyb, ntab);
                w += spanv * (ga * bs + gd * (1. - bs));
            }
            W[(size_t)y * nxp + j] = w;
        }
    }
    // V = -int_{-1}^{y} dU/dx dy (cbflow.f:241-300): x Fourier coefficients
    // (Nyquist dropped), Chebyshev coefficients in y, times -i alpha,
    // integrated in y with V(-1) = 0, back to physical points
    std::vector<double> cs(2 * (size_t)(ny - 1));
    for (size_t q = 0; q < cs.size(); q++)
        cs[q] = std::cos(kPi * (double)q / (double)(ny - 1));
    std::vector<double> cx(nxp), sx(nxp);
    for (int j = 0; j < nxp; j++) {
        cx[j] = std::cos(2. * kPi * j / nxp);
        sx[j] = std::sin(2. * kPi * j / nxp);
    }
    std::vector<double> hr((size_t)ny * hp), hi((size_t)ny * hp);
    for (int y = 0; y < ny; y++)
        for (int k = 0; k < hp; k++) {
            double ar = 0., ai = 0.;
            for (int j = 0; j < nxp; j++) {
                const int q = (int)(((long)k * j) % nxp);
                ar += U[(size_t)y * nxp + j] * cx[q];
                ai -= U[(size_t)y * nxp + j] * sx[q];
            }
            hr[(size_t)y * hp + k] = ar / (double)nxp;
            hi[(size_t)y * hp + k] = ai / (double)nxp;
        }
    std::vector<double> col(ny), ar(ny), ai(ny), br(ny), bi(ny), vr(ny), vi(ny);
    std::vector<double> Vr((size_t)ny * hp), Vi((size_t)ny * hp);
    for (int k = 0; k < hp && !parallel_type(b->fltype); k++) {
        const double alfa = 2. * kPi / xl * (double)k;
        for (int y = 0; y < ny; y++)
            col[y] = hr[(size_t)y * hp + k];
        cheb_coeffs(ny, col.data(), ar.data(), cs.data());
        for (int y = 0; y < ny; y++)
            col[y] = hi[(size_t)y * hp + k];
        cheb_coeffs(ny, col.data(), ai.data(), cs.data());
        // -i alpha (ar + i ai) = alpha ai - i alpha ar
        for (int q = 0; q < ny; q++) {
            const double t = alfa * ai[q];
            ai[q] = -alfa * ar[q];
            ar[q] = t;
        }
        // Chebyshev integral (icheb.f): b_k = (a_{k-1} - a_{k+1})/(2k),
        // b_1 = a_0 - a_2/2, b_{n-1} = a_{n-2}/(2(n-1)); b_0 from V(-1) = 0
        auto integ = [&](const std::vector<double> &a, std::vector<double> &bb) {
            for (int q = 2; q < ny - 1; q++)
                bb[q] = (a[q - 1] - a[q + 1]) * (.5 / (double)q);
            bb[ny - 1] = a[ny - 2] * (.5 / (double)(ny - 1));
            bb[1] = a[0] - .5 * a[2];
            double c0 = 0.;
            for (int q = 1; q < ny; q++)
                c0 += (q & 1) ? bb[q] : -bb[q];
            bb[0] = c0;
        };
        integ(ar, br);
        integ(ai, bi);
        cheb_values(ny, br.data(), vr.data(), cs.data());
        cheb_values(ny, bi.data(), vi.data(), cs.data());
        for (int y = 0; y < ny; y++) {
            Vr[(size_t)y * hp + k] = vr[y];
            Vi[(size_t)y * hp + k] = vi[y];
        }
    }
    for (int y = 0; y < ny && !parallel_type(b->fltype); y++)
        for (int j = 0; j < nxp; j++) {
            double s = Vr[(size_t)y * hp];
            for (int k = 1; k < hp; k++) {
                const int q = (int)(((long)k * j) % nxp);
                s += 2. * (Vr[(size_t)y * hp + k] * cx[q] - Vi[(size_t)y * hp + k] * sx[q]);
            }
            V[(size_t)y * nxp + j] = s;
        }
    const size_t pl = (size_t)ny * nxp;
    for (size_t i = 0; i < pl; i++) {
        base[i] = U[i];
        base[pl + i] = V[i];
        base[2 * pl + i] = W[i];
    }
    // d/dy at y = +1 per physical x as a weighted sum of the values,
    // w_y = 2/(n-1) sum_k k^2 c_y c_k cos(pi k y/(n-1)) (c = 1/2 at both
    // ends; T_k'(1) = k^2), in the order of preprbl.f:88-121, then x Fourier
    // coefficients / nxp for kx < nx/2 (bflow.f:80-105).  The weights reach
    // ~ny^2/3 while the sum is ~1e-5 of U for a boundary layer, so rounding
    // of the summands (~1e-16 U each) is what limits these derivatives.
    if (bjr || bji || du0upp) {
        const double pf = 3.1415926535897932385;
        std::vector<double> wd(ny, 0.);
        for (int i = 0; i < ny; i++) {
            const double ci = (i == 0 || i == ny - 1) ? .5 : 1.;
            for (int y = 0; y < ny; y++) {
                const double cy = (y == 0 || y == ny - 1) ? .5 : 1.;
                wd[y] += (double)i * (double)i * cy * ci *
                         std::cos(pf * (double)(i * y) / (double)(ny - 1)) *
                         (2. / (double)(ny - 1));
            }
        }
        std::vector<double> d((size_t)3 * nxp, 0.);
        for (int c = 0; c < 3; c++)
            for (int y = 0; y < ny; y++)
                for (int j = 0; j < nxp; j++)
                    d[(size_t)c * nxp + j] += base[(size_t)c * pl + (size_t)y * nxp + j] * wd[y];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < nxh; k++) {
                double sr = 0., si = 0.;
                for (int j = 0; j < nxp; j++) {
                    const int q = (int)(((long)k * j) % nxp);
                    sr += d[(size_t)c * nxp + j] * cx[q];
                    si -= d[(size_t)c * nxp + j] * sx[q];
                }
                if (bjr)
                    bjr[(size_t)c * nxh + k] = sr / (double)nxp;
                if (bji)
                    bji[(size_t)c * nxh + k] = si / (double)nxp;
                if (du0upp && c == 0 && k == 0)
                    *du0upp = sr / (double)nxp;
            }
    }
    return 0;
}

int parallel_force(const simson_baseflow_t *b, const double *tab, int ntab, double cdev,
                   double t, double *force)
{
    if (!b || !tab || !force || ntab < 4 || b->ny < 5 || !(b->re > 0.) || !(b->dstar > 0.) ||
        !(b->dstar2 > 0.) || !parallel_type(b->fltype))
        return SIMSON_E_ARG;
    const int ny = b->ny;
    const double re = b->re, rlam = b->rlam;
    const double ymax = tab[(size_t)(ntab - 1) * 8];
    const double dyb = ymax / (double)(ntab - 1);
    const double x0 = b->dstar * b->dstar * re * (rlam + 1.) / (b->dstar2 * b->dstar2 * 2.);
    // nonlinbl.f:771-795 at the boundary layer's position xblc = x0 + cdev t
    // (bla.f:1826), without the unnormalised transforms' factor nxp nzp
    const double xblc = x0 + cdev * t;
    if (!(xblc > 0.))
        return SIMSON_E_ARG;
    const double xx = std::pow(xblc / x0, rlam);
    const double h1u = cdev * (rlam / x0) * std::pow(xblc / x0, rlam - 1.);
    const double h3u = .5 * (rlam + 1.) / xblc;
    for (int y = 0; y < ny; y++) {
        const double ybl = 1. + std::cos(kPi * (double)y / (double)(ny - 1));
        const double e = ybl * std::sqrt((rlam + 1.) * re * xx / (2. * xblc));
        const double h2u = cdev * .5 * (rlam - 1.) * std::pow(xx, 1.5) *
                           std::sqrt(.5 * (rlam + 1.) * re) * std::pow(1. / xblc, 1.5) * ybl;
        const double d1f = interp4(e, tab + 2, 8, dyb, ntab);
        const double d2f = interp4(e, tab + 3, 8, dyb, ntab);
        const double d3f = interp4(e, tab + 4, 8, dyb, ntab);
        force[y] = d1f * h1u + d2f * h2u - d3f * xx * xx * h3u;
        force[ny + y] = 0.;
        force[2 * ny + y] = 0.;
        if (b->fltype == -2)
            force[2 * ny + y] = -b->spanv * h3u * interp4(e, tab + 7, 8, dyb, ntab);
    }
    return 0;
}

}  // namespace sg
