// Localized volume force of bla.i loctyp = 3 (host side).
//
// locf.f:358-410 adds to the nonlinear term, on the physical points of the
// 3/2 x grid and at every z, the force
//   F_i(x, y, t) = amp_i exp(-((1 + y)/yscale)^2) f(x) sin(tomeg tc),
//   f(x) = ((x - xloc0)/xscale)^xtype exp(-((x - xloc0)/xscale)^2),
// with x wrapped into the box (x from the box centre plus xsc) and tc the
// time at the start of the substep.  Uniform in z, it lives in the kz = 0
// plane: per kx < nx/2 its x Fourier coefficient (over the nxp points,
// divided by nxp) times the y profile.  The library keeps that pattern as
// Chebyshev coefficients in y (what the column solver adds to H) and scales
// it by sin(tomeg tc) each substep.
#include <cmath>
#include <vector>

#include "sg_internal.h"

namespace sg {

int local_force_pattern(int nx, int ny, double xl, const simson_locf_t *lf, double2 *phys)
{
    if (nx <= 0 || nx % 4 != 0 || ny < 3 || !(xl > 0.) || !lf || !(lf->xscale > 0.) ||
        !(lf->yscale > 0.) || !(lf->xtype == 0. || lf->xtype == 1.))
        return SIMSON_E_ARG;
    const int nxp = 3 * nx / 2, h = nxp / 2, nxh = nx / 2;
    const double pi = 3.14159265358979323846;
    // f(x) at the even and odd points of the 3/2 grid (xc11, xc22 of locf.f)
    std::vector<double> fx(nxp);
    auto fof = [&](double xc) {
        xc -= (double)(int)((xc + xl / 2.) / xl) * xl;
        const double r = (xc - lf->xloc0) / lf->xscale;
        return std::pow(r, lf->xtype) * std::exp(-r * r);
    };
    for (int x = 1; x <= h; x++) {
        fx[2 * (x - 1)] = fof((double)(2 * x - 1 - h - 1) / (double)nxp * xl + lf->xsc);
        fx[2 * (x - 1) + 1] = fof((double)(2 * x - h - 1) / (double)nxp * xl + lf->xsc);
    }
    std::vector<double2> fk(nxh);
    for (int k = 0; k < nxh; k++) {
        double sr = 0., si = 0.;
        for (int j = 0; j < nxp; j++) {
            const double a = 2. * pi * (double)(((long)k * j) % nxp) / (double)nxp;
            sr += fx[j] * std::cos(a);
            si -= fx[j] * std::sin(a);
        }
        fk[k] = make_double2(sr / (double)nxp, si / (double)nxp);
    }
    const double amp[3] = {lf->ampx, lf->ampy, lf->ampz};
    for (int y = 0; y < ny; y++) {
        const double yc = 1. + std::cos(pi * (double)y / (double)(ny - 1));
        const double fy = std::exp(-(yc / lf->yscale) * (yc / lf->yscale));
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < nxh; k++)
                phys[((size_t)c * ny + y) * nxh + k] =
                    make_double2(amp[c] * fy * fk[k].x, amp[c] * fy * fk[k].y);
    }
    return 0;
}

}  // namespace sg
