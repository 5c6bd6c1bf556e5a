// Blowing and suction strip at the lower wall, bla.i wbci = 1 (host side).
//
// cwallbc.f:53-115 sets, on the physical points of the 3/2 grid,
//   v_w = max(s P, s P damp),  P = amp f(x) cos(zbet z),
//   s = sin(tomeg tc) (1 when tomeg = 0),
//   f(x) = S((x - xstart)/xrise) - S((x - xend)/xfall + 1)
// (S the smooth step of step.f), and transforms it forward and truncates
// it to the stored modes (cwallbc.f:380-407).  For s >= 0 the wall field is
// s max(P, P damp), for s < 0 it is s min(P, P damp), so both patterns are
// transformed once here and the step only scales one of them by s.
#include <cmath>
#include <vector>

#include "sg_internal.h"

namespace sg {

int wall_spectra(int nx, int nz, double xl, double zl, const simson_wall_t *w, double2 *pmax,
                 double2 *pmin)
{
    if (nx <= 0 || nz <= 0 || xl <= 0. || zl <= 0. || w->xrise <= 0. || w->xfall <= 0.)
        return SIMSON_E_ARG;
    const int nxp = 3 * nx / 2, nzp = 3 * nz / 2, h = nxp / 2, nxh = nx / 2;
    double xstart = w->xstart, xend = w->xend;
    if (xend >= xl / 2.) {
        xend -= xl;
        if (xstart >= xl / 2.)
            xstart -= xl;
    }
    // f(x) at the even and odd points (xc11, xc22 of cwallbc.f)
    std::vector<double> fx(nxp);
    for (int x = 1; x <= h; x++) {
        double xc11 = (double)(2 * x - 1 - h - 1) / (double)nxp * xl + w->xsc;
        xc11 -= (int)((xc11 + xl / 2.) / xl) * xl;
        double xc22 = (double)(2 * x - h - 1) / (double)nxp * xl + w->xsc;
        xc22 -= (int)((xc22 + xl / 2.) / xl) * xl;
        fx[2 * (x - 1)] = smooth_step((xc11 - xstart) / w->xrise) -
                          smooth_step((xc11 - xend) / w->xfall + 1.);
        fx[2 * (x - 1) + 1] = smooth_step((xc22 - xstart) / w->xrise) -
                              smooth_step((xc22 - xend) / w->xfall + 1.);
    }
    const double pi = 3.14159265358979323846;
    // x transform of both patterns row by row, kx < nx/2, then z
    std::vector<double2> rowp((size_t)nzp * nxh), rowm((size_t)nzp * nxh);
    std::vector<double> cx(nxp), sx(nxp), cz(nzp), sz(nzp);
    for (int j = 0; j < nxp; j++) {
        cx[j] = std::cos(2. * pi * j / nxp);
        sx[j] = std::sin(2. * pi * j / nxp);
    }
    for (int j = 0; j < nzp; j++) {
        cz[j] = std::cos(2. * pi * j / nzp);
        sz[j] = std::sin(2. * pi * j / nzp);
    }
    for (int z = 0; z < nzp; z++) {
        const double zc = zl * (double)(z + 1 - nzp / 2 - 1) / (double)nzp + w->zsc;
        const double cb = std::cos(w->zbet * zc);
        for (int kx = 0; kx < nxh; kx++) {
            double ar = 0., ai = 0., br = 0., bi = 0.;
            for (int j = 0; j < nxp; j++) {
                const double p = w->amp * fx[j] * cb;
                const double pmx = std::fmax(p, p * w->damp), pmn = std::fmin(p, p * w->damp);
                const int m = (int)(((long)kx * j) % nxp);
                ar += pmx * cx[m];
                ai -= pmx * sx[m];
                br += pmn * cx[m];
                bi -= pmn * sx[m];
            }
            rowp[(size_t)z * nxh + kx] = make_double2(ar, ai);
            rowm[(size_t)z * nxh + kx] = make_double2(br, bi);
        }
    }
    const double norm = 1. / ((double)nxp * (double)nzp);
    for (int zs = 0; zs < nz; zs++) {
        for (int kx = 0; kx < nxh; kx++) {
            double2 a = make_double2(0., 0.), b = make_double2(0., 0.);
            if (zs != nz / 2) {
                const int kz = zs < nz / 2 ? zs : zs + nzp - nz;     // padded row
                for (int z = 0; z < nzp; z++) {
                    const int m = (int)(((long)kz * z) % nzp);
                    const double c = cz[m], sn = -sz[m];
                    const double2 p = rowp[(size_t)z * nxh + kx], q = rowm[(size_t)z * nxh + kx];
                    a.x += p.x * c - p.y * sn;
                    a.y += p.x * sn + p.y * c;
                    b.x += q.x * c - q.y * sn;
                    b.y += q.x * sn + q.y * c;
                }
            }
            pmax[(size_t)zs * nxh + kx] = make_double2(a.x * norm, a.y * norm);
            pmin[(size_t)zs * nxh + kx] = make_double2(b.x * norm, b.y * norm);
        }
    }
    return 0;
}

}  // namespace sg
