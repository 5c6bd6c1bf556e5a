// Fringe function of bla's spatial simulations (host side): the strength
// lambda(x) of the forcing -lambda (u - U) toward the base flow that
// fring.f and fringp.f add to the nonlinear term, at the physical points of
// the 3/2-dealiased x grid (ffun.f:37-75, without the Orr-Sommerfeld modes).
#include <cmath>

#include "sg_internal.h"

namespace sg {

// lam[j], j < nxp = 3 nx/2, at x_j = -xl/2 + j xl/nxp + xsc.  As in
// ffun.f, only the even-point index range xstart..xend (1-based) of the
// fringe is set, the rest stays zero; a fringe that wraps across the box
// end is the union of the two ranges, each raised by fmax.
int fringe_function(int nx, double xl, const simson_fringe_t *f, double *lam)
{
    const int nxp = 3 * nx / 2, h = nxp / 2;
    if (nx <= 0 || xl <= 0. || f->fmax <= 0. || f->frise <= 0. || f->ffall <= 0. ||
        f->fstart >= f->fend || std::fabs(f->fstart) > .5 * xl || std::fabs(f->fend) > .5 * xl)
        return SIMSON_E_ARG;
    const double r = 0.499999999999999999;       // as ffun.f:24 (0.5 in FP64)
    const double dnx = xl / nxp, xsc = f->xsc;
    const double fstc = f->fstart + (int)((xsc - f->fstart) / xl + r) * xl;
    const double fenc = f->fend + (int)((xsc - f->fend) / xl + r) * xl;
    const int xstart = (int)(h * (fstc - xsc + xl / 2.) / xl) + 1;
    const int xend = (int)(h * (fenc - xsc + xl / 2.) / xl) + 1;
    const bool wrap = xstart > xend;
    for (int j = 0; j < nxp; j++)
        lam[j] = 0.;
    for (int i = wrap ? 0 : 1; i <= 1; i++) {
        const int xst = i == 0 ? 1 : xstart;
        const int xen = i == 0 ? xend : (wrap ? h : xend);
        for (int x = std::max(1, xst); x <= std::min(h, xen); x++) {
            const double xc1 = (double)(2 * x - 1 - h - 1) * dnx + xsc, xc2 = xc1 + dnx;
            double f1 = f->fmax * (smooth_step((xc1 - fstc) / f->frise) -
                                   smooth_step((xc1 - fenc) / f->ffall + 1.));
            double f2 = f->fmax * (smooth_step((xc2 - fstc) / f->frise) -
                                   smooth_step((xc2 - fenc) / f->ffall + 1.));
            if (wrap) {
                f1 += f->fmax;
                f2 += f->fmax;
            }
            lam[2 * (x - 1)] = f1;
            lam[2 * (x - 1) + 1] = f2;
        }
    }
    return 0;
}

}  // namespace sg
