// ADM-RT large-eddy model of bla (iles = 1, ihighorder = 2): the filter
// operators, built on the host once per configuration.
//
// The relaxation term of nonlinbl.f:920-966 subtracts chi nxp nzp g_i from
// H_i, where g = u - G_xz (u - H_y^N u) is the high-pass filtered velocity
// computed at the end of linearbl.f:1871-1962 (and of prhs.f:95-195):
//   * H_y: the implicit compact high-pass filter in y of hpf_y.f: rows 2 ..
//     ny-1 solve A (H_y u) = (A - B) u for the pentadiagonal implicit
//     stencil A and the explicit stencil B of init_filter.f (rows 2 and
//     ny-1 use one-sided five-point stencils, whose fifth entry is
//     eliminated with the next row), rows 1 and ny are zero; N = iord + 1;
//   * G_xz: the spectral transfer function lpfxz = 1 - (1 - Gx Gz)^N of
//     bla.f:1494-1518, Gx and Gz the Fourier transforms of the homogeneous
//     five-point filter.
// The filter coefficients are those of Stolz, Adams and Kleiser (Phys.
// Fluids 13, 2001) as init_filter.f:399-625 evaluates them: second-order
// moment conditions, the free coefficient set by minimum dispersion
// (interior) or a zero third moment (the one-sided stencils), and the
// implicit part scaled so that the transfer function is 1/2 at the cutoff.
//
// H_y^N is linear and the same for every (kx, kz) column, so the library
// applies L = I - H_y^N as a dense ny x ny matrix on the matrix cores (the
// transform GEMM kernel), built here by running the filter on the columns
// of the identity exactly as hpf_y.f runs it on the velocity.
#include <cmath>
#include <cstring>
#include <vector>

#include "sg_internal.h"

namespace sg {

namespace {

// interior (centered) stencil at offsets (-dmm, -dm, 0, dp, dpp):
// explicit b[5], implicit a[5] (init_filter.f:399-498, i3mom = 0)
void centered(double dmm, double dm, double dp, double dpp, double cutoff, double *b, double *a)
{
    // free coefficient of minimum dispersion (init_filter.f:417-440)
    const double s3 =
        2. * dpp * pow(dmm, 3) * dm * dm + 2. * dpp * dpp * dmm * dmm * dm * dm -
        2. * dpp * pow(dmm, 4) * dm - 2. * dpp * dpp * pow(dmm, 3) * dm +
        2. * dpp * pow(dmm, 3) * dp * dp + 2. * dpp * dpp * dmm * dmm * dp * dp +
        2. * dpp * pow(dmm, 4) * dp + 2. * dpp * dpp * pow(dmm, 3) * dp +
        pow(dp, 3) * dm * dm * dmm - dp * pow(dm, 4) * dmm + 2. * dp * dp * dm * dm * dmm * dmm +
        dp * pow(dm, 3) * dmm * dmm - pow(dp, 3) * dm * dm * dpp + dp * pow(dm, 4) * dpp +
        2. * dp * dp * dm * dm * dpp * dpp + dp * pow(dm, 3) * dpp * dpp +
        dm * pow(dp, 4) * dmm - pow(dm, 3) * dp * dp * dmm + dm * pow(dp, 3) * dmm * dmm -
        dm * pow(dp, 4) * dpp + pow(dm, 3) * dp * dp * dpp + dm * pow(dp, 3) * dpp * dpp;
    const double den =
        2. * dp * pow(dpp, 4) * dm + 2. * pow(dmm, 3) * dp * dp * dm -
        2. * dp * dp * dm * dm * dmm * dmm - 2. * dmm * pow(dm, 4) * dpp +
        2. * dpp * dpp * dmm * dmm * dm * dm - 2. * dmm * dmm * pow(dp, 3) * dpp +
        2. * pow(dmm, 4) * dp * dm - 2. * dmm * pow(dm, 3) * dpp * dpp +
        2. * dmm * pow(dp, 3) * dpp * dpp - 2. * pow(dmm, 3) * dm * dm * dp +
        2. * dmm * dmm * pow(dm, 3) * dpp + 2. * pow(dpp, 3) * dm * dm * dp -
        2. * pow(dpp, 3) * dp * dp * dm + 2. * dpp * dpp * dmm * dmm * dp * dp -
        2. * dmm * pow(dp, 4) * dpp + 4. * dpp * dpp * pow(dmm, 4) -
        2. * dp * dp * dm * dm * dpp * dpp + dmm * dmm * pow(dp, 4) + dm * dm * pow(dpp, 4) -
        2. * pow(dmm, 3) * pow(dm, 3) + dpp * dpp * pow(dp, 4) + 4. * pow(dpp, 4) * dmm * dmm +
        dmm * dmm * pow(dm, 4) + dpp * dpp * pow(dm, 4) + 8. * pow(dpp, 3) * pow(dmm, 3) +
        2. * pow(dpp, 3) * pow(dm, 3) + pow(dmm, 4) * dm * dm + 2. * pow(dmm, 3) * pow(dp, 3) -
        2. * pow(dpp, 3) * pow(dp, 3) + pow(dmm, 4) * dp * dp + dp * dp * pow(dpp, 4) +
        4. * dmm * dp * dp * dpp * dm * dm + 2. * dmm * dp * dp * dm * dpp * dpp -
        2. * dmm * dm * dm * dp * dpp * dpp + 2. * dmm * dmm * dp * dpp * dm * dm +
        4. * dmm * dmm * dp * dm * dpp * dpp - 2. * dmm * dmm * dm * dpp * dp * dp;
    const double bpp = -0.5 * (s3 / den);
    // zeroth to second moments (init_filter.f:449-453)
    const double bp = -0.5 * (dm * dm + 2. * dmm * bpp * dpp - dmm * dm + 2. * bpp * dpp * dpp) /
                      (dp * dp - dm * dm + dmm * dp + dmm * dm);
    const double bm = 0.5 - bp;
    const double bmm = (dpp * bpp + dp * bp - dm * bm) / dmm;
    const double b0 = 1. - (bpp + bp + bm + bmm);
    b[0] = bmm; b[1] = bm; b[2] = b0; b[3] = bp; b[4] = bpp;
    // implicit part: scale rfact so that |B/A| = 1/2 at k = pi cutoff
    // (init_filter.f:456-483)
    const double rk = M_PI * cutoff;
    const double gre = bmm * cos(-2. * rk) + bm * cos(-rk) + b0 + bp * cos(rk) + bpp * cos(2. * rk);
    const double gim = bmm * sin(-2. * rk) + bm * sin(-rk) + bp * sin(rk) + bpp * sin(2. * rk);
    const double qa = (gre - 1) * (gre - 1) + gim * gim, qb = 2 * (gre - 1);
    const double qc = 1 - 4 * gre * gre - 4 * gim * gim;
    const double rf = (-qb - sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
    a[0] = rf * bmm; a[1] = rf * bm; a[3] = rf * bp; a[4] = rf * bpp;
    a[2] = 1. - (a[4] + a[3] + a[1] + a[0]);
}

// one-sided stencil at offsets (-dm, 0, dp, d2, d3) with a zero third
// moment (init_filter.f:529-625, i3mom = 1)
void one_sided(double dm, double dp, double d2, double d3, double cutoff, double *b, double *a)
{
    const double b3 =
        0.5 * dp * dm * (-dp * dm - d2 * dp + d2 * d2 + d2 * dm) /
        (-dm * dm * dp * dp - dp * dp * d2 * d3 - dp * dp * d2 * dm + d3 * d3 * dp * dp +
         dp * d2 * d2 * d3 + dp * d2 * d2 * dm - pow(d3, 3) * dp + dp * d3 * d2 * dm +
         dm * dm * d2 * dp - dp * dm * d3 * d3 - d2 * d2 * d3 * d3 + pow(d3, 3) * d2 -
         dm * d2 * d2 * d3 + pow(d3, 3) * dm - d3 * d2 * dm * dm + d3 * d3 * dm * dm);
    const double bp = 0.5 * (dm * dm - 2. * dm * dm * b3 - 2. * d2 * b3 * d3 + d2 * dm -
                             2. * d2 * dm * b3 + 2. * b3 * d3 * d3) /
                      (-dp * dp + dm * dm + d2 * dp + d2 * dm);
    const double bm = 0.5 - bp - b3;
    const double b2 = -(b3 * d3 + bp * dp - bm * dm) / d2;
    const double b0 = 1. - (b3 + b2 + bp + bm);
    b[0] = bm; b[1] = b0; b[2] = bp; b[3] = b2; b[4] = b3;
    const double rk = M_PI * cutoff;
    const double gre = bm * cos(-rk) + b0 + bp * cos(rk) + b2 * cos(2. * rk) + b3 * cos(3. * rk);
    const double gim = bm * sin(-rk) + bp * sin(rk) + b2 * sin(2. * rk) + b3 * sin(3. * rk);
    const double qa = (gre - 1) * (gre - 1) + gim * gim, qb = 2 * (gre - 1);
    const double qc = 1 - 4 * gre * gre - 4 * gim * gim;
    const double rf = (-qb - sqrt(qb * qb - 4 * qa * qc)) / (2 * qa);
    a[0] = rf * bm; a[2] = rf * bp; a[3] = rf * b2; a[4] = rf * b3;
    a[1] = 1. - (a[4] + a[3] + a[2] + a[0]);
}

// transfer function of the homogeneous filter at k = pi t / (m - 1),
// t = 0 .. m-1 (bla.f:1498 passes m = n/2 + 1; init_filter.f:100-146)
std::vector<double> spectral(int m, const double *b, const double *a)
{
    std::vector<double> g(m);
    for (int t = 0; t < m; t++) {
        const double rk = t * M_PI / (m - 1);
        const double ra = b[0] * cos(-2. * rk) + b[1] * cos(-rk) + b[2] + b[3] * cos(rk) +
                          b[4] * cos(2. * rk);
        const double rc = a[0] * cos(-2. * rk) + a[1] * cos(-rk) + a[2] + a[3] * cos(rk) +
                          a[4] * cos(2. * rk);
        const double rb = b[0] * sin(-2. * rk) + b[1] * sin(-rk) + b[3] * sin(rk) + b[4] * sin(2. * rk);
        const double rd = a[0] * sin(-2. * rk) + a[1] * sin(-rk) + a[3] * sin(rk) + a[4] * sin(2. * rk);
        g[t] = (ra * rc + rb * rd) / (rc * rc + rd * rd);
    }
    return g;
}

}  // namespace

int les_filter(int nx, int ny, int nz, double cutoff, int iord, double *lpf, double *L)
{
    if (nx < 4 || nz < 2 || ny < 7 || ny % 2 == 0 || iord < 0 || iord > 5 || !(cutoff > 0.) ||
        cutoff > 1.)
        return SIMSON_E_ARG;
    const int N = iord + 1, nxh = nx / 2;
    // homogeneous directions: uniform stencil (2, 1, 1, 2)
    double bh[5], ah[5];
    centered(2., 1., 1., 2., cutoff, bh, ah);
    const std::vector<double> gx = spectral(nxh + 1, bh, ah), gz = spectral(nz / 2 + 1, bh, ah);
    for (int z = 0; z < nz; z++) {
        const int kz = z <= nz / 2 ? z : nz - z;
        for (int x = 0; x < nxh; x++)
            lpf[(size_t)z * nxh + x] = 1. - pow(1. - gx[x] * gz[kz], N);
    }
    // wall-normal direction on the Gauss-Lobatto points (y = +1 first)
    std::vector<double> eta(ny);
    for (int j = 0; j < ny; j++)
        eta[j] = cos(M_PI * j / (ny - 1));
    std::vector<double> B(5 * (size_t)ny, 0.), A(5 * (size_t)ny, 0.);   // [row][5]
    B[0] = A[0] = 1.;
    B[5 * (ny - 1) + 4] = A[5 * (ny - 1) + 4] = 1.;
    one_sided(eta[0] - eta[1], eta[1] - eta[2], eta[1] - eta[3], eta[1] - eta[4], cutoff, &B[5],
              &A[5]);
    for (int k = 2; k <= ny - 3; k++)
        centered(eta[k - 2] - eta[k], eta[k - 1] - eta[k], eta[k] - eta[k + 1],
                 eta[k] - eta[k + 2], cutoff, &B[5 * k], &A[5 * k]);
    {
        // mirror of the one-sided stencil (coeff_filter_r3_l1)
        const int k = ny - 2;
        double bt[5], at[5];
        one_sided(eta[k] - eta[k + 1], eta[k - 1] - eta[k], eta[k - 2] - eta[k],
                  eta[k - 3] - eta[k], cutoff, bt, at);
        for (int l = 0; l < 5; l++) {
            B[5 * k + l] = bt[4 - l];
            A[5 * k + l] = at[4 - l];
        }
    }
    // pentadiagonal band of A (entries j-2 .. j+2 of row j) with the fifth
    // entries of rows 2 and ny-1 eliminated by rows 3 and ny-2
    // (init_filter.f:305-343)
    const double f3 = -A[5 * 1 + 4] / A[5 * 2 + 4];
    const double fm3 = -A[5 * (ny - 2) + 0] / A[5 * (ny - 3) + 0];
    std::vector<double> d2(ny, 0.), d1(ny, 0.), d0(ny, 0.), u1(ny, 0.), u2(ny, 0.);
    d0[0] = 1.;
    d0[ny - 1] = 1.;
    d1[1] = A[5 + 0] + A[10 + 0] * f3;
    d0[1] = A[5 + 1] + A[10 + 1] * f3;
    u1[1] = A[5 + 2] + A[10 + 2] * f3;
    u2[1] = A[5 + 3] + A[10 + 3] * f3;
    {
        const int k = ny - 2;
        d2[k] = A[5 * k + 1] + A[5 * (k - 1) + 1] * fm3;
        d1[k] = A[5 * k + 2] + A[5 * (k - 1) + 2] * fm3;
        d0[k] = A[5 * k + 3] + A[5 * (k - 1) + 3] * fm3;
        u1[k] = A[5 * k + 4] + A[5 * (k - 1) + 4] * fm3;
    }
    for (int k = 2; k <= ny - 3; k++) {
        d2[k] = A[5 * k];
        d1[k] = A[5 * k + 1];
        d0[k] = A[5 * k + 2];
        u1[k] = A[5 * k + 3];
        u2[k] = A[5 * k + 4];
    }
    // LU factors without pivoting, unit upper diagonal (bandsolver.f:70-140)
    const int n = ny;
    d2[0] = d2[1] = d1[0] = 0.;
    u1[n - 1] = u2[n - 2] = u2[n - 1] = 0.;
    for (int i = 0; i < n; i++) {
        if (i >= 2)
            d1[i] -= d2[i] * u1[i - 2];
        if (i >= 1)
            d0[i] -= (i >= 2 ? d2[i] * u2[i - 2] : 0.) + d1[i] * u1[i - 1];
        if (d0[i] == 0.)
            return SIMSON_E_ARG;
        const double h = 1. / d0[i];
        u1[i] = (u1[i] - (i >= 1 ? d1[i] * u2[i - 1] : 0.)) * h;
        u2[i] *= h;
    }
    // H_y^N on the columns of the identity, then L = I - H_y^N
    std::vector<double> X((size_t)n * n, 0.), rs(n);
    for (int j = 0; j < n; j++)
        X[(size_t)j * n + j] = 1.;                 // X[col][row]
    for (int rep = 0; rep < N; rep++)
        for (int c = 0; c < n; c++) {
            double *d = &X[(size_t)c * n];
            std::fill(rs.begin(), rs.end(), 0.);
            for (int l = 0; l < 5; l++) {
                rs[1] += (A[5 + l] - B[5 + l]) * d[l];
                rs[n - 2] += (A[5 * (n - 2) + l] - B[5 * (n - 2) + l]) * d[n - 5 + l];
            }
            for (int l = 0; l < 5; l++)
                for (int j = 2; j <= n - 3; j++)
                    rs[j] += (A[5 * j + l] - B[5 * j + l]) * d[j - 2 + l];
            rs[1] += rs[2] * f3;
            rs[n - 2] += rs[n - 3] * fm3;
            // forward and back substitution (bandsolver.f:166-206)
            rs[0] *= 1. / d0[0];
            rs[1] = (rs[1] - d1[1] * rs[0]) * (1. / d0[1]);
            for (int j = 2; j < n; j++)
                rs[j] = (rs[j] - d2[j] * rs[j - 2] - d1[j] * rs[j - 1]) * (1. / d0[j]);
            d[n - 1] = rs[n - 1];
            d[n - 2] = rs[n - 2] - u1[n - 2] * d[n - 1];
            for (int j = n - 3; j >= 0; j--)
                d[j] = rs[j] - u1[j] * d[j + 1] - u2[j] * d[j + 2];
        }
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            L[(size_t)r * n + c] = (r == c ? 1. : 0.) - X[(size_t)c * n + r];
    return 0;
}

}  // namespace sg
