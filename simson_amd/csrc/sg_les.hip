// ADM-RT large-eddy model of bla (iles = 1, ihighorder = 2): the filter
// operators, built on the host once per configuration.
//
// The relaxation term of nonlinbl.f:920-966 subtracts chi nxp nzp g_i from
// H_i, where g = u - G_xz (u - H_y^N u) is the high-pass filtered velocity
// computed at the end of linearbl.f:1871-1962 (and of prhs.f:95-195):
//   * H_y: the implicit compact high-pass filter in y of hpf_y.f,
//     H_y = A^-1 (A - B) with the implicit (A) and explicit (B) five-point
//     row stencils of the low-pass filter on the Gauss-Lobatto points; the
//     wall rows of H_y are zero; N = iord + 1;
//   * G_xz: the spectral transfer function lpfxz = 1 - (1 - Gx Gz)^N of
//     bla.f:1494-1518, Gx and Gz the transfer functions of the uniform
//     stencil.
//
// The stencils are the filters of Stolz, Adams and Kleiser (Phys. Fluids 13,
// 2001), defined by conditions on the weights b_j at grid offsets j and
// signed distances p_j from the filtered point:
//   sum b_j = 1, sum b_j p_j = 0, sum b_j p_j^2 = 0, sum (-1)^j b_j = 0
// (second-order moments, and B(pi) = 0 for B(k) = sum b_j e^{ijk}), and for
// the fifth weight either minimum dispersion -- the least integral over
// [0, pi] of (Im B)^2, proportional to sum_m (b_m - b_-m)^2 -- on the
// centred interior rows, or a zero third moment on the rows next to the
// walls, whose stencils reach one point to one side and three to the other.
// The implicit weights are a_j = r b_j (j != 0), a_0 = 1 - sum_{j != 0} a_j,
// so that A = 1 - r + r B, with r the smaller root of the half-transfer
// condition |B(kc)|^2 = |A(kc)|^2 / 4 at kc = pi cutoff.  Here the centred
// rows solve the Lagrange (KKT) system of the constrained minimisation and
// the wall rows the 5 x 5 moment system, by Gaussian elimination with
// partial pivoting on distances scaled by the stencil width.
//
// H_y^N is linear and the same for every (kx, kz) column, so the library
// applies L = I - H_y^N as a dense ny x ny matrix on the matrix cores (the
// transform GEMM kernel); H_y is formed here by a dense pivoted solve.
#include <cmath>
#include <complex>
#include <vector>

#include "sg_internal.h"

namespace sg {

namespace {

// solve M x = r (n x n, row-major, overwritten) with partial pivoting;
// r receives x; false if singular
bool gauss_solve(int n, int nrhs, std::vector<double> &M, std::vector<double> &r)
{
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int i = c + 1; i < n; i++)
            if (std::fabs(M[(size_t)i * n + c]) > std::fabs(M[(size_t)piv * n + c]))
                piv = i;
        if (M[(size_t)piv * n + c] == 0.)
            return false;
        if (piv != c) {
            for (int k = 0; k < n; k++)
                std::swap(M[(size_t)c * n + k], M[(size_t)piv * n + k]);
            for (int k = 0; k < nrhs; k++)
                std::swap(r[(size_t)c * nrhs + k], r[(size_t)piv * nrhs + k]);
        }
        const double inv = 1. / M[(size_t)c * n + c];
        for (int i = c + 1; i < n; i++) {
            const double f = M[(size_t)i * n + c] * inv;
            if (f == 0.)
                continue;
            for (int k = c; k < n; k++)
                M[(size_t)i * n + k] -= f * M[(size_t)c * n + k];
            for (int k = 0; k < nrhs; k++)
                r[(size_t)i * nrhs + k] -= f * r[(size_t)c * nrhs + k];
        }
    }
    for (int c = n - 1; c >= 0; c--) {
        const double inv = 1. / M[(size_t)c * n + c];
        for (int k = 0; k < nrhs; k++) {
            double v = r[(size_t)c * nrhs + k];
            for (int i = c + 1; i < n; i++)
                v -= M[(size_t)c * n + i] * r[(size_t)i * nrhs + k];
            r[(size_t)c * nrhs + k] = v * inv;
        }
    }
    return true;
}

// Explicit and implicit weights b[5], a[5] of one row stencil: points at
// grid offsets off[q] and signed distances p[q]; a zero third moment
// (wall rows) or minimum dispersion (centred rows, offsets -2..2).
bool row_stencil(const int *off, const double *p, bool third, double cutoff, double *b,
                 double *a)
{
    double h = 0.;
    for (int q = 0; q < 5; q++)
        h = std::fmax(h, std::fabs(p[q]));
    double s[5];
    for (int q = 0; q < 5; q++)
        s[q] = p[q] / h;
    // constraint rows: moments 0..2, the alternating sum, (moment 3)
    double Cm[5][5];
    for (int q = 0; q < 5; q++) {
        Cm[0][q] = 1.;
        Cm[1][q] = s[q];
        Cm[2][q] = s[q] * s[q];
        Cm[3][q] = (off[q] & 1) ? -1. : 1.;
        Cm[4][q] = s[q] * s[q] * s[q];
    }
    if (third) {
        std::vector<double> M(25), r(5, 0.);
        for (int i = 0; i < 5; i++)
            for (int q = 0; q < 5; q++)
                M[i * 5 + q] = Cm[i][q];
        r[0] = 1.;
        if (!gauss_solve(5, 1, M, r))
            return false;
        for (int q = 0; q < 5; q++)
            b[q] = r[q];
    } else {
        // minimise |D b|^2 (D b = b_m - b_-m, m = 1, 2) subject to the four
        // constraints: [2 D^T D, C^T; C, 0] [b; mu] = [0; e_0]
        double D[2][5] = {};
        for (int q = 0; q < 5; q++)
            for (int m = 1; m <= 2; m++) {
                if (off[q] == m)
                    D[m - 1][q] = 1.;
                if (off[q] == -m)
                    D[m - 1][q] = -1.;
            }
        std::vector<double> M(81, 0.), r(9, 0.);
        for (int i = 0; i < 5; i++)
            for (int k = 0; k < 5; k++)
                M[i * 9 + k] = 2. * (D[0][i] * D[0][k] + D[1][i] * D[1][k]);
        for (int c = 0; c < 4; c++)
            for (int q = 0; q < 5; q++) {
                M[q * 9 + 5 + c] = Cm[c][q];
                M[(5 + c) * 9 + q] = Cm[c][q];
            }
        r[5] = 1.;
        if (!gauss_solve(9, 1, M, r))
            return false;
        for (int q = 0; q < 5; q++)
            b[q] = r[q];
    }
    // half-transfer condition at the cutoff for A = 1 - r + r B
    std::complex<double> B(0., 0.);
    const double kc = M_PI * cutoff;
    for (int q = 0; q < 5; q++)
        B += b[q] * std::exp(std::complex<double>(0., kc * off[q]));
    const double qa = std::norm(B - 1.), qb = 2. * (B.real() - 1.),
                 qc = 1. - 4. * std::norm(B);
    const double disc = qb * qb - 4. * qa * qc;
    if (!(qa > 0.) || !(disc >= 0.))
        return false;
    const double rf = (-qb - std::sqrt(disc)) / (2. * qa);
    double rest = 0.;
    for (int q = 0; q < 5; q++) {
        a[q] = off[q] == 0 ? 0. : rf * b[q];
        rest += a[q];
    }
    for (int q = 0; q < 5; q++)
        if (off[q] == 0)
            a[q] = 1. - rest;
    return true;
}

// transfer function Re(B/A) of the uniform stencil at k = pi t / (m - 1),
// t = 0 .. m-1 (bla.f:1498 samples it so)
bool uniform_transfer(int m, double cutoff, std::vector<double> &g)
{
    static const int off[5] = {-2, -1, 0, 1, 2};
    const double p[5] = {-2., -1., 0., 1., 2.};
    double b[5], a[5];
    if (!row_stencil(off, p, false, cutoff, b, a))
        return false;
    g.assign(m, 0.);
    for (int t = 0; t < m; t++) {
        const double k = t * M_PI / (m - 1);
        std::complex<double> Bk(0., 0.), Ak(0., 0.);
        for (int q = 0; q < 5; q++) {
            const std::complex<double> e = std::exp(std::complex<double>(0., k * off[q]));
            Bk += b[q] * e;
            Ak += a[q] * e;
        }
        g[t] = (Bk / Ak).real();
    }
    return true;
}

}  // namespace

int les_filter(int nx, int ny, int nz, double cutoff, int iord, double *lpf, double *L)
{
    // cutoff = 1 puts the half-transfer point on the grid Nyquist
    // wavenumber, where B = 0: the implicit factor is then r = 1, A = B and
    // the transfer function B/A is undefined at k = pi
    if (nx < 4 || nz < 2 || ny < 7 || ny % 2 == 0 || iord < 0 || iord > 5 || !(cutoff > 0.) ||
        !(cutoff < 1.))
        return SIMSON_E_ARG;
    const int N = iord + 1, nxh = nx / 2, n = ny;
    std::vector<double> gx, gz;
    if (!uniform_transfer(nxh + 1, cutoff, gx) || !uniform_transfer(nz / 2 + 1, cutoff, gz))
        return SIMSON_E_ARG;
    for (int z = 0; z < nz; z++) {
        const int kz = z <= nz / 2 ? z : nz - z;
        for (int x = 0; x < nxh; x++)
            lpf[(size_t)z * nxh + x] = 1. - std::pow(1. - gx[x] * gz[kz], N);
    }
    // A and A - B on the Gauss-Lobatto points (y = +1 first); wall rows of
    // A are those of the identity and of A - B zero
    std::vector<double> eta(n);
    for (int j = 0; j < n; j++)
        eta[j] = std::cos(M_PI * j / (n - 1));
    std::vector<double> A((size_t)n * n, 0.), R((size_t)n * n, 0.);
    A[0] = 1.;
    A[(size_t)n * n - 1] = 1.;
    for (int k = 1; k < n - 1; k++) {
        const int c0 = k == 1 ? 0 : (k == n - 2 ? n - 5 : k - 2);
        int off[5];
        double p[5], b[5], a[5];
        for (int q = 0; q < 5; q++) {
            off[q] = c0 + q - k;
            p[q] = eta[k] - eta[c0 + q];
        }
        if (!row_stencil(off, p, k == 1 || k == n - 2, cutoff, b, a))
            return SIMSON_E_ARG;
        for (int q = 0; q < 5; q++) {
            A[(size_t)k * n + c0 + q] = a[q];
            R[(size_t)k * n + c0 + q] = a[q] - b[q];
        }
    }
    // H_y = A^-1 (A - B), then H_y^N and L = I - H_y^N
    if (!gauss_solve(n, n, A, R))
        return SIMSON_E_ARG;
    std::vector<double> P(R), T((size_t)n * n);
    for (int rep = 1; rep < N; rep++) {
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double v = 0.;
                for (int k = 0; k < n; k++)
                    v += R[(size_t)i * n + k] * P[(size_t)k * n + j];
                T[(size_t)i * n + j] = v;
            }
        P.swap(T);
    }
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            L[(size_t)r * n + c] = (r == c ? 1. : 0.) - P[(size_t)r * n + c];
    return 0;
}

int les_filter_secondary(int nx, int nz, double cutoff, int iord, double *lpf2)
{
    // the homogeneous filters sampled on nx/4 + 1 and nz/4 + 1 points
    // (bla.f:1520-1543): the transfer function of a grid with half the
    // wavenumbers, zero above kx = nx/4 and |kz| = nz/4
    if (nx < 8 || nz < 4 || iord < 0 || iord > 5 || !(cutoff > 0.) || !(cutoff < 1.))
        return SIMSON_E_ARG;
    const int N = iord + 1, nxh = nx / 2, mx = nx / 4, mz = nz / 4;
    std::vector<double> gx, gz;
    if (!uniform_transfer(mx + 1, cutoff, gx) || !uniform_transfer(mz + 1, cutoff, gz))
        return SIMSON_E_ARG;
    for (int z = 0; z < nz; z++) {
        const int kz = z <= nz / 2 ? z : nz - z;
        for (int x = 0; x < nxh; x++) {
            const double f = (x <= mx && kz <= mz) ? gx[x] * gz[kz] : 0.;
            lpf2[(size_t)z * nxh + x] = 1. - std::pow(1. - f, N);
        }
    }
    return 0;
}

}  // namespace sg
