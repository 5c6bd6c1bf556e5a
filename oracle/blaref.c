/*
 * blaref.c -- CPU restatement (FP64) of the Simson `bla` time step for the
 * temporal plane channel.  TEST INFRASTRUCTURE ONLY (see blaref.h): it is
 * the oracle the HIP step is checked against and is never on the product
 * path.
 *
 * Written from the mathematics of the reference routines; each function
 * names the reference lines whose behaviour it restates.  Transforms are
 * plain DFTs (mixed-radix recursion) and direct Chebyshev sums, not FFTPACK,
 * so agreement with the reference is to rounding, not bitwise.
 */
#include "blaref.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

#define PI 3.14159265358979323846

/* ------------------------------------------------------------------ */
/* DFT: y_k = sum_j x_j exp(sign*2*pi*i*j*k/n), n = 2^a 3^b 5^c ...   */
/* ------------------------------------------------------------------ */

static int smallest_factor(int n)
{
    for (int p = 2; p * p <= n; p++)
        if (n % p == 0)
            return p;
    return n;
}

static cplx *roots_table(int n)
{
    cplx *w = malloc(sizeof(cplx) * (size_t)n);
    for (int t = 0; t < n; t++) {
        /* reduce the angle to keep the table accurate near pi/2, pi ... */
        long num = 4L * t, q = num / n, r = num % n;  /* 2 pi t/n = q*pi/2 + 2 pi r/(4n) */
        double a = 2.0 * PI * (double)r / (4.0 * (double)n);
        double c = cos(a), s = sin(a);
        switch (q & 3) {
        case 0: w[t] = c + I * s; break;
        case 1: w[t] = -s + I * c; break;
        case 2: w[t] = -c - I * s; break;
        default: w[t] = s - I * c; break;
        }
    }
    return w;
}

/* recursive mixed-radix: x strided by s, y contiguous (length n) */
static void dft_rec(int n, const cplx *x, int s, cplx *y, int sign, int N,
                    const cplx *W)
{
    if (n == 1) {
        y[0] = x[0];
        return;
    }
    int p = smallest_factor(n), m = n / p, step = N / n;
    for (int r = 0; r < p; r++)
        dft_rec(m, x + (size_t)r * s, s * p, y + (size_t)r * m, sign, N, W);
    cplx in[64], out[64];
    if (p > 64)
        abort();
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++)
            in[r] = y[(size_t)r * m + k];
        for (int q = 0; q < p; q++) {
            cplx acc = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)(k + q * m) * r) % n;
                cplx w = W[(e * step) % N];
                acc += in[r] * (sign > 0 ? w : conj(w));
            }
            out[q] = acc;
        }
        for (int q = 0; q < p; q++)
            y[k + (size_t)q * m] = out[q];
    }
}

typedef struct {
    int n;
    cplx *w;
    cplx *buf;
} dftplan;

static void plan_init(dftplan *pl, int n)
{
    pl->n = n;
    pl->w = roots_table(n);
    pl->buf = NULL;
}

static void plan_free(dftplan *pl)
{
    free(pl->w);
}

/* in place on contiguous x with scratch y (length n) */
static void dft_inplace(const dftplan *pl, cplx *x, cplx *y, int sign)
{
    dft_rec(pl->n, x, 1, y, sign, pl->n, pl->w);
    memcpy(x, y, sizeof(cplx) * (size_t)pl->n);
}

void orc_cdft(int n, double *re, double *im, int sign)
{
    dftplan pl;
    plan_init(&pl, n);
    cplx *x = malloc(sizeof(cplx) * (size_t)n), *y = malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; j++)
        x[j] = re[j] + I * im[j];
    dft_inplace(&pl, x, y, sign);
    for (int j = 0; j < n; j++) {
        re[j] = creal(x[j]);
        im[j] = cimag(x[j]);
    }
    free(x);
    free(y);
    plan_free(&pl);
}

/* vrfftb (vecfft.txt:214-280): Hermitian half spectrum -> real signal.
 * Only the real parts of X_0 and X_{n/2} enter. */
static void rbwd(const dftplan *pl, const cplx *X, double *x, cplx *t1, cplx *t2)
{
    int n = pl->n, h = n / 2;
    t1[0] = creal(X[0]);
    t1[h] = creal(X[h]);
    for (int k = 1; k < h; k++) {
        t1[k] = X[k];
        t1[n - k] = conj(X[k]);
    }
    dft_inplace(pl, t1, t2, +1);
    for (int m = 0; m < n; m++)
        x[m] = creal(t1[m]);
}

/* vrfftf (vecfft.txt:134-170): real signal -> X_0..X_{n/2} */
static void rfwd(const dftplan *pl, const double *x, cplx *X, cplx *t1, cplx *t2)
{
    int n = pl->n, h = n / 2;
    for (int m = 0; m < n; m++)
        t1[m] = x[m];
    dft_inplace(pl, t1, t2, -1);
    for (int k = 0; k <= h; k++)
        X[k] = t1[k];
    X[0] = creal(X[0]);
    X[h] = creal(X[h]);
}

void orc_rdft_bwd(int n, const double *xr, const double *xi, double *x)
{
    dftplan pl;
    plan_init(&pl, n);
    cplx *X = malloc(sizeof(cplx) * (size_t)(n / 2 + 1));
    cplx *t1 = malloc(sizeof(cplx) * (size_t)n), *t2 = malloc(sizeof(cplx) * (size_t)n);
    for (int k = 0; k <= n / 2; k++)
        X[k] = xr[k] + I * xi[k];
    rbwd(&pl, X, x, t1, t2);
    free(X);
    free(t1);
    free(t2);
    plan_free(&pl);
}

void orc_rdft_fwd(int n, const double *x, double *xr, double *xi)
{
    dftplan pl;
    plan_init(&pl, n);
    cplx *X = malloc(sizeof(cplx) * (size_t)(n / 2 + 1));
    cplx *t1 = malloc(sizeof(cplx) * (size_t)n), *t2 = malloc(sizeof(cplx) * (size_t)n);
    rfwd(&pl, x, X, t1, t2);
    for (int k = 0; k <= n / 2; k++) {
        xr[k] = creal(X[k]);
        xi[k] = cimag(X[k]);
    }
    free(X);
    free(t1);
    free(t2);
    plan_free(&pl);
}

/* ------------------------------------------------------------------ */
/* Chebyshev operations                                                */
/* ------------------------------------------------------------------ */

/* cos(pi*m/(n-1)) for m in [0, 2(n-1)) */
static double *cheb_cos_table(int n)
{
    int M = 2 * (n - 1);
    double *c = malloc(sizeof(double) * (size_t)M);
    cplx *w = roots_table(M);      /* exp(2 pi i m / M) = exp(i pi m/(n-1)) */
    for (int m = 0; m < M; m++)
        c[m] = creal(w[m]);
    free(w);
    return c;
}

/* vchbf (vecfft.txt:554-561): X_i = sum_k c_k c_i x_k cos(k i pi/(n-1)),
 * c_0 = c_{n-1} = 1/2; multiplied by `scale`. */
static void chbf_tab(int n, const double *ct, const double *x, double *out, double scale)
{
    int M = 2 * (n - 1);
    for (int i = 0; i < n; i++) {
        double acc = 0.5 * (x[0] + ((i & 1) ? -x[n - 1] : x[n - 1]));
        for (int k = 1; k < n - 1; k++)
            acc += x[k] * ct[(long)k * i % M];
        if (i == 0 || i == n - 1)
            acc *= 0.5;
        out[i] = acc * scale;
    }
}

/* vchbb (vecfft.txt:626-631): x_i = sum_k X_k cos(k i pi/(n-1)) */
static void chbb_tab(int n, const double *ct, const double *X, double *out)
{
    int M = 2 * (n - 1);
    for (int i = 0; i < n; i++) {
        double acc = 0.0;
        for (int k = 0; k < n; k++)
            acc += X[k] * ct[(long)k * i % M];
        out[i] = acc;
    }
}

void orc_chbf(int n, const double *x, double *out)
{
    double *ct = cheb_cos_table(n);
    chbf_tab(n, ct, x, out, 1.0);
    free(ct);
}

void orc_chbb(int n, const double *x, double *out)
{
    double *ct = cheb_cos_table(n);
    chbb_tab(n, ct, x, out);
    free(ct);
}

/* dcheb.f:24-36: coefficients of d/dy.  b_{n-1} = 0, b_{n-2} = 2(n-1)a_{n-1},
 * b_k = b_{k+2} + 2(k+1) a_{k+1}, then b_0 halved (c_0 = 2). */
static void dcheb(int n, const double *a, double *b)
{
    double bk2 = 0.0, bk1 = 0.0;      /* b_{k+2}, b_{k+1} */
    b[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; k--) {
        double bk = bk2 + 2.0 * (double)(k + 1) * a[k + 1];
        b[k] = bk;
        bk2 = bk1;
        bk1 = bk;
    }
    b[0] *= 0.5;
}

void orc_dcheb(int n, const double *a, double *b)
{
    dcheb(n, a, b);
}

/* icheb.f:24-36: integral with b_0 = intcon. */
void orc_icheb(int n, const double *a, double *b, double intcon)
{
    for (int i = 2; i < n - 1; i++)
        b[i] = (a[i - 1] - a[i + 1]) * (0.5 / (double)i);
    b[n - 1] = a[n - 2] * (0.5 / (double)(n - 1));
    b[1] = a[0] - 0.5 * a[2];
    b[0] = intcon;
}

/* Chebyshev tau system for u'' - lam u = f (setmacrhs.f:23-74 builds it,
 * trid.f:120-199 solves it).  Unknowns a_0..a_{n-1}.  Rows:
 *   sum_{k even} a_k = bc_e,  sum_{k odd} a_k = bc_o,
 *   for k = 2..n-1 (g = f + lam*a, the u'' coefficients, truncated to
 *   degree n-3):
 *     a_k = c_{k-2} g_{k-2}/(4k(k-1)) - [k<=n-3] g_k/(2(k^2-1))
 *           + [k<=n-5] g_{k+2}/(4k(k+1)),   c_0 = 2, c_j = 1.
 * Each parity class is a tridiagonal chain plus one full row; it is
 * reduced bottom-up to a_k = rho_k - gam_k a_{k-2}, closed with the BC row
 * and substituted forward. */
static void tau_solve(int n, double lam, const double *f, double bce, double bco,
                      double *a, double *gam, double *rho)
{
    for (int par = 0; par < 2; par++) {
        int k0 = par;
        int klast = (n - 1 - par) % 2 == 0 ? n - 1 : n - 2;
        /* bottom-up elimination over k = klast, klast-2, ..., k0+2 */
        double g_up = 0.0, r_up = 0.0;    /* gam, rho of row k+2 */
        for (int k = klast; k >= k0 + 2; k -= 2) {
            double kk = (double)k;
            double ckm2 = (k - 2 == 0) ? 2.0 : 1.0;
            double L = -ckm2 * lam / (4.0 * kk * (kk - 1.0));
            double D = 1.0 + (k <= n - 3 ? lam / (2.0 * (kk * kk - 1.0)) : 0.0);
            double U = (k <= n - 5) ? -lam / (4.0 * kk * (kk + 1.0)) : 0.0;
            double R = ckm2 * f[k - 2] / (4.0 * kk * (kk - 1.0));
            if (k <= n - 3)
                R -= f[k] / (2.0 * (kk * kk - 1.0));
            if (k <= n - 5)
                R += f[k + 2] / (4.0 * kk * (kk + 1.0));
            double piv = D - U * g_up;
            double gk = L / piv;
            double rk = (R - U * r_up) / piv;
            gam[k] = gk;
            rho[k] = rk;
            g_up = gk;
            r_up = rk;
        }
        /* a_k = P_k + Q_k a_{k0}; accumulate the BC row */
        double P = 0.0, Q = 1.0, sP = 0.0, sQ = 1.0;
        for (int k = k0 + 2; k <= klast; k += 2) {
            P = rho[k] - gam[k] * P;
            Q = -gam[k] * Q;
            sP += P;
            sQ += Q;
        }
        double bc = par == 0 ? bce : bco;
        double a0 = (bc - sP) / sQ;
        a[k0] = a0;
        for (int k = k0 + 2; k <= klast; k += 2)
            a[k] = rho[k] - gam[k] * a[k - 2];
    }
}

void orc_tau_solve(int n, double lam, const double *f, double bc_e, double bc_o,
                   double *u)
{
    double *g = malloc(sizeof(double) * (size_t)n), *r = malloc(sizeof(double) * (size_t)n);
    tau_solve(n, lam, f, bc_e, bc_o, u, g, r);
    free(g);
    free(r);
}

/* ------------------------------------------------------------------ */
/* State                                                               */
/* ------------------------------------------------------------------ */

struct orc {
    int nx, ny, nz, nxh, nxp, nzp;
    double re, xl, zl, rot;
    int cflux;
    double mflux, px;
    double u0low, u0upp, w0low, w0upp;
    double *alfa, *beta, *eta, *deta, *ct;
    dftplan px_plan, pz_plan;
    struct les_s *les;                 /* ADM-RT filters (orc_set_les) or NULL */
    double *gur, *gui;                 /* [3][nz][ny][nxh] high-pass velocities */
    double les_chi;                    /* chi/dstar of the relaxation term */
    double *fr, *fi;                   /* [3][nz][ny][nxh] external force or NULL */
    double *flam, *fbase;              /* fringe [nxp] and base flow [3][ny][nxp] or NULL */
    double *wvr, *wvi;                 /* lower-wall v [nz][nxh] (spectral) or NULL */
    int ibc;                           /* 0 (no-slip), 100 or 101 (free stream above) */
    double *bjr, *bji;                 /* [3][nxh] U_i (ibc 100) or dU_i/dy (101) at y = 1 */
    double du0upp;                     /* its mean (kx = 0) streamwise part */
    double *mf;                        /* [3][ny] mean force or NULL */
};

struct les_s;
static void les_free(struct les_s *L);
static void les_gu(orc_t *o, const double *ur, const double *ui);

orc_t *orc_create(int nx, int ny, int nz, double re, double xl, double zl,
                  double rot, int cflux, double retau)
{
    orc_t *o = calloc(1, sizeof(orc_t));
    o->nx = nx;
    o->ny = ny;
    o->nz = nz;
    o->nxh = nx / 2;
    o->nxp = nx + nx / 2;           /* par.f:86, nfxd = 1 */
    o->nzp = nz + nz / 2;           /* nfzd = 1 */
    o->re = re;
    o->xl = xl;
    o->zl = zl;
    o->rot = rot;
    o->cflux = cflux;
    if (cflux) {                    /* bla.f:1629-1630 */
        o->mflux = 4. / 3.;
        o->px = -2. / re;
    } else {                        /* bla.f:1645 */
        o->mflux = 0.;
        o->px = -(retau / re) * (retau / re);
    }
    /* preprbl.f:34-55 */
    o->alfa = malloc(sizeof(double) * (size_t)o->nxh);
    for (int x = 0; x < o->nxh; x++)
        o->alfa[x] = 2. * PI / xl * (double)x;
    o->beta = malloc(sizeof(double) * (size_t)nz);
    o->beta[0] = 0.;
    for (int z = 1; z <= nz / 2; z++) {
        o->beta[z] = 2. * PI / zl * (double)z;
        o->beta[nz - z] = -2. * PI / zl * (double)z;
    }
    o->eta = malloc(sizeof(double) * (size_t)ny);
    o->ct = cheb_cos_table(ny);
    for (int y = 0; y < ny; y++)
        o->eta[y] = o->ct[y];
    o->deta = malloc(sizeof(double) * (size_t)ny);
    o->deta[0] = 1. - o->eta[1];
    o->deta[ny - 1] = o->deta[0];
    for (int y = 1; y < ny - 1; y++)
        o->deta[y] = (o->eta[y - 1] - o->eta[y + 1]) * .5;
    plan_init(&o->px_plan, o->nxp);
    plan_init(&o->pz_plan, o->nzp);
    return o;
}

void orc_destroy(orc_t *o)
{
    if (!o)
        return;
    free(o->alfa);
    free(o->beta);
    free(o->eta);
    free(o->deta);
    free(o->ct);
    plan_free(&o->px_plan);
    plan_free(&o->pz_plan);
    les_free(o->les);
    free(o->fr);
    free(o->fi);
    free(o->flam);
    free(o->fbase);
    free(o->wvr);
    free(o->wvi);
    free(o->bjr);
    free(o->bji);
    free(o->mf);
    free(o->gur);
    free(o->gui);
    free(o);
}

double orc_get_px(const orc_t *o) { return o->px; }
void orc_set_px(orc_t *o, double px) { o->px = px; }

#define IDX(o, s, z, y, x) ((((size_t)(s) * (o)->nz + (z)) * (o)->ny + (y)) * (o)->nxh + (x))

/* ------------------------------------------------------------------ */
/* Step 2: nonlinear terms per y plane (nonlinbl.f)                    */
/* ------------------------------------------------------------------ */

/* storage z index -> row of the padded nzp grid (getxz.f:41-61) */
static int zrow(const orc_t *o, int z)
{
    return z < o->nz / 2 ? z : z + o->nzp - o->nz;
}

static double nonlin_plane(orc_t *o, double *ur, double *ui, int y, int do_cfl)
{
    int nxh = o->nxh, nz = o->nz, nxp = o->nxp, nzp = o->nzp, hx = nxp / 2 + 1;
    size_t box = (size_t)nzp * hx;
    cplx *B = calloc(6 * box, sizeof(cplx));          /* [f][zp][kx] */
    double *R = malloc(sizeof(double) * 6 * (size_t)nzp * nxp);   /* [f][zp][xp] */
    int nmax = nxp > nzp ? nxp : nzp;
    cplx *t1 = malloc(sizeof(cplx) * (size_t)nmax), *t2 = malloc(sizeof(cplx) * (size_t)nmax);
    cplx *col = malloc(sizeof(cplx) * (size_t)nzp);
    /* fields in padded box: 0 u, 1 v, 2 w, 3 om_x, 4 om_y, 5 om_z
     * (getxz with ipad=1, nonlinbl.f:229-247); the storage oddball
     * z = nz/2 lands in the zeroed band and is dropped. */
    static const int slot_of[6] = {0, 1, 2, 3, -1, 4};
    for (int f = 0; f < 6; f++) {
        if (slot_of[f] < 0)
            continue;
        for (int z = 0; z < nz; z++) {
            if (z == nz / 2)
                continue;
            int zp = zrow(o, z);
            for (int x = 0; x < nxh; x++) {
                size_t s = IDX(o, slot_of[f], z, y, x);
                B[f * box + (size_t)zp * hx + x] = ur[s] + I * ui[s];
            }
        }
    }
    /* omega_y = i beta u - i alpha w  (nonlinbl.f:308-355) */
    for (int z = 0; z < nz; z++) {
        if (z == nz / 2)
            continue;
        int zp = zrow(o, z);
        for (int x = 0; x < nxh; x++) {
            cplx uu = B[0 * box + (size_t)zp * hx + x], ww = B[2 * box + (size_t)zp * hx + x];
            B[4 * box + (size_t)zp * hx + x] = I * o->beta[z] * uu - I * o->alfa[x] * ww;
        }
    }
    /* Fourier synthesis (fft2db.f): z complex over the nx/2 columns, then
     * x half-complex -> real for every row. */
    for (int f = 0; f < 6; f++) {
        for (int x = 0; x < nxh; x++) {
            for (int zp = 0; zp < nzp; zp++)
                col[zp] = B[f * box + (size_t)zp * hx + x];
            dft_inplace(&o->pz_plan, col, t1, +1);
            for (int zp = 0; zp < nzp; zp++)
                B[f * box + (size_t)zp * hx + x] = col[zp];
        }
        for (int zp = 0; zp < nzp; zp++)
            rbwd(&o->px_plan, B + f * box + (size_t)zp * hx,
                 R + ((size_t)f * nzp + zp) * nxp, t1, t2);
    }
    /* CFL (boxcfl.f:29-76) */
    double cflp = 0.;
    if (do_cfl) {
        double dyinv = 1. / o->deta[y];      /* (ny-1)/(nyp-1) = 1 */
        if (o->ny >= 17) {
            double cap = (double)(o->ny - 1) * (double)(o->ny - 1) / 10. / 3.14;
            if (cap < dyinv)
                dyinv = cap;
        }
        double dxi = (double)(o->nx - 2) / o->xl;
        double dzi = (double)(o->nz - 2) / o->zl;
        for (size_t p = 0; p < (size_t)nzp * nxp; p++) {
            double c = fabs(R[p]) * dxi + fabs(R[(size_t)nzp * nxp + p]) * dyinv +
                       fabs(R[2 * (size_t)nzp * nxp + p]) * dzi;
            if (c > cflp)
                cflp = c;
        }
    }
    /* H = u x (omega + 2 Omega e_z) in rotational form (advection.f:52-72) */
    size_t np = (size_t)nzp * nxp;
    for (size_t p = 0; p < np; p++) {
        double u = R[p], v = R[np + p], w = R[2 * np + p];
        double ox = R[3 * np + p], oy = R[4 * np + p], oz = R[5 * np + p] + 2. * o->rot;
        double h1 = v * oz - w * oy;
        double h2 = w * ox - u * oz;
        double h3 = u * oy - v * ox;
        /* fringe damping toward the base flow (fring.f: no forced waves) */
        if (o->flam) {
            int xp = (int)(p % (size_t)nxp);
            double lam = o->flam[xp];
            const double *b = o->fbase + (size_t)y * nxp + xp;
            size_t bs = (size_t)o->ny * nxp;
            h1 -= lam * (u - b[0]);
            h2 -= lam * (v - b[bs]);
            h3 -= lam * (w - b[2 * bs]);
        }
        R[p] = h1;
        R[np + p] = h2;
        R[2 * np + p] = h3;
    }
    /* Fourier analysis (fft2df.f) and truncation into slots 0-2 (putxz.f,
     * including the z oddball row, which lands in storage z = nz/2). */
    for (int f = 0; f < 3; f++) {
        for (int zp = 0; zp < nzp; zp++)
            rfwd(&o->px_plan, R + ((size_t)f * nzp + zp) * nxp,
                 B + f * box + (size_t)zp * hx, t1, t2);
        for (int x = 0; x < nxh; x++) {
            for (int zp = 0; zp < nzp; zp++)
                col[zp] = B[f * box + (size_t)zp * hx + x];
            dft_inplace(&o->pz_plan, col, t1, -1);
            for (int z = 0; z < nz; z++) {
                int zp = zrow(o, z);
                size_t s = IDX(o, f, z, y, x);
                ur[s] = creal(col[zp]);
                ui[s] = cimag(col[zp]);
                /* ADM-RT relaxation (nonlinbl.f:920-966), scaled by nxp nzp
                 * as the transforms are unnormalised; the oddball row gets
                 * none (getxz.f zeroes it in the filtered field) */
                if (o->les && z != nz / 2) {
                    double c = o->les_chi * (double)nxp * (double)nzp;
                    ur[s] -= c * o->gur[s];
                    ui[s] -= c * o->gui[s];
                }
                /* eddy-viscosity SGS force sigma_i (nonlinbl.f:1011-1050):
                 * subtracted with the factor nxp nzp of the unnormalised
                 * transforms, not on the oddball row (getxz.f with ipad = 1
                 * zeroes it) */
                if (o->fr && z != nz / 2) {
                    double c = (double)nxp * (double)nzp;
                    ur[s] -= o->fr[s] * c;
                    ui[s] -= o->fi[s] * c;
                }
            }
        }
        /* parallel-flow forcing of temporal boundary layers on the mean
         * mode of H_1 and H_3 (nonlinbl.f:771-795), times nxp nzp */
        if (o->mf && f != 1)
            ur[IDX(o, f, 0, y, 0)] += o->mf[(size_t)f * o->ny + y] * (double)nxp * (double)nzp;
    }
    free(B);
    free(R);
    free(t1);
    free(t2);
    free(col);
    return cflp;
}

double orc_nonlin(orc_t *o, double *ur, double *ui, int do_cfl)
{
    double cfl = 0.;
    /* the filtered field of the current velocities (computed by the
     * reference at the end of the previous linearbl or in prhs, from the
     * same velocities) */
    if (o->les)
        les_gu(o, ur, ui);
#pragma omp parallel for schedule(dynamic) reduction(max : cfl)
    for (int y = 0; y < o->ny; y++) {
        double c = nonlin_plane(o, ur, ui, y, do_cfl);
        if (c > cfl)
            cfl = c;
    }
    return cfl;
}

/* ------------------------------------------------------------------ */
/* Step 3: linear solve per z plane (linearbl.f, tau method, ibc = 0)  */
/* ------------------------------------------------------------------ */

typedef struct {
    double *v[48];
    double *gam, *rho;
} work_t;

static void work_alloc(work_t *w, int n)
{
    for (int i = 0; i < 48; i++)
        w->v[i] = calloc((size_t)n, sizeof(double));
    w->gam = calloc((size_t)n, sizeof(double));
    w->rho = calloc((size_t)n, sizeof(double));
}

static void work_free(work_t *w)
{
    for (int i = 0; i < 48; i++)
        free(w->v[i]);
    free(w->gam);
    free(w->rho);
}

static double mass_flux(int n, const double *u)
{
    /* linearbl.f:1367-1370: integral over [-1,1] of sum u_k T_k */
    double m = 0.;
    for (int k = 0; k < n; k += 2)
        m -= u[k] * 2. / (double)(k * k - 1);
    return m;
}

/* ibc = 101 (evald.f, evalbc.f:142-159, solvbc.f, linearbl.f:946-1150):
 * the upper boundary y = +1 is a free stream, du_i/dy = dU_i/dy of the base
 * flow there (the kz = 0 plane; zero elsewhere), the lower one the no-slip
 * wall.  Three homogeneous solutions that vanish with their slope at the
 * wall are added to the particular one (v_p, om_p, which has v = dv/dy = 0
 * at both walls):
 *   vh1 = v_h with each parity class scaled to unit slope at +1
 *         (dvh1/dy(+1) = 2), so vh1(+-1) = 0, dvh1/dy(-1) = 0;
 *   vh2 = the k^2 Helmholtz solution with v = 1 per class (v(+1) = 2,
 *         v(-1) = 0), its wall slopes cancelled with v_h;
 *   omh = phi_h, the lam^2 solution with 1 per class (om(+1) = 2).
 * By continuity and the definition of omega_y, du/dy and dw/dy at +1 fix
 *   d2v/dy2 = -i (alpha U' + beta W'),  d om/dy = i (beta U' - alpha W'),
 * and dv/dy = V' directly, so v -= c1 vh1 + c2 vh2, om -= c3 omh with
 *   c1 = -V'/2, c2 = (d2v_p - c1 A1 - d2v*)/A2, c3 = (dom_p - dom*)/B,
 * A1, A2 = d2/dy2 of vh1, vh2 at +1 and B = domh/dy there.  The reference
 * run of tests/golden/blasius_64x33x8.npz pins this form (to 1e-13 after one
 * step). */
static void free_stream_bc(orc_t *o, int z, int x, double alfa, double beta, double *vr,
                           double *vi, double *omr, double *omi, const double *phh,
                           const double *vh, const double *Sh, work_t *W)
{
    int n = o->ny;
    double *vh1 = W->v[43], *vh2 = W->v[44], *t1 = W->v[45], *t2 = W->v[46];
    double *zero = W->v[39];
    double k2 = alfa * alfa + beta * beta;
    /* vh2 and its wall slopes cancelled with v_h, per class */
    tau_solve(n, k2, zero, 1., 1., vh2, W->gam, W->rho);
    dcheb(n, vh2, t1);
    double S2[2] = {0., 0.};
    for (int k = 0; k < n; k++)
        S2[k & 1] += t1[k];
    for (int k = 0; k < n; k++) {
        int c = k & 1;
        vh2[k] += (-S2[1 - c] / Sh[1 - c]) * vh[k];
        vh1[k] = vh[k] * (1. / Sh[1 - c]);
    }
    /* derivatives at y = +1: sums of the derivative coefficients */
    double A1 = 0., A2 = 0., B = 0., d2r = 0., d2i = 0., d1r = 0., d1i = 0.;
    dcheb(n, vh1, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        A1 += t2[k];
    dcheb(n, vh2, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        A2 += t2[k];
    dcheb(n, phh, t1);
    for (int k = 0; k < n; k++)
        B += t1[k];
    dcheb(n, vr, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        d2r += t2[k];
    dcheb(n, vi, t1);
    dcheb(n, t1, t2);
    for (int k = 0; k < n; k++)
        d2i += t2[k];
    dcheb(n, omr, t1);
    for (int k = 0; k < n; k++)
        d1r += t1[k];
    dcheb(n, omi, t1);
    for (int k = 0; k < n; k++)
        d1i += t1[k];
    /* targets from the base flow's U', V', W' (kz = 0 plane) */
    double Ur = 0., Ui = 0., Vr = 0., Vi = 0., Wr = 0., Wi = 0.;
    if (z == 0 && o->bjr) {
        Ur = o->bjr[x];
        Ui = o->bji[x];
        Vr = o->bjr[o->nxh + x];
        Vi = o->bji[o->nxh + x];
        Wr = o->bjr[2 * o->nxh + x];
        Wi = o->bji[2 * o->nxh + x];
    }
    double sr = alfa * Ur + beta * Wr, si = alfa * Ui + beta * Wi;   /* d2v* = -i s */
    double tr = beta * Ur - alfa * Wr, ti = beta * Ui - alfa * Wi;   /* dom* = i t */
    double c1r = -.5 * Vr, c1i = -.5 * Vi;
    double c2r = (d2r - c1r * A1 - si) / A2, c2i = (d2i - c1i * A1 + sr) / A2;
    double c3r = (d1r + ti) / B, c3i = (d1i - tr) / B;
    for (int k = 0; k < n; k++) {
        vr[k] -= c1r * vh1[k] + c2r * vh2[k];
        vi[k] -= c1i * vh1[k] + c2i * vh2[k];
        omr[k] -= c3r * phh[k];
        omi[k] -= c3i * phh[k];
    }
}

static void linear_plane(orc_t *o, double *ur, double *ui, double *puw, int z,
                         double an, double bn, double anp1, double bnp1, work_t *W)
{
    int n = o->ny, nxh = o->nxh;
    double re = o->re;
    double cn = 2. / ((double)o->nxp * (double)(n - 1) * (double)o->nzp);
    double beta = o->beta[z];
    double c1 = 2. * re / (an + bn) * an;           /* linearbl.f:304 */
    double c1n = 2. * re / (anp1 + bnp1);           /* linearbl.f:1729 (tau) */
    double c2 = 2. * re * bnp1 / (anp1 + bnp1);     /* linearbl.f:1721 */
    double *tmp = W->v[47];
    for (int x = 0; x < nxh; x++) {
        int k0 = (z == 0 && x == 0);
        double alfa = o->alfa[x];
        double k2 = alfa * alfa + beta * beta;
        double lam2 = k2 + 2. * re / (an + bn);
        double k2i = k0 ? 0. : 1. / k2;
        double *h1r = W->v[0], *h1i = W->v[1], *h2r = W->v[2], *h2i = W->v[3];
        double *h3r = W->v[4], *h3i = W->v[5];
        double *pvr = W->v[6], *pvi = W->v[7], *pwr = W->v[8], *pwi = W->v[9];
        double *hvr = W->v[10], *hvi = W->v[11], *hwr = W->v[12], *hwi = W->v[13];
        double *fvr = W->v[14], *fvi = W->v[15], *fwr = W->v[16], *fwi = W->v[17];
        double *phr = W->v[18], *phi = W->v[19], *phh = W->v[20];
        double *omr = W->v[21], *omi = W->v[22];
        double *vr = W->v[23], *vi = W->v[24], *vh = W->v[25];
        double *dvr = W->v[26], *dvi = W->v[27], *dvh = W->v[28];
        double *domr = W->v[29], *domi = W->v[30];
        double *ur_ = W->v[31], *ui_ = W->v[32], *wr_ = W->v[33], *wi_ = W->v[34];
        double *oxr = W->v[35], *oxi = W->v[36], *ozr = W->v[37], *ozi = W->v[38];
        double *zero = W->v[39];
        double *gr = W->v[40], *gi = W->v[41];
        /* gather and transform H to Chebyshev space (linearbl.f:262-273) */
        double *hh[6] = {h1r, h1i, h2r, h2i, h3r, h3i};
        for (int c = 0; c < 3; c++) {
            for (int y = 0; y < n; y++) {
                tmp[y] = ur[IDX(o, c, z, y, x)];
                W->v[42][y] = ui[IDX(o, c, z, y, x)];
            }
            chbf_tab(n, o->ct, tmp, hh[2 * c], cn);
            chbf_tab(n, o->ct, W->v[42], hh[2 * c + 1], cn);
        }
        for (int y = 0; y < n; y++) {
            pvr[y] = ur[IDX(o, 5, z, y, x)];
            pvi[y] = ui[IDX(o, 5, z, y, x)];
            pwr[y] = ur[IDX(o, 6, z, y, x)];
            pwi[y] = ui[IDX(o, 6, z, y, x)];
            zero[y] = 0.;
        }
        /* h_omega = i beta H1 - i alpha H3, f_omega = p_omega - c1 h_omega,
         * h_v = -(d/dy(i alpha H1 + i beta H3) + k^2 H2), f_v = p_v - c1 h_v
         * (linearbl.f:325-361) */
        for (int y = 0; y < n; y++) {
            hwr[y] = -beta * h1i[y] + alfa * h3i[y];
            hwi[y] = beta * h1r[y] - alfa * h3r[y];
            fwr[y] = pwr[y] - c1 * hwr[y];
            fwi[y] = pwi[y] - c1 * hwi[y];
            gr[y] = -alfa * h1i[y] - beta * h3i[y];
            gi[y] = alfa * h1r[y] + beta * h3r[y];
        }
        dcheb(n, gr, tmp);
        for (int y = 0; y < n; y++)
            hvr[y] = -(tmp[y] + k2 * h2r[y]);
        dcheb(n, gi, tmp);
        for (int y = 0; y < n; y++) {
            hvi[y] = -(tmp[y] + k2 * h2i[y]);
            fvr[y] = pvr[y] - c1 * hvr[y];
            fvi[y] = pvi[y] - c1 * hvi[y];
        }
        double fuw1[1024], fuw2[1024], huw1[1024], huw2[1024];
        if (k0) {                                   /* linearbl.f:303-319 */
            for (int y = 0; y < n; y++) {
                fuw1[y] = puw[y] - c1 * h1r[y];
                fuw2[y] = puw[n + y] - c1 * h3r[y];
                huw1[y] = h1r[y];
                huw2[y] = h3r[y];
                fvr[y] = fvi[y] = fwr[y] = fwi[y] = 0.;
            }
            fuw1[0] += 2. * re * o->px;
        }
        /* ibc = 100 (evalbc.f:142-159 with the values bu1jr(:, i, 1) of
         * bflow.f:80-84): u_i(+1) = U_i(+1) of the base flow on the kz = 0
         * plane (zero elsewhere, a wall), so v(+1) = V, and by continuity
         * and the definition of omega_y dv/dy(+1) = -i (alpha U + beta W),
         * omega_y(+1) = i (beta U - alpha W); the mean mode keeps its
         * Dirichlet values below */
        double dUr = 0., dUi = 0., dVr = 0., dVi = 0., dOr = 0., dOi = 0.;
        if (o->ibc == 100 && z == 0 && o->bjr && !k0) {
            double Ur = o->bjr[x], Ui = o->bji[x];
            double Wr = o->bjr[2 * nxh + x], Wi = o->bji[2 * nxh + x];
            dVr = o->bjr[nxh + x];
            dVi = o->bji[nxh + x];
            dUr = alfa * Ui + beta * Wi;           /* dv/dy(+1) */
            dUi = -(alfa * Ur + beta * Wr);
            dOr = -(beta * Ui - alfa * Wi);        /* omega_y(+1) */
            dOi = beta * Ur - alfa * Wr;
        }
        /* Helmholtz solves with lam^2 (linearbl.f:625-626): particular
         * phi (bc 0,0), homogeneous phi_h (bc 1,1), omega_y (bc 0,0, or
         * its free-stream value above) */
        tau_solve(n, lam2, fvr, 0., 0., phr, W->gam, W->rho);
        tau_solve(n, lam2, fvi, 0., 0., phi, W->gam, W->rho);
        tau_solve(n, lam2, zero, 1., 1., phh, W->gam, W->rho);
        tau_solve(n, lam2, fwr, .5 * dOr, .5 * dOr, omr, W->gam, W->rho);
        tau_solve(n, lam2, fwi, .5 * dOi, .5 * dOi, omi, W->gam, W->rho);
        /* v from phi with k^2 (linearbl.f:1000-1001); an imposed lower-wall
         * v_w (wbci = 1) enters as the Dirichlet values v(1) = 0,
         * v(-1) = v_w: (v_w/2, -v_w/2) for the even and odd classes
         * (linearbl.f:458-469, 944-945) */
        double bwr = 0., bwi = 0.;
        if (o->wvr && !k0) {
            bwr = .5 * o->wvr[(size_t)z * nxh + x];
            bwi = .5 * o->wvi[(size_t)z * nxh + x];
        }
        tau_solve(n, k2, phr, bwr + .5 * dVr, -bwr + .5 * dVr, vr, W->gam, W->rho);
        tau_solve(n, k2, phi, bwi + .5 * dVi, -bwi + .5 * dVi, vi, W->gam, W->rho);
        tau_solve(n, k2, phh, 0., 0., vh, W->gam, W->rho);
        /* superpose the homogeneous solution per parity so that dv/dy = 0
         * at both walls (linearbl.f:1023-1070), or dv/dy(+1) = dUr + i dUi
         * with ibc = 100: half of it in each class */
        dcheb(n, vr, dvr);
        dcheb(n, vi, dvi);
        dcheb(n, vh, dvh);
        double Sr[2] = {0, 0}, Si[2] = {0, 0}, Sh[2] = {0, 0};
        for (int y = 0; y < n; y++) {
            Sr[y & 1] += dvr[y];
            Si[y & 1] += dvi[y];
            Sh[y & 1] += dvh[y];
        }
        double cr[2], ci[2];
        for (int s = 0; s < 2; s++) {
            cr[s] = (.5 * dUr - Sr[1 - s]) * (1. / Sh[1 - s]);
            ci[s] = (.5 * dUi - Si[1 - s]) * (1. / Sh[1 - s]);
        }
        for (int y = 0; y < n; y++) {
            vr[y] += cr[y & 1] * vh[y];
            vi[y] += ci[y & 1] * vh[y];
        }
        /* ibc = 0: the particular solutions satisfy u=v=w=0 at the walls,
         * so the 3x3 boundary system (evald/evalbc/solvbc) has a zero RHS
         * and adds nothing. */
        if (o->ibc == 101 && !k0)
            free_stream_bc(o, z, x, alfa, beta, vr, vi, omr, omi, phh, vh, Sh, W);
        /* dv, phi = D^2 v - k^2 v (linearbl.f:1151-1166), d omega */
        dcheb(n, vr, dvr);
        dcheb(n, dvr, tmp);
        for (int y = 0; y < n; y++)
            phr[y] = tmp[y] - k2 * vr[y];
        dcheb(n, vi, dvi);
        dcheb(n, dvi, tmp);
        for (int y = 0; y < n; y++)
            phi[y] = tmp[y] - k2 * vi[y];
        dcheb(n, omr, domr);
        dcheb(n, omi, domi);
        /* u, w, omega_x, omega_z (linearbl.f:1190-1212) */
        for (int y = 0; y < n; y++) {
            ur_[y] = (-alfa * dvi[y] + beta * omi[y]) * k2i;
            ui_[y] = (alfa * dvr[y] - beta * omr[y]) * k2i;
            wr_[y] = (-alfa * omi[y] - beta * dvi[y]) * k2i;
            wi_[y] = (alfa * omr[y] + beta * dvr[y]) * k2i;
            ozr[y] = (-beta * domi[y] + alfa * phi[y]) * k2i;
            ozi[y] = (beta * domr[y] - alfa * phr[y]) * k2i;
            oxr[y] = (-alfa * domi[y] - beta * phi[y]) * k2i;
            oxi[y] = (alfa * domr[y] + beta * phr[y]) * k2i;
        }
        if (k0) {
            /* mean flow (linearbl.f:1216-1692, tau branch, ibc = 0) */
            double *u0 = W->v[43], *w0 = W->v[44], *du0 = W->v[45], *dw0 = W->v[46];
            double u0upp = o->u0upp, u0low = o->u0low, w0upp = o->w0upp, w0low = o->w0low;
            if (o->ibc == 100 || o->ibc == 101) {
                /* with a base flow: U = 1 above (with ibc = 101 replaced by
                 * the free-stream condition below), 0 below
                 * (bla.f:1360-1365) */
                u0upp = 1.;
                u0low = w0upp = w0low = 0.;
            }
            double bu_e = .5 * (u0upp + u0low), bu_o = .5 * (u0upp - u0low);
            double bw_e = .5 * (w0upp + w0low), bw_o = .5 * (w0upp - w0low);
            tau_solve(n, lam2, fuw1, bu_e, bu_o, u0, W->gam, W->rho);
            tau_solve(n, lam2, fuw2, bw_e, bw_o, w0, W->gam, W->rho);
            if (o->cflux) {
                /* secant step on the pressure gradient (linearbl.f:1355-1417) */
                double m1 = mass_flux(n, u0);
                fuw1[0] += 2. * re * o->px;
                tau_solve(n, lam2, fuw1, bu_e, bu_o, u0, W->gam, W->rho);
                double m2 = mass_flux(n, u0);
                fuw1[0] -= 4. * re * o->px;
                o->px = o->px * (1. + (o->mflux - m1) / (m2 - m1));
                fuw1[0] += 2. * re * o->px;
                tau_solve(n, lam2, fuw1, bu_e, bu_o, u0, W->gam, W->rho);
            }
            if (o->ibc == 101) {
                /* free stream above (linearbl.f:1473-1487): add the lam^2
                 * homogeneous solution phi_h (1 per class: 2 at y = +1, 0 at
                 * the wall) so that du0/dy(+1) = du0upp, dw0/dy(+1) = 0 */
                double *t1 = W->v[42];
                double d0u = 0., d0w = 0., dh = 0.;
                dcheb(n, u0, t1);
                for (int y = 0; y < n; y++)
                    d0u += t1[y];
                dcheb(n, w0, t1);
                for (int y = 0; y < n; y++)
                    d0w += t1[y];
                dcheb(n, phh, t1);
                for (int y = 0; y < n; y++)
                    dh += t1[y];
                double cu = (d0u - o->du0upp) / dh, cw = d0w / dh;
                for (int y = 0; y < n; y++) {
                    u0[y] -= cu * phh[y];
                    w0[y] -= cw * phh[y];
                }
            }
            dcheb(n, u0, du0);
            dcheb(n, w0, dw0);
            for (int y = 0; y < n; y++) {
                oxr[y] = dw0[y];
                ozr[y] = -du0[y];
                ur_[y] = u0[y];
                wr_[y] = w0[y];
                vr[y] = 0.;
                vi[y] = 0.;
                ui_[y] = wi_[y] = oxi[y] = ozi[y] = 0.;
            }
            /* the mean v is the mean wall velocity, constant in y
             * (linearbl.f:1663-1670) */
            if (o->wvr)
                vr[0] = o->wvr[0];
            /* next partial RHS for the mean flow (linearbl.f:1734-1746) */
            dcheb(n, du0, tmp);
            for (int y = 0; y < n; y++)
                puw[y] = -tmp[y] - c1n * u0[y] - c2 * huw1[y];
            dcheb(n, dw0, tmp);
            for (int y = 0; y < n; y++)
                puw[n + y] = -tmp[y] - c1n * w0[y] - c2 * huw2[y];
        }
        /* next partial RHS (linearbl.f:1761-1784, tau branch) */
        double k2m = k2 - c1n;
        double *d1 = W->v[42];
        dcheb(n, phr, d1);
        dcheb(n, d1, tmp);
        for (int y = 0; y < n; y++)
            pvr[y] = -tmp[y] + k2m * phr[y] - c2 * hvr[y];
        dcheb(n, phi, d1);
        dcheb(n, d1, tmp);
        for (int y = 0; y < n; y++)
            pvi[y] = -tmp[y] + k2m * phi[y] - c2 * hvi[y];
        dcheb(n, domr, tmp);
        for (int y = 0; y < n; y++)
            pwr[y] = -tmp[y] + k2m * omr[y] - c2 * hwr[y];
        dcheb(n, domi, tmp);
        for (int y = 0; y < n; y++)
            pwi[y] = -tmp[y] + k2m * omi[y] - c2 * hwi[y];
        /* back to physical y and store (linearbl.f:1830-1867) */
        double *outs[10] = {ur_, ui_, vr, vi, wr_, wi_, oxr, oxi, ozr, ozi};
        for (int c = 0; c < 5; c++) {
            chbb_tab(n, o->ct, outs[2 * c], tmp);
            for (int y = 0; y < n; y++)
                ur[IDX(o, c, z, y, x)] = tmp[y];
            chbb_tab(n, o->ct, outs[2 * c + 1], tmp);
            for (int y = 0; y < n; y++)
                ui[IDX(o, c, z, y, x)] = tmp[y];
        }
        for (int y = 0; y < n; y++) {
            ur[IDX(o, 5, z, y, x)] = pvr[y];
            ui[IDX(o, 5, z, y, x)] = pvi[y];
            ur[IDX(o, 6, z, y, x)] = pwr[y];
            ui[IDX(o, 6, z, y, x)] = pwi[y];
        }
    }
}

void orc_linear(orc_t *o, double *ur, double *ui, double *puw,
                double an, double bn, double anp1, double bnp1)
{
#pragma omp parallel
    {
        work_t W;
        work_alloc(&W, o->ny);
#pragma omp for schedule(dynamic)
        for (int z = 0; z < o->nz; z++) {
            if (z == o->nz / 2)          /* linearbl.f:227, z oddball */
                continue;
            linear_plane(o, ur, ui, puw, z, an, bn, anp1, bnp1, &W);
        }
        work_free(&W);
    }
}

/* ------------------------------------------------------------------ */
/* prhs (prhs.f:93-709)                                                */
/* ------------------------------------------------------------------ */

static void prhs_plane(orc_t *o, double *ur, double *ui, double *puw, int z,
                       double an, work_t *W)
{
    int n = o->ny, nxh = o->nxh;
    double re = o->re, beta = o->beta[z];
    double cn = 2. / (double)(n - 1);
    double *tmp = W->v[47];
    for (int x = 0; x < nxh; x++) {
        double alfa = o->alfa[x];
        double k2 = alfa * alfa + beta * beta;
        double lam = k2 - 2. * re / an;
        double *U[6];
        for (int c = 0; c < 6; c++)
            U[c] = W->v[c];
        for (int c = 0; c < 3; c++) {
            for (int y = 0; y < n; y++) {
                W->v[40][y] = ur[IDX(o, c, z, y, x)];
                W->v[41][y] = ui[IDX(o, c, z, y, x)];
            }
            chbf_tab(n, o->ct, W->v[40], U[2 * c], cn);
            chbf_tab(n, o->ct, W->v[41], U[2 * c + 1], cn);
        }
        double *uR = U[0], *uI = U[1], *vR = U[2], *vI = U[3], *wR = U[4], *wI = U[5];
        double *oyr = W->v[6], *oyi = W->v[7], *oxr = W->v[8], *oxi = W->v[9];
        double *ozr = W->v[10], *ozi = W->v[11], *phr = W->v[12], *phi = W->v[13];
        double *d1 = W->v[14], *d2 = W->v[15];
        if (z == 0 && x == 0) {
            dcheb(n, uR, d1);
            dcheb(n, d1, d2);
            for (int y = 0; y < n; y++)
                puw[y] = -2. * re / an * uR[y] - d2[y];
            dcheb(n, wR, d1);
            dcheb(n, d1, d2);
            for (int y = 0; y < n; y++)
                puw[n + y] = -2. * re / an * wR[y] - d2[y];
        }
        for (int y = 0; y < n; y++) {
            oyr[y] = -uI[y] * beta + wI[y] * alfa;
            oyi[y] = uR[y] * beta - wR[y] * alfa;
        }
        dcheb(n, wR, oxr);
        dcheb(n, wI, oxi);
        dcheb(n, uR, ozr);
        dcheb(n, uI, ozi);
        for (int y = 0; y < n; y++) {
            oxr[y] = oxr[y] + vI[y] * beta;
            oxi[y] = oxi[y] - vR[y] * beta;
            ozr[y] = -ozr[y] - vI[y] * alfa;
            ozi[y] = -ozi[y] + vR[y] * alfa;
        }
        /* phi = D^2 v - k^2 v, p_omega = -D^2 omega_y + lam omega_y,
         * p_v = -D^2 phi + lam phi (prhs.f:595-672) */
        double *pvr = W->v[16], *pvi = W->v[17], *pwr = W->v[18], *pwi = W->v[19];
        dcheb(n, vR, d1);
        dcheb(n, d1, phr);
        dcheb(n, vI, d1);
        dcheb(n, d1, phi);
        dcheb(n, oyr, d1);
        dcheb(n, d1, pwr);
        dcheb(n, oyi, d1);
        dcheb(n, d1, pwi);
        for (int y = 0; y < n; y++) {
            phr[y] = phr[y] - k2 * vR[y];
            phi[y] = phi[y] - k2 * vI[y];
            pwr[y] = -pwr[y] + lam * oyr[y];
            pwi[y] = -pwi[y] + lam * oyi[y];
        }
        dcheb(n, phr, d1);
        dcheb(n, d1, pvr);
        dcheb(n, phi, d1);
        dcheb(n, d1, pvi);
        for (int y = 0; y < n; y++) {
            pvr[y] = -pvr[y] + lam * phr[y];
            pvi[y] = -pvi[y] + lam * phi[y];
        }
        double *outs[4] = {oxr, oxi, ozr, ozi};
        for (int c = 0; c < 2; c++) {
            chbb_tab(n, o->ct, outs[2 * c], tmp);
            for (int y = 0; y < n; y++)
                ur[IDX(o, 3 + c, z, y, x)] = tmp[y];
            chbb_tab(n, o->ct, outs[2 * c + 1], tmp);
            for (int y = 0; y < n; y++)
                ui[IDX(o, 3 + c, z, y, x)] = tmp[y];
        }
        for (int y = 0; y < n; y++) {
            ur[IDX(o, 5, z, y, x)] = pvr[y];
            ui[IDX(o, 5, z, y, x)] = pvi[y];
            ur[IDX(o, 6, z, y, x)] = pwr[y];
            ui[IDX(o, 6, z, y, x)] = pwi[y];
        }
    }
}

void orc_prhs(orc_t *o, double *ur, double *ui, double *puw, double an)
{
    /* u0upp etc. from the wall values of the mean mode (rdiscbl.f:319-323,
     * thresholded as in bla.f:1041-1044) */
    double v[4] = {ur[IDX(o, 0, 0, 0, 0)], ur[IDX(o, 0, 0, o->ny - 1, 0)],
                   ur[IDX(o, 2, 0, 0, 0)], ur[IDX(o, 2, 0, o->ny - 1, 0)]};
    for (int i = 0; i < 4; i++)
        if (fabs(v[i]) < 1e-8)
            v[i] = 0.;
    o->u0upp = v[0];
    o->u0low = v[1];
    o->w0upp = v[2];
    o->w0low = v[3];
#pragma omp parallel
    {
        work_t W;
        work_alloc(&W, o->ny);
#pragma omp for schedule(dynamic)
        for (int z = 0; z < o->nz; z++)
            prhs_plane(o, ur, ui, puw, z, an, &W);
        work_free(&W);
    }
}

/* ------------------------------------------------------------------ */
/* Driver (bla.f:1465-1477, 1790-2610 with fixed dt)                   */
/* ------------------------------------------------------------------ */

int orc_run(orc_t *o, double *ur, double *ui, double *puw, int nsteps,
            double dt, int icfl, double *cfl_out)
{
    static const double anrk[3] = {8. / 15., 5. / 12., 3. / 4.};
    static const double bnrk[3] = {0., -17. / 60., -5. / 12.};
    const int nst = 3;
    double dtn = dt, an = dtn * anrk[0], bn, anp1 = an, bnp1 = 0., dtnp1 = dtn;
    orc_prhs(o, ur, ui, puw, anp1);
    int status = 0;
    for (int it = 1; it <= nst * nsteps; it++) {
        int do_cfl = icfl > 0 && ((it - 1) % icfl) == 0;
        double cflp = orc_nonlin(o, ur, ui, do_cfl);
        if (do_cfl) {
            double cfl = (cflp * PI + 2. * fabs(o->rot)) * dtn;   /* wcflbl.f:87 */
            if (cfl_out)
                *cfl_out = cfl;
            if (cfl > 10.)
                return 324234;
        }
        an = anp1;
        bn = bnp1;
        anp1 = dtnp1 * anrk[it % nst];
        bnp1 = dtnp1 * bnrk[it % nst];
        orc_linear(o, ur, ui, puw, an, bn, anp1, bnp1);
    }
    return status;
}

/* ------------------------------------------------------------------ */
/* Pressure (bla.f:2486-2527: nonlinp.f then linearp.f)                */
/* ------------------------------------------------------------------ */

/* Tau solve of u'' - lam u = f with the Neumann boundary rows of
 * setmac.f (neumann = 1): sum_{k even} k^2 a_k = bc_e,
 * sum_{k odd} k^2 a_k = bc_o, i.e. du/dy at y = +-1 (T_k'(1) = k^2).  The
 * interior rows are those of tau_solve (setmac.f:48-87 and crhsc.f build the
 * same system as setmacrhs.f).  When a class's boundary row vanishes after
 * elimination (lam = 0, even class: the constant is free) its first
 * coefficient is set to zero, as tridc.f:127-131. */
static void tau_solve_neumann(int n, double lam, const double *f, double bce, double bco,
                              double *a, double *gam, double *rho)
{
    for (int par = 0; par < 2; par++) {
        int k0 = par;
        int klast = (n - 1 - par) % 2 == 0 ? n - 1 : n - 2;
        double g_up = 0.0, r_up = 0.0;
        for (int k = klast; k >= k0 + 2; k -= 2) {
            double kk = (double)k;
            double ckm2 = (k - 2 == 0) ? 2.0 : 1.0;
            double L = -ckm2 * lam / (4.0 * kk * (kk - 1.0));
            double D = 1.0 + (k <= n - 3 ? lam / (2.0 * (kk * kk - 1.0)) : 0.0);
            double U = (k <= n - 5) ? -lam / (4.0 * kk * (kk + 1.0)) : 0.0;
            double R = ckm2 * f[k - 2] / (4.0 * kk * (kk - 1.0));
            if (k <= n - 3)
                R -= f[k] / (2.0 * (kk * kk - 1.0));
            if (k <= n - 5)
                R += f[k + 2] / (4.0 * kk * (kk + 1.0));
            double piv = D - U * g_up;
            gam[k] = L / piv;
            rho[k] = (R - U * r_up) / piv;
            g_up = gam[k];
            r_up = rho[k];
        }
        double P = 0.0, Q = 1.0, w0 = (double)(k0 * k0);
        double sP = 0.0, sQ = w0;
        for (int k = k0 + 2; k <= klast; k += 2) {
            P = rho[k] - gam[k] * P;
            Q = -gam[k] * Q;
            sP += (double)k * (double)k * P;
            sQ += (double)k * (double)k * Q;
        }
        double bc = par == 0 ? bce : bco;
        double a0 = sQ == 0.0 ? 0.0 : (bc - sP) / sQ;
        a[k0] = a0;
        for (int k = k0 + 2; k <= klast; k += 2)
            a[k] = rho[k] - gam[k] * a[k - 2];
    }
}

void orc_tau_solve_neumann(int n, double lam, const double *f, double bc_e, double bc_o,
                           double *u)
{
    double *g = malloc(sizeof(double) * (size_t)n), *r = malloc(sizeof(double) * (size_t)n);
    tau_solve_neumann(n, lam, f, bc_e, bc_o, u, g, r);
    free(g);
    free(r);
}

/* nonlinp.f for one y plane: H = u x (om + 2 Omega e_z) and W = |u|^2 in
 * physical space, forward transforms, then D = i(alpha H1 + beta H3); the
 * results (unnormalised, times nxp nzp) go to q[3][z][y][x]: D, H2, W
 * (putxz.f into slots 4, 5, 8 in the reference, z oddball row included). */
static void nonlinp_plane(orc_t *o, const double *ur, const double *ui, int y, double *qr,
                          double *qi)
{
    int nxh = o->nxh, nz = o->nz, nxp = o->nxp, nzp = o->nzp, hx = nxp / 2 + 1;
    size_t box = (size_t)nzp * hx;
    cplx *B = calloc(6 * box, sizeof(cplx));
    double *R = malloc(sizeof(double) * 6 * (size_t)nzp * nxp);
    int nmax = nxp > nzp ? nxp : nzp;
    cplx *t1 = malloc(sizeof(cplx) * (size_t)nmax), *t2 = malloc(sizeof(cplx) * (size_t)nmax);
    cplx *col = malloc(sizeof(cplx) * (size_t)nzp);
    static const int slot_of[6] = {0, 1, 2, 3, -1, 4};
    for (int f = 0; f < 6; f++) {
        if (slot_of[f] < 0)
            continue;
        for (int z = 0; z < nz; z++) {
            if (z == nz / 2)
                continue;
            int zp = zrow(o, z);
            for (int x = 0; x < nxh; x++) {
                size_t s = IDX(o, slot_of[f], z, y, x);
                B[f * box + (size_t)zp * hx + x] = ur[s] + I * ui[s];
            }
        }
    }
    for (int z = 0; z < nz; z++) {
        if (z == nz / 2)
            continue;
        int zp = zrow(o, z);
        for (int x = 0; x < nxh; x++) {
            cplx uu = B[0 * box + (size_t)zp * hx + x], ww = B[2 * box + (size_t)zp * hx + x];
            B[4 * box + (size_t)zp * hx + x] = I * o->beta[z] * uu - I * o->alfa[x] * ww;
        }
    }
    for (int f = 0; f < 6; f++) {
        for (int x = 0; x < nxh; x++) {
            for (int zp = 0; zp < nzp; zp++)
                col[zp] = B[f * box + (size_t)zp * hx + x];
            dft_inplace(&o->pz_plan, col, t1, +1);
            for (int zp = 0; zp < nzp; zp++)
                B[f * box + (size_t)zp * hx + x] = col[zp];
        }
        for (int zp = 0; zp < nzp; zp++)
            rbwd(&o->px_plan, B + f * box + (size_t)zp * hx,
                 R + ((size_t)f * nzp + zp) * nxp, t1, t2);
    }
    /* nonlinp.f: products and |u|^2 (fields 0-2 H, 3 W) */
    size_t np = (size_t)nzp * nxp;
    for (size_t p = 0; p < np; p++) {
        double u = R[p], v = R[np + p], w = R[2 * np + p];
        double ox = R[3 * np + p], oy = R[4 * np + p], oz = R[5 * np + p] + 2. * o->rot;
        R[p] = v * oz - w * oy;
        R[np + p] = w * ox - u * oz;
        R[2 * np + p] = u * oy - v * ox;
        R[3 * np + p] = u * u + v * v + w * w;
        /* the same fringe term (fringp.f) */
        if (o->flam) {
            int xp = (int)(p % (size_t)nxp);
            double lam = o->flam[xp];
            const double *b = o->fbase + (size_t)y * nxp + xp;
            size_t bs = (size_t)o->ny * nxp;
            R[p] -= lam * (u - b[0]);
            R[np + p] -= lam * (v - b[bs]);
            R[2 * np + p] -= lam * (w - b[2 * bs]);
        }
    }
    cplx *S = malloc(sizeof(cplx) * 4 * (size_t)nz * nxh);   /* [f][z][x] */
    for (int f = 0; f < 4; f++) {
        for (int zp = 0; zp < nzp; zp++)
            rfwd(&o->px_plan, R + ((size_t)f * nzp + zp) * nxp,
                 B + f * box + (size_t)zp * hx, t1, t2);
        for (int x = 0; x < nxh; x++) {
            for (int zp = 0; zp < nzp; zp++)
                col[zp] = B[f * box + (size_t)zp * hx + x];
            dft_inplace(&o->pz_plan, col, t1, -1);
            for (int z = 0; z < nz; z++)
                S[((size_t)f * nz + z) * nxh + x] = col[zrow(o, z)];
        }
    }
    for (int z = 0; z < nz; z++)
        for (int x = 0; x < nxh; x++) {
            cplx h1 = S[((size_t)0 * nz + z) * nxh + x], h3 = S[((size_t)2 * nz + z) * nxh + x];
            cplx d = I * (o->alfa[x] * h1 + o->beta[z] * h3);
            cplx h2 = S[((size_t)1 * nz + z) * nxh + x], w = S[((size_t)3 * nz + z) * nxh + x];
            size_t s0 = IDX(o, 0, z, y, x), s1 = IDX(o, 1, z, y, x), s2 = IDX(o, 2, z, y, x);
            qr[s0] = creal(d);
            qi[s0] = cimag(d);
            qr[s1] = creal(h2);
            qi[s1] = cimag(h2);
            qr[s2] = creal(w);
            qi[s2] = cimag(w);
        }
    free(S);
    free(B);
    free(R);
    free(t1);
    free(t2);
    free(col);
}

/* linearp.f for one z plane (not the oddball): Neumann Poisson problem
 * (D^2 - k^2) phi = dH2/dy + i(alpha H1 + beta H3) with
 * dphi/dy(+-1) = H2 + (v'' - k^2 v)/Re at the walls (the time-derivative
 * term of linearp.f vanishes for the no-slip walls of this path), the
 * k = 0 mode from the integral of H2 with phi(-1) = E(-1) (icheb.f), then
 * p = phi - E with E = |u|^2 / 2; p at physical y into pr, pi [z][y][x]. */
static void linearp_plane(orc_t *o, const double *ur, const double *ui, const double *qr,
                          const double *qi, int z, double *pr, double *pi)
{
    int n = o->ny, nxh = o->nxh;
    double c4 = 2. / ((double)o->nxp * (double)(n - 1) * (double)o->nzp);
    double cw = 1. / ((double)o->nxp * (double)o->nzp);
    double *buf = malloc(sizeof(double) * 16 * (size_t)n);
    double *dr = buf, *di = buf + n, *hr = buf + 2 * n, *hi = buf + 3 * n;
    double *er = buf + 4 * n, *ei = buf + 5 * n, *fr = buf + 6 * n, *fi = buf + 7 * n;
    double *pcr = buf + 8 * n, *pci = buf + 9 * n, *g = buf + 10 * n, *r = buf + 11 * n;
    double *tr = buf + 12 * n, *ti = buf + 13 * n, *vr = buf + 14 * n, *vi = buf + 15 * n;
    double beta = o->beta[z];
    for (int x = 0; x < nxh; x++) {
        double k2 = o->alfa[x] * o->alfa[x] + beta * beta;
        for (int y = 0; y < n; y++) {
            tr[y] = qr[IDX(o, 0, z, y, x)];
            ti[y] = qi[IDX(o, 0, z, y, x)];
        }
        chbf_tab(n, o->ct, tr, dr, c4);
        chbf_tab(n, o->ct, ti, di, c4);
        double v1r = qr[IDX(o, 1, z, 0, x)] * cw, v1i = qi[IDX(o, 1, z, 0, x)] * cw;
        double v2r = qr[IDX(o, 1, z, n - 1, x)] * cw, v2i = qi[IDX(o, 1, z, n - 1, x)] * cw;
        for (int y = 0; y < n; y++) {
            tr[y] = qr[IDX(o, 1, z, y, x)];
            ti[y] = qi[IDX(o, 1, z, y, x)];
        }
        chbf_tab(n, o->ct, tr, hr, c4);
        chbf_tab(n, o->ct, ti, hi, c4);
        double bc0 = qr[IDX(o, 2, z, n - 1, x)] * cw * 0.5;
        for (int y = 0; y < n; y++) {
            tr[y] = qr[IDX(o, 2, z, y, x)];
            ti[y] = qi[IDX(o, 2, z, y, x)];
        }
        chbf_tab(n, o->ct, tr, er, 0.5 * c4);
        chbf_tab(n, o->ct, ti, ei, 0.5 * c4);
        /* f = dH2/dy + D */
        dcheb(n, hr, fr);
        dcheb(n, hi, fi);
        for (int k = 0; k < n; k++) {
            fr[k] += dr[k];
            fi[k] += di[k];
        }
        /* v'' at the walls from the Chebyshev coefficients of v
         * (the weights wd1 of preprbl.f:91-113) */
        for (int y = 0; y < n; y++) {
            tr[y] = ur[IDX(o, 1, z, y, x)];
            ti[y] = ui[IDX(o, 1, z, y, x)];
        }
        chbf_tab(n, o->ct, tr, vr, 2. / (double)(n - 1));
        chbf_tab(n, o->ct, ti, vi, 2. / (double)(n - 1));
        double ddr1 = 0., ddi1 = 0., ddr2 = 0., ddi2 = 0.;
        for (int k = 0; k < n; k++) {
            double kk = (double)k, w2 = kk * kk * (kk * kk - 1.) / 3.;
            double sg = (k & 1) ? -1. : 1.;
            ddr1 += w2 * vr[k];
            ddi1 += w2 * vi[k];
            ddr2 += sg * w2 * vr[k];
            ddi2 += sg * w2 * vi[k];
        }
        double re = o->re;
        v1r += (ddr1 - k2 * ur[IDX(o, 1, z, 0, x)]) / re;
        v1i += (ddi1 - k2 * ui[IDX(o, 1, z, 0, x)]) / re;
        v2r += (ddr2 - k2 * ur[IDX(o, 1, z, n - 1, x)]) / re;
        v2i += (ddi2 - k2 * ui[IDX(o, 1, z, n - 1, x)]) / re;
        tau_solve_neumann(n, k2, fr, 0.5 * (v1r - v2r), 0.5 * (v1r + v2r), pcr, g, r);
        tau_solve_neumann(n, k2, fi, 0.5 * (v1i - v2i), 0.5 * (v1i + v2i), pci, g, r);
        if (z == 0 && x == 0) {
            /* mean pressure: integral of H2 with phi(-1) = E(-1) */
            hr[n - 1] = 0.;
            orc_icheb(n, hr, pcr, bc0);
            for (int k = 1; k + 1 < n; k += 2)
                pcr[0] += pcr[k] - pcr[k + 1];
            for (int k = 0; k < n; k++)
                pci[k] = 0.;
        }
        for (int k = 0; k < n; k++) {
            pcr[k] -= er[k];
            pci[k] -= ei[k];
        }
        chbb_tab(n, o->ct, pcr, tr);
        chbb_tab(n, o->ct, pci, ti);
        for (int y = 0; y < n; y++) {
            pr[((size_t)z * n + y) * nxh + x] = tr[y];
            pi[((size_t)z * n + y) * nxh + x] = ti[y];
        }
    }
    free(buf);
}

void orc_pressure(orc_t *o, const double *ur, const double *ui, double *pr, double *pi)
{
    size_t F = (size_t)o->nz * o->ny * o->nxh;
    double *qr = malloc(sizeof(double) * 3 * F), *qi = malloc(sizeof(double) * 3 * F);
#pragma omp parallel for schedule(dynamic)
    for (int y = 0; y < o->ny; y++)
        nonlinp_plane(o, ur, ui, y, qr, qi);
#pragma omp parallel for schedule(dynamic)
    for (int z = 0; z < o->nz; z++) {
        if (z == o->nz / 2) {
            /* the oddball plane keeps what nonlinp stored in slot 8 */
            for (int y = 0; y < o->ny; y++)
                for (int x = 0; x < o->nxh; x++) {
                    pr[((size_t)z * o->ny + y) * o->nxh + x] = qr[IDX(o, 2, z, y, x)];
                    pi[((size_t)z * o->ny + y) * o->nxh + x] = qi[IDX(o, 2, z, y, x)];
                }
            continue;
        }
        linearp_plane(o, ur, ui, qr, qi, z, pr, pi);
    }
    free(qr);
    free(qi);
}

/* ------------------------------------------------------------------ */
/* ADM-RT LES (iles = 1, ihighorder = 2): filters of init_filter.f,    */
/* high-pass velocity of linearbl.f:1871-1962 / prhs.f:95-195 and the  */
/* relaxation term of nonlinbl.f:920-966                               */
/* ------------------------------------------------------------------ */

/* The filter operators themselves come from the caller (oracle/les_ref.py
 * derives them from their defining conditions): lpf [nz][nxh], the x-z
 * transfer function of bla.f:1503-1518, and L = I - H_y^N [ny][ny], so that
 * u - H_y^N u = L u. */
struct les_s {
    int ny;
    double *lpf;
    double *L;
};

static void les_free(struct les_s *L)
{
    if (!L)
        return;
    free(L->lpf);
    free(L->L);
    free(L);
}

/* External SGS force for the next orc_nonlin calls (sigma of the
 * eddy-viscosity models, sgs_stress.f, at physical y): fr, fi
 * [3][nz][ny][nxh] are copied; NULL switches it off. */
void orc_set_force(orc_t *o, const double *fr, const double *fi)
{
    size_t n = (size_t)3 * o->nz * o->ny * o->nxh;
    if (!fr || !fi) {
        free(o->fr);
        free(o->fi);
        o->fr = o->fi = NULL;
        return;
    }
    if (!o->fr) {
        o->fr = malloc(n * sizeof(double));
        o->fi = malloc(n * sizeof(double));
    }
    memcpy(o->fr, fr, n * sizeof(double));
    memcpy(o->fi, fi, n * sizeof(double));
}

/* Lower-wall vertical velocity of wbci = 1 for the next orc_linear calls
 * (the spectral v_w of cwallbc.f, [nz][nxh] real and imaginary parts);
 * NULL switches it off. */
void orc_set_wall(orc_t *o, const double *vr, const double *vi)
{
    size_t n = (size_t)o->nz * o->nxh;
    if (!vr || !vi) {
        free(o->wvr);
        free(o->wvi);
        o->wvr = o->wvi = NULL;
        return;
    }
    if (!o->wvr) {
        o->wvr = malloc(n * sizeof(double));
        o->wvi = malloc(n * sizeof(double));
    }
    memcpy(o->wvr, vr, n * sizeof(double));
    memcpy(o->wvi, vi, n * sizeof(double));
}

/* Fringe forcing of spatial simulations (fring.f, fringp.f without forced
 * waves): H_i -= lam(x) (u_i - U_i(x, y)) at the physical points of the 3/2
 * x grid, lam [nxp], base [3][ny][nxp]; NULL switches it off. */
void orc_set_fringe(orc_t *o, const double *lam, const double *base)
{
    free(o->flam);
    free(o->fbase);
    o->flam = o->fbase = NULL;
    if (!lam || !base)
        return;
    size_t nb = (size_t)3 * o->ny * o->nxp;
    o->flam = malloc((size_t)o->nxp * sizeof(double));
    o->fbase = malloc(nb * sizeof(double));
    memcpy(o->flam, lam, (size_t)o->nxp * sizeof(double));
    memcpy(o->fbase, base, nb * sizeof(double));
}

int orc_set_les(orc_t *o, const double *lpf, const double *L, double chi)
{
    les_free(o->les);
    o->les = NULL;
    free(o->gur);
    free(o->gui);
    o->gur = o->gui = NULL;
    if (!lpf || !L)
        return 0;
    struct les_s *F = calloc(1, sizeof(struct les_s));
    size_t nl = (size_t)o->nz * o->nxh, nm = (size_t)o->ny * o->ny;
    F->ny = o->ny;
    F->lpf = malloc(sizeof(double) * nl);
    F->L = malloc(sizeof(double) * nm);
    memcpy(F->lpf, lpf, sizeof(double) * nl);
    memcpy(F->L, L, sizeof(double) * nm);
    o->les = F;
    o->les_chi = chi;                   /* chi/dstar, dstar = 1 (bla.f:2220) */
    size_t fs = (size_t)o->nz * o->ny * o->nxh;
    o->gur = malloc(sizeof(double) * 3 * fs);
    o->gui = malloc(sizeof(double) * 3 * fs);
    return 0;
}

/* high-pass filtered velocities g = u - lpf (u - H_y^N u) = u - lpf L u of
 * slots 0-2 (linearbl.f:1871-1962, ihighorder = 2) into o->gur, o->gui */
static void les_gu(orc_t *o, const double *ur, const double *ui)
{
    const struct les_s *F = o->les;
    int ny = o->ny, nxh = o->nxh;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < 3; s++)
        for (int z = 0; z < o->nz; z++)
            for (int x = 0; x < nxh; x++) {
                double lp = F->lpf[(size_t)z * nxh + x];
                for (int part = 0; part < 2; part++) {
                    const double *u = part ? ui : ur;
                    double *g = part ? o->gui : o->gur;
                    for (int y = 0; y < ny; y++) {
                        double lu = 0.;
                        for (int k = 0; k < ny; k++)
                            lu += F->L[(size_t)y * ny + k] * u[IDX(o, s, z, k, x)];
                        g[IDX(o, s, z, y, x)] = u[IDX(o, s, z, y, x)] - lp * lu;
                    }
                }
            }
}

int orc_set_bc(orc_t *o, int ibc, const double *bjr, const double *bji, double du0upp)
{
    if (ibc != 0 && ibc != 100 && ibc != 101)
        return -1;
    o->ibc = ibc;
    o->du0upp = du0upp;
    free(o->bjr);
    free(o->bji);
    o->bjr = o->bji = NULL;
    if (bjr) {
        size_t m = 3 * (size_t)o->nxh;
        o->bjr = malloc(sizeof(double) * m);
        o->bji = malloc(sizeof(double) * m);
        memcpy(o->bjr, bjr, sizeof(double) * m);
        memcpy(o->bji, bji, sizeof(double) * m);
    }
    return 0;
}

void orc_set_mean_force(orc_t *o, const double *f)
{
    free(o->mf);
    o->mf = NULL;
    if (!f)
        return;
    o->mf = malloc(sizeof(double) * 3 * (size_t)o->ny);
    memcpy(o->mf, f, sizeof(double) * 3 * (size_t)o->ny);
}
