/*
 * blaref -- CPU restatement (FP64) of one Simson `bla` RK3-Crank-Nicolson
 * substep for the temporal plane channel (fltype 1, ibc 0, tau method).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle the HIP library is checked
 * against; only tests, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product path never calls it.
 *
 * Main storage follows the reference (bla/bla.f:67, bla/par.f:111-120):
 * ur, ui are double arrays [slot][z][y][x] (Fortran ur(x,y,z,slot)) with
 * x < nx/2, y < ny (Gauss-Lobatto points, y=0 is the upper wall), z < nz and
 * slots 0-2 velocities u,v,w; 3-4 vorticities omega_x, omega_z; 5-6 the
 * partial right-hand sides p_v, p_omega (Chebyshev space).
 */
#ifndef BLAREF_H
#define BLAREF_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc_t;

/* nx, nz even; ny odd with ny-1 factoring into 2,3,5.  px/mflux follow
 * bla.f:1622-1662 (cflux: mflux=4/3, px=-2/re; else px=-(retau/re)^2). */
orc_t *orc_create(int nx, int ny, int nz, double re, double xl, double zl,
                  double rot, int cflux, double retau);
void orc_destroy(orc_t *o);
double orc_get_px(const orc_t *o);
void orc_set_px(orc_t *o, double px);

/* prhs (bla/prhs.f:93-709): partial RHS for the first stage with weight an,
 * vorticities omega_x, omega_z, and the k=0 partial RHS puw[2][ny]. */
void orc_prhs(orc_t *o, double *ur, double *ui, double *puw, double an);

/* Step 2 (bla.f:2251-2288, nonlinbl.f): nonlinear terms H_1..3 into slots
 * 0-2 for every y plane.  Returns the CFL partial max (boxcfl.f) over all
 * planes when do_cfl is non-zero, else 0. */
double orc_nonlin(orc_t *o, double *ur, double *ui, int do_cfl);

/* Step 3 (bla.f:2433-2475, linearbl.f): advance all z planes except the
 * z-oddball one; updates px under constant flux. */
void orc_linear(orc_t *o, double *ur, double *ui, double *puw,
                double an, double bn, double anp1, double bnp1);

/* Whole driver: nsteps full RK3 steps with fixed dt starting from a field
 * whose slots 0-2 hold velocities (prhs is called first).  Mirrors the RK
 * weight bookkeeping of bla.f:1465-1477,2384-2397.  cfl_out (optional)
 * receives CFL*dt (wcflbl.f:87) of the last evaluation; icfl as in bla.i
 * after rparambl's adjustment (<=0 disables). Returns 0 or 324234 if
 * CFL*dt exceeded 10. */
int orc_run(orc_t *o, double *ur, double *ui, double *puw, int nsteps,
            double dt, int icfl, double *cfl_out);

/* Pressure of the current state (bla.f:2486-2527: nonlinp.f over all y
 * planes, then linearp.f over all z planes), the slot-8 content that
 * wdiscp.f writes: pr, pi [z][y][x] at physical y.  Slots 0-4 of ur/ui must
 * hold u, v, w, omega_x, omega_z. */
void orc_pressure(orc_t *o, const double *ur, const double *ui, double *pr, double *pi);

/* ADM-RT LES (sgs.i iles = 1, ihighorder = 2): the relaxation term of
 * nonlinbl.f:920-966 with the filter operators lpf [nz][nx/2] (x-z transfer
 * function, bla.f:1503-1518) and L = I - H_y^(iord+1) [ny][ny] (wall-normal
 * low-pass, hpf_y.f) given by the caller, and chi/dstar (dstar = 1);
 * lpf = NULL turns it off.  Returns 0. */
int orc_set_les(orc_t *o, const double *lpf, const double *L, double chi);

/* ---- building blocks exposed for unit tests ---- */
/* External SGS force sigma [3][nz][ny][nxh] (physical y) subtracted from H
 * (times nxp nzp, not on the z-oddball row) in later orc_nonlin calls, as
 * nonlinbl.f:1011-1050 does for the eddy-viscosity models; NULL: off. */
void orc_set_force(orc_t *o, const double *fr, const double *fi);

/* Fringe forcing (fring.f, fringp.f): lam [nxp] at the physical points of
 * the 3/2 x grid, base flow [3][ny][nxp]; NULL: off. */
void orc_set_fringe(orc_t *o, const double *lam, const double *base);

/* Lower-wall v_w [nz][nxh] (spectral, wbci = 1) for the v solves of later
 * orc_linear calls; NULL: off. */
void orc_set_wall(orc_t *o, const double *vr, const double *vi);

/* Boundary conditions at y = +1 (evalbc.f): ibc 0 (no-slip, the default) or
 * 101 (free stream: du_i/dy = dU_i/dy of the base flow; the lower wall stays
 * no-slip).  bjr/bji [3][nxh]: dU_i/dy at y = +1 per kx of the kz = 0 plane
 * (bflow.f's bu1jr(:, i, 2)), NULL for zero; du0upp its kx = 0 streamwise
 * part for the mean flow.  Returns 0, or -1 for another ibc. */
int orc_set_bc(orc_t *o, int ibc, const double *bjr, const double *bji, double du0upp);

/* Mean force f [3][ny] (physical y) added to the kx = kz = 0 mode of H_1
 * and H_3 in every nonlinear step, as nonlinbl.f:771-795 adds the
 * parallel-flow forcing of temporal boundary layers; NULL: off. */
void orc_set_mean_force(orc_t *o, const double *f);

/* complex DFT, y_k = sum_j x_j exp(sign 2 pi i jk/n), in place on (re,im) */
void orc_cdft(int n, double *re, double *im, int sign);
/* vrfftb convention: half spectrum X[0..n/2] (re,im) -> x[0..n-1] real */
void orc_rdft_bwd(int n, const double *xr, const double *xi, double *x);
/* vrfftf convention: x[0..n-1] -> X[0..n/2] */
void orc_rdft_fwd(int n, const double *x, double *xr, double *xi);
/* vchbf / vchbb on one sequence of length n (unnormalised) */
void orc_chbf(int n, const double *x, double *out);
void orc_chbb(int n, const double *x, double *out);
/* dcheb / icheb on one sequence of n coefficients */
void orc_dcheb(int n, const double *a, double *b);
void orc_icheb(int n, const double *a, double *b, double intcon);
/* Chebyshev tau solve of u'' - lam u = f with Sum_even u = bc_e,
 * Sum_odd u = bc_o (setmacrhs.f + trid.f tau branch). */
void orc_tau_solve(int n, double lam, const double *f, double bc_e,
                   double bc_o, double *u);
/* the same with the Neumann rows Sum_even k^2 u_k = bc_e, Sum_odd k^2 u_k =
 * bc_o (setmac.f neumann = 1, tridc.f) */
void orc_tau_solve_neumann(int n, double lam, const double *f, double bc_e,
                           double bc_o, double *u);

#ifdef __cplusplus
}
#endif
#endif
