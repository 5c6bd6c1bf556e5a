/*
 * simson_gpu.h -- C ABI of the MI355X Simson `bla` time step.
 *
 * The reference has no plugin or FFI layer: its step is the body of the
 * main loop of bla/bla.f:1790-2885, i.e. per RK substep one call of
 * nonlinbl per y plane (bla.f:2251-2288) followed by one call of linearbl
 * per z plane (bla.f:2433-2475), with prhs once at start-up (bla.f:1562).
 * These entry points replace exactly those calls; a host that keeps
 * bla's par.f/bla.i parameters, its .u file reader/writer and its RK
 * bookkeeping (bla.f:1465-1477, 2384-2397) calls them instead.
 *
 * Conventions
 *  - All host arrays are FP64 in the reference main-storage layout
 *    ur(memnx, memny, memnz, memnxyz) / ui(...) of bla.f:67, i.e. C arrays
 *    [slot][z][y][x] with x < nx/2, y < ny, z < nz.  Slots: 0-2 u,v,w;
 *    3-4 omega_x, omega_z; 5-6 partial right-hand sides p_v, p_omega.
 *  - puw is the k=0 partial right-hand side puw(ny, 2) of bla.f:131, a C
 *    array [2][ny].
 *  - Status: 0 on success; otherwise the reference stop code the reference
 *    would have raised (e.g. 324234 for CFL*dt > 10, wcflbl.f:97) or a
 *    negative SIMSON_E_* code for errors of this library.
 *  - The state lives on the GPU between calls.  No host pointer is kept
 *    after a call returns.  A handle is used from one host thread at a time.
 */
#ifndef SIMSON_GPU_H
#define SIMSON_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define SIMSON_E_ARG (-1)       /* invalid argument / unsupported setting */
#define SIMSON_E_HIP (-2)       /* HIP runtime error */
#define SIMSON_E_NODEV (-3)     /* no HIP device */

#define SIMSON_FP64 0           /* all arithmetic in FP64 */
/* Mixed precision: the x-z FFTs, the nonlinear products and the y
 * Chebyshev transforms run in FP32 (transforms on FP32 matrix cores with
 * FP32 accumulation) on FP32 copies of the fields; the Chebyshev-tau solves,
 * the kx = kz = 0 mean flow, the pressure-gradient / flux update and the
 * RK3-CN partial right-hand sides p_v, p_omega, puw stay FP64.  Uploads,
 * prhs and downloads are FP64 (downloaded velocities carry FP32 rounding).
 * Requires nzp = 3nz/2 and nxp/2 = 3nx/4 among the specialised FFT lengths
 * and ny among the matrix-core transform sizes (ny <= 127 with
 * ny = 1 mod 2, or ny = 8K + 1 up to 257 with the parity-split form). */
#define SIMSON_MIXED 1

typedef struct simson_params {
    int nx, ny, nz;      /* par.f nx, ny (= nyp, nfyd = 0), nz; nfxd = nfzd = 1 */
    int precision;       /* SIMSON_FP64 or SIMSON_MIXED */
    int ibc;             /* bla.i ibc: 0 (no-slip walls), 100 (free
                            stream u_i = U_i above) or 101 (du_i/dy =
                            dU_i/dy above), no-slip below; 100 and 101
                            need cflux = 0, see simson_gpu_free_stream */
    int cflux;           /* bla.i cflux (constant mass flux) */
    double re;           /* Reynolds number from the .u header */
    double xl, zl;       /* box lengths (bla.i xl, .u zl) */
    double rot;          /* bla.i rot */
    double retau;        /* bla.i retau (used when cflux == 0) */
} simson_params_t;

typedef struct simson_gpu simson_gpu_t;

/* Allocate the device state for one channel configuration.
 * Replaces the set-up of bla.f:1080-1082 (preprbl) and the pressure
 * gradient initialisation of bla.f:1622-1662. */
int simson_gpu_init(const simson_params_t *p, simson_gpu_t **out);

/* Upload nslots (3 or 7) slots of main storage, the k=0 partial RHS puw
 * (may be NULL when nslots == 3) and the pressure gradient px.  With
 * nslots == 3 the wall values u0low/u0upp/w0low/w0upp are taken from the
 * mean mode as rdiscbl.f:319-323 and bla.f:1041-1044 do. */
int simson_gpu_upload(simson_gpu_t *g, const double *ur, const double *ui,
                      int nslots, const double *puw, double px);

/* prhs (bla.f:1562, prhs.f:93-709): partial RHS, omega_x, omega_z and puw
 * from the velocities, for a first stage of weight an. */
int simson_gpu_prhs(simson_gpu_t *g, double an);

/* Step 2 of a substep alone: nonlinbl over all y planes (bla.f:2251-2288).
 * Slots 0-2 receive H_1..3 (times nxp*nzp, unnormalised as in the
 * reference).  When do_cfl is non-zero, *cfl_out receives the boxcfl
 * partial max (synchronises). */
int simson_gpu_nonlin(simson_gpu_t *g, int do_cfl, double *cfl_out);

/* The boxcfl maximum of the current velocity field without stepping it
 * (the CFL evaluation of getdt.f that sets the first step of a run with a
 * variable time step, bla.f:1407-1415): *cfl receives the maximum, so that
 * the CFL number of a step dt is (*cfl pi + 2 |rot|) dt as wcflbl.f:87
 * forms it.  A NaN in the field gives +infinity.  The state is unchanged.
 * Synchronises. */
int simson_gpu_cfl(simson_gpu_t *g, double *cfl);

/* Step 3 of a substep alone: linearbl over all z planes (bla.f:2433-2475)
 * with the weights an, bn, anp1, bnp1.  px_out (may be NULL) receives the
 * pressure gradient afterwards (synchronises). */
int simson_gpu_linear(simson_gpu_t *g, double an, double bn, double anp1,
                      double bnp1, double *px_out);

/* One RK substep: nonlinbl over all y planes (bla.f:2251-2288), then
 * linearbl over all z planes (bla.f:2433-2475) with the weights an, bn of
 * this substep and anp1, bnp1 of the next (bla.f:2384-2397).  When do_cfl
 * is non-zero, *cfl_out receives the partial CFL max of boxcfl.f (multiply
 * by pi and dt as wcflbl.f:87) and the call synchronises.  px_out (may be
 * NULL) receives the pressure gradient after the substep (synchronises). */
int simson_gpu_substep(simson_gpu_t *g, double an, double bn, double anp1,
                       double bnp1, int do_cfl, double *cfl_out, double *px_out);

/* nsteps full RK3 steps at fixed dt (bla.f:1465-1477 and the substep loop
 * 1790-2610 with vart = .false., nst = 3), prhs included when first != 0.
 * CFL*dt is evaluated every icfl substeps (icfl <= 0: never) and the run
 * returns 324234 if it exceeded 10 (checked when the run ends); *cfl_out
 * receives the last CFL*dt.  Launches are asynchronous until the end. */
int simson_gpu_run(simson_gpu_t *g, int nsteps, double dt, int icfl, int first,
                   double *cfl_out);

/* Download nslots slots of main storage, puw and px (any may be NULL). */
int simson_gpu_download(simson_gpu_t *g, double *ur, double *ui, int nslots,
                        double *puw, double *px);

/* Plane statistics of the amplitude file from the current velocity state:
 * the per-y sums of boxamp.f for every y plane (the call in nonlinbl.f:430,
 * which bla makes at the first substep of each iamp-th step), into amp, a
 * C array [20][ny] whose row i is amp(:, i+1) of boxamp.f; rows 13, 19 and 20
 * are zero.  The integration over y and the amp.stat line are the host's
 * (wampbl.f).  Synchronises. */
int simson_gpu_amp(simson_gpu_t *g, double *amp);

/* The same plane statistics without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-y sums go to a fixed scratch array of the library,
 * [20][257] doubles (41 KB) per device, instead of the handle's buffer;
 * ny above 257 returns SIMSON_E_ARG.  All calls on one GPU share that
 * array, so calls that overlap on one GPU (from several threads or
 * handles) race; it keeps the last call's contents.  Synchronises. */
int simson_gpu_amp_noalloc(simson_gpu_t *g, double *amp);

/* x-y statistics of boxxys.f (the per-plane call of bla.f:2089-2105, made
 * at the start of every ixys-th step from t = txys on) for the current
 * velocity state, accumulated into xys, a C array [nxys][ny][nx] whose
 * row i is xys(:, :, i+1) of the reference (x in physical order):
 *   1-3 u, v, w; 4-6 u^2, v^2, w^2; 7-9 om_x, om_y, om_z;
 *   10-12 their squares; 13-15 uv, uw, vw (boxxys.f:540-636),
 * each summed over the nz physical z points with weight dtn/nz, after the
 * lower-wall velocity u0low, w0low is subtracted from the mean mode and the
 * frame is shifted by (xs, zs) as in xzsh.f.  As in boxxys.f, indices 1-12
 * need nxys >= 12 and 13-15 need nxys = 15; 12 <= nxys <= 15 (the higher
 * statistics of the reference's list are not computed: -1).  The caller
 * keeps the sum of the weights (sumw) and divides by it when it writes the
 * file (wxys.f).  Synchronises. */
int simson_gpu_xys(simson_gpu_t *g, double dtn, double xs, double zs, int nxys, double *xys);

/* Pressure of the current state: nonlinp over all y planes and linearp over
 * all z planes (bla.f:2486-2527, nonlinp.f, linearp.f with the Neumann tau
 * solver of setmac.f/crhsc.f/tridc.f), i.e. what bla puts in slot 8 and
 * wdiscp.f writes to <name>.p.  pr, pi receive p at physical y as C arrays
 * [z][y][x]; the z-oddball plane holds the transform of u.u there, as in
 * the reference.  The wall term -(v - v_prev)/dt of linearp.f is zero for the
 * no-slip walls.  Needs ibc = 0 and nzp and nxp/2 among the specialised FFT
 * lengths.  Allocates its work buffers at the first call; synchronises. */
int simson_gpu_pressure(simson_gpu_t *g, double *pr, double *pi);

/* Large-eddy models of bla when bla.i sets sgs (sgs.i, rparamles.f:62-81):
 * - iles 1, ADM-RT: the relaxation term of nonlinbl.f:920-966 with the
 *   high-pass filtered velocity of linearbl.f:1871-1962 (ihighorder = 2,
 *   the filters of init_filter.f and bla.f:1494-1518).  cutoff =
 *   1/cutoff_inv of sgs.i (0 < cutoff <= 1), 0 <= iord <= 5, and chi is the
 *   relaxation rate chi/dstar of bla.f:2220 (dstar = 1 for the channel).
 * - iles 2, Smagorinsky eddy viscosity: the stresses of les.f (called per
 *   plane from bla.f:1945-1998 at the start of every substep) with a fixed
 *   coefficient cs, or with cs = 0 the dynamic procedure (Germano/Lilly,
 *   2-D spectral cutoff test filter, averaging in z) clipped by ineg (0:
 *   none; 1: C >= 0; 2: nu + nu_t >= 0; rparamles.f:36-39), and their
 *   divergence sgs_stress.f:62-131 subtracted from H (nonlinbl.f:1011-1050).
 *   The reference compiles it in only with iiles_smag = 1 (bla.f:28).
 *   Computed in FP64 in both precision forms.
 * - iles 3, the Smagorinsky model on the high-pass filtered velocity
 *   (les.f:180-215, 404-620, 883-1380): as iles 2 with the strain of the
 *   velocities and vorticities filtered as for ADM-RT (cutoff, iord) and,
 *   with cs = 0, B_ij of the dynamic procedure from the secondary filter
 *   of linearbl.f:2067-2256.  FP64 in both precision forms.
 * iles 0 turns the model off.  Call after init, before the first step;
 * needs nzp among the specialised z transform lengths. */
typedef struct simson_les {
    int iles;            /* 0: no model; 1: ADM-RT; 2: Smagorinsky; 3: HPF Smagorinsky */
    int iord;            /* sgs.i iord (iles 1, 3) */
    double cutoff;       /* 1 / sgs.i cutoff_inv (iles 1, 3) */
    double chi;          /* sgs.i chi divided by dstar (iles 1) */
    double cs;           /* sgs.i cs (iles 2, 3; 0: dynamic procedure) */
    int ineg;            /* sgs.i ineg (iles 2, 3) */
} simson_les_t;

int simson_gpu_les(simson_gpu_t *g, const simson_les_t *p);

/* Fringe region of spatial simulations (bla.i spat, rparambl.f:547-720):
 * the forcing -lambda(x) (u_i - U_i(x, y)) toward the base flow U that
 * fring.f adds to the nonlinear term of every nonlinbl call
 * (nonlinbl.f:678-693) and fringp.f to that of the pressure
 * (nonlinp.f:172-174), with the fringe function lambda of ffun.f (no
 * forced waves, OS modes or streaks).  Lengths in the channel scaling of
 * bla.i; xsc is the frame position xs - cnp1 u0low of bla.f:1819 (0 for a
 * fixed frame); fstart < fend within -xl/2 .. xl/2, fmax, frise, ffall > 0
 * (rparambl.f:1898-1913). */
typedef struct simson_fringe {
    double fmax;         /* bla.i fmax */
    double fstart;       /* bla.i fstart */
    double fend;         /* bla.i fend */
    double frise;        /* bla.i frise */
    double ffall;        /* bla.i ffall */
    double xsc;          /* frame position */
} simson_fringe_t;

/* Switch the fringe on with base, a C array [3][ny][3nx/2] of the base-flow
 * velocities u, v, w at the physical points x_j = -xl/2 + j xl/nxp of the
 * 3/2 x grid (bu1/bu2 of bflow.f interleaved: even j from bu1, odd from
 * bu2), or off with f = NULL.  For the spatial Poiseuille flow of fltype 4
 * the base flow is u = 1 - y^2 (cbflow.f:117-131); the pressure gradient is
 * zero (init with cflux = 0, retau = 0; bla.f:1657-1662).  Needs nxp/2
 * among the specialised x-row lengths.  Synchronises. */
int simson_gpu_fringe(simson_gpu_t *g, const simson_fringe_t *f, const double *base);

/* Blowing and suction at the lower wall, bla.i wbci = 1 (rparambl.f:
 * 1206-1232): the vertical velocity v(x, z, y = -1) = max(s P, s P damp),
 * P = amp f(x) cos(zbet z), s = sin(tomeg tc) (1 when tomeg = 0), f the
 * smooth strip of cwallbc.f:53-115 between xstart and xend, imposed on the
 * v solve of every linear step (linearbl.f:458-469) at the time tc of the
 * start of the substep (bla.f:1797, 2405), with the mean v set to the
 * mean wall velocity (linearbl.f:1663-1670).  Coordinates in the channel
 * scaling, x and z measured from the box centre (xsc, zsc the frame
 * position, 0 for a fixed frame); the CFL estimate then keeps the full 1/dy
 * near the walls (boxcfl.f:38-43). */
typedef struct simson_wall {
    double amp;          /* wp1 */
    double damp;         /* wp2: factor on the suction part */
    double xstart;       /* wpds1 */
    double xend;         /* wpds2 */
    double xrise;        /* wpds3 */
    double xfall;        /* wpds4 */
    double zbet;         /* wpdds1: spanwise wavenumber */
    double tomeg;        /* wpdds2: angular frequency */
    double xsc, zsc;     /* frame position */
} simson_wall_t;

/* Switch the wall forcing on (w) or off (NULL).  Synchronises. */
int simson_gpu_wall(simson_gpu_t *g, const simson_wall_t *w);

/* The spectral wall velocity the library imposes at time tc, vr, vi
 * [nz][nx/2] (cwallbc.f:53-115, 380-407), without a device.  Returns 0 or
 * SIMSON_E_ARG. */
int simson_wall_velocity(int nx, int nz, double xl, double zl, const simson_wall_t *w,
                         double tc, double *vr, double *vi);

/* Falkner-Skan-Cooke similarity table (bls/fsc.f; the fsc.dat that rbla.f
 * reads): on eta_i = ymax i/(n-1), i < n, the rows
 *   tab[8 i ..] = eta, f, f', f'', f''', g, g', g''
 * of f''' = -f f'' - beta (1 - f'^2), beta = 2m/(m+1), f(0) = f'(0) = 0,
 * f'(ymax) = 1 (shooting with Newton's method to |f'(ymax) - 1| <= eps, one
 * RK4 step per interval as fsc.f), and g'' = -f g', g(0) = 0, g(ymax) = 1.
 * dstar2 = ymax - f(ymax), the displacement thickness in eta.  Host only;
 * returns 0 or SIMSON_E_ARG (bad input, m < -0.0904 or no convergence). */
int simson_fsc_table(double m, int n, double ymax, double eps, double *tab, double *dstar2);

/* Boundary-layer base flows of cbflow.f: spatial (161-300; fltype 6
 * Blasius, 7 Falkner-Skan with m = rlam, 8 Falkner-Skan-Cooke with spanv)
 * and temporal (58-84; fltype 3, -1, -2: the profile at x0 at every x),
 * lengths in the units of the .u file (re and xl as stored there, dstar =
 * 2/yl). */
typedef struct simson_baseflow {
    int nx, ny;          /* par.f nx, ny */
    int fltype;          /* 6, 7, 8 (spatial) or 3, -1, -2 (temporal) */
    double xl, re;       /* box length and Reynolds number of the .u file */
    double dstar;        /* .u dstar: displacement thickness at the box
                            centre in the units of y + 1 */
    double dstar2;       /* the table's displacement thickness in eta */
    double bstart, blength;   /* .u blending zone */
    double rlam, spanv;  /* m and the spanwise scale (fltype 7, 8) */
    double xsc;          /* frame position (0 for a fixed box) */
    double u0low, w0low; /* lower-wall velocities */
} simson_baseflow_t;

/* The base flow base [3][ny][3nx/2] (u, v, w at x_j = -xl/2 + j xl/nxp +
 * xsc and eta_y = cos(pi y/(ny-1)): the layout simson_gpu_fringe takes),
 * from the similarity table tab [ntab][8] of simson_fsc_table (or fsc.dat);
 * v from continuity with v(-1) = 0.  With bjr/bji [3][nx/2] and du0upp not
 * NULL, also the data of simson_gpu_free_stream: dU_i/dy at y = +1 as x
 * Fourier coefficients / nxp (bflow.f:80-105) and their kx = 0 streamwise
 * value.  Host only; returns 0 or SIMSON_E_ARG. */
int simson_base_flow(const simson_baseflow_t *b, const double *tab, int ntab, double *base,
                     double *bjr, double *bji, double *du0upp);

/* The forcing that keeps a temporal boundary layer's parallel profile from
 * diffusing (nonlinbl.f:771-795; fltype 3, -1, -2): force [3][ny] at
 * eta_y = cos(pi y/(ny-1)), the x, y, z components of the mean force at the
 * boundary layer's position x0 + cdev t (bla.f:1826), from the similarity
 * table tab [ntab][8].  Host only; returns 0 or SIMSON_E_ARG. */
int simson_parallel_force(const simson_baseflow_t *b, const double *tab, int ntab, double cdev,
                          double t, double *force);

/* A mean force f [3][ny] (physical y, e.g. from simson_parallel_force)
 * added to the kx = kz = 0 mode of the nonlinear term H in every following
 * substep, as nonlinbl.f:771-795 adds it (the y component does not enter:
 * the mean v is the wall's); NULL switches it off.  A time-dependent force
 * (cdev != 0) is the caller's to update between substeps.  Synchronises. */
int simson_gpu_mean_force(simson_gpu_t *g, const double *force);

/* Localized volume force of bla.i loctyp = 3 (locf.f:358-410), in the
 * units of the .u file (lengths times dstar, amplitudes and the frequency
 * over dstar, as rescale.f scales bla.i's fpdds/fpds):
 *   F_i = amp_i exp(-((1 + y)/yscale)^2) f(x) sin(tomeg tc),
 *   f(x) = ((x - xloc0)/xscale)^xtype exp(-((x - xloc0)/xscale)^2),
 * x from the box centre (plus xsc), uniform in z, added to the nonlinear
 * term at the time tc of the start of each substep. */
typedef struct simson_locf {
    double ampx, ampy, ampz;   /* fpdds1..3 */
    double xscale, xloc0;      /* fpds1, fpds2 */
    double yscale;             /* fpds3 */
    double xtype;              /* fp1: 0 or 1 */
    double tomeg;              /* fpdds4 */
    double xsc;                /* frame position */
} simson_locf_t;

/* Switch the localized force on (lf) or off (NULL); it is applied to every
 * later substep with the time of simson_gpu_set_time / simson_gpu_run.
 * Synchronises. */
int simson_gpu_local_force(simson_gpu_t *g, const simson_locf_t *lf);

/* The force's spatial pattern without a device: pr/pi [3][ny][nx/2], the
 * x Fourier coefficients (over the 3/2 grid, divided by nxp) of the kz = 0
 * plane at eta_y = cos(pi y/(ny-1)), before the factor sin(tomeg tc).
 * Returns 0 or SIMSON_E_ARG. */
int simson_local_force_pattern(int nx, int ny, double xl, const simson_locf_t *lf, double *pr,
                               double *pi);

/* Free-stream data of ibc = 101 or 100 (replaces evalbc.f:142-159's use of
 * bflow.f's bu1jr and du0upp): bjr/bji [3][nx/2], component-major, the
 * base flow's dU/dy, dV/dy, dW/dy at y = +1 per kx (Fourier coefficients
 * normalised by 1/nxp, bflow.f:80-105), imposed on the kz = 0 plane, or both
 * NULL for zero; du0upp the kx = 0 value of dU/dy there, the mean mode's
 * condition (linearbl.f:1516-1531).  With ibc = 101 the mean mode's wall
 * values are U = 1 above (then replaced by du0upp) and 0 below, as bla.f:
 * 1360-1365 sets them for every base flow.  Until called, b = 0 and
 * du0upp = 0.  With ibc = 100 bjr/bji hold instead the base flow's values
 * U, V, W at y = +1 (bu1jr(:, i, 1)) and du0upp is not used: u_i(+1) =
 * U_i(+1) on the kz = 0 plane, and the mean mode keeps U = 1 above.
 * Returns SIMSON_E_ARG unless init had ibc = 100 or 101.  Synchronises. */
int simson_gpu_free_stream(simson_gpu_t *g, const double *bjr, const double *bji, double du0upp);

/* The time t of the start of the next substep, the tc of bla.f:1797 that
 * time-dependent boundary forcing uses (0 after init).  simson_gpu_run
 * advances it by c_k dt per substep and leaves it at the start of the next
 * step; a host that calls simson_gpu_linear itself sets it before each
 * call. */
int simson_gpu_set_time(simson_gpu_t *g, double t);

/* The fringe function lambda at the 3nx/2 physical points of the 3/2 x
 * grid (ffun.f:37-75), without a device.  Returns 0 or SIMSON_E_ARG. */
int simson_fringe_function(int nx, double xl, const simson_fringe_t *f, double *lam);

/* The ADM-RT filter operators of one grid, without a device: lpf [nz][nx/2]
 * is the x-z transfer function lpfxz of bla.f:1503-1518 and L [ny][ny] the
 * wall-normal low-pass I - H_y^(iord+1), H_y the compact high-pass filter of
 * hpf_y.f on the Gauss-Lobatto points.  Returns 0 or SIMSON_E_ARG. */
int simson_les_filter(int nx, int ny, int nz, double cutoff, int iord, double *lpf, double *L);

/* The secondary x-z transfer function lpfxz2 [nz][nx/2] of the dynamic HPF
 * model (bla.f:1520-1543: the homogeneous filters of init_filter.f sampled
 * on nx/4 + 1 and nz/4 + 1 points, zero above), without a device.  Returns
 * 0 or SIMSON_E_ARG. */
int simson_les_filter_secondary(int nx, int nz, double cutoff, int iord, double *lpf2);

/* ---- Learned subgrid closure (no counterpart in the reference: it is the
 * learned replacement of the tau_ij model of les.f / sgs_stress.f:62-131) ----
 * A multilayer perceptron 10 -> 64 -> 64 -> 6 (GELU) maps the filtered
 * velocity-gradient tensor d ubar_i/dx_j (9 values) and y+ to the six
 * subgrid stresses tau_11, 12, 13, 22, 23, 33 (inputs and outputs
 * standardised by the caller).  Training and inference run on the BF16
 * matrix cores with FP32 accumulation; parameters and Adam state are FP32
 * on the device.  Loss: mean squared error over points and outputs. */
typedef struct simson_mlp simson_mlp_t;

/* New network with seeded Glorot-uniform weights and zero biases. */
int simson_mlp_create(unsigned long long seed, simson_mlp_t **out);
/* Number of FP32 parameters (layout: W1 [16][64] (rows 10-15 zero), b1 [64],
 * W2 [64][64], b2 [64], W3 [64][16] (columns 6-15 zero), b3 [16]). */
int simson_mlp_nparams(void);
/* Copy the parameters out (get) and/or in (set, which also resets Adam). */
int simson_mlp_params(simson_mlp_t *h, float *get, const float *set);
/* Keep a training set on the device: X [n][10], T [n][6] (rows shuffled by
 * the caller; batches are consecutive tiles of 16 rows). */
int simson_mlp_set_data(simson_mlp_t *h, const float *X, const float *T, int n);
/* steps Adam steps (lr, beta 0.9 / 0.999) on batches of `batch` rows
 * (a multiple of 16), cycling through the training set; losses (may be
 * NULL) receives each step's batch loss.  Synchronises. */
int simson_mlp_train(simson_mlp_t *h, int steps, int batch, float lr, float *losses);
/* Gradient of the loss over the n rows of X [n][10], T [n][6] (host
 * arrays) with respect to the parameters, and the loss; no update. */
int simson_mlp_grad(simson_mlp_t *h, const float *X, const float *T, int n, float *grad,
                    float *loss);
/* Y [n][6] = network(X [n][10]) (host arrays). */
int simson_mlp_infer(simson_mlp_t *h, const float *X, int n, float *Y);
const char *simson_mlp_last_error(const simson_mlp_t *h);
void simson_mlp_destroy(simson_mlp_t *h);

/* Use a trained closure network inside the step, in place of the subgrid
 * model of les.f: at every substep the nine filtered velocity gradients
 * d u_i/dx_j (from u, v, w and the vorticities) and y+ of every point of the
 * dealiased x-z grid go through the network, and the divergence of its
 * stresses, sigma_i = d tau_ij/dx_j (sgs_stress.f:62-131), is subtracted
 * from H_i as nonlinbl.f:1011-1050 subtracts the eddy-viscosity force.
 * params: the network (simson_mlp_params layout), NULL turns the closure off.
 * Features: x_k = g_k in_scale[k] + in_shift[k] for the gradients (k < 9),
 * yfeat[y] for y+ (already standardised); stresses tau_o = net_o
 * out_scale[o] + out_shift[o] (o: 11 12 13 22 23 33).  Turns ADM-RT off.
 * Needs the specialised FFT lengths and ny <= 128. */
int simson_gpu_closure(simson_gpu_t *g, const float *params, const double *in_scale,
                       const double *in_shift, const double *yfeat, const double *out_scale,
                       const double *out_shift);

/* Wait for all queued work. */
int simson_gpu_sync(simson_gpu_t *g);

/* Per-kernel timing with HIP events on the library's stream.  While
 * enabled, every launch is bracketed by events; simson_gpu_timing_read
 * fills up to cap entries of per-kernel-family totals (ms) and launch
 * counts and returns the number of families.  Names are static strings. */
int simson_gpu_timing_enable(simson_gpu_t *g, int on);
int simson_gpu_timing_read(simson_gpu_t *g, const char **names, double *total_ms,
                           long *count, int cap);

/* Last HIP error text of this handle (static storage, never NULL). */
const char *simson_gpu_last_error(const simson_gpu_t *g);

void simson_gpu_finalize(simson_gpu_t *g);

#ifdef __cplusplus
}
#endif
#endif
