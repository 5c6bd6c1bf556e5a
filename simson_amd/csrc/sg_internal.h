// Internal declarations of the MI355X Simson step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cmath>
#include <vector>

#include "simson_gpu.h"

namespace sg {

// The smooth step of step.f: 0 for x <= 0.02, 1 for x > 0.98, and
// 1 / (1 + exp(1/(x - 1) + 1/x)) between (host side).
inline double smooth_step(double x)
{
    if (x <= 0.02)
        return 0.;
    if (x <= 0.98)
        return 1. / (1. + std::exp(1. / (x - 1.) + 1. / x));
    return 1.;
}

constexpr int kMaxRad = 16;
constexpr int kNSlot = 7;          // host-visible slots: u v w om_x om_z p_v p_om
constexpr int kNPhys = 5;          // slots kept at physical y on the device
constexpr int kScratch = 16;       // per-column scratch arrays of the start-up kernels

// Maximum for the CFL reductions with NaN taken as +infinity: a field that
// has blown up must fail the CFL check, and fmax alone drops a NaN operand.
// Once a partial maximum is +infinity, fmax keeps it.
template <typename T>
__device__ __forceinline__ T cfl_max(T m, T v)
{
    return v != v ? (T)__builtin_inf() : fmax(m, v);
}

// Radix schedule and twiddle table of one FFT length.
struct FftPlan {
    int n = 0;
    int nrad = 0;
    int rad[kMaxRad] = {};
    const double2 *tw = nullptr;   // exp(+2 pi i t / tabn), t < tabn
    int tabn = 0;
};

// Which planes a batched Chebyshev GEMM reads and writes.
// Physical planes ("row" layout, main storage) are [slot][z][y][x] complex;
// Chebyshev planes ("column" layout) are [slot][z][x][k] complex, so that
// each (kx,kz) column is contiguous for the per-column solver.
struct GemmIO {
    const double *in;
    double *out;
    int in_slot[8];
    int out_slot[8];
    int nf;          // number of fields
    int nz;          // planes per field
    int ny;          // Chebyshev length
    int nxh;         // complex columns per plane
    int in_col;      // input in column layout
    int out_col;     // output in column layout
    int skip_z;      // plane to skip (z-oddball) or -1
};

// Factorisation of the tau Helmholtz operator u'' - lam u for one lam per
// column, [z][x][k]: gamma_k (bottom-up elimination multiplier) and
// 1/pivot_k; for the lam^2 = k^2 + 2 Re/(a+b) stages also the homogeneous
// wall-normal velocity v_h and its parity sums sum_k k^2 v_h,k.
struct LamFactor {
    double2 *gp = nullptr;     // (gamma, 1/pivot)
    double *vh = nullptr;      // v_h (lam^2 stages only)
    double2 *sh = nullptr;     // [z][x] per-parity sums (lam^2 stages only)
    // free stream above (ibc = 101; lam^2 stages only, else null): the
    // columns (phi_h, v_h2) and per column (A1, A2, B, 0) = d2v_h1/dy2,
    // d2v_h2/dy2 and dphi_h/dy at y = +1
    double2 *hb = nullptr;
    double *hc = nullptr;
};

// Arguments of the per-(kx,kz)-column kernels.
struct ColArgs {
    double2 *U;           // physical slots [5][nz][ny][nxh]
    double2 *C;           // Chebyshev columns [5][nz][nxh][ny]
    float2 *C32;          // the same in FP32 (mixed form)
    double2 *Pp;          // partial RHS columns [2][nz][nxh][ny]
    double *W;            // start-up scratch [kScratch][ny][nz*nxh]
    double *puw;          // [2][ny]
    double *px;           // device scalar
    const double *alfa;   // [nxh]
    const double *beta;   // [nz]
    int ny, nz, nxh;
    double re;
    double mflux;
    int cflux;
    double u0upp, u0low, w0upp, w0low;
    // imposed lower-wall v (wbci = 1): per column v(-1) = vscale vwall[t]
    // (vwall [nz][nxh] or null), and the mean v = v0 (linearbl.f:458-469,
    // 1663-1670)
    const double2 *vwall;
    double vscale, v0;
    // ibc = 101: du_i/dy at y = +1 follows the base flow, bj [3][nxh] its
    // dU/dy, dV/dy, dW/dy there per kx of the kz = 0 plane (null: zero),
    // du0upp the mean streamwise value (evalbc.f:142-159, linearbl.f:1500-1531)
    int ibc;
    const double2 *bj;
    double du0upp;
    // mean forcing of temporal boundary layers (nonlinbl.f:771-795) as the
    // Chebyshev coefficients [2][ny] it adds to H1 and H3 of the kx = kz = 0
    // column, or null
    const double *mforce;
    // localized volume force (loctyp 3): Chebyshev coefficients [3][nxh][ny]
    // of the kz = 0 plane, added to H_i times lscale, or null
    const double2 *lforce;
    double lscale;
};

// Step-2 launch arguments.
// V is double2 (FP64 step) or float2 (FP32 step 2 of the mixed form).
template <typename V>
struct Step2Args {
    const V *U;           // physical slots [5][nz][ny][nxh]
    V *Uw;                // H_1..3 into slots 0-2 of the same layout
    V *P;                 // [6][ny][nzp][nxh]
    const double *alfa, *beta, *deta;
    double *cfl_blk;
    int nx, ny, nz, nxh, nxp, nzp;
    double rot;
    double xl, zl;
    int do_cfl;
    // split_mean: the kx = kz = 0 mode (the mean profiles U, W, Om_x, Om_z
    // of each y plane) stays out of the transformed fields and enters the
    // products as exact per-plane coefficients, so that FP32 fields carry
    // the fluctuations at full FP32 relative precision (mixed form)
    int split_mean;
    // ADM-RT relaxation (nonlinbl.f:920-966), applied by the z-analysis
    // kernel: H_i -= chi_n (u_i - lpf(x, z) Wf_i) on every storage row but the
    // z oddball, with u_i the velocity still in Uw slot i, Wf_i = L_y u_i
    // ([3][nz][ny][nxh]) and chi_n = chi nxp nzp
    int les;
    const V *Wf;
    const double *lpf;
    double chi_n;
    // learned closure (sgs_stress.f:62-131 with the network's stresses):
    // H_i -= sigma_i = i alpha tau_i1 + d tau_i2/dy + i beta tau_i3, from the
    // unnormalised stress spectra TAU [6][nz][ny][nxh] (11 12 13 22 23 33)
    // and DT = d/dy of tau_12, tau_22, tau_23 [3][nz][ny][nxh]
    int clos;
    const V *TAU, *DT;
    // fringe forcing of spatial simulations (fring.f, fringp.f): H_i -=
    // lambda(x) (u_i - U_i(x, y)) at the physical points of the 3/2 x grid,
    // packed as the x rows (even points in .x, odd in .y): flam [nxp/2],
    // fbase [3][ny][nxp/2]; applied by the four-step x-row kernels
    int fringe;
    const V *flam, *fbase;
    // no wall cap on the CFL's 1/dy (boxcfl.f:38-43 cap it only when no
    // wall velocity is imposed)
    int nocap;
};
typedef Step2Args<double2> NonlinArgs;

// Learned closure evaluated on the x rows (sg_xrow.hip): network parameters
// (FP32, the simson_mlp layout with the offsets below), the affine maps of the
// nine gradient features (x_n = g in_a + in_b), the y+ feature per y plane
// (standardised), and the map of the six outputs (tau = y_n out_a + out_b)
struct ClosArgs {
    const float *W;
    int ow2, ow3, ob1, ob2, ob3;
    float in_a[9], in_b[9];
    const float *yfeat;
    float out_a[6], out_b[6];
};
template <typename V>
bool launch_xclos(const Step2Args<V> &a, const V *wx, const ClosArgs &c, hipStream_t s);
// the nine velocity-gradient fields d u_i / d x_j of a y plane, padded and
// synthesised in z, into a.P as [9][ny][nzp][nx/2]
template <typename V>
bool launch_gsyn4(const Step2Args<V> &a, const V *tw, hipStream_t s);

// spectra [nz][nx/2] of the wbci = 1 wall velocity patterns max(P, P damp)
// and min(P, P damp), P = amp f(x) cos(zbet z) on the 3/2 grid (host)
int wall_spectra(int nx, int nz, double xl, double zl, const simson_wall_t *w, double2 *pmax,
                 double2 *pmin);

// fringe function of ffun.f at the nxp physical points of the 3/2 x grid
// (host); 0 or SIMSON_E_ARG
int fringe_function(int nx, double xl, const simson_fringe_t *f, double *lam);

// ADM-RT filter operators (host): lpf [nz][nx/2] and L = I - H_y^(iord+1)
// [ny][ny]; 0 or SIMSON_E_ARG
int les_filter(int nx, int ny, int nz, double cutoff, int iord, double *lpf, double *L);
// the secondary x-z transfer function of the dynamic HPF model [nz][nx/2]
int les_filter_secondary(int nx, int nz, double cutoff, int iord, double *lpf2);

void launch_zsyn(const NonlinArgs &a, const FftPlan &pz, hipStream_t s);
void launch_xrow(const NonlinArgs &a, const FftPlan &pm, const double2 *wx, hipStream_t s);
void launch_zana(const NonlinArgs &a, const FftPlan &pz, hipStream_t s);
int  xrow_blocks(const NonlinArgs &a);
// four-step x-row kernel (register DFTs, in-place LDS passes); false when
// nxp/2 has no specialisation.  Blocks written to cfl_blk: xrow4_blocks.
template <typename V>
bool launch_xrow4(const Step2Args<V> &a, const V *wx, hipStream_t s);
template <typename V>
int xrow4_blocks(const Step2Args<V> &a);
// four-step z kernels; false when nzp has no specialisation
template <typename V>
bool launch_zsyn4(const Step2Args<V> &a, const V *tw, hipStream_t s);
template <typename V>
bool launch_zana4(const Step2Args<V> &a, const V *tw, hipStream_t s, int nf = 3);
// the pressure variant of the x-row kernel (H and W = u.u, nonlinp.f)
bool launch_xrow4_press(const Step2Args<double2> &a, const double2 *wx, hipStream_t s);
void launch_cfl_reduce(const double *blk, int n, double *out, hipStream_t s);
// linearp.f per (kx,kz) column (pressure), see sg_press.hip
int  pressure_max_ny();
void launch_pressure_cols(const double2 *Qc, const double2 *gk, const double *alfa,
                          const double *beta, int nz, int nxh, int ny, double re,
                          double nxpnzp, double *scratch, double2 *out, hipStream_t s);
// x-y statistics 1-15 of boxxys.f for the current field, times c1 = dtn/nz,
// into out [15][ny][nx]; Q [6][ny][nz][nxh] is work space
struct XysArgs {
    const double2 *U;
    double2 *Q;
    double *out;
    const double *alfa, *beta;
    int ny, nz, nxh;
    double c1, u0low, w0low, xs, zs;
    int shift;
};
// eddy-viscosity LES (sg_smag.hip): stresses into TAU [6][nz][ny][nxh],
// scaled by nxp nzp like the learned closure's; Q [12][ny][nz][nxh], SS
// [7][ny][nz][nxh] and (dynamic) W [21][ny][nz][nx/4] are work space
struct SmagArgs {
    const double2 *U;
    double2 *Q, *SS, *W, *TAU;
    double *part;                 // [ny][nzb][2][nx] partial z sums (dynamic)
    const double *alfa, *beta, *dlt2;
    // HPF model (iles 3): LW = L u of slots 0-4 [5][nz][ny][nxh] (the ADM-RT
    // wall-normal low-pass), and the primary and secondary x-z transfer
    // functions [nz][nxh]; the strain comes from u - lpf LW, the dynamic
    // procedure's B_ij from u - lpf2 LW
    const double2 *LW;
    const double *lpf, *lpf2;
    int hpf;
    int nqf;                      // fields of Q: 12, or 21 with the secondary gradients
    int nx, ny, nz, nxh, nq;
    int dynamic, ineg, rt;
    double u0low, w0low, cs2, tw2, re, hscale;
};
bool smag_fits(int nz, int nxh, size_t lds_max);
int smag_zblocks(int nz, int nxh, size_t lds_max);
int smag_prepare(size_t lds_max);
// pm's twiddle table exp(+2 pi i t / nx), t < nx, also serves the packing of
// the real x transforms
void launch_smag(SmagArgs a, const FftPlan &pz, const FftPlan &pm, size_t lds_max,
                 hipStream_t s);
bool xys_fits(int nz, int nxh);
void launch_xys(const XysArgs &a, const FftPlan &pz, const FftPlan &pm, const double2 *wx,
                hipStream_t s);
// per-y amplitude sums of boxamp.f into amp[20][ny]
void launch_amp(const double2 *U, const double *alfa, const double *beta, int nz, int ny, int nxh,
                double *amp, hipStream_t s);

// A y transform applied to every plane: the dense matrix [ny][ny] (FP64,
// and FP32 in the mixed form) and, where ny has a parity-split kernel and
// the matrix the parity structure of the Gauss-Lobatto grid, its packed
// half-size blocks (sg_cheb.hip)
struct ChebOp {
    const double *A = nullptr;
    const float *A32 = nullptr;
    const double *S = nullptr;
    const float *S32 = nullptr;
    int K = 0;        // split kernel size, ny = 8 (K - 1) + 1; 0: none
    int flip = 0;     // the operator swaps parities (a y derivative)
};
// FP64 transform: split kernel, else the dense MFMA kernel (ny <= 128),
// else a VALU product
void launch_cheb_gemm(const ChebOp &op, const GemmIO &io, hipStream_t s);
// FP32 transform (mixed precision; in/out replace io.in/io.out); false
// when ny has neither an FP32 split nor a dense MFMA kernel
bool launch_cheb_f32(const ChebOp &op, const float *in, float *out, const GemmIO &io,
                     hipStream_t s);
// MFMA form of the dense transform for ny <= 128; false when ny is outside it
bool launch_cheb_mfma(const double *A, const GemmIO &io, hipStream_t s);
bool launch_cheb_mfma_f32(const float *A, const float *in, float *out, const GemmIO &io,
                          hipStream_t s);
int  cheb_mfma_ks(int ny);
// parity split (host): packed blocks of A for the given layouts and the
// kernel size K, or 0 (no split kernel for ny, or no parity structure)
int  cheb_split(const double *A, int ny, int in_col, int out_col, std::vector<double> &M,
                int *flip);
bool launch_cheb_par(const double *M, int K, int flip, const GemmIO &io, hipStream_t s);
bool launch_cheb_par_f32(const float *M, int K, int flip, const float *in, float *out,
                         const GemmIO &io, hipStream_t s);
// an FP32 transform exists for ny (split kernel or dense MFMA kernel)
bool cheb_f32_ok(int ny);
// nzp / nxp/2 with a four-step step-2 kernel (needed by the FP32 path)
bool has_zfft4(int nzp);
bool has_xrow4(int m);
// FP64 <-> FP32 copies of n complex values
void launch_narrow(const double2 *src, float2 *dst, size_t n, hipStream_t s);
void launch_widen(const float2 *src, double2 *dst, size_t n, hipStream_t s);
// the kx = kz = 0 column of slots 0-4 of U from FP64 Chebyshev columns C
void launch_mean_synth(const double *Tb, const double2 *C, double2 *U, int ny, int nz, int nxh,
                       hipStream_t s);
void launch_prhs(const ColArgs &a, double an, hipStream_t s);

// tau factorisations: k^2 once (also kept as kfT [2][ny][ncol], gamma and
// 1/pivot planes), lam^2 for one value of a+b (uses a.W planes 0-4 as work
// space)
void launch_factor_k2(const ColArgs &a, LamFactor kf, double *kfT, hipStream_t s);
void launch_factor_lam(const ColArgs &a, double apb, LamFactor lf, const double *kfT,
                       hipStream_t s);
// one linear substep, one wavefront per column
// (fp32_cols: the Chebyshev columns are C32, else C)
// allow the generic step-2 kernels dynamic LDS up to lds_max bytes
int nonlin_prepare(int lds_max);
int fsc_table(double m, int n, double ymax, double eps, double *tab, double *dstar2);
int parallel_force(const simson_baseflow_t *b, const double *tab, int ntab, double cdev,
                   double t, double *force);
int local_force_pattern(int nx, int ny, double xl, const simson_locf_t *lf, double2 *phys);
int spatial_base_flow(const simson_baseflow_t *b, const double *tab, int ntab, double *base,
                      double *bjr, double *bji, double *du0upp);
int  launch_tau(const ColArgs &a, LamFactor lf, LamFactor kf, double an, double bn,
                double anp1, double bnp1, bool fp32_cols, hipStream_t s);

}  // namespace sg
