// Internal declarations of the MI355X Simson step library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "simson_gpu.h"

namespace sg {

constexpr int kMaxRad = 16;
constexpr int kNSlot = 7;          // u v w om_x om_z p_v p_om
constexpr int kTauArrays = 28;     // per-column scratch arrays of the tau kernel

// Radix schedule and twiddle table of one FFT length.
struct FftPlan {
    int n = 0;
    int nrad = 0;
    int rad[kMaxRad] = {};
    const double2 *tw = nullptr;   // exp(+2 pi i t / tabn), t < tabn
    int tabn = 0;
};

// Which planes a batched Chebyshev GEMM reads and writes.  Plane (f, z) of
// a [slot][z][ny][ncols] array starts at ((slot*nz)+z)*ny*ncols doubles.
struct GemmIO {
    const double *in;
    double *out;
    int in_slot[8];
    int out_slot[8];
    int nf;          // number of fields
    int nz;          // planes per field
    int ny;          // rows (Chebyshev length)
    int ncols;       // doubles per row (2 * nx/2)
    int skip_z;      // plane to skip (z-oddball) or -1
};

// Arguments of the per-(kx,kz)-column solver kernels.
struct ColArgs {
    double2 *U;           // main storage [7][nz][ny][nxh]
    double2 *C;           // Chebyshev planes [5][nz][ny][nxh]
    double *W;            // scratch [kTauArrays][ny][nz*nxh]
    double *puw;          // [2][ny]
    double *px;           // device scalar
    double *mf;           // mean-flow scratch [8][ny]
    const double *alfa;   // [nxh]
    const double *beta;   // [nz]
    int ny, nz, nxh;
    double re;
    double mflux;
    int cflux;
    double u0upp, u0low, w0upp, w0low;
};

// Step-2 launch arguments.
struct NonlinArgs {
    const double2 *U;
    double2 *Uw;
    double2 *P;           // [6][ny][nzp][nxh]
    const double *alfa, *beta, *deta;
    double *cfl_blk;
    int nx, ny, nz, nxh, nxp, nzp;
    double rot;
    double xl, zl;
    int do_cfl;
};

void launch_zsyn(const NonlinArgs &a, const FftPlan &pz, hipStream_t s);
void launch_xrow(const NonlinArgs &a, const FftPlan &pm, const double2 *wx, hipStream_t s);
void launch_zana(const NonlinArgs &a, const FftPlan &pz, hipStream_t s);
int  xrow_blocks(const NonlinArgs &a);
void launch_cfl_reduce(const double *blk, int n, double *out, hipStream_t s);

void launch_cheb_gemm(const double *A, const GemmIO &io, hipStream_t s);
void launch_tau(const ColArgs &a, double an, double bn, double anp1, double bnp1,
                hipStream_t s);
void launch_prhs(const ColArgs &a, double an, hipStream_t s);

}  // namespace sg
