// Step 3 support kernels (bla/linearbl.f, tau method) and the start-up
// partial RHS (bla/prhs.f:93-709):
//
//   k_gemm       batched y Chebyshev transforms as dense products with the
//                (ny x ny) transform matrix: analysis of H (vchbf + scaling,
//                linearbl.f:262-273), synthesis of u,v,w,om_x,om_z (vchbb,
//                linearbl.f:1830-1846); physical planes are [y][x], Chebyshev
//                planes are stored column-major [x][k] for the column solver;
//   k_factor_*   the tau Helmholtz factorisations (gamma, 1/pivot) per column
//                for lam = k^2 and for lam^2 = k^2 + 2 Re/(a+b), plus the
//                homogeneous v_h (setmacrhs.f + trid.f), computed only when
//                a+b takes a new value (every substep with a variable dt);
//   k_prhs       one thread per column: om_x, om_z, p_v, p_omega and puw from
//                the velocities at start-up.
#include "sg_internal.h"
#include "sg_tau.h"

namespace sg {

namespace {

// ---------------------------------------------------------------------
// batched Chebyshev transform: Y[b] = A (ny x ny) * X[b] (ny x 2 nxh)
// ---------------------------------------------------------------------
constexpr int GBM = 64, GBN = 64, GBK = 16;

__device__ __forceinline__ size_t gidx(int col, int j, int n, int ny, int nc)
{
    // element (row j, column n = 2x + reim) of one plane
    return col ? ((size_t)(n >> 1) * ny + j) * 2 + (n & 1) : (size_t)j * nc + n;
}

__global__ void __launch_bounds__(256) k_gemm(const double *__restrict__ A, GemmIO io)
{
    __shared__ double As[GBK][GBM + 1];
    __shared__ double Xs[GBK][GBN + 1];
    __shared__ double Ys[GBM][GBN + 1];
    const int b = blockIdx.z;
    const int f = b / io.nz, z = b - f * io.nz;
    if (z == io.skip_z)
        return;
    const int ny = io.ny, nc = 2 * io.nxh;
    const size_t plane = (size_t)ny * nc;
    const double *X = io.in + ((size_t)io.in_slot[f] * io.nz + z) * plane;
    double *Y = io.out + ((size_t)io.out_slot[f] * io.nz + z) * plane;
    const int m0 = blockIdx.y * GBM, n0 = blockIdx.x * GBN;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    double acc[4][4];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            acc[i][j] = 0.0;
    for (int k0 = 0; k0 < ny; k0 += GBK) {
        for (int i = 0; i < 4; i++) {
            const int e = tid + 256 * i;
            const int mm = e / GBK, kk = e - mm * GBK;
            const int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < ny && gk < ny) ? A[(size_t)gm * ny + gk] : 0.0;
            int k2, nn;
            if (io.in_col) {          // k fastest within one complex column
                nn = ((e >> 5) << 1) | (e & 1);
                k2 = (e >> 1) & 15;
            } else {
                k2 = e / GBN;
                nn = e - k2 * GBN;
            }
            const int gk2 = k0 + k2, gn = n0 + nn;
            Xs[k2][nn] = (gk2 < ny && gn < nc) ? X[gidx(io.in_col, gk2, gn, ny, nc)] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GBK; kk++) {
            double av[4], xv[4];
            for (int i = 0; i < 4; i++)
                av[i] = As[kk][ty + 16 * i];
            for (int j = 0; j < 4; j++)
                xv[j] = Xs[kk][tx + 16 * j];
            for (int i = 0; i < 4; i++)
                for (int j = 0; j < 4; j++)
                    acc[i][j] = fma(av[i], xv[j], acc[i][j]);
        }
        __syncthreads();
    }
    if (!io.out_col) {
        for (int i = 0; i < 4; i++) {
            const int gm = m0 + ty + 16 * i;
            if (gm >= ny)
                continue;
            for (int j = 0; j < 4; j++) {
                const int gn = n0 + tx + 16 * j;
                if (gn < nc)
                    Y[(size_t)gm * nc + gn] = acc[i][j];
            }
        }
        return;
    }
    // column-major output: stage the tile so that stores run along k
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
            Ys[ty + 16 * i][tx + 16 * j] = acc[i][j];
    __syncthreads();
    for (int e = tid; e < GBM * GBN; e += 256) {
        const int reim = e & 1, kk = (e >> 1) & 63, xx = e >> 7;
        const int nn = 2 * xx + reim;
        const int gm = m0 + kk, gn = n0 + nn;
        if (gm < ny && gn < nc)
            Y[gidx(1, gm, gn, ny, nc)] = Ys[kk][nn];
    }
}

// views of one column (k = 0..ny-1)
struct ColView {
    double *p;
    size_t s;
    __device__ __forceinline__ double &operator[](int k) const { return p[(size_t)k * s]; }
};

// dcheb.f:24-36
__device__ void dcheb_seq(int n, ColView a, ColView b)
{
    double bk2 = 0.0, bk1 = 0.0;
    b[n - 1] = 0.0;
    for (int k = n - 2; k >= 1; k--) {
        const double bk = bk2 + 2.0 * (double)(k + 1) * a[k + 1];
        b[k] = bk;
        bk2 = bk1;
        bk1 = bk;
    }
    b[0] = 0.5 * (bk2 + 2.0 * a[1]);
}

// Tau solve with stored factors (gamma, 1/pivot) per element, see sg_tau.h.
// (gamma, 1/pivot) of one column as two strided arrays
struct GpStrided {
    ColView x, y;
    __device__ __forceinline__ double g(int k) const { return x[k]; }
    __device__ __forceinline__ double ip(int k) const { return y[k]; }
};

template <typename GP>
__device__ void seq_solve(int n, double lam, ColView f, bool fzero, double bce, double bco,
                          const GP &gp, ColView a, ColView rho)
{
    for (int c = 0; c < 2; c++) {
        const int klast = ((n - 1 - c) & 1) == 0 ? n - 1 : n - 2;
        double r_up = 0.0;
#pragma unroll 8
        for (int k = klast; k >= c + 2; k -= 2) {
            const double R = fzero ? 0.0 : tau_rhs(n, k, f[k - 2], f[k],
                                                   k + 2 < n ? f[k + 2] : 0.0);
            const double U = tau_up(n, k, lam);
            const double rk = (R - U * r_up) * gp.ip(k);
            rho[k] = rk;
            r_up = rk;
        }
        double P = 0.0, Q = 1.0, sP = 0.0, sQ = 1.0;
#pragma unroll 8
        for (int k = c + 2; k <= klast; k += 2) {
            const double g = gp.g(k);
            P = rho[k] - g * P;
            Q = -g * Q;
            sP += P;
            sQ += Q;
        }
        double prev = ((c == 0 ? bce : bco) - sP) / sQ;
        a[c] = prev;
#pragma unroll 8
        for (int k = c + 2; k <= klast; k += 2) {
            prev = rho[k] - gp.g(k) * prev;
            a[k] = prev;
        }
    }
}

// gamma_k, 1/pivot_k for lam (bottom-up elimination of each parity chain)
__device__ void factor_seq(int n, double lam, ColView gx, ColView gy)
{
    for (int c = 0; c < 2; c++) {
        const int klast = ((n - 1 - c) & 1) == 0 ? n - 1 : n - 2;
        double g_up = 0.0;
        for (int k = klast; k >= c + 2; k -= 2) {
            const double piv = tau_diag(n, k, lam) - tau_up(n, k, lam) * g_up;
            const double g = tau_low(k, lam) / piv;
            gx[k] = g;
            gy[k] = 1.0 / piv;
            g_up = g;
        }
        gx[c] = 0.0;
        gy[c] = 0.0;
    }
}

// Both factorisation kernels run one thread per column with every
// per-column array in a [k][column] layout (stride ncol), so that the
// lanes of a wave touch consecutive addresses; k_cols_to_rows then writes
// the tau kernel's [column][k] layout through LDS tiles.

// lam = k^2: the factors in both layouts (kf and kfT [k][column])
__global__ void __launch_bounds__(256) k_factor_k2(ColArgs a, double *kfT)
{
    const int ncol = a.nz * a.nxh;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol)
        return;
    const int z = t / a.nxh, x = t - z * a.nxh;
    const double k2 = a.alfa[x] * a.alfa[x] + a.beta[z] * a.beta[z];
    const size_t plane = (size_t)a.ny * ncol;
    factor_seq(a.ny, k2, ColView{kfT + t, (size_t)ncol}, ColView{kfT + plane + t, (size_t)ncol});
}

// lam^2 = k^2 + 2 Re/(a+b): factors and the homogeneous v_h with its parity
// wall sums, into the scratch W (0-1: factors, 2: phi_h, 3: rho, 4: v_h)
__global__ void __launch_bounds__(64) k_factor_lam(ColArgs a, double apb, LamFactor lf,
                                                   const double *kfT)
{
    const int ncol = a.nz * a.nxh;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol)
        return;
    const int z = t / a.nxh, x = t - z * a.nxh;
    const int n = a.ny;
    const size_t plane = (size_t)n * ncol;
    const double k2 = a.alfa[x] * a.alfa[x] + a.beta[z] * a.beta[z];
    const double lam2 = k2 + 2. * a.re / apb;
    auto W = [&](int i) { return ColView{a.W + (size_t)i * plane + t, (size_t)ncol}; };
    const GpStrided gp{W(0), W(1)};
    factor_seq(n, lam2, gp.x, gp.y);
    // homogeneous solutions: phi_h (bc 1,1) with lam^2, then v_h (bc 0,0)
    // with k^2 (linearbl.f:625-626, 1000-1001)
    const ColView phh = W(2), rho = W(3), vh = W(4);
    seq_solve(n, lam2, phh, true, 1., 1., gp, phh, rho);
    const GpStrided kgp{ColView{const_cast<double *>(kfT) + t, (size_t)ncol},
                        ColView{const_cast<double *>(kfT) + plane + t, (size_t)ncol}};
    seq_solve(n, k2, phh, false, 0., 0., kgp, vh, rho);
    double s[2] = {0., 0.};
#pragma unroll 8
    for (int k = 0; k < n; k++)
        s[k & 1] += (double)k * (double)k * vh[k];
    lf.sh[t] = make_double2(s[0], s[1]);
    if (!lf.hc)
        return;
    // free stream above (ibc = 101, linearbl.f:946-1098): v_h2 solves the
    // k^2 problem with v = 1 per parity class (v(+1) = 2, v(-1) = 0), its
    // wall slopes cancelled with v_h; v_h1 is v_h with each class scaled
    // to unit slope at y = +1.  A1, A2: d2/dy2 at +1 of v_h1, v_h2
    // (T_k''(1) = k^2 (k^2 - 1)/3); B: dphi_h/dy at +1 (T_k'(1) = k^2).
    const ColView vh2 = W(5);
    seq_solve(n, k2, phh, true, 1., 1., kgp, vh2, rho);
    double s2[2] = {0., 0.};
#pragma unroll 8
    for (int k = 0; k < n; k++)
        s2[k & 1] += (double)k * (double)k * vh2[k];
    const double r2[2] = {-s2[0] / s[0], -s2[1] / s[1]};
    const double r1[2] = {1. / s[0], 1. / s[1]};
    double A1 = 0., A2 = 0., B = 0.;
#pragma unroll 8
    for (int k = 0; k < n; k++) {
        const double kk = (double)k * (double)k, w2 = kk * (kk - 1.) * (1. / 3.);
        const double v2 = vh2[k] + r2[k & 1] * vh[k];
        vh2[k] = v2;
        A1 += w2 * (vh[k] * r1[k & 1]);
        A2 += w2 * v2;
        B += kk * phh[k];
    }
    lf.hc[4 * t] = A1;
    lf.hc[4 * t + 1] = A2;
    lf.hc[4 * t + 2] = B;
    lf.hc[4 * t + 3] = 0.;
}

// [k][column] -> [column][k]: out2 (double2 from the planes a, b) and,
// when out1 is given, out1 from the plane c
constexpr int kTT = 32;
__global__ void __launch_bounds__(256) k_cols_to_rows(const double *__restrict__ pa,
                                                      const double *__restrict__ pb,
                                                      const double *__restrict__ pc,
                                                      double2 *__restrict__ out2,
                                                      double *__restrict__ out1, int n, int ncol)
{
    __shared__ double ta[kTT][kTT + 1], tb[kTT][kTT + 1], tc[kTT][kTT + 1];
    const int c0 = blockIdx.x * kTT, k0 = blockIdx.y * kTT;
    const int tx = threadIdx.x & (kTT - 1), ty = threadIdx.x / kTT;   // 32 x 8
    for (int r = ty; r < kTT; r += 8) {
        const int k = k0 + r, c = c0 + tx;
        const bool ok = k < n && c < ncol;
        const size_t i = (size_t)k * ncol + c;
        ta[r][tx] = ok ? pa[i] : 0.;
        tb[r][tx] = ok ? pb[i] : 0.;
        if (out1)
            tc[r][tx] = ok ? pc[i] : 0.;
    }
    __syncthreads();
    for (int r = ty; r < kTT; r += 8) {
        const int c = c0 + r, k = k0 + tx;
        if (c < ncol && k < n) {
            const size_t o = (size_t)c * n + k;
            out2[o] = make_double2(ta[tx][r], tb[tx][r]);
            if (out1)
                out1[o] = tc[tx][r];
        }
    }
}

__global__ void __launch_bounds__(256) k_prhs(ColArgs a, double an)
{
    const int ncol = a.nz * a.nxh;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncol)
        return;
    const int z = t / a.nxh, x = t - z * a.nxh;
    const int n = a.ny;
    const double re = a.re;
    const double alfa = a.alfa[x], beta = a.beta[z];
    const double k2 = alfa * alfa + beta * beta;
    const double lam = k2 - 2. * re / an;
    const size_t fs = (size_t)a.nz * a.nxh * n;        // slot stride (complex)
    double *cb = (double *)(a.C + (size_t)t * n);       // column t of slot 0
    auto cv = [&](int s, int c) { return ColView{cb + 2 * s * fs + c, 2}; };
    auto W = [&](int i) { return ColView{a.W + (size_t)i * n * ncol + t, (size_t)ncol}; };
    ColView uR = cv(0, 0), uI = cv(0, 1), vR = cv(1, 0), vI = cv(1, 1), wR = cv(2, 0),
            wI = cv(2, 1);
    ColView oyr = W(0), oyi = W(1), dwr = W(2), dwi = W(3), dur = W(4), dui = W(5);
    ColView t1 = W(6), t2 = W(7), phr = W(8), phi = W(9), pwr = W(10), pwi = W(11);
    ColView pvr = W(12), pvi = W(13);
    if (z == 0 && x == 0) {              // prhs.f:534-546
        dcheb_seq(n, uR, t1);
        dcheb_seq(n, t1, t2);
        for (int k = 0; k < n; k++)
            a.puw[k] = -2. * re / an * uR[k] - t2[k];
        dcheb_seq(n, wR, t1);
        dcheb_seq(n, t1, t2);
        for (int k = 0; k < n; k++)
            a.puw[n + k] = -2. * re / an * wR[k] - t2[k];
    }
    for (int k = 0; k < n; k++) {       // prhs.f:563-573
        oyr[k] = -uI[k] * beta + wI[k] * alfa;
        oyi[k] = uR[k] * beta - wR[k] * alfa;
    }
    dcheb_seq(n, wR, dwr);
    dcheb_seq(n, wI, dwi);
    dcheb_seq(n, uR, dur);
    dcheb_seq(n, uI, dui);
    // phi = D^2 v - k^2 v ; p_omega = -D^2 om_y + lam om_y ; p_v = -D^2 phi + lam phi
    dcheb_seq(n, vR, t1);
    dcheb_seq(n, t1, phr);
    dcheb_seq(n, vI, t1);
    dcheb_seq(n, t1, phi);
    for (int k = 0; k < n; k++) {
        phr[k] = phr[k] - k2 * vR[k];
        phi[k] = phi[k] - k2 * vI[k];
    }
    dcheb_seq(n, oyr, t1);
    dcheb_seq(n, t1, t2);
    for (int k = 0; k < n; k++)
        pwr[k] = -t2[k] + lam * oyr[k];
    dcheb_seq(n, oyi, t1);
    dcheb_seq(n, t1, t2);
    for (int k = 0; k < n; k++)
        pwi[k] = -t2[k] + lam * oyi[k];
    dcheb_seq(n, phr, t1);
    dcheb_seq(n, t1, t2);
    for (int k = 0; k < n; k++)
        pvr[k] = -t2[k] + lam * phr[k];
    dcheb_seq(n, phi, t1);
    dcheb_seq(n, t1, t2);
    for (int k = 0; k < n; k++)
        pvi[k] = -t2[k] + lam * phi[k];
    double2 *C3 = a.C + 3 * fs + (size_t)t * n;
    double2 *C4 = a.C + 4 * fs + (size_t)t * n;
    double2 *P0 = a.Pp + (size_t)t * n;
    double2 *P1 = a.Pp + fs + (size_t)t * n;
    for (int k = 0; k < n; k++) {       // prhs.f:578-590 then storage
        const double vr = vR[k], vi = vI[k];
        C3[k] = make_double2(dwr[k] + vi * beta, dwi[k] - vr * beta);
        C4[k] = make_double2(-dur[k] - vi * alfa, -dui[k] + vr * alfa);
        P0[k] = make_double2(pvr[k], pvi[k]);
        P1[k] = make_double2(pwr[k], pwi[k]);
    }
}

}  // namespace

void launch_cheb_gemm(const ChebOp &op, const GemmIO &io, hipStream_t s)
{
    if (op.S && launch_cheb_par(op.S, op.K, op.flip, io, s))
        return;
    const double *A = op.A;
    if (launch_cheb_mfma(A, io, s))
        return;
    dim3 grid((2 * io.nxh + GBN - 1) / GBN, (io.ny + GBM - 1) / GBM, io.nf * io.nz);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, s, A, io);
}

void launch_prhs(const ColArgs &a, double an, hipStream_t s)
{
    const int ncol = a.nz * a.nxh;
    hipLaunchKernelGGL(k_prhs, dim3((ncol + 255) / 256), dim3(256), 0, s, a, an);
}

void launch_factor_k2(const ColArgs &a, LamFactor kf, double *kfT, hipStream_t s)
{
    const int ncol = a.nz * a.nxh, n = a.ny;
    const size_t plane = (size_t)n * ncol;
    hipLaunchKernelGGL(k_factor_k2, dim3((ncol + 255) / 256), dim3(256), 0, s, a, kfT);
    dim3 tg((ncol + kTT - 1) / kTT, (n + kTT - 1) / kTT);
    hipLaunchKernelGGL(k_cols_to_rows, tg, dim3(256), 0, s, kfT, kfT + plane,
                       (const double *)nullptr, kf.gp, (double *)nullptr, n, ncol);
}

void launch_factor_lam(const ColArgs &a, double apb, LamFactor lf, const double *kfT,
                       hipStream_t s)
{
    const int ncol = a.nz * a.nxh, n = a.ny;
    const size_t plane = (size_t)n * ncol;
    // one thread per column in blocks of one wave: the columns are few
    // (nz nx/2) and each is a latency-bound chain, so they are spread over
    // as many CUs as possible
    hipLaunchKernelGGL(k_factor_lam, dim3((ncol + 63) / 64), dim3(64), 0, s, a, apb, lf, kfT);
    dim3 tg((ncol + kTT - 1) / kTT, (n + kTT - 1) / kTT);
    hipLaunchKernelGGL(k_cols_to_rows, tg, dim3(256), 0, s, a.W, a.W + plane, a.W + 4 * plane,
                       lf.gp, lf.vh, n, ncol);
    if (lf.hb)
        hipLaunchKernelGGL(k_cols_to_rows, tg, dim3(256), 0, s, a.W + 2 * plane,
                           a.W + 5 * plane, (const double *)nullptr, lf.hb, (double *)nullptr, n,
                           ncol);
}

}  // namespace sg
