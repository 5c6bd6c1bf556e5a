// Step 2 of the bla substep on the GPU: the 3/2-dealiased nonlinear term
// H = u x (omega + 2 Omega e_z) for every y plane (bla/nonlinbl.f, called
// per plane from bla/bla.f:2251-2288).
//
// Three kernels per substep, all planes at once:
//   k_zsyn  gather u,v,w,om_x,om_z of a y plane from main storage, form
//           om_y = i beta u - i alpha w, pad to nzp modes (getxz.f) and run
//           the inverse complex z transform (fft2db.f:33) in LDS;
//   k_xrow  per (y, z) row: packed inverse real x transform of the six
//           fields (fft2db.f:45), the rotational-form product
//           (advection.f:52-72), the CFL partial max (boxcfl.f), and the
//           forward real x transform (fft2df.f:32);
//   k_zana  forward complex z transform (fft2df.f:37) and truncation back to
//           nz modes into slots 0-2 (putxz.f, z oddball row included).
#include "sg_fft.h"
#include "sg_internal.h"

namespace sg {

namespace {

__device__ __forceinline__ size_t uidx(int slot, int z, int y, int x, int nz, int ny, int nxh)
{
    return (((size_t)slot * nz + z) * ny + y) * nxh + x;
}

// storage z index of padded row zp, or -1 for the zeroed dealias band
// (getxz.f:41-83: rows nz/2 .. nzp-nz/2 are zero, including the oddball)
__device__ __forceinline__ int zsrc(int zp, int nz, int nzp)
{
    if (zp < nz / 2)
        return zp;
    if (zp > nzp - nz / 2)
        return zp - (nzp - nz);
    return -1;
}

__global__ void __launch_bounds__(256) k_zsyn(NonlinArgs a, FftPlan pz, int B)
{
    extern __shared__ double2 lds[];
    const int nzp = a.nzp, nz = a.nz, nxh = a.nxh, ny = a.ny;
    double2 *buf0 = lds, *buf1 = lds + (size_t)B * nzp;
    const int f = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    for (int idx = threadIdx.x; idx < nzp * B; idx += blockDim.x) {
        const int zp = idx / B, b = idx - zp * B;
        double2 v = make_double2(0., 0.);
        const int z = zsrc(zp, nz, nzp);
        if (b < nb && z >= 0) {
            const int x = x0 + b;
            if (f == 4) {
                const double2 u = a.U[uidx(0, z, y, x, nz, ny, nxh)];
                const double2 w = a.U[uidx(2, z, y, x, nz, ny, nxh)];
                const double be = a.beta[z], al = a.alfa[x];
                v = make_double2(-be * u.y + al * w.y, be * u.x - al * w.x);
            } else {
                const int slot = f < 4 ? f : 4;
                v = a.U[uidx(slot, z, y, x, nz, ny, nxh)];
            }
        }
        buf0[b * nzp + zp] = v;
    }
    __syncthreads();
    double2 *res = lds_fft(buf0, buf1, nzp, B, pz, +1.0);
    double2 *P = a.P + (((size_t)f * ny + y) * nzp) * nxh + x0;
    for (int idx = threadIdx.x; idx < nzp * B; idx += blockDim.x) {
        const int zp = idx / B, b = idx - zp * B;
        if (b < nb)
            P[(size_t)zp * nxh + b] = res[b * nzp + zp];
    }
}

__global__ void __launch_bounds__(256) k_xrow(NonlinArgs a, FftPlan pm, const double2 *__restrict__ wx,
                                              int RW)
{
    extern __shared__ double2 lds[];
    const int M = a.nxp / 2, nxh = a.nxh, nzp = a.nzp, ny = a.ny;
    double2 *A = lds, *Bf = lds + (size_t)6 * RW * M;
    const int zp0 = blockIdx.x * RW, y = blockIdx.y;
    const int nr = min(RW, nzp - zp0);
    // half spectrum X_k (k < nx/2, zero above) -> packed Z_k, k < M
    for (int idx = threadIdx.x; idx < 6 * RW * M; idx += blockDim.x) {
        const int f = idx / (RW * M);
        const int rem = idx - f * RW * M;
        const int r = rem / M, k = rem - r * M;
        double2 Z = make_double2(0., 0.);
        if (r < nr) {
            const double2 *row = a.P + (((size_t)f * ny + y) * nzp + zp0 + r) * nxh;
            double2 Xk = k < nxh ? row[k] : make_double2(0., 0.);
            if (k == 0)
                Xk.y = 0.;            // only Re X_0 enters vrfftb
            const int km = M - k;     // X_M (k = 0) is beyond nx/2: zero
            double2 Xm = (k > 0 && km < nxh) ? row[km] : make_double2(0., 0.);
            Xm.y = -Xm.y;
            const double2 s = cadd(Xk, Xm), d = csub(Xk, Xm);
            const double2 w = wx[k];
            Z = cadd(s, mul_is(cmul(w, d), 1.0));
        }
        A[(f * RW + r) * M + k] = Z;
    }
    __syncthreads();
    double2 *res = lds_fft(A, Bf, M, 6 * RW, pm, +1.0);
    double2 *oth = (res == A) ? Bf : A;
    // products on the physical row (even points in .x, odd points in .y)
    double cflmax = 0.;
    double dyinv = 1.0 / a.deta[y];
    if (a.ny >= 17) {
        const double cap = (double)(a.ny - 1) * (double)(a.ny - 1) / 10. / 3.14;
        dyinv = fmin(dyinv, cap);
    }
    const double dxi = (double)(a.nx - 2) / a.xl;
    const double dzi = (double)(a.nz - 2) / a.zl;
    const double r2 = 2.0 * a.rot;
    for (int idx = threadIdx.x; idx < RW * M; idx += blockDim.x) {
        const int r = idx / M, n = idx - r * M;
        const double2 u = res[(0 * RW + r) * M + n], v = res[(1 * RW + r) * M + n];
        const double2 w = res[(2 * RW + r) * M + n], ox = res[(3 * RW + r) * M + n];
        const double2 oy = res[(4 * RW + r) * M + n], oz = res[(5 * RW + r) * M + n];
        if (a.do_cfl && r < nr) {
            const double ce = fabs(u.x) * dxi + fabs(v.x) * dyinv + fabs(w.x) * dzi;
            const double co = fabs(u.y) * dxi + fabs(v.y) * dyinv + fabs(w.y) * dzi;
            cflmax = fmax(cflmax, fmax(ce, co));
        }
        double2 h1, h2, h3;
        h1.x = v.x * (oz.x + r2) - w.x * oy.x;
        h2.x = w.x * ox.x - u.x * (oz.x + r2);
        h3.x = u.x * oy.x - v.x * ox.x;
        h1.y = v.y * (oz.y + r2) - w.y * oy.y;
        h2.y = w.y * ox.y - u.y * (oz.y + r2);
        h3.y = u.y * oy.y - v.y * ox.y;
        res[(0 * RW + r) * M + n] = h1;
        res[(1 * RW + r) * M + n] = h2;
        res[(2 * RW + r) * M + n] = h3;
    }
    __syncthreads();
    double2 *fin = lds_fft(res, oth, M, 3 * RW, pm, -1.0);
    // packed spectrum -> X_k = E_k + exp(-2 pi i k/nxp) O_k for k < nx/2
    for (int idx = threadIdx.x; idx < 3 * RW * nxh; idx += blockDim.x) {
        const int h = idx / (RW * nxh);
        const int rem = idx - h * RW * nxh;
        const int r = rem / nxh, k = rem - r * nxh;
        if (r < nr) {
            const double2 Zk = fin[(h * RW + r) * M + k];
            const double2 Zm = conjd(fin[(h * RW + r) * M + (k == 0 ? 0 : M - k)]);
            const double2 E = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y + Zm.y));
            const double2 dd = csub(Zk, Zm);
            const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);   // (Zk - conj Zm)/(2i)
            const double2 X = cadd(E, cmul(conjd(wx[k]), O));
            a.P[(((size_t)h * ny + y) * nzp + zp0 + r) * nxh + k] = X;
        }
    }
    if (a.do_cfl) {
        // block max of the CFL partial
        for (int off = 32; off > 0; off >>= 1)
            cflmax = fmax(cflmax, __shfl_xor(cflmax, off));
        __shared__ double wmax[8];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        if (lane == 0)
            wmax[wv] = cflmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.;
            for (int i = 0; i < (int)(blockDim.x >> 6); i++)
                m = fmax(m, wmax[i]);
            a.cfl_blk[blockIdx.y * gridDim.x + blockIdx.x] = m;
        }
    }
}

__global__ void __launch_bounds__(256) k_zana(NonlinArgs a, FftPlan pz, int B)
{
    extern __shared__ double2 lds[];
    const int nzp = a.nzp, nz = a.nz, nxh = a.nxh, ny = a.ny;
    double2 *buf0 = lds, *buf1 = lds + (size_t)B * nzp;
    const int h = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    const double2 *P = a.P + (((size_t)h * ny + y) * nzp) * nxh + x0;
    for (int idx = threadIdx.x; idx < nzp * B; idx += blockDim.x) {
        const int zp = idx / B, b = idx - zp * B;
        buf0[b * nzp + zp] = b < nb ? P[(size_t)zp * nxh + b] : make_double2(0., 0.);
    }
    __syncthreads();
    double2 *res = lds_fft(buf0, buf1, nzp, B, pz, -1.0);
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        if (b < nb) {
            const int zp = z < nz / 2 ? z : z + nzp - nz;
            a.Uw[uidx(h, z, y, x0 + b, nz, ny, nxh)] = res[b * nzp + zp];
        }
    }
}

__global__ void k_cfl_reduce(const double *blk, int n, double *out)
{
    double m = 0.;
    for (int i = threadIdx.x; i < n; i += blockDim.x)
        m = fmax(m, blk[i]);
    for (int off = 32; off > 0; off >>= 1)
        m = fmax(m, __shfl_xor(m, off));
    __shared__ double wmax[4];
    if ((threadIdx.x & 63) == 0)
        wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.;
        for (int i = 0; i < (int)(blockDim.x >> 6); i++)
            r = fmax(r, wmax[i]);
        out[0] = r;                   // latest evaluation
        out[1] = fmax(out[1], r);     // running max since the last reset
    }
}

int zblock(int nxh, int nzp)
{
    // columns per block so that two LDS buffers stay within 64 KiB
    int B = 65536 / (2 * nzp * (int)sizeof(double2));
    if (B < 1)
        B = 1;
    if (B > 16)
        B = 16;
    if (B > nxh)
        B = nxh;
    return B;
}

int xrows(int nxp, int nzp)
{
    const int M = nxp / 2;
    int RW = 65536 / (2 * 6 * M * (int)sizeof(double2));
    if (RW < 1)
        RW = 1;
    if (RW > 8)
        RW = 8;
    if (RW > nzp)
        RW = nzp;
    return RW;
}

}  // namespace

void launch_zsyn(const NonlinArgs &a, const FftPlan &pz, hipStream_t s)
{
    const int B = zblock(a.nxh, a.nzp);
    dim3 grid((a.nxh + B - 1) / B, 6, a.ny);
    const size_t lds = 2 * (size_t)B * a.nzp * sizeof(double2);
    hipLaunchKernelGGL(k_zsyn, grid, dim3(256), lds, s, a, pz, B);
}

int xrow_blocks(const NonlinArgs &a)
{
    const int RW = xrows(a.nxp, a.nzp);
    return ((a.nzp + RW - 1) / RW) * a.ny;
}

void launch_xrow(const NonlinArgs &a, const FftPlan &pm, const double2 *wx, hipStream_t s)
{
    const int RW = xrows(a.nxp, a.nzp);
    dim3 grid((a.nzp + RW - 1) / RW, a.ny);
    const size_t lds = 2 * (size_t)6 * RW * (a.nxp / 2) * sizeof(double2);
    hipLaunchKernelGGL(k_xrow, grid, dim3(256), lds, s, a, pm, wx, RW);
}

void launch_zana(const NonlinArgs &a, const FftPlan &pz, hipStream_t s)
{
    const int B = zblock(a.nxh, a.nzp);
    dim3 grid((a.nxh + B - 1) / B, 3, a.ny);
    const size_t lds = 2 * (size_t)B * a.nzp * sizeof(double2);
    hipLaunchKernelGGL(k_zana, grid, dim3(256), lds, s, a, pz, B);
}

void launch_cfl_reduce(const double *blk, int n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_cfl_reduce, dim3(1), dim3(256), 0, s, blk, n, out);
}

}  // namespace sg
