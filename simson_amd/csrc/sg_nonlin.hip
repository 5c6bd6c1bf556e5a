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
    if (a.ny >= 17 && !a.nocap) {
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
            cflmax = cfl_max(cfl_max(cflmax, ce), co);
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

// one block of kCflT threads; each thread keeps kCflU loads in flight
constexpr int kCflT = 1024, kCflU = 8;

__global__ void __launch_bounds__(kCflT) k_cfl_reduce(const double *blk, int n, double *out)
{
    double m = 0.;
    for (int i0 = threadIdx.x; i0 < n; i0 += kCflT * kCflU) {
        double v[kCflU];
#pragma unroll
        for (int u = 0; u < kCflU; u++) {
            const int i = i0 + u * kCflT;
            v[u] = i < n ? blk[i] : 0.;
        }
#pragma unroll
        for (int u = 0; u < kCflU; u++)
            m = fmax(m, v[u]);
    }
    for (int off = 32; off > 0; off >>= 1)
        m = fmax(m, __shfl_xor(m, off));
    __shared__ double wmax[kCflT / 64];
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

// ---------------------------------------------------------------------
// Size-specialised versions: the transform length is a compile-time
// constant (radix schedule unrolled, index arithmetic by constants), the
// twiddle table is staged in LDS and the z kernels pad each transform to
// N+1 entries so that the column gather/scatter does not hit one LDS bank.
// ---------------------------------------------------------------------
constexpr int kLdsCap = 65536;

// columns per z block: power of two <= 16 whose two buffers and the
// twiddle table fit in kLdsCap (at least 1)
constexpr int zb_of(int n)
{
    int B = 16;
    while (B > 1 && (2 * B * (n + 1) + n) * 16 > kLdsCap)
        B /= 2;
    return B;
}

// rows per x block (six fields, two buffers, the length-2m twiddles)
constexpr int rw_of(int m)
{
    int RW = 8;
    while (RW > 1 && (2 * 6 * RW * m + 2 * m) * 16 > kLdsCap)
        RW--;
    return RW;
}

template <int NZP>
__global__ void __launch_bounds__(256) k_zsyn_t(NonlinArgs a, const double2 *__restrict__ twg)
{
    constexpr int B = zb_of(NZP), SB = NZP + 1;
    extern __shared__ double2 lds[];
    double2 *tw = lds, *buf0 = lds + NZP, *buf1 = buf0 + B * SB;
    const int nz = a.nz, nxh = a.nxh, ny = a.ny;
    const int f = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    for (int i = threadIdx.x; i < NZP; i += blockDim.x)
        tw[i] = twg[i];
    for (int idx = threadIdx.x; idx < NZP * B; idx += blockDim.x) {
        const int zp = idx / B, b = idx - zp * B;
        double2 v = make_double2(0., 0.);
        const int z = zsrc(zp, nz, NZP);
        if (b < nb && z >= 0) {
            const int x = x0 + b;
            if (f == 4) {
                const double2 u = a.U[uidx(0, z, y, x, nz, ny, nxh)];
                const double2 w = a.U[uidx(2, z, y, x, nz, ny, nxh)];
                const double be = a.beta[z], al = a.alfa[x];
                v = make_double2(-be * u.y + al * w.y, be * u.x - al * w.x);
            } else {
                v = a.U[uidx(f < 4 ? f : 4, z, y, x, nz, ny, nxh)];
            }
        }
        buf0[b * SB + zp] = v;
    }
    __syncthreads();
    double2 *res = fft_t<NZP, NZP, SB, +1>(buf0, buf1, B, tw);
    double2 *P = a.P + (((size_t)f * ny + y) * NZP) * nxh + x0;
    for (int idx = threadIdx.x; idx < NZP * B; idx += blockDim.x) {
        const int zp = idx / B, b = idx - zp * B;
        if (b < nb)
            P[(size_t)zp * nxh + b] = res[b * SB + zp];
    }
}

template <int NZP>
__global__ void __launch_bounds__(256) k_zana_t(NonlinArgs a, const double2 *__restrict__ twg)
{
    constexpr int B = zb_of(NZP), SB = NZP + 1;
    extern __shared__ double2 lds[];
    double2 *tw = lds, *buf0 = lds + NZP, *buf1 = buf0 + B * SB;
    const int nz = a.nz, nxh = a.nxh, ny = a.ny;
    const int h = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    for (int i = threadIdx.x; i < NZP; i += blockDim.x)
        tw[i] = twg[i];
    const double2 *P = a.P + (((size_t)h * ny + y) * NZP) * nxh + x0;
    for (int idx = threadIdx.x; idx < NZP * B; idx += blockDim.x) {
        const int zp = idx / B, b = idx - zp * B;
        buf0[b * SB + zp] = b < nb ? P[(size_t)zp * nxh + b] : make_double2(0., 0.);
    }
    __syncthreads();
    double2 *res = fft_t<NZP, NZP, SB, -1>(buf0, buf1, B, tw);
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        if (b < nb) {
            const int zp = z < nz / 2 ? z : z + NZP - nz;
            a.Uw[uidx(h, z, y, x0 + b, nz, ny, nxh)] = res[b * SB + zp];
        }
    }
}

template <int M>
__global__ void __launch_bounds__(256) k_xrow_t(NonlinArgs a, const double2 *__restrict__ twg)
{
    constexpr int RW = rw_of(M);
    extern __shared__ double2 lds[];
    double2 *tw = lds;                        // exp(+2 pi i t/(2M)), t < 2M
    double2 *A = lds + 2 * M, *Bf = A + 6 * RW * M;
    const int nxh = a.nxh, nzp = a.nzp, ny = a.ny;
    const int zp0 = blockIdx.x * RW, y = blockIdx.y;
    const int nr = min(RW, nzp - zp0);
    for (int i = threadIdx.x; i < 2 * M; i += blockDim.x)
        tw[i] = twg[i];
    __syncthreads();
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
                Xk.y = 0.;
            const int km = M - k;
            double2 Xm = (k > 0 && km < nxh) ? row[km] : make_double2(0., 0.);
            Xm.y = -Xm.y;
            const double2 s = cadd(Xk, Xm), d = csub(Xk, Xm);
            Z = cadd(s, mul_is(cmul(tw[k], d), 1.0));
        }
        A[(f * RW + r) * M + k] = Z;
    }
    __syncthreads();
    double2 *res = fft_t<M, 2 * M, M, +1>(A, Bf, 6 * RW, tw);
    double2 *oth = (res == A) ? Bf : A;
    double cflmax = 0.;
    double dyinv = 1.0 / a.deta[y];
    if (a.ny >= 17 && !a.nocap)
        dyinv = fmin(dyinv, (double)(a.ny - 1) * (double)(a.ny - 1) / 10. / 3.14);
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
            cflmax = cfl_max(cfl_max(cflmax, ce), co);
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
    double2 *fin = fft_t<M, 2 * M, M, -1>(res, oth, 3 * RW, tw);
    for (int idx = threadIdx.x; idx < 3 * RW * nxh; idx += blockDim.x) {
        const int h = idx / (RW * nxh);
        const int rem = idx - h * RW * nxh;
        const int r = rem / nxh, k = rem - r * nxh;
        if (r < nr) {
            const double2 Zk = fin[(h * RW + r) * M + k];
            const double2 Zm = conjd(fin[(h * RW + r) * M + (k == 0 ? 0 : M - k)]);
            const double2 E = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y + Zm.y));
            const double2 dd = csub(Zk, Zm);
            const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);
            const double2 X = cadd(E, cmul(conjd(tw[k]), O));
            a.P[(((size_t)h * ny + y) * nzp + zp0 + r) * nxh + k] = X;
        }
    }
    if (a.do_cfl) {
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

template <int NZP>
void zsyn_t(const NonlinArgs &a, const double2 *tw, hipStream_t s)
{
    constexpr int B = zb_of(NZP);
    dim3 grid((a.nxh + B - 1) / B, 6, a.ny);
    const size_t lds = (size_t)(2 * B * (NZP + 1) + NZP) * sizeof(double2);
    hipLaunchKernelGGL(k_zsyn_t<NZP>, grid, dim3(256), lds, s, a, tw);
}

template <int NZP>
void zana_t(const NonlinArgs &a, const double2 *tw, hipStream_t s)
{
    constexpr int B = zb_of(NZP);
    dim3 grid((a.nxh + B - 1) / B, 3, a.ny);
    const size_t lds = (size_t)(2 * B * (NZP + 1) + NZP) * sizeof(double2);
    hipLaunchKernelGGL(k_zana_t<NZP>, grid, dim3(256), lds, s, a, tw);
}

template <int M>
void xrow_t(const NonlinArgs &a, const double2 *tw, hipStream_t s)
{
    constexpr int RW = rw_of(M);
    dim3 grid((a.nzp + RW - 1) / RW, a.ny);
    const size_t lds = (size_t)(2 * 6 * RW * M + 2 * M) * sizeof(double2);
    hipLaunchKernelGGL(k_xrow_t<M>, grid, dim3(256), lds, s, a, tw);
}

// lengths with a specialised kernel whose buffers fit the LDS cap
#define SG_ZLENS(X) X(12) X(24) X(48) X(72) X(96) X(192) X(576)
#define SG_XLENS(X) X(12) X(24) X(48) X(96) X(192)

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
    switch (a.nzp) {
#define SG_CASE(N)                     \
    case N:                            \
        zsyn_t<N>(a, pz.tw, s);        \
        return;
        SG_ZLENS(SG_CASE)
#undef SG_CASE
    default:
        break;
    }
    const int B = zblock(a.nxh, a.nzp);
    dim3 grid((a.nxh + B - 1) / B, 6, a.ny);
    const size_t lds = 2 * (size_t)B * a.nzp * sizeof(double2);
    hipLaunchKernelGGL(k_zsyn, grid, dim3(256), lds, s, a, pz, B);
}

int nonlin_prepare(int lds_max)
{
    const void *fns[] = {(const void *)k_zsyn, (const void *)k_zana, (const void *)k_xrow};
    for (const void *fn : fns)
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds_max) !=
            hipSuccess)
            return -1;
    return 0;
}

int xrow_rows(const NonlinArgs &a)
{
    switch (a.nxp / 2) {
#define SG_CASE(N)                     \
    case N:                            \
        return rw_of(N);
        SG_XLENS(SG_CASE)
#undef SG_CASE
    default:
        return xrows(a.nxp, a.nzp);
    }
}

int xrow_blocks(const NonlinArgs &a)
{
    const int RW = xrow_rows(a);
    return ((a.nzp + RW - 1) / RW) * a.ny;
}

void launch_xrow(const NonlinArgs &a, const FftPlan &pm, const double2 *wx, hipStream_t s)
{
    switch (a.nxp / 2) {
#define SG_CASE(N)                     \
    case N:                            \
        xrow_t<N>(a, wx, s);           \
        return;
        SG_XLENS(SG_CASE)
#undef SG_CASE
    default:
        break;
    }
    const int RW = xrows(a.nxp, a.nzp);
    dim3 grid((a.nzp + RW - 1) / RW, a.ny);
    const size_t lds = 2 * (size_t)6 * RW * (a.nxp / 2) * sizeof(double2);
    hipLaunchKernelGGL(k_xrow, grid, dim3(256), lds, s, a, pm, wx, RW);
}

void launch_zana(const NonlinArgs &a, const FftPlan &pz, hipStream_t s)
{
    switch (a.nzp) {
#define SG_CASE(N)                     \
    case N:                            \
        zana_t<N>(a, pz.tw, s);        \
        return;
        SG_ZLENS(SG_CASE)
#undef SG_CASE
    default:
        break;
    }
    const int B = zblock(a.nxh, a.nzp);
    dim3 grid((a.nxh + B - 1) / B, 3, a.ny);
    const size_t lds = 2 * (size_t)B * a.nzp * sizeof(double2);
    hipLaunchKernelGGL(k_zana, grid, dim3(256), lds, s, a, pz, B);
}

void launch_cfl_reduce(const double *blk, int n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_cfl_reduce, dim3(1), dim3(kCflT), 0, s, blk, n, out);
}

}  // namespace sg
