// Plane statistics of bla's amplitude file (boxamp.f, called per y plane
// from nonlinbl.f:430): per-y sums over the Fourier modes of the stored
// velocities and vorticities, one workgroup per y plane.
//
// With weight w = 2 for kx > 0 and 1 for kx = 0 (the kx < 0 half is the
// conjugate), over all modes except kx = kz = 0 and the z oddball kz = -nz/2:
//   amp(1..3)  sum w |u|^2, |v|^2, |w|^2
//   amp(4..6)  sum w |om_x|^2, |om_y|^2, |om_z|^2   (om_y = i beta u - i alpha w)
//   amp(7)     sum w |om_y|^2 / (alpha^2 + beta^2)
//   amp(8)     sum w Re(u conj v)
//   amp(14,15) sum w Re(u conj w), Re(v conj w)
//   amp(16..18) sum w |u|^2, |v|^2, |w|^2
// and from the (0,0) mode amp(9..12) = Re u, Re w, Re om_x, Re om_z.
// amp(13), amp(19), amp(20) are left zero (the writer fills amp(13)).
//
// The x-y statistics of boxxys.f (called per y plane from bla.f:2089-2105)
// on the unpadded nx x nz grid, two kernels per call:
//   k_xys_z  gather u, v, w, om_x, om_y, om_z of a y plane (getxz.f with
//            ipad = 0: the z oddball zeroed), subtract the lower-wall
//            velocity from the mean mode and apply the frame shift
//            exp(-i (alpha xs + beta zs)) (boxxys.f:382-412, xzsh.f), then
//            the inverse complex z transform of length nz (vcfftb) in LDS;
//   k_xys_x  per y plane: the inverse real x transform of length nx
//            (vrfftb, packed as nx/2 complex points) of every z row and
//            the z sums of the one-point statistics 1-15 (boxxys.f:540-636)
//            u_i, u_i^2, om_i, om_i^2, uv, uw, vw, weighted by dtn/nz.
#include <algorithm>

#include "sg_fft.h"
#include "sg_internal.h"

namespace sg {

namespace {

constexpr int kAmpT = 256;

__global__ void __launch_bounds__(kAmpT) k_amp(const double2 *__restrict__ U,
                                               const double *__restrict__ alfa,
                                               const double *__restrict__ beta, int nz, int ny,
                                               int nxh, double *__restrict__ amp)
{
    const int y = blockIdx.x, tid = threadIdx.x;
    const size_t F = (size_t)nz * ny * nxh;
    double s[10];
    for (int i = 0; i < 10; i++)
        s[i] = 0.;
    for (int m = tid; m < nz * nxh; m += kAmpT) {
        const int z = m / nxh, x = m - z * nxh;
        if (z == nz / 2 || (x == 0 && z == 0))
            continue;
        const size_t o = ((size_t)z * ny + y) * nxh + x;
        const double2 u = U[o], v = U[F + o], w = U[2 * F + o];
        const double2 ox = U[3 * F + o], oz = U[4 * F + o];
        const double al = alfa[x], be = beta[z];
        const double oyr = -be * u.y + al * w.y, oyi = be * u.x - al * w.x;
        const double wt = x > 0 ? 2. : 1.;
        const double uu = u.x * u.x + u.y * u.y, vv = v.x * v.x + v.y * v.y;
        const double ww = w.x * w.x + w.y * w.y, oy2 = oyr * oyr + oyi * oyi;
        s[0] += wt * uu;
        s[1] += wt * vv;
        s[2] += wt * ww;
        s[3] += wt * (ox.x * ox.x + ox.y * ox.y);
        s[4] += wt * oy2;
        s[5] += wt * (oz.x * oz.x + oz.y * oz.y);
        s[6] += wt * oy2 / (al * al + be * be);
        s[7] += wt * (u.x * v.x + u.y * v.y);
        s[8] += wt * (u.x * w.x + u.y * w.y);
        s[9] += wt * (v.x * w.x + v.y * w.y);
    }
    __shared__ double red[kAmpT / 64][10];
    const int lane = tid & 63, wv = tid >> 6;
    for (int i = 0; i < 10; i++) {
        double a = s[i];
        for (int off = 32; off > 0; off >>= 1)
            a += __shfl_xor(a, off);
        if (lane == 0)
            red[wv][i] = a;
    }
    __syncthreads();
    if (tid == 0) {
        double t[10];
        for (int i = 0; i < 10; i++) {
            t[i] = 0.;
            for (int j = 0; j < kAmpT / 64; j++)
                t[i] += red[j][i];
        }
        double *a = amp;                         // [20][ny], column i holds amp(., i+1)
        for (int i = 0; i < 6; i++)
            a[i * ny + y] = t[i];
        a[6 * ny + y] = t[6];
        a[7 * ny + y] = t[7];
        const size_t o0 = (size_t)y * nxh;       // the (0,0) mode
        a[8 * ny + y] = U[o0].x;
        a[9 * ny + y] = U[2 * F + o0].x;
        a[10 * ny + y] = U[3 * F + o0].x;
        a[11 * ny + y] = U[4 * F + o0].x;
        a[12 * ny + y] = 0.;
        a[13 * ny + y] = t[8];
        a[14 * ny + y] = t[9];
        a[15 * ny + y] = t[0];
        a[16 * ny + y] = t[1];
        a[17 * ny + y] = t[2];
        a[18 * ny + y] = 0.;
        a[19 * ny + y] = 0.;
    }
}

constexpr int kXysT = 256;
constexpr int kXysLds = 2048;      // complex values per LDS buffer (two buffers)
constexpr int kXysN = 15;

// field f of plane y (0-2 velocities, 3 om_x, 4 om_y, 5 om_z) into
// Q [6][ny][nz][nxh], transformed to physical z
__global__ void __launch_bounds__(kXysT) k_xys_z(XysArgs a, FftPlan pz, int B)
{
    extern __shared__ double2 lds[];
    const int nz = a.nz, nxh = a.nxh, ny = a.ny;
    const size_t F = (size_t)nz * ny * nxh;
    double2 *buf0 = lds, *buf1 = lds + (size_t)B * nz;
    const int f = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        double2 v = make_double2(0., 0.);
        if (b < nb && z != nz / 2) {
            const int x = x0 + b;
            const size_t o = ((size_t)z * ny + y) * nxh + x;
            const double al = a.alfa[x], be = a.beta[z];
            if (f == 4) {
                const double2 u = a.U[o], w = a.U[2 * F + o];
                v = make_double2(-be * u.y + al * w.y, be * u.x - al * w.x);
            } else {
                v = a.U[(size_t)(f < 4 ? f : 4) * F + o];
                if (x == 0 && z == 0) {
                    if (f == 0)
                        v.x -= a.u0low;
                    if (f == 2)
                        v.x -= a.w0low;
                }
            }
            if (a.shift) {
                double sn, cs;
                sincos(-a.xs * al - a.zs * be, &sn, &cs);
                v = make_double2(v.x * cs - v.y * sn, v.y * cs + v.x * sn);
            }
        }
        buf0[b * nz + z] = v;
    }
    __syncthreads();
    double2 *res = lds_fft(buf0, buf1, nz, B, pz, +1.0);
    double2 *Q = a.Q + (((size_t)f * ny + y) * nz) * nxh + x0;
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        if (b < nb)
            Q[(size_t)z * nxh + b] = res[b * nz + z];
    }
}

// one y plane per block: rows of RW z values at a time through the packed
// inverse x transform, then the z sums; G groups of M threads split the
// rows and are combined in a fixed order at the end
__global__ void __launch_bounds__(kXysT) k_xys_x(XysArgs a, FftPlan pm,
                                                 const double2 *__restrict__ wx, int RW)
{
    extern __shared__ double2 lds[];
    const int M = a.nxh, nz = a.nz, ny = a.ny, y = blockIdx.x;
    double2 *A = lds, *Bf = lds + (size_t)6 * RW * M;
    const int G = kXysT / M;
    const int tid = threadIdx.x, n = tid % M, grp = tid / M;
    double acc[kXysN][2];
#pragma unroll
    for (int i = 0; i < kXysN; i++)
        acc[i][0] = acc[i][1] = 0.;
    for (int z0 = 0; z0 < nz; z0 += RW) {
        const int nr = min(RW, nz - z0);
        // half spectrum X_k, k < nx/2 -> packed Z_k (vrfftb: only Re X_0,
        // no Nyquist term)
        for (int idx = tid; idx < 6 * RW * M; idx += blockDim.x) {
            const int f = idx / (RW * M);
            const int rem = idx - f * RW * M;
            const int r = rem / M, k = rem - r * M;
            double2 Z = make_double2(0., 0.);
            if (r < nr) {
                const double2 *row = a.Q + (((size_t)f * ny + y) * nz + z0 + r) * M;
                double2 Xk = row[k];
                if (k == 0)
                    Xk.y = 0.;
                double2 Xm = k > 0 ? row[M - k] : make_double2(0., 0.);
                Xm.y = -Xm.y;
                const double2 sm = cadd(Xk, Xm), d = csub(Xk, Xm);
                Z = cadd(sm, mul_is(cmul(wx[k], d), 1.0));
            }
            A[(f * RW + r) * M + k] = Z;
        }
        __syncthreads();
        const double2 *res = lds_fft(A, Bf, M, 6 * RW, pm, +1.0);
        if (grp < G) {
            for (int r = grp; r < nr; r += G) {
                double2 q[6];
#pragma unroll
                for (int f = 0; f < 6; f++)
                    q[f] = res[(f * RW + r) * M + n];
#pragma unroll
                for (int e = 0; e < 2; e++) {
                    double v[6];
#pragma unroll
                    for (int f = 0; f < 6; f++)
                        v[f] = e ? q[f].y : q[f].x;
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        acc[i][e] += v[i];
                        acc[3 + i][e] += v[i] * v[i];
                        acc[6 + i][e] += v[3 + i];
                        acc[9 + i][e] += v[3 + i] * v[3 + i];
                    }
                    acc[12][e] += v[0] * v[1];
                    acc[13][e] += v[0] * v[2];
                    acc[14][e] += v[1] * v[2];
                }
            }
        }
        __syncthreads();
    }
    // fixed-order combination of the groups through LDS
    double *red = (double *)lds;
    if (grp < G) {
#pragma unroll
        for (int i = 0; i < kXysN; i++) {
            red[((size_t)grp * kXysN + i) * 2 * M + 2 * n] = acc[i][0];
            red[((size_t)grp * kXysN + i) * 2 * M + 2 * n + 1] = acc[i][1];
        }
    }
    __syncthreads();
    const int nx = 2 * M;
    for (int idx = tid; idx < kXysN * nx; idx += blockDim.x) {
        const int i = idx / nx, x = idx - i * nx;
        double t = 0.;
        for (int gq = 0; gq < G; gq++)
            t += red[((size_t)gq * kXysN + i) * nx + x];
        a.out[((size_t)i * ny + y) * nx + x] = t * a.c1;
    }
}

}  // namespace

int xys_zblock(int nz, int nxh)
{
    return max(1, min(nxh, kXysLds / nz));
}

int xys_rows(int nz, int nxh)
{
    return max(1, min(nz, kXysLds / (6 * nxh)));
}

bool xys_fits(int nz, int nxh)
{
    // one z column and six x rows in a buffer, at least one thread per
    // packed x point (the group sums take at most kXysT * 2 * kXysN doubles,
    // within the two buffers)
    return nz <= kXysLds && 6 * nxh <= kXysLds && nxh <= kXysT;
}

void launch_xys(const XysArgs &a, const FftPlan &pz, const FftPlan &pm, const double2 *wx,
                hipStream_t s)
{
    const int B = xys_zblock(a.nz, a.nxh);
    const size_t lz = (size_t)2 * B * a.nz * sizeof(double2);
    hipLaunchKernelGGL(k_xys_z, dim3((a.nxh + B - 1) / B, 6, a.ny), dim3(kXysT), lz, s, a, pz,
                       B);
    const int RW = xys_rows(a.nz, a.nxh);
    const size_t lx = std::max((size_t)2 * 6 * RW * a.nxh * sizeof(double2),
                               (size_t)(kXysT / a.nxh) * kXysN * 2 * a.nxh * sizeof(double));
    hipLaunchKernelGGL(k_xys_x, dim3(a.ny), dim3(kXysT), lx, s, a, pm, wx, RW);
}

void launch_amp(const double2 *U, const double *alfa, const double *beta, int nz, int ny, int nxh,
                double *amp, hipStream_t s)
{
    hipLaunchKernelGGL(k_amp, dim3(ny), dim3(kAmpT), 0, s, U, alfa, beta, nz, ny, nxh, amp);
}

}  // namespace sg
