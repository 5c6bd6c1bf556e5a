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

}  // namespace

void launch_amp(const double2 *U, const double *alfa, const double *beta, int nz, int ny, int nxh,
                double *amp, hipStream_t s)
{
    hipLaunchKernelGGL(k_amp, dim3(ny), dim3(kAmpT), 0, s, U, alfa, beta, nz, ny, nxh, amp);
}

}  // namespace sg
