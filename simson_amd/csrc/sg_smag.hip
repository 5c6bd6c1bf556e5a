// Eddy-viscosity LES of bla (sgs.i iles = 2, les.f with iwale = 0, no
// scalar): the stresses tau_ij = -2 nu_t S_ij of the Smagorinsky model with
// a fixed coefficient (C = cs^2) or the dynamic procedure (Germano/Lilly:
// C(x, y) = 1/2 <M_ij L_ij>_z / <M_ij M_ij>_z with the 2-D spectral cutoff
// test filter of filter_cutoff.f), on the unpadded nx x nz grid of every y
// plane, as les.f does (called per plane from bla.f:1945-1998 at the start
// of each substep).
//
// sgs.i iles = 3 (the HPF model, les.f:180-215, 404-620, 883-1380) is the
// same pipeline with the strain of the high-pass filtered velocities and
// vorticities u - lpf L u that linearbl.f:1871-2066 stores (L the ADM-RT
// wall-normal low-pass, lpf its x-z transfer function; L u comes from a
// y-transform product over the five slots before the pipeline), no
// u0low/w0low shift, and for the dynamic procedure B_ij from the secondary
// filter u - lpf2 L u of linearbl.f:2067-2256 (the x-z filter of a grid
// with half the wavenumbers, bla.f:1520-1543).
//
// Kernels (all y planes at once):
//   k_smag_z   u, v, w and the nine velocity gradients (les.f:278-350, from
//              the stored velocities and om_x, om_z) synthesised in z
//              (vcfftb, length nz) into Q [12][ny][nz][nx/2]; with the
//              dynamic HPF model also the nine gradients of the secondary
//              filtered fields (Q [21][...]);
//   k_smag_xa  per row: inverse real x transform (vrfftb), S_ij and |S|
//              (les.f:627-645) kept at physical x-z in SS [7][ny][nz][nx/2]
//              (even points in .x, odd in .y); for the dynamic procedure the
//              21 fields to be test-filtered (-u_i u_j, -S_ij |S| Delta^2,
//              S_ij, u_i: les.f:717-780), forward x transform and the x
//              cutoff kx < nx/4, into W [21][ny][nz][nx/4];
//   k_smag_zf  the z part of the test filter: forward z transform, cutoff
//              |kz| <= nz/4, inverse, scaling 1/(nx nz) (filter_cutoff.f);
//   k_smag_xc1 per block of z rows: inverse x transform of the filtered
//              fields, L_ij, M_ij and the partial z sums of M.L and M.M
//              (les.f:784-880, 1380-1416);
//   k_smag_c   per y plane: the clipped coefficient C(x) from the partial
//              sums, added in a fixed order (les.f:1418-1510);
//   k_smag_xc2 per block of z rows: nu_t and tau_ij / (nx nz) with the
//              forward x transform (les.f:1530-1620), into Q as
//              [6][ny][nz][nx/2];
//   k_smag_za  forward z transform (vcfftf), the oddball row zeroed
//              (les.f:1630-1645), scaled by nxp nzp to the normalisation of
//              the step's nonlinear term, into TAU [6][nz][ny][nx/2].
// The force sigma_i = d tau_ij/dx_j (sgs_stress.f) and its subtraction
// from H are the learned closure's path (the D-matrix product and k_zana4).
#include <algorithm>

#include "sg_fft.h"
#include "sg_internal.h"

namespace sg {

namespace {

constexpr int kT = 256;
constexpr int kLdsZ = 2048;        // complex values per LDS buffer of the z kernels
constexpr int kMaxNx = 512;
constexpr int kIB = 4;             // items per thread whose global loads issue together

__device__ __forceinline__ size_t qidx(int f, int y, int z, int x, int ny, int nz, int nx)
{
    return (((size_t)f * ny + y) * nz + z) * nx + x;
}

__device__ __forceinline__ size_t sidx(int s, int z, int y, int x, int nz, int ny, int nxh)
{
    return (((size_t)s * nz + z) * ny + y) * nxh + x;
}

// |S| = sqrt(2 S_ij S_ij) for the order 11 12 13 22 23 33
__device__ __forceinline__ double sabs_of(const double *s)
{
    return sqrt(2. * (s[0] * s[0] + s[3] * s[3] + s[5] * s[5]) +
                4. * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4]));
}

// S_ij (order 11 12 13 22 23 33) of the gradients g[3 i + j] = du_i/dx_j
__device__ __forceinline__ void strain_of(const double *g, double *s)
{
    s[0] = g[0];
    s[1] = .5 * (g[3] + g[1]);
    s[2] = .5 * (g[6] + g[2]);
    s[3] = g[4];
    s[4] = .5 * (g[5] + g[7]);
    s[5] = g[8];
}

// the plan with its twiddle table copied to LDS at dst (tabn values)
__device__ FftPlan plan_in_lds(const FftPlan &p, double2 *dst)
{
    for (int i = threadIdx.x; i < p.tabn; i += blockDim.x)
        dst[i] = p.tw[i];
    __syncthreads();
    FftPlan q = p;
    q.tw = dst;
    return q;
}

__global__ void __launch_bounds__(kT) k_smag_z(SmagArgs a, FftPlan pz, int B)
{
    extern __shared__ double2 lds[];
    const int nz = a.nz, nxh = a.nxh, ny = a.ny;
    const size_t F = (size_t)nz * ny * nxh;
    double2 *buf0 = lds, *buf1 = lds + (size_t)B * nz;
    const FftPlan pl = plan_in_lds(pz, lds + (size_t)2 * B * nz);
    const int fq = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    // output fields 12..20 are the gradients 3..11 of the secondary filtered
    // fields; sel 0 reads the stored fields, 1 the primary high-pass filtered
    // and 2 the secondary ones
    const int f = fq < 12 ? fq : fq - 9;
    const int sel = fq < 3 ? 0 : (fq < 12 ? (a.hpf ? 1 : 0) : 2);
    const double *lp = sel == 1 ? a.lpf : a.lpf2;
    // the fields this block reads: fa, and fb for the y derivatives of u and
    // w (p1 is loaded from fa again where it is not needed)
    int fa = f, fb = f;
    if (f >= 3) {
        const int i = (f - 3) / 3, j = (f - 3) - 3 * i;
        fa = fb = i;
        if (j == 1) {
            fa = i == 0 ? 1 : (i == 1 ? 0 : 3);
            fb = i == 0 ? 4 : (i == 1 ? 2 : 1);
        }
    }
    const int tot = nz * B;
    for (int base = threadIdx.x; base < tot; base += kIB * kT) {
        double2 p0[kIB], p1[kIB];
        double al[kIB], be[kIB];
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = min(base + i * kT, tot - 1);
            const int z = idx / B, x = min(x0 + idx - z * B, nxh - 1);
            const size_t o = sidx(0, z, y, x, nz, ny, nxh);
            p0[i] = a.U[fa * F + o];
            p1[i] = a.U[fb * F + o];
            if (sel) {
                const double l = lp[(size_t)z * nxh + x];
                const double2 w0 = a.LW[fa * F + o], w1 = a.LW[fb * F + o];
                p0[i] = make_double2(p0[i].x - l * w0.x, p0[i].y - l * w0.y);
                p1[i] = make_double2(p1[i].x - l * w1.x, p1[i].y - l * w1.y);
            }
            al[i] = a.alfa[x];
            be[i] = a.beta[z];
        }
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = base + i * kT;
            if (idx >= tot)
                continue;
            const int z = idx / B, b = idx - z * B;
            double2 v = make_double2(0., 0.);
            if (b < nb && z != nz / 2) {
                const int x = x0 + b;
                if (f < 3) {
                    v = p0[i];
                    if (x == 0 && z == 0 && !a.hpf) {
                        if (f == 0)
                            v.x -= a.u0low;
                        if (f == 2)
                            v.x -= a.w0low;
                    }
                } else {
                    const int ii = (f - 3) / 3, j = (f - 3) - 3 * ii;
                    if (j != 1) {
                        const double k = j == 0 ? al[i] : be[i];
                        const double2 w = p0[i];
                        v = make_double2(-k * w.y, k * w.x);
                    } else if (ii == 0) {       // du/dy = i alpha v - om_z
                        const double2 vv = p0[i], oz = p1[i];
                        v = make_double2(-al[i] * vv.y - oz.x, al[i] * vv.x - oz.y);
                    } else if (ii == 1) {       // dv/dy = -i alpha u - i beta w
                        const double2 uu = p0[i], ww = p1[i];
                        v = make_double2(al[i] * uu.y + be[i] * ww.y,
                                         -al[i] * uu.x - be[i] * ww.x);
                    } else {                    // dw/dy = om_x + i beta v
                        const double2 ox = p0[i], vv = p1[i];
                        v = make_double2(ox.x - be[i] * vv.y, ox.y + be[i] * vv.x);
                    }
                }
            }
            buf0[b * nz + z] = v;
        }
    }
    __syncthreads();
    const double2 *res = lds_fft(buf0, buf1, nz, B, pl, +1.0);
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        if (b < nb)
            a.Q[qidx(fq, y, z, x0 + b, ny, nz, nxh)] = res[b * nz + z];
    }
}

// half spectrum row (k < nk given, zero above) -> packed Z_k for the
// inverse real transform of length 2M (only Re X_0, no Nyquist term).
// pack_load reads X_k and X_{M-k} through clamped addresses (so that the
// loads of several items can issue before any is used); pack_vals forms Z_k.
__device__ __forceinline__ void pack_load(const double2 *row, int k, int M, int nk,
                                          double2 &Xk, double2 &Xm)
{
    const int km = M - k;
    const double2 p = row[min(k, nk - 1)], q = row[min(km, nk - 1)];
    const double2 z = make_double2(0., 0.);
    Xk = k < nk ? p : z;
    Xm = (k > 0 && km < nk) ? q : z;
}

__device__ __forceinline__ double2 pack_vals(double2 Xk, double2 Xm, int k, const double2 *wx)
{
    if (k == 0)
        Xk.y = 0.;
    Xm.y = -Xm.y;
    return cadd(cadd(Xk, Xm), mul_is(cmul(wx[k], csub(Xk, Xm)), 1.0));
}

// NF packed rows [f][RW][M] of a [f][ny][nz][nk] spectrum into A (rows
// r >= nr zero); the global loads of kIB items per thread issue together
__device__ __forceinline__ void pack_rows(double2 *A, const double2 *src, int NF, int RW, int nr,
                                          int y, int z0, int ny, int nz, int M, int nk,
                                          const double2 *wx)
{
    const int tot = NF * RW * M;
    for (int base = threadIdx.x; base < tot; base += kIB * kT) {
        double2 xk[kIB], xm[kIB];
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = min(base + i * kT, tot - 1);
            const int f = idx / (RW * M);
            const int rem = idx - f * RW * M;
            const int r = rem / M, k = rem - r * M;
            pack_load(src + qidx(f, y, min(z0 + r, nz - 1), 0, ny, nz, nk), k, M, nk, xk[i],
                      xm[i]);
        }
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = base + i * kT;
            if (idx < tot) {
                const int f = idx / (RW * M);
                const int rem = idx - f * RW * M;
                const int r = rem / M, k = rem - r * M;
                A[(f * RW + r) * M + k] =
                    r < nr ? pack_vals(xk[i], xm[i], k, wx) : make_double2(0., 0.);
            }
        }
    }
}

// packed forward transform result -> X_k, k < M
__device__ __forceinline__ double2 unpack_row(const double2 *fin, int k, int M,
                                              const double2 *wx)
{
    const double2 Zk = fin[k];
    const double2 Zm = conjd(fin[k == 0 ? 0 : M - k]);
    const double2 E = make_double2(0.5 * (Zk.x + Zm.x), 0.5 * (Zk.y + Zm.y));
    const double2 dd = csub(Zk, Zm);
    const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);
    return cadd(E, cmul(conjd(wx[k]), O));
}

__global__ void __launch_bounds__(kT) k_smag_xa(SmagArgs a, FftPlan pm, int RW)
{
    extern __shared__ double2 lds[];
    const int M = a.nxh, nz = a.nz, ny = a.ny, nq = a.nq;
    const int NF = a.dynamic ? 21 : 12;      // capacity of each buffer in rows of fields
    double2 *A = lds, *Bf = lds + (size_t)NF * RW * M;
    // the x twiddles exp(+2 pi i t / nx) of the plan are also the half-angle
    // factors of the real-transform packing; both are read from LDS
    const FftPlan pl = plan_in_lds(pm, lds + (size_t)2 * NF * RW * M);
    const int z0 = blockIdx.x * RW, y = blockIdx.y;
    const int nr = min(RW, nz - z0), NQ = a.nqf;
    pack_rows(A, a.Q, NQ, RW, nr, y, z0, ny, nz, M, M, pl.tw);
    __syncthreads();
    double2 *res = lds_fft(A, Bf, M, NQ * RW, pl, +1.0);
    double2 *oth = res == A ? Bf : A;
    const double d2 = a.dlt2[y];
    for (int idx = threadIdx.x; idx < RW * M; idx += blockDim.x) {
        const int r = idx / M, n = idx - r * M;
        double2 q[12];
#pragma unroll
        for (int f = 0; f < 12; f++)
            q[f] = res[(f * RW + r) * M + n];
        // even (e = 0) and odd (e = 1) physical points of the packed row
        double o[2][21], sv[2][7];
#pragma unroll
        for (int e = 0; e < 2; e++) {
            double u[3], g[9];
#pragma unroll
            for (int i = 0; i < 3; i++)
                u[i] = e ? q[i].y : q[i].x;
#pragma unroll
            for (int i = 0; i < 9; i++)
                g[i] = e ? q[3 + i].y : q[3 + i].x;
            // g[3 i + j] = du_i/dx_j
            double *sk = sv[e];
            strain_of(g, sk);
            sk[6] = sabs_of(sk);
            const double dr = sk[6] * d2;
            // B_ij: the strain itself, or that of the secondary filtered
            // fields (dynamic HPF model)
            double sb[6];
            if (NQ == 21) {
#pragma unroll
                for (int i = 0; i < 9; i++) {
                    const double2 t = res[((12 + i) * RW + r) * M + n];
                    g[i] = e ? t.y : t.x;
                }
                strain_of(g, sb);
            } else {
#pragma unroll
                for (int k = 0; k < 6; k++)
                    sb[k] = sk[k];
            }
            const int pa[6] = {0, 1, 2, 1, 2, 2}, pb[6] = {0, 0, 0, 1, 1, 2};
#pragma unroll
            for (int k = 0; k < 6; k++) {
                o[e][k] = -u[pa[k]] * u[pb[k]];
                o[e][6 + k] = -sk[k] * dr;
                o[e][12 + k] = sb[k];
            }
#pragma unroll
            for (int i = 0; i < 3; i++)
                o[e][18 + i] = u[i];
        }
        if (r < nr) {
#pragma unroll
            for (int k = 0; k < 7; k++)
                a.SS[qidx(k, y, z0 + r, n, ny, nz, M)] = make_double2(sv[0][k], sv[1][k]);
        }
        if (a.dynamic) {
#pragma unroll
            for (int f = 0; f < 21; f++)
                oth[(f * RW + r) * M + n] = make_double2(o[0][f], o[1][f]);
        }
    }
    if (!a.dynamic)
        return;
    __syncthreads();
    double2 *fin = lds_fft(oth, res, M, 21 * RW, pl, -1.0);
    for (int idx = threadIdx.x; idx < 21 * RW * nq; idx += blockDim.x) {
        const int f = idx / (RW * nq);
        const int rem = idx - f * RW * nq;
        const int r = rem / nq, k = rem - r * nq;
        if (r < nr)
            a.W[qidx(f, y, z0 + r, k, ny, nz, nq)] =
                unpack_row(fin + (f * RW + r) * M, k, M, pl.tw);
    }
}

// z part of the test filter on the nx/4 kept columns, scaled by 1/(nx nz)
__global__ void __launch_bounds__(kT) k_smag_zf(SmagArgs a, FftPlan pz, int B)
{
    extern __shared__ double2 lds[];
    const int nz = a.nz, ny = a.ny, nq = a.nq;
    double2 *buf0 = lds, *buf1 = lds + (size_t)B * nz;
    const FftPlan pl = plan_in_lds(pz, lds + (size_t)2 * B * nz);
    const int f = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nq - x0);
    double2 *W = a.W + qidx(f, y, 0, x0, ny, nz, nq);
    const int tot = nz * B;
    for (int base = threadIdx.x; base < tot; base += kIB * kT) {
        double2 v[kIB];
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = min(base + i * kT, tot - 1);
            const int z = idx / B, b = idx - z * B;
            v[i] = W[(size_t)z * nq + min(b, nb - 1)];
        }
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = base + i * kT;
            if (idx < tot) {
                const int z = idx / B, b = idx - z * B;
                buf0[b * nz + z] = b < nb ? v[i] : make_double2(0., 0.);
            }
        }
    }
    __syncthreads();
    double2 *sp = lds_fft(buf0, buf1, nz, B, pl, -1.0);
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx % nz;
        if (z > nz / 4 && z < nz - nz / 4)
            sp[idx] = make_double2(0., 0.);
    }
    __syncthreads();
    double2 *oth = sp == buf0 ? buf1 : buf0;
    const double2 *res = lds_fft(sp, oth, nz, B, pl, +1.0);
    const double sc = (double)a.nx * (double)nz;
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        if (b < nb) {
            const double2 v = res[b * nz + z];
            W[(size_t)z * nq + b] = make_double2(v.x / sc, v.y / sc);
        }
    }
}

// dynamic procedure, pass 1: the RW rows z0.. of plane y (block (bz, y),
// z0 = RW bz): inverse x transform of the filtered fields, L_ij, M_ij, and
// the partial z sums of M.L and M.M per x point into part [ny][nzb][2][nx]
__global__ void __launch_bounds__(kT) k_smag_xc1(SmagArgs a, FftPlan pm, int RW)
{
    extern __shared__ double2 lds[];
    __shared__ double red[2 * 2 * kT];
    const int M = a.nxh, nz = a.nz, ny = a.ny, nq = a.nq;
    const int bz = blockIdx.x, y = blockIdx.y, nzb = gridDim.x;
    const int nx = 2 * M, tid = threadIdx.x;
    const int z0 = bz * RW, nr = min(RW, nz - z0);
    double2 *A = lds, *Bf = lds + (size_t)21 * RW * M;
    const FftPlan pl = plan_in_lds(pm, lds + (size_t)2 * 21 * RW * M);
    const int G = kT / M, n = tid % M, grp = tid / M;
    double num[2] = {0., 0.}, den[2] = {0., 0.};
    const double gd2 = a.tw2 * a.dlt2[y];
    // the filtered rows keep only kx < nq: one item at a time, with no load
    // for the modes above
    for (int idx = tid; idx < 21 * RW * M; idx += blockDim.x) {
        const int f = idx / (RW * M);
        const int rem = idx - f * RW * M;
        const int r = rem / M, k = rem - r * M;
        double2 Z = make_double2(0., 0.);
        if (r < nr) {
            const double2 *row = a.W + qidx(f, y, z0 + r, 0, ny, nz, nq);
            const int km = M - k;
            const double2 Xk = k < nq ? row[k] : make_double2(0., 0.);
            const double2 Xm = (k > 0 && km < nq) ? row[km] : make_double2(0., 0.);
            Z = pack_vals(Xk, Xm, k, pl.tw);
        }
        A[(f * RW + r) * M + k] = Z;
    }
    __syncthreads();
    const double2 *res = lds_fft(A, Bf, M, 21 * RW, pl, +1.0);
    if (grp < G) {
        for (int r = grp; r < nr; r += G) {
            double2 q[21];
#pragma unroll
            for (int f = 0; f < 21; f++)
                q[f] = res[(f * RW + r) * M + n];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                double v[21];
#pragma unroll
                for (int f = 0; f < 21; f++)
                    v[f] = e ? q[f].y : q[f].x;
                const int pa[6] = {0, 1, 2, 1, 2, 2}, pb[6] = {0, 0, 0, 1, 1, 2};
                double L[6], Mm[6];
                const double dr = gd2 * sabs_of(v + 12);
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    L[k] = v[k] + v[18 + pa[k]] * v[18 + pb[k]];
                    Mm[k] = v[6 + k] + v[12 + k] * dr;
                }
                num[e] += Mm[0] * L[0] + Mm[3] * L[3] + Mm[5] * L[5] +
                          2. * (Mm[1] * L[1] + Mm[2] * L[2] + Mm[4] * L[4]);
                den[e] += Mm[0] * Mm[0] + Mm[3] * Mm[3] + Mm[5] * Mm[5] +
                          2. * (Mm[1] * Mm[1] + Mm[2] * Mm[2] + Mm[4] * Mm[4]);
            }
        }
        red[(grp * 2 + 0) * 2 * M + 2 * n] = num[0];
        red[(grp * 2 + 0) * 2 * M + 2 * n + 1] = num[1];
        red[(grp * 2 + 1) * 2 * M + 2 * n] = den[0];
        red[(grp * 2 + 1) * 2 * M + 2 * n + 1] = den[1];
    }
    __syncthreads();
    double *out = a.part + ((size_t)y * nzb + bz) * 2 * nx;
    for (int idx = tid; idx < 2 * nx; idx += blockDim.x) {
        const int c = idx / nx, x = idx - c * nx;
        double t = 0.;
        for (int gq = 0; gq < G; gq++)
            t += red[(gq * 2 + c) * nx + x];
        out[idx] = t;
    }
}

// the coefficient C(x, y) of the dynamic procedure from the partial sums,
// added in a fixed order and clipped by ineg (les.f:1418-1510), one block
// per y plane; written over the first partial row part[y][0][0][x]
__global__ void __launch_bounds__(kT) k_smag_c(SmagArgs a, int nzb)
{
    const int ny = a.ny, nx = 2 * a.nxh, y = blockIdx.x;
    double *pp = a.part + (size_t)y * nzb * 2 * nx;
    for (int x = threadIdx.x; x < nx; x += blockDim.x) {
        double cn = 0., cd = 0.;
        for (int b = 0; b < nzb; b++) {
            cn += pp[(size_t)b * 2 * nx + x];
            cd += pp[(size_t)b * 2 * nx + nx + x];
        }
        double c = .5 * cn / cd;
        if (a.ineg == 1)
            c = fmax(c, 0.);
        if (a.ineg != 0 && (y == 0 || y == ny - 1))
            c = 0.;
        pp[x] = c;
    }
}

// pass 2 (block (bz, y), rows RT bz..): C(x, y) (dynamic, from k_smag_c)
// or cs^2, then nu_t, tau_ij / (nx nz) and the forward x transform into Q
// as [6][ny][nz][nx/2]
__global__ void __launch_bounds__(kT) k_smag_xc2(SmagArgs a, FftPlan pm, int RT, int nzb)
{
    extern __shared__ double2 lds[];
    __shared__ double cx[kMaxNx];
    const int M = a.nxh, nz = a.nz, ny = a.ny, y = blockIdx.y;
    const int nx = 2 * M, tid = threadIdx.x;
    const int z0 = blockIdx.x * RT, nr = min(RT, nz - z0);
    const double d2 = a.dlt2[y];
    for (int x = tid; x < nx; x += blockDim.x)
        cx[x] = a.dynamic ? a.part[(size_t)y * nzb * 2 * nx + x] : a.cs2;
    __syncthreads();
    double2 *A = lds, *Bf = lds + (size_t)6 * RT * M;
    const FftPlan pl = plan_in_lds(pm, lds + (size_t)2 * 6 * RT * M);
    const double gd = 1. / ((double)nx * (double)nz);
    for (int idx = tid; idx < RT * M; idx += blockDim.x) {
        const int r = idx / M, n = idx - r * M;
        double2 t[6];
#pragma unroll
        for (int k = 0; k < 6; k++)
            t[k] = make_double2(0., 0.);
        if (r < nr) {
            const double2 sa = a.SS[qidx(6, y, z0 + r, n, ny, nz, M)];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                double nu = cx[2 * n + e] * (e ? sa.y : sa.x) * d2;
                if (a.ineg == 2 && nu < -1. / a.re)
                    nu = -1. / a.re;
                const double d1 = -2. * nu * gd;
#pragma unroll
                for (int k = 0; k < 6; k++) {
                    const double2 s = a.SS[qidx(k, y, z0 + r, n, ny, nz, M)];
                    (e ? t[k].y : t[k].x) = d1 * (e ? s.y : s.x);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 6; k++)
            A[(k * RT + r) * M + n] = t[k];
    }
    __syncthreads();
    const double2 *fin = lds_fft(A, Bf, M, 6 * RT, pl, -1.0);
    for (int idx = tid; idx < 6 * RT * M; idx += blockDim.x) {
        const int k = idx / (RT * M);
        const int rem = idx - k * RT * M;
        const int r = rem / M, kx = rem - r * M;
        if (r < nr)
            a.Q[qidx(k, y, z0 + r, kx, ny, nz, M)] =
                unpack_row(fin + (k * RT + r) * M, kx, M, pl.tw);
    }
}

// forward z transform of the six stresses, scaled to the nonlinear term's
// normalisation, the z oddball row zeroed
__global__ void __launch_bounds__(kT) k_smag_za(SmagArgs a, FftPlan pz, int B)
{
    extern __shared__ double2 lds[];
    const int nz = a.nz, ny = a.ny, nxh = a.nxh;
    double2 *buf0 = lds, *buf1 = lds + (size_t)B * nz;
    const FftPlan pl = plan_in_lds(pz, lds + (size_t)2 * B * nz);
    const int f = blockIdx.y, y = blockIdx.z, x0 = blockIdx.x * B;
    const int nb = min(B, nxh - x0);
    const int tot = nz * B;
    for (int base = threadIdx.x; base < tot; base += kIB * kT) {
        double2 v[kIB];
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = min(base + i * kT, tot - 1);
            const int z = idx / B, b = idx - z * B;
            v[i] = a.Q[qidx(f, y, z, x0 + min(b, nb - 1), ny, nz, nxh)];
        }
#pragma unroll
        for (int i = 0; i < kIB; i++) {
            const int idx = base + i * kT;
            if (idx < tot) {
                const int z = idx / B, b = idx - z * B;
                buf0[b * nz + z] = b < nb ? v[i] : make_double2(0., 0.);
            }
        }
    }
    __syncthreads();
    const double2 *res = lds_fft(buf0, buf1, nz, B, pl, -1.0);
    for (int idx = threadIdx.x; idx < nz * B; idx += blockDim.x) {
        const int z = idx / B, b = idx - z * B;
        if (b < nb) {
            double2 v = res[b * nz + z];
            if (z == nz / 2)
                v = make_double2(0., 0.);
            a.TAU[sidx(f, z, y, x0 + b, nz, ny, nxh)] =
                make_double2(v.x * a.hscale, v.y * a.hscale);
        }
    }
}

int zblock(int nz, int ncol) { return std::max(1, std::min(ncol, kLdsZ / nz)); }

}  // namespace

// rows of z per block: as many as fit in kRowLds of LDS (several blocks
// per CU), at least one (up to lds_max)
constexpr size_t kRowLds = 48 * 1024;

int smag_rows(int nxh, int nfield, size_t lds_max)
{
    const size_t per = (size_t)2 * nfield * nxh * sizeof(double2);
    return (int)std::max<size_t>(1, std::min(lds_max, kRowLds) / per);
}

int smag_zblocks(int nz, int nxh, size_t lds_max)
{
    const int RC = std::min(nz, smag_rows(nxh, 21, lds_max));
    return (nz + RC - 1) / RC;
}

bool smag_fits(int nz, int nxh, size_t lds_max)
{
    return nz <= kLdsZ && 2 * nxh <= kMaxNx && nxh <= kT &&
           (size_t)(2 * 21 * nxh + 2 * nxh) * sizeof(double2) <= lds_max &&
           (size_t)(2 * kLdsZ + nz) * sizeof(double2) <= lds_max;
}

int smag_prepare(size_t lds_max)
{
    const void *fns[] = {(const void *)k_smag_z,   (const void *)k_smag_xa,
                         (const void *)k_smag_zf,  (const void *)k_smag_xc1,
                         (const void *)k_smag_xc2, (const void *)k_smag_za};
    for (const void *fn : fns)
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max) !=
            hipSuccess)
            return -1;
    return 0;
}

void launch_smag(SmagArgs a, const FftPlan &pz, const FftPlan &pm,
                 size_t lds_max, hipStream_t s)
{
    const int B = zblock(a.nz, a.nxh);
    const size_t lz = ((size_t)2 * B * a.nz + a.nz) * sizeof(double2);
    const size_t twx = (size_t)a.nx * sizeof(double2);
    hipLaunchKernelGGL(k_smag_z, dim3((a.nxh + B - 1) / B, a.nqf, a.ny), dim3(kT), lz, s, a, pz,
                       B);
    const int NF = a.dynamic ? 21 : 12;
    const int RA = std::min(a.nz, smag_rows(a.nxh, NF, lds_max));
    hipLaunchKernelGGL(k_smag_xa, dim3((a.nz + RA - 1) / RA, a.ny), dim3(kT),
                       (size_t)2 * NF * RA * a.nxh * sizeof(double2) + twx, s, a, pm, RA);
    if (a.dynamic) {
        const int Bq = zblock(a.nz, a.nq);
        hipLaunchKernelGGL(k_smag_zf, dim3((a.nq + Bq - 1) / Bq, 21, a.ny), dim3(kT),
                           ((size_t)2 * Bq * a.nz + a.nz) * sizeof(double2), s, a, pz, Bq);
    }
    if (a.dynamic) {
        const int RC = std::min(a.nz, smag_rows(a.nxh, 21, lds_max));
        hipLaunchKernelGGL(k_smag_xc1, dim3((a.nz + RC - 1) / RC, a.ny), dim3(kT),
                           (size_t)2 * 21 * RC * a.nxh * sizeof(double2) + twx, s, a, pm,
                           RC);
    }
    if (a.dynamic)
        hipLaunchKernelGGL(k_smag_c, dim3(a.ny), dim3(kT), 0, s, a,
                           smag_zblocks(a.nz, a.nxh, lds_max));
    const int RT = std::min(a.nz, smag_rows(a.nxh, 6, lds_max));
    hipLaunchKernelGGL(k_smag_xc2, dim3((a.nz + RT - 1) / RT, a.ny), dim3(kT),
                       (size_t)2 * 6 * RT * a.nxh * sizeof(double2) + twx, s, a, pm, RT,
                       smag_zblocks(a.nz, a.nxh, lds_max));
    hipLaunchKernelGGL(k_smag_za, dim3((a.nxh + B - 1) / B, 6, a.ny), dim3(kT), lz, s, a, pz, B);
}

}  // namespace sg
