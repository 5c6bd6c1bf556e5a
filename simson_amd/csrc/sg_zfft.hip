// z passes of step 2 (bla/nonlinbl.f) as four-step FFTs in LDS.
//
//   k_zsyn4  gather u, v, w, om_x, om_z of a y plane, form
//            om_y = i beta u - i alpha w, pad to nzp modes with the zeroed
//            dealiasing band (getxz.f:41-83), inverse complex z transform
//            (fft2db.f:33); with split_mean the kx = kz = 0 mode is left
//            out (the x-row kernel adds the mean profiles itself);
//   k_zana4  forward complex z transform (fft2df.f:37) and truncation back
//            to nz modes, z oddball row included (putxz.f:34-54).
//
// A block owns B x columns of one (field, y) plane; each column is one
// length-N (N = nzp = N1 N2) transform done as two in-place LDS passes of
// register DFTs, as in the x-row kernel:
//   inverse: pass 1 (for n1 < N1, N2-point DFT of the stride-N1 column,
//            twiddle exp(+2 pi i n1 k2 / N)), pass 2 (for k2 < N2, N1-point
//            DFT of block k2); position N1 k2 + k1 then holds zp = k2 + N2 k1;
//   forward: the input is loaded into that digit-reversed order, pass A
//            (N1-point DFT of block a, twiddle exp(-2 pi i a c / N)) and
//            pass B (N2-point DFT of column c) leave natural order.
// Columns are padded by one slot per N1 block.
#include "sg_cplx.h"
#include "sg_internal.h"

namespace sg {

namespace {

// N = N1 N2 with N1 the smallest divisor >= sqrt(N) (contiguous blocks)
constexpr int z_n1(int n)
{
    int best = n;
    for (int d = n; d >= 1; d--)
        if (n % d == 0 && d * d >= n)
            best = d;
    return best;
}

template <typename V, int N>
struct ZPlan {
    static constexpr int N1 = z_n1(N);
    static constexpr int N2 = N / N1;
    static constexpr int SCOL = N + N / N1;
    static constexpr int budget = 40 * 1024;
    static constexpr int per_col = SCOL * (int)sizeof(V);
    static constexpr int raw = (budget - N * (int)sizeof(V)) / per_col;
    static constexpr int B = raw >= 16 ? 16 : (raw >= 8 ? 8 : (raw >= 4 ? 4 : (raw >= 2 ? 2 : 1)));
    static constexpr int tasks = B * (N1 > N2 ? N1 : N2);
    static constexpr int NT = tasks >= 256 ? 256 : ((tasks + 63) / 64) * 64;
    static constexpr size_t lds = (size_t)(N + B * SCOL) * sizeof(V);
    // blocks of consecutive x groups that share one 128-byte line of the
    // x-contiguous planes
    static constexpr int G0 = 128 / (B * (int)sizeof(V));
    static constexpr int G = G0 < 1 ? 1 : (G0 > 8 ? 8 : G0);
};

// Logical (x group, field, y) of a block.  The hardware hands the block of
// linear id p to XCD p mod 8, each XCD with its own L2; the G blocks of
// consecutive x groups that read or write one cache line are given linear
// ids 8 apart (p = 8 G q + 8 s + xcd -> logical 8 G q + G xcd + s), so
// that they run on one XCD at about the same time and the line is fetched
// from memory once.  The last partial chunk keeps the identity mapping.
template <int G>
__device__ __forceinline__ void block_xfy(int &bx, int &f, int &y)
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int total = gx * gy * gridDim.z;
    int l = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (G > 1) {
        constexpr int C = 8 * G;
        const int q = l / C, r = l - q * C;
        if ((q + 1) * C <= total)
            l = q * C + (r & 7) * G + (r >> 3);
    }
    bx = l % gx;
    const int rest = l / gx;
    f = rest % gy;
    y = rest / gy;
}

__device__ __forceinline__ size_t uidx(int slot, int z, int y, int x, int nz, int ny, int nxh)
{
    return (((size_t)slot * nz + z) * ny + y) * nxh + x;
}

// digit-reversed position of zp (zp = k2 + N2 k1 -> N1 k2 + k1), padded
template <int N1, int N2>
__device__ __forceinline__ int dr_off(int zp)
{
    const int k2 = zp % N2, k1 = zp / N2;
    return (N1 + 1) * k2 + k1;
}

// gather of the velocity-gradient field f = 3 i + j (d u_i / d x_j) from
// the storage slots u, v, w, om_x, om_z: x and z derivatives are i alpha,
// i beta; the y derivatives follow from the vorticities and continuity,
// du/dy = i alpha v - om_z, dv/dy = -(i alpha u + i beta w),
// dw/dy = om_x + i beta v.  Term t of field f: slot, multiplier (0: 1,
// 1: i alpha, 2: i beta) and sign; slot -1: no second term.
struct GradTerm {
    signed char s1, m1, g1, s2, m2, g2;
};
__constant__ GradTerm kGrad[9] = {
    {0, 1, 1, -1, 0, 0}, {1, 1, 1, 4, 0, -1}, {0, 2, 1, -1, 0, 0},
    {1, 1, 1, -1, 0, 0}, {0, 1, -1, 2, 2, -1}, {1, 2, 1, -1, 0, 0},
    {2, 1, 1, -1, 0, 0}, {3, 0, 1, 1, 2, 1},  {2, 2, 1, -1, 0, 0}};

template <typename V>
__device__ __forceinline__ V grad_mul(V v, int m, int sg, typename ScalarOf<V>::T al,
                                      typename ScalarOf<V>::T be)
{
    typedef typename ScalarOf<V>::T T;
    const T k = m == 1 ? al : be;
    V r = m == 0 ? v : vmk<V>(-k * v.y, k * v.x);
    return sg < 0 ? vmk<V>(-r.x, -r.y) : r;
}

// MODE 0: the six step-2 fields u, v, w, om_x, om_y, om_z; MODE 1: the nine
// velocity-gradient fields of the learned closure (into a.P as [9][ny][N][nxh])
template <typename V, int N, int MODE = 0>
__global__ void __launch_bounds__((ZPlan<V, N>::NT)) k_zsyn4(Step2Args<V> a, const V *__restrict__ twg)
{
    typedef typename ScalarOf<V>::T T;
    using P = ZPlan<V, N>;
    constexpr int N1 = P::N1, N2 = P::N2, SCOL = P::SCOL, B = P::B, NT = P::NT;
    constexpr int S1 = N1 + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V *tw = reinterpret_cast<V *>(lds_raw);     // exp(+2 pi i t / N)
    V *A = tw + N;                              // [B][SCOL]
    const int nz = a.nz, nxh = a.nxh, ny = a.ny, tid = threadIdx.x;
    int bx, f, y;
    block_xfy<P::G>(bx, f, y);
    const int x0 = bx * B;
    const int nb = min(B, nxh - x0);
    for (int i = tid; i < N; i += NT)
        tw[i] = twg[i];
    // thread: column b = tid % B, padded rows zp = tid / B + ZS j; the loads
    // of a batch of NBT rows use clamped indices and issue back to back
    static_assert(NT % B == 0, "block must cover whole rows");
    constexpr int ZS = NT / B, NI = (N + ZS - 1) / ZS, NBT = NI < 6 ? NI : 6;
    const int b = tid % B, zq = tid / B;
    const int xc = min(x0 + b, nxh - 1);
    const T al = (T)a.alfa[xc];
    const bool om_y = MODE == 0 && f == 4;      // block-uniform
    const GradTerm gt = MODE == 1 ? kGrad[f] : GradTerm{0, 0, 0, -1, 0, 0};
    const bool two = MODE == 1 ? gt.s2 >= 0 : om_y;
#pragma unroll
    for (int j0 = 0; j0 < NI; j0 += NBT) {
        V u[NBT], w[NBT];
        T bt[NBT];
#pragma unroll
        for (int q = 0; q < NBT; q++) {
            const int zp = min(zq + ZS * (j0 + q), N - 1);
            // storage row of padded mode zp; rows nz/2 .. N-nz/2 stay zero
            const int z = zp < nz / 2 ? zp : (zp > N - nz / 2 ? zp - (N - nz) : 0);
            // MODE 0: fields u, v, w, om_x, om_y, om_z from storage slots
            // 0-3, (0, 2), 4
            const int s1 = MODE == 1 ? gt.s1 : (om_y ? 0 : (f < 4 ? f : 4));
            const int s2 = MODE == 1 ? gt.s2 : 2;
            u[q] = a.U[uidx(s1, z, y, xc, nz, ny, nxh)];
            if (two)
                w[q] = a.U[uidx(s2, z, y, xc, nz, ny, nxh)];
            bt[q] = (T)a.beta[z];
        }
#pragma unroll
        for (int q = 0; q < NBT; q++) {
            const int zp = zq + ZS * (j0 + q);
            if (j0 + q >= NI || zp >= N)
                continue;
            const int z = zp < nz / 2 ? zp : (zp > N - nz / 2 ? zp - (N - nz) : -1);
            V v = u[q];
            if (MODE == 1) {
                v = grad_mul(u[q], gt.m1, gt.g1, al, bt[q]);
                if (two) {
                    const V v2 = grad_mul(w[q], gt.m2, gt.g2, al, bt[q]);
                    v = vadd(v, v2);
                }
            } else if (om_y) {
                v = vmk<V>(-bt[q] * u[q].y + al * w[q].y, bt[q] * u[q].x - al * w[q].x);
            }
            // with split_mean the kx = kz = 0 mode enters the products
            // directly (step-2 fields only); the mean v (nonzero only with
            // an imposed wall velocity) stays in the field
            if (b >= nb || z < 0 ||
                (MODE == 0 && a.split_mean && f != 1 && z == 0 && x0 + b == 0))
                v = vmk<V>(0, 0);
            A[b * SCOL + zp + zp / N1] = v;
        }
    }
    __syncthreads();
    for (int t = tid; t < B * N1; t += NT) {
        const int b = t / N1, n1 = t - b * N1;
        V *col = A + b * SCOL + n1;
        V v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++)
            v[n2] = col[S1 * n2];
        reg_dft<V, N2, +1, N>(v, tw);
#pragma unroll
        for (int k2 = 1; k2 < N2; k2++)
            if (n1 > 0)
                v[k2] = vmul(v[k2], tw[n1 * k2]);
#pragma unroll
        for (int k2 = 0; k2 < N2; k2++)
            col[S1 * k2] = v[k2];
    }
    __syncthreads();
    for (int t = tid; t < B * N2; t += NT) {
        const int b = t / N2, k2 = t - b * N2;
        V *blk = A + b * SCOL + S1 * k2;
        V v[N1];
#pragma unroll
        for (int i = 0; i < N1; i++)
            v[i] = blk[i];
        reg_dft<V, N1, +1, N>(v, tw);
#pragma unroll
        for (int i = 0; i < N1; i++)
            blk[i] = v[i];
    }
    __syncthreads();
    V *Pp = a.P + (((size_t)f * ny + y) * N) * nxh + x0;
    for (int idx = tid; idx < N * B; idx += NT) {
        const int zp = idx / B, b = idx - zp * B;
        if (b < nb)
            Pp[(size_t)zp * nxh + b] = A[b * SCOL + dr_off<N1, N2>(zp)];
    }
}

template <typename V, int N>
__global__ void __launch_bounds__((ZPlan<V, N>::NT)) k_zana4(Step2Args<V> a, const V *__restrict__ twg)
{
    using P = ZPlan<V, N>;
    constexpr int N1 = P::N1, N2 = P::N2, SCOL = P::SCOL, B = P::B, NT = P::NT;
    constexpr int S1 = N1 + 1;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V *tw = reinterpret_cast<V *>(lds_raw);
    V *A = tw + N;
    const int nz = a.nz, nxh = a.nxh, ny = a.ny, tid = threadIdx.x;
    int bx, h, y;
    block_xfy<P::G>(bx, h, y);
    const int x0 = bx * B;
    const int nb = min(B, nxh - x0);
    for (int i = tid; i < N; i += NT)
        tw[i] = twg[i];
    const V *Pp = a.P + (((size_t)h * ny + y) * N) * nxh + x0;
    {
        static_assert(NT % B == 0, "block must cover whole rows");
        constexpr int ZS = NT / B, NI = (N + ZS - 1) / ZS, NBT = NI < 8 ? NI : 8;
        const int b = tid % B, zq = tid / B, bc = min(b, nb - 1);
#pragma unroll
        for (int j0 = 0; j0 < NI; j0 += NBT) {
            V v[NBT];
#pragma unroll
            for (int q = 0; q < NBT; q++)
                v[q] = Pp[(size_t)min(zq + ZS * (j0 + q), N - 1) * nxh + bc];
#pragma unroll
            for (int q = 0; q < NBT; q++) {
                const int zp = zq + ZS * (j0 + q);
                if (j0 + q < NI && zp < N)
                    A[b * SCOL + dr_off<N1, N2>(zp)] = b < nb ? v[q] : vmk<V>(0, 0);
            }
        }
    }
    __syncthreads();
    for (int t = tid; t < B * N2; t += NT) {
        const int b = t / N2, k2 = t - b * N2;
        V *blk = A + b * SCOL + S1 * k2;
        V v[N1];
#pragma unroll
        for (int i = 0; i < N1; i++)
            v[i] = blk[i];
        reg_dft<V, N1, -1, N>(v, tw);
#pragma unroll
        for (int c = 0; c < N1; c++) {
            if (c > 0 && k2 > 0) {
                V wt = tw[k2 * c];
                wt.y = -wt.y;
                v[c] = vmul(v[c], wt);
            }
            blk[c] = v[c];
        }
    }
    __syncthreads();
    for (int t = tid; t < B * N1; t += NT) {
        const int b = t / N1, c = t - b * N1;
        V *col = A + b * SCOL + c;
        V v[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; n2++)
            v[n2] = col[S1 * n2];
        reg_dft<V, N2, -1, N>(v, tw);
#pragma unroll
        for (int d = 0; d < N2; d++)
            col[S1 * d] = v[d];
    }
    __syncthreads();
    if (!a.les && !a.clos) {
        for (int idx = tid; idx < nz * B; idx += NT) {
            const int z = idx / B, b = idx - z * B;
            if (b < nb) {
                const int zp = z < nz / 2 ? z : z + N - nz;
                a.Uw[uidx(h, z, y, x0 + b, nz, ny, nxh)] = A[b * SCOL + zp + zp / N1];
            }
        }
        return;
    }
    // with the ADM-RT relaxation or the learned closure's force: thread
    // column b = tid % B, rows z = tid / B + ZS j; the operands (u still in
    // Uw and Wf, or two stress spectra and a y derivative) are read in
    // batches
    typedef typename ScalarOf<V>::T T;
    constexpr int ZS = NT / B;
    const int b = tid % B, zq = tid / B, xc = min(x0 + b, nxh - 1);
    const T cn = (T)a.chi_n, al = (T)a.alfa[xc];
    // tau_i1 and tau_i3 of row h of the stress tensor (11 12 13 22 23 33)
    const int t1 = h, t3 = h == 0 ? 2 : (h == 1 ? 4 : 5);
    for (int z0 = zq; z0 < nz; z0 += 4 * ZS) {
        V u[4], w[4], e[4];
        T be[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int z = min(z0 + q * ZS, nz - 1);
            if (a.les) {
                u[q] = a.Uw[uidx(h, z, y, xc, nz, ny, nxh)];
                w[q] = a.Wf[uidx(h, z, y, xc, nz, ny, nxh)];
            } else {
                u[q] = a.TAU[uidx(t1, z, y, xc, nz, ny, nxh)];
                w[q] = a.TAU[uidx(t3, z, y, xc, nz, ny, nxh)];
                e[q] = a.DT[uidx(h, z, y, xc, nz, ny, nxh)];
                be[q] = (T)a.beta[z];
            }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int z = z0 + q * ZS;
            if (z >= nz || b >= nb)
                continue;
            const int zp = z < nz / 2 ? z : z + N - nz;
            V v = A[b * SCOL + zp + zp / N1];
            if (z != nz / 2) {
                if (a.les) {
                    const T lp = (T)a.lpf[(size_t)z * nxh + xc];
                    v.x -= cn * (u[q].x - lp * w[q].x);
                    v.y -= cn * (u[q].y - lp * w[q].y);
                } else {
                    v.x -= -al * u[q].y + e[q].x - be[q] * w[q].y;
                    v.y -= al * u[q].x + e[q].y + be[q] * w[q].x;
                }
            }
            a.Uw[uidx(h, z, y, xc, nz, ny, nxh)] = v;
        }
    }
}

}  // namespace

#define SG_Z4LENS(X) X(12) X(24) X(48) X(72) X(96) X(144) X(192) X(288) X(384) X(576)

bool has_zfft4(int nzp)
{
#define SG_CASE(N)  \
    if (nzp == N)   \
        return true;
    SG_Z4LENS(SG_CASE)
#undef SG_CASE
    return false;
}

template <typename V>
bool launch_zsyn4(const Step2Args<V> &a, const V *tw, hipStream_t s)
{
    switch (a.nzp) {
#define SG_CASE(N)                                                                       \
    case N: {                                                                            \
        using P = ZPlan<V, N>;                                                           \
        dim3 grid((a.nxh + P::B - 1) / P::B, 6, a.ny);                                   \
        hipLaunchKernelGGL((k_zsyn4<V, N>), grid, dim3(P::NT), P::lds, s, a, tw);        \
        return true;                                                                     \
    }
        SG_Z4LENS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

template <typename V>
bool launch_gsyn4(const Step2Args<V> &a, const V *tw, hipStream_t s)
{
    switch (a.nzp) {
#define SG_CASE(N)                                                                       \
    case N: {                                                                            \
        using P = ZPlan<V, N>;                                                           \
        dim3 grid((a.nxh + P::B - 1) / P::B, 9, a.ny);                                   \
        hipLaunchKernelGGL((k_zsyn4<V, N, 1>), grid, dim3(P::NT), P::lds, s, a, tw);     \
        return true;                                                                     \
    }
        SG_Z4LENS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

template <typename V>
bool launch_zana4(const Step2Args<V> &a, const V *tw, hipStream_t s, int nf)
{
    switch (a.nzp) {
#define SG_CASE(N)                                                                       \
    case N: {                                                                            \
        using P = ZPlan<V, N>;                                                           \
        dim3 grid((a.nxh + P::B - 1) / P::B, nf, a.ny);                                  \
        hipLaunchKernelGGL((k_zana4<V, N>), grid, dim3(P::NT), P::lds, s, a, tw);        \
        return true;                                                                     \
    }
        SG_Z4LENS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

template bool launch_zsyn4<double2>(const Step2Args<double2> &, const double2 *, hipStream_t);
template bool launch_zsyn4<float2>(const Step2Args<float2> &, const float2 *, hipStream_t);
template bool launch_gsyn4<double2>(const Step2Args<double2> &, const double2 *, hipStream_t);
template bool launch_gsyn4<float2>(const Step2Args<float2> &, const float2 *, hipStream_t);
template bool launch_zana4<double2>(const Step2Args<double2> &, const double2 *, hipStream_t,
                                    int);
template bool launch_zana4<float2>(const Step2Args<float2> &, const float2 *, hipStream_t, int);

}  // namespace sg
