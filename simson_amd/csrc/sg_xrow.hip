// x-row pass of step 2 (bla/nonlinbl.f) as a four-step FFT in LDS.
//
// Per (y, zp) row and field the half spectrum X_k, k < nx/2, is packed into
// M = nxp/2 complex values Z (the real transform of length nxp as one
// complex transform of length M, physical point pairs (x_2n, x_2n+1) in
// (re, im)), transformed, multiplied pointwise (advection.f:52-72, with the
// CFL partial of boxcfl.f) and transformed back (fft2db.f:45, fft2df.f:32).
//
// With M = M1 M2 every length-M transform is two in-place LDS passes of
// register DFTs:
//   inverse  pass 1  for each n1 < M1: M2-point DFT of the stride-M1 column
//                    n1 + M1 n2, times exp(+2 pi i n1 k2 / M), stored back
//                    at n1 + M1 k2;
//            pass 2  for each k2 < M2: M1-point DFT of the contiguous block
//                    M1 k2 + n1; position M1 k2 + k1 then holds the physical
//                    pair n = k2 + M2 k1 (a digit-reversed order);
//   forward  pass A  for each k2: M1-point DFT of block k2 (the same
//                    digit-reversed positions), times exp(-2 pi i k2 c / M);
//            pass B  for each c < M1: M2-point DFT of column c, which leaves
//                    the spectrum in natural order.
// Pass 2, the products and pass A touch only one block per task and run
// fused in registers; each pass reads and writes the same LDS slots, so one
// buffer serves all passes.  Rows are padded by one slot per M1 block.
#include <cstdint>

#include "sg_cplx.h"
#include "sg_internal.h"

namespace sg {

namespace {

// M1 x M2: from M = 288 (the split form below) M2 = 16, so that the
// column DFTs of passes 1 and B stay small; the M1-point block DFTs then
// run one field at a time
template <int M>
struct XPlan {
    static constexpr int M1 = (M >= 288 && M % 16 == 0)
                                  ? M / 16
                                  : (M % 4 == 0) ? ((M > 96 && M % 8 == 0) ? 8 : 4)
                                                 : ((M % 2 == 0) ? 2 : 3);
    static constexpr int M2 = M / M1;
    static constexpr int SROW = M + M / M1;
    static constexpr int TABN = 2 * M;
};

// SPLIT (from M = 288): the inverse pass 2, the products and the forward
// pass A run as three LDS passes (per field and block, per physical point,
// per output field and block) instead of one fused task per block, which
// would hold seven M1-value arrays in registers (one wave per SIMD at
// these lengths) with only RW M2 tasks to spread over the block.
template <typename V, int M>
struct XBlock {
    static constexpr bool SPLIT = M >= 288;
    static constexpr int budget = 52 * 1024;
    static constexpr int per_row = 6 * XPlan<M>::SROW * (int)sizeof(V);
    static constexpr int raw = (budget - XPlan<M>::TABN * (int)sizeof(V)) / per_row;
    static constexpr int RW = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
    static constexpr int tasks1 = 6 * RW * XPlan<M>::M1;
    static constexpr int tasks2 = SPLIT ? RW * M : RW * XPlan<M>::M2;
    static constexpr int tmax = tasks1 > tasks2 ? tasks1 : tasks2;
    static constexpr int NT = tmax >= 256 ? 256 : ((tmax + 63) / 64) * 64;
    static constexpr size_t lds = (size_t)(XPlan<M>::TABN + 6 * RW * XPlan<M>::SROW) * sizeof(V);
};

// PRESS: the pressure variant (nonlinp.f) also forms W = u.u on the
// physical row and transforms it back as a fourth output field.
template <typename V, int M, bool PRESS>
__global__ void __launch_bounds__((XBlock<V, M>::NT)) k_xrow4(Step2Args<V> a, const V *__restrict__ twg)
{
    typedef typename ScalarOf<V>::T T;
    constexpr int M1 = XPlan<M>::M1, M2 = XPlan<M>::M2, SROW = XPlan<M>::SROW;
    constexpr int TABN = XPlan<M>::TABN;
    constexpr int RW = XBlock<V, M>::RW, NT = XBlock<V, M>::NT;
    constexpr int S1 = M1 + 1;                   // padded block stride
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V *tw = reinterpret_cast<V *>(lds_raw);      // exp(+2 pi i t / (2M))
    V *A = tw + TABN;                            // [6][RW][SROW]
    const int nxh = a.nxh, nzp = a.nzp, ny = a.ny;
    const int zp0 = blockIdx.x * RW, y = blockIdx.y;
    const int nr = min(RW, nzp - zp0);
    const int tid = threadIdx.x;
    for (int i = tid; i < TABN; i += NT)
        tw[i] = twg[i];
    __syncthreads();

    // half spectrum X_k (k < nx/2, zero above) -> packed Z_k, k < M.  With
    // M <= NT each thread keeps one k (its twiddle, flags and LDS offset)
    // and walks the (field, row) pairs in batches of NB: the loads of a batch
    // are unconditional (clamped addresses, zeroed by select), so they issue
    // back to back instead of one round trip per pair.
    constexpr int NXH = 2 * M / 3;               // nx/2 (nxp = 3 nx/2 = 2M)
    constexpr int TPK = NT / M;                  // threads per k
    if constexpr (TPK >= 1) {
        constexpr int NPAIR = 6 * RW, NJ = (NPAIR + TPK - 1) / TPK;
        constexpr int NB = NJ < 8 ? NJ : 8;
        const int k = tid < TPK * M ? tid % M : 0;
        const int p0 = tid < TPK * M ? tid / M : NPAIR;
        const V twk = tw[k];
        const int km = M - k;
        const bool hk = k < NXH, hm = k > 0 && km < NXH;
        const int ik = hk ? k : 0, im = hm ? km : 0;
        const int ko = k + k / M1;
#pragma unroll
        for (int j0 = 0; j0 < NJ; j0 += NB) {
            V xk[NB], xm[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int p = min(p0 + TPK * (j0 + b), NPAIR - 1);
                const int f = p / RW, r = min(p - f * RW, nr - 1);
                const V *row = a.P + (((size_t)f * ny + y) * nzp + zp0 + r) * NXH;
                xk[b] = row[ik];
                xm[b] = row[im];
            }
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int p = p0 + TPK * (j0 + b);
                if (j0 + b >= NJ || p >= NPAIR)
                    continue;
                const int r = p % RW;
                V Xk = hk ? xk[b] : vmk<V>(0, 0);
                if (k == 0)
                    Xk.y = 0;                    // only Re X_0 enters vrfftb
                V Xm = hm ? xm[b] : vmk<V>(0, 0);
                Xm.y = -Xm.y;
                const V sm = vadd(Xk, Xm), d = vsub(Xk, Xm);
                V Z = vadd(sm, vmul_is(vmul(twk, d), (T)1));
                if (r >= nr)
                    Z = vmk<V>(0, 0);
                A[p * SROW + ko] = Z;
            }
        }
    } else {
        for (int idx = tid; idx < 6 * RW * M; idx += NT) {
            const int f = idx / (RW * M);
            const int rem = idx - f * RW * M;
            const int r = rem / M, k = rem - r * M;
            V Z = vmk<V>(0, 0);
            if (r < nr) {
                const V *row = a.P + (((size_t)f * ny + y) * nzp + zp0 + r) * nxh;
                V Xk = k < nxh ? row[k] : vmk<V>(0, 0);
                if (k == 0)
                    Xk.y = 0;
                const int km = M - k;            // X_M (k = 0) is beyond nx/2
                V Xm = (k > 0 && km < nxh) ? row[km] : vmk<V>(0, 0);
                Xm.y = -Xm.y;
                const V sm = vadd(Xk, Xm), d = vsub(Xk, Xm);
                Z = vadd(sm, vmul_is(vmul(tw[k], d), (T)1));
            }
            A[(f * RW + r) * SROW + k + k / M1] = Z;
        }
    }
    __syncthreads();

    // inverse pass 1: six fields
    for (int t = tid; t < 6 * RW * M1; t += NT) {
        const int fr = t / M1, n1 = t - fr * M1;
        V *row = A + fr * SROW + n1;
        V v[M2];
#pragma unroll
        for (int n2 = 0; n2 < M2; n2++)
            v[n2] = row[S1 * n2];
        reg_dft<V, M2, +1, TABN>(v, tw);
#pragma unroll
        for (int k2 = 1; k2 < M2; k2++)
            if (n1 > 0)
                v[k2] = vmul(v[k2], tw[2 * n1 * k2]);
#pragma unroll
        for (int k2 = 0; k2 < M2; k2++)
            row[S1 * k2] = v[k2];
    }
    __syncthreads();

    // inverse pass 2, products, forward pass A (one block per task)
    T cflmax = 0;
    T dyinv = (T)(1.0 / a.deta[y]);
    if (a.ny >= 17 && !a.nocap)
        dyinv = fmin(dyinv, (T)((double)(a.ny - 1) * (double)(a.ny - 1) / 10. / 3.14));
    const T dxi = (T)((double)(a.nx - 2) / a.xl);
    const T dzi = (T)((double)(a.nz - 2) / a.zl);
    const T r2 = (T)(2.0 * a.rot);
    // mean profiles of this plane when the fields hold only fluctuations
    // (storage slots 0 u, 2 w, 3 om_x, 4 om_z at kx = kz = 0; Om_y = 0, and
    // the mean v, nonzero only with an imposed wall velocity, stays in v)
    T Um = 0, Wm = 0, Oxm = 0, Ozm = 0;
    if (a.split_mean) {
        const size_t F = (size_t)a.nz * ny * nxh, yo = (size_t)y * nxh;
        Um = a.U[yo].x;
        Wm = a.U[2 * F + yo].x;
        Oxm = a.U[3 * F + yo].x;
        Ozm = a.U[4 * F + yo].x;
    }
    constexpr int FS = RW * SROW;                // field stride in A
    if constexpr (XBlock<V, M>::SPLIT) {
        // inverse pass 2: an M1-point DFT per field and block, in place
        for (int t = tid; t < 6 * RW * M2; t += NT) {
            const int fr = t / M2, k2 = t - fr * M2;
            V *blk = A + fr * SROW + S1 * k2;
            V q[M1];
#pragma unroll
            for (int i = 0; i < M1; i++)
                q[i] = blk[i];
            reg_dft<V, M1, +1, TABN>(q, tw);
#pragma unroll
            for (int i = 0; i < M1; i++)
                blk[i] = q[i];
        }
        __syncthreads();
        // products at each physical point n = k2 + M2 i (slot S1 k2 + i of
        // its row), H into the slots of u, v, w (and u.u into om_x's); the
        // same expressions as the fused form below
        const T zr = Ozm + r2;
        for (int t = tid; t < RW * M; t += NT) {
            const int r = t / M, p = t - r * M;
            const int k2 = p / M1, i = p - k2 * M1;
            V *pt = A + r * SROW + S1 * k2 + i;
            const V u = pt[0], v = pt[FS], w = pt[2 * FS];
            const V ox = pt[3 * FS], oy = pt[4 * FS], oz = pt[5 * FS];
            if (a.do_cfl && r < nr) {
                const T ce = fabs(Um + u.x) * dxi + fabs(v.x) * dyinv + fabs(Wm + w.x) * dzi;
                const T co = fabs(Um + u.y) * dxi + fabs(v.y) * dyinv + fabs(Wm + w.y) * dzi;
                cflmax = cfl_max(cfl_max(cflmax, ce), co);
            }
            V h2 = vmk<V>((Wm + w.x) * ox.x + w.x * Oxm, (Wm + w.y) * ox.y + w.y * Oxm);
            V h3 = vmk<V>(-v.x * (Oxm + ox.x), -v.y * (Oxm + ox.y));
            V h1 = vmk<V>(-(Wm + w.x) * oy.x, -(Wm + w.y) * oy.y);
            h3 = vmk<V>(h3.x + (Um + u.x) * oy.x, h3.y + (Um + u.y) * oy.y);
            const T zx = oz.x + zr, zy = oz.y + zr;
            h1 = vmk<V>(v.x * zx + h1.x, v.y * zy + h1.y);
            h2 = vmk<V>(h2.x - (Um * oz.x + u.x * zx), h2.y - (Um * oz.y + u.y * zy));
            if (a.fringe) {
                const int n = k2 + M2 * i;
                const V lam = a.flam[n];
                const V b0 = a.fbase[(size_t)y * M + n];
                const V b1 = a.fbase[((size_t)ny + y) * M + n];
                const V b2 = a.fbase[((size_t)2 * ny + y) * M + n];
                h1 = vmk<V>(h1.x - lam.x * (Um + u.x - b0.x), h1.y - lam.y * (Um + u.y - b0.y));
                h2 = vmk<V>(h2.x - lam.x * (v.x - b1.x), h2.y - lam.y * (v.y - b1.y));
                h3 = vmk<V>(h3.x - lam.x * (Wm + w.x - b2.x), h3.y - lam.y * (Wm + w.y - b2.y));
            }
            if constexpr (PRESS)
                pt[3 * FS] = vmk<V>(u.x * u.x + v.x * v.x + w.x * w.x,
                                    u.y * u.y + v.y * v.y + w.y * w.y);
            pt[0] = h1;
            pt[FS] = h2;
            pt[2 * FS] = h3;
        }
        __syncthreads();
        // forward pass A per output field and block
        constexpr int NOA = PRESS ? 4 : 3;
        for (int t = tid; t < NOA * RW * M2; t += NT) {
            const int hr = t / M2, k2 = t - hr * M2;
            V *blk = A + hr * SROW + S1 * k2;
            V q[M1];
#pragma unroll
            for (int i = 0; i < M1; i++)
                q[i] = blk[i];
            reg_dft<V, M1, -1, TABN>(q, tw);
#pragma unroll
            for (int c = 0; c < M1; c++) {
                if (c > 0 && k2 > 0) {
                    V wt = tw[2 * k2 * c];
                    wt.y = -wt.y;
                    q[c] = vmul(q[c], wt);
                }
                blk[c] = q[c];
            }
        }
    } else
    for (int t = tid; t < RW * M2; t += NT) {
        const int r = t / M2, k2 = t - r * M2;
        V *blk = A + r * SROW + S1 * k2;         // field f at blk + f * RW * SROW
        V u[M1], v[M1], w[M1], o[M1], h1[M1], h2[M1], h3[M1];
#pragma unroll
        for (int i = 0; i < M1; i++) {
            u[i] = blk[i];
            v[i] = blk[FS + i];
            w[i] = blk[2 * FS + i];
        }
        reg_dft<V, M1, +1, TABN>(u, tw);
        reg_dft<V, M1, +1, TABN>(v, tw);
        reg_dft<V, M1, +1, TABN>(w, tw);
        V hw[PRESS ? M1 : 1];
        if constexpr (PRESS) {
#pragma unroll
            for (int i = 0; i < M1; i++)
                hw[i] = vmk<V>(u[i].x * u[i].x + v[i].x * v[i].x + w[i].x * w[i].x,
                               u[i].y * u[i].y + v[i].y * v[i].y + w[i].y * w[i].y);
        }
        if (a.do_cfl && r < nr) {
#pragma unroll
            for (int i = 0; i < M1; i++) {
                const T ce = fabs(Um + u[i].x) * dxi + fabs(v[i].x) * dyinv +
                             fabs(Wm + w[i].x) * dzi;
                const T co = fabs(Um + u[i].y) * dxi + fabs(v[i].y) * dyinv +
                             fabs(Wm + w[i].y) * dzi;
                cflmax = cfl_max(cfl_max(cflmax, ce), co);
            }
        }
        // H = u x (om + 2 Omega e_z) (advection.f:52-72).  With the mean
        // split off, u = Um + u' etc.; the mean x mean part (nonzero only in
        // the kx = kz = 0 mode of H_2, which the v equation does not use
        // there) is left out and every remaining term is O(fluctuation):
        //   H1 = v (om_z + Ozm + 2Om) - (Wm + w) om_y
        //   H2 = (Wm + w) om_x + w Oxm - (Um + u) om_z - u (Ozm + 2Om)
        //   H3 = (Um + u) om_y - v (Oxm + om_x)
        // (all means zero: the plain rotational form).
        const T zr = Ozm + r2;
        // om_x
#pragma unroll
        for (int i = 0; i < M1; i++)
            o[i] = blk[3 * FS + i];
        reg_dft<V, M1, +1, TABN>(o, tw);
#pragma unroll
        for (int i = 0; i < M1; i++) {
            h2[i] = vmk<V>((Wm + w[i].x) * o[i].x + w[i].x * Oxm,
                           (Wm + w[i].y) * o[i].y + w[i].y * Oxm);
            h3[i] = vmk<V>(-v[i].x * (Oxm + o[i].x), -v[i].y * (Oxm + o[i].y));
        }
        // om_y
#pragma unroll
        for (int i = 0; i < M1; i++)
            o[i] = blk[4 * FS + i];
        reg_dft<V, M1, +1, TABN>(o, tw);
#pragma unroll
        for (int i = 0; i < M1; i++) {
            h1[i] = vmk<V>(-(Wm + w[i].x) * o[i].x, -(Wm + w[i].y) * o[i].y);
            h3[i] = vmk<V>(h3[i].x + (Um + u[i].x) * o[i].x, h3[i].y + (Um + u[i].y) * o[i].y);
        }
        // om_z + 2 Omega
#pragma unroll
        for (int i = 0; i < M1; i++)
            o[i] = blk[5 * FS + i];
        reg_dft<V, M1, +1, TABN>(o, tw);
#pragma unroll
        for (int i = 0; i < M1; i++) {
            const T zx = o[i].x + zr, zy = o[i].y + zr;
            h1[i] = vmk<V>(v[i].x * zx + h1[i].x, v[i].y * zy + h1[i].y);
            h2[i] = vmk<V>(h2[i].x - (Um * o[i].x + u[i].x * zx),
                           h2[i].y - (Um * o[i].y + u[i].y * zy));
        }
        // fringe damping toward the base flow (fring.f, fringp.f): point
        // n = k2 + M2 i of the row (the order the two inverse passes leave)
        if (a.fringe) {
#pragma unroll
            for (int i = 0; i < M1; i++) {
                const int n = k2 + M2 * i;
                const V lam = a.flam[n];
                const V b0 = a.fbase[(size_t)y * M + n];
                const V b1 = a.fbase[((size_t)ny + y) * M + n];
                const V b2 = a.fbase[((size_t)2 * ny + y) * M + n];
                h1[i] = vmk<V>(h1[i].x - lam.x * (Um + u[i].x - b0.x),
                               h1[i].y - lam.y * (Um + u[i].y - b0.y));
                h2[i] = vmk<V>(h2[i].x - lam.x * (v[i].x - b1.x),
                               h2[i].y - lam.y * (v[i].y - b1.y));
                h3[i] = vmk<V>(h3[i].x - lam.x * (Wm + w[i].x - b2.x),
                               h3[i].y - lam.y * (Wm + w[i].y - b2.y));
            }
        }
        // forward pass A
        reg_dft<V, M1, -1, TABN>(h1, tw);
        reg_dft<V, M1, -1, TABN>(h2, tw);
        reg_dft<V, M1, -1, TABN>(h3, tw);
        if constexpr (PRESS)
            reg_dft<V, M1, -1, TABN>(hw, tw);
#pragma unroll
        for (int c = 0; c < M1; c++) {
            if (c > 0 && k2 > 0) {
                V wt = tw[2 * k2 * c];
                wt.y = -wt.y;
                h1[c] = vmul(h1[c], wt);
                h2[c] = vmul(h2[c], wt);
                h3[c] = vmul(h3[c], wt);
                if constexpr (PRESS)
                    hw[c] = vmul(hw[c], wt);
            }
            blk[c] = h1[c];
            blk[FS + c] = h2[c];
            blk[2 * FS + c] = h3[c];
            if constexpr (PRESS)
                blk[3 * FS + c] = hw[c];
        }
    }
    __syncthreads();

    // forward pass B: three (four) fields, natural order
    constexpr int NOUT = PRESS ? 4 : 3;
    for (int t = tid; t < NOUT * RW * M1; t += NT) {
        const int hr = t / M1, c = t - hr * M1;
        V *row = A + hr * SROW + c;
        V v[M2];
#pragma unroll
        for (int n2 = 0; n2 < M2; n2++)
            v[n2] = row[S1 * n2];
        reg_dft<V, M2, -1, TABN>(v, tw);
#pragma unroll
        for (int d = 0; d < M2; d++)
            row[S1 * d] = v[d];
    }
    __syncthreads();

    // packed spectrum -> X_k = E_k + exp(-2 pi i k/nxp) O_k for k < nx/2
    auto unpack = [&](int h, int r, int k, V twk) {
        const V *row = A + (h * RW + r) * SROW;
        const int km = k == 0 ? 0 : M - k;
        const V Zk = row[k + k / M1];
        const V Zm = vconj(row[km + km / M1]);
        const V E = vmk<V>((T)0.5 * (Zk.x + Zm.x), (T)0.5 * (Zk.y + Zm.y));
        const V dd = vsub(Zk, Zm);
        const V O = vmk<V>((T)0.5 * dd.y, (T)-0.5 * dd.x);
        a.P[(((size_t)h * ny + y) * nzp + zp0 + r) * NXH + k] = vadd(E, vmul(vconj(twk), O));
    };
    constexpr int TPX = NT / NXH;                // threads per output k
    if constexpr (TPX >= 1) {
        constexpr int NQ = NOUT * RW, NJ = (NQ + TPX - 1) / TPX;
        const int k = tid % NXH, q0 = tid / NXH;
        const V twk = tw[k];
#pragma unroll
        for (int j = 0; j < NJ; j++) {
            const int q = q0 + TPX * j;
            if (tid >= TPX * NXH || q >= NQ)
                continue;
            const int h = q / RW, r = q - h * RW;
            if (r < nr)
                unpack(h, r, k, twk);
        }
    } else {
        for (int idx = tid; idx < NOUT * RW * NXH; idx += NT) {
            const int h = idx / (RW * NXH);
            const int rem = idx - h * RW * NXH;
            const int r = rem / NXH, k = rem - r * NXH;
            if (r < nr)
                unpack(h, r, k, tw[k]);
        }
    }
    if (a.do_cfl) {
        double m = (double)cflmax;
        for (int off = 32; off > 0; off >>= 1)
            m = fmax(m, __shfl_xor(m, off));
        __shared__ double wmax[NT / 64];
        const int lane = tid & 63, wv = tid >> 6;
        if (lane == 0)
            wmax[wv] = m;
        __syncthreads();
        if (tid == 0) {
            double b = 0.;
            for (int i = 0; i < NT / 64; i++)
                b = fmax(b, wmax[i]);
            a.cfl_blk[blockIdx.y * gridDim.x + blockIdx.x] = b;
        }
    }
}

template <typename V, int M, bool PRESS = false>
void xrow4(const Step2Args<V> &a, const V *tw, hipStream_t s)
{
    using B = XBlock<V, M>;
    dim3 grid((a.nzp + B::RW - 1) / B::RW, a.ny);
    hipLaunchKernelGGL((k_xrow4<V, M, PRESS>), grid, dim3(B::NT), B::lds, s, a, tw);
}

// ---------------------------------------------------------------------
// Learned closure (sgs network) on the x rows: the nine velocity-gradient
// fields of a (y, zp) block are synthesised in x, the network maps the
// gradients and y+ of every physical point to the six subgrid stresses on
// the BF16 matrix cores, and the stresses are transformed back.  Fields in
// a.P: [9][ny][nzp][nx/2] in, stresses into fields 0-5 of the same rows.
// ---------------------------------------------------------------------

template <typename V, int M>
struct CBlock {
    static constexpr int NF = 9;
    static constexpr int budget = 52 * 1024;
    static constexpr int per_row = NF * XPlan<M>::SROW * (int)sizeof(V);
    static constexpr int raw = (budget - XPlan<M>::TABN * (int)sizeof(V)) / per_row;
    static constexpr int RW = raw < 1 ? 1 : (raw > 8 ? 8 : raw);
    static constexpr int tasks1 = NF * RW * XPlan<M>::M1;
    static constexpr int tasks2 = NF * RW * XPlan<M>::M2;
    static constexpr int tmax = tasks1 > tasks2 ? tasks1 : tasks2;
    static constexpr int NT = tmax >= 256 ? 256 : ((tmax + 63) / 64) * 64;
    static constexpr size_t lds = (size_t)(XPlan<M>::TABN + NF * RW * XPlan<M>::SROW) * sizeof(V);
};

typedef short cs4 __attribute__((ext_vector_type(4)));
typedef float cf4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

// exact GELU (the training kernels use the same; a rational erf with two
// transcendental instructions measured slower than the device erff)
__device__ __forceinline__ float gelu_f(float z)
{
    return 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
}

template <typename V, int M>
__global__ void __launch_bounds__((CBlock<V, M>::NT))
    k_xclos(Step2Args<V> a, const V *__restrict__ twg, ClosArgs c)
{
    typedef typename ScalarOf<V>::T T;
    constexpr int M1 = XPlan<M>::M1, M2 = XPlan<M>::M2, SROW = XPlan<M>::SROW;
    constexpr int TABN = XPlan<M>::TABN, NF = CBlock<V, M>::NF;
    constexpr int RW = CBlock<V, M>::RW, NT = CBlock<V, M>::NT;
    constexpr int S1 = M1 + 1, NXH = 2 * M / 3, FS = RW * SROW;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V *tw = reinterpret_cast<V *>(lds_raw);
    V *A = tw + TABN;                            // [NF][RW][SROW]
    const int nzp = a.nzp, ny = a.ny;
    const int zp0 = blockIdx.x * RW, y = blockIdx.y;
    const int nr = min(RW, nzp - zp0);
    const int tid = threadIdx.x;
    for (int i = tid; i < TABN; i += NT)
        tw[i] = twg[i];
    __syncthreads();
    // half spectra -> packed Z_k (as k_xrow4, for NF fields)
    for (int idx = tid; idx < NF * RW * M; idx += NT) {
        const int f = idx / (RW * M);
        const int rem = idx - f * RW * M;
        const int r = rem / M, k = rem - r * M;
        V Z = vmk<V>(0, 0);
        if (r < nr) {
            const V *row = a.P + (((size_t)f * ny + y) * nzp + zp0 + r) * NXH;
            V Xk = k < NXH ? row[k] : vmk<V>(0, 0);
            if (k == 0)
                Xk.y = 0;
            const int km = M - k;
            V Xm = (k > 0 && km < NXH) ? row[km] : vmk<V>(0, 0);
            Xm.y = -Xm.y;
            const V sm = vadd(Xk, Xm), d = vsub(Xk, Xm);
            Z = vadd(sm, vmul_is(vmul(tw[k], d), (T)1));
        }
        A[(f * RW + r) * SROW + k + k / M1] = Z;
    }
    __syncthreads();
    // inverse pass 1 (columns) and pass 2 (blocks) of every field
    for (int t = tid; t < NF * RW * M1; t += NT) {
        const int fr = t / M1, n1 = t - fr * M1;
        V *row = A + fr * SROW + n1;
        V v[M2];
#pragma unroll
        for (int n2 = 0; n2 < M2; n2++)
            v[n2] = row[S1 * n2];
        reg_dft<V, M2, +1, TABN>(v, tw);
#pragma unroll
        for (int k2 = 1; k2 < M2; k2++)
            if (n1 > 0)
                v[k2] = vmul(v[k2], tw[2 * n1 * k2]);
#pragma unroll
        for (int k2 = 0; k2 < M2; k2++)
            row[S1 * k2] = v[k2];
    }
    __syncthreads();
    for (int t = tid; t < NF * RW * M2; t += NT) {
        const int fr = t / M2, k2 = t - fr * M2;
        V *blk = A + fr * SROW + S1 * k2;
        V v[M1];
#pragma unroll
        for (int i = 0; i < M1; i++)
            v[i] = blk[i];
        reg_dft<V, M1, +1, TABN>(v, tw);
#pragma unroll
        for (int i = 0; i < M1; i++)
            blk[i] = v[i];
    }
    __syncthreads();
    // the network on tiles of 16 physical points (8 packed positions, re and
    // im); features-as-rows products: Z^T = W^T X^T, so each layer's result
    // (lane l: unit 4(l/16)+r of point l%16) is the next layer's B operand
    {
        const int l = tid & 63, wv = tid >> 6, nwv = NT / 64;
        const int lr = l & 15, lg = l >> 4;
        cs4 w1[4], w2[4][4], w3[4];
        float b1[4][4], b2[4][4], b3[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
#pragma unroll
            for (int j = 0; j < 4; j++) {
                w1[mb][j] = to_bf16(c.W[(4 * lg + j) * 64 + 16 * mb + lr]);
                w3[mb][j] = to_bf16(c.W[c.ow3 + (16 * mb + 4 * lg + j) * 16 + lr]);
            }
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    w2[mb][kb][j] = to_bf16(c.W[c.ow2 + (16 * kb + 4 * lg + j) * 64 + 16 * mb + lr]);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                b1[mb][r] = c.W[c.ob1 + 16 * mb + 4 * lg + r];
                b2[mb][r] = c.W[c.ob2 + 16 * mb + 4 * lg + r];
            }
        }
#pragma unroll
        for (int r = 0; r < 4; r++)
            b3[r] = c.W[c.ob3 + 4 * lg + r];
        const float yf = c.yfeat[y];
        constexpr int NTILE = RW * M / 8;
        for (int t = wv; t < NTILE; t += nwv) {
            const int r = t / (M / 8), pos = 8 * (t % (M / 8)) + (lr >> 1);
            const int off = r * SROW + pos + pos / M1;
            const bool im = lr & 1;
            cs4 xb;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int k = 4 * lg + j;
                float x = 0.f;
                if (k < 9) {
                    const V v = A[k * FS + off];
                    x = (float)(im ? v.y : v.x) * c.in_a[k] + c.in_b[k];
                } else if (k == 9) {
                    x = yf;
                }
                xb[j] = to_bf16(x);
            }
            cs4 h1[4], h2[4];
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                cf4 z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w1[mb], xb, cf4{0.f, 0.f, 0.f, 0.f},
                                                                 0, 0, 0);
#pragma unroll
                for (int q = 0; q < 4; q++)
                    h1[mb][q] = to_bf16(gelu_f(z[q] + b1[mb][q]));
            }
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
                cf4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kb = 0; kb < 4; kb++)
                    z = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w2[mb][kb], h1[kb], z, 0, 0, 0);
#pragma unroll
                for (int q = 0; q < 4; q++)
                    h2[mb][q] = to_bf16(gelu_f(z[q] + b2[mb][q]));
            }
            cf4 yo = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < 4; kb++)
                yo = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(w3[kb], h2[kb], yo, 0, 0, 0);
            // every lane read its inputs before any lane of the wave writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int o = 4 * lg + q;
                if (o < 6) {
                    const T tau = (T)((yo[q] + b3[q]) * c.out_a[o] + c.out_b[o]);
                    V &dst = A[o * FS + off];
                    if (im)
                        dst.y = tau;
                    else
                        dst.x = tau;
                }
            }
        }
    }
    __syncthreads();
    // forward pass A (blocks) and pass B (columns) of the six stresses
    for (int t = tid; t < 6 * RW * M2; t += NT) {
        const int fr = t / M2, k2 = t - fr * M2;
        V *blk = A + fr * SROW + S1 * k2;
        V v[M1];
#pragma unroll
        for (int i = 0; i < M1; i++)
            v[i] = blk[i];
        reg_dft<V, M1, -1, TABN>(v, tw);
#pragma unroll
        for (int cc = 0; cc < M1; cc++) {
            if (cc > 0 && k2 > 0) {
                V wt = tw[2 * k2 * cc];
                wt.y = -wt.y;
                v[cc] = vmul(v[cc], wt);
            }
            blk[cc] = v[cc];
        }
    }
    __syncthreads();
    for (int t = tid; t < 6 * RW * M1; t += NT) {
        const int hr = t / M1, cc = t - hr * M1;
        V *row = A + hr * SROW + cc;
        V v[M2];
#pragma unroll
        for (int n2 = 0; n2 < M2; n2++)
            v[n2] = row[S1 * n2];
        reg_dft<V, M2, -1, TABN>(v, tw);
#pragma unroll
        for (int d = 0; d < M2; d++)
            row[S1 * d] = v[d];
    }
    __syncthreads();
    for (int idx = tid; idx < 6 * RW * NXH; idx += NT) {
        const int h = idx / (RW * NXH);
        const int rem = idx - h * RW * NXH;
        const int r = rem / NXH, k = rem - r * NXH;
        if (r >= nr)
            continue;
        const V *row = A + (h * RW + r) * SROW;
        const int km = k == 0 ? 0 : M - k;
        const V Zk = row[k + k / M1];
        const V Zm = vconj(row[km + km / M1]);
        const V E = vmk<V>((T)0.5 * (Zk.x + Zm.x), (T)0.5 * (Zk.y + Zm.y));
        const V dd = vsub(Zk, Zm);
        const V O = vmk<V>((T)0.5 * dd.y, (T)-0.5 * dd.x);
        a.P[(((size_t)h * ny + y) * nzp + zp0 + r) * NXH + k] = vadd(E, vmul(vconj(tw[k]), O));
    }
}

template <typename V, int M>
bool xclos(const Step2Args<V> &a, const V *tw, const ClosArgs &c, hipStream_t s)
{
    using B = CBlock<V, M>;
    if constexpr (B::lds > 65536) {
        // nine fields of one row exceed 64 KiB from M = 384 (FP64): allow
        // the kernel the larger dynamic LDS once (gfx950 has 160 KiB)
        static const bool ok = hipFuncSetAttribute((const void *)k_xclos<V, M>,
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)B::lds) == hipSuccess;
        if (!ok)
            return false;
    }
    dim3 grid((a.nzp + B::RW - 1) / B::RW, a.ny);
    hipLaunchKernelGGL((k_xclos<V, M>), grid, dim3(B::NT), B::lds, s, a, tw, c);
    return true;
}

}  // namespace

#define SG_X4LENS(X) X(12) X(24) X(48) X(96) X(192) X(288) X(384)

bool has_xrow4(int m)
{
#define SG_CASE(N) \
    if (m == N)    \
        return true;
    SG_X4LENS(SG_CASE)
#undef SG_CASE
    return false;
}

template <typename V>
bool launch_xrow4(const Step2Args<V> &a, const V *wx, hipStream_t s)
{
    switch (a.nxp / 2) {
#define SG_CASE(N)                \
    case N:                       \
        xrow4<V, N>(a, wx, s);    \
        return true;
        SG_X4LENS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

template <typename V>
int xrow4_blocks(const Step2Args<V> &a)
{
    switch (a.nxp / 2) {
#define SG_CASE(N)                                                         \
    case N:                                                                \
        return ((a.nzp + XBlock<V, N>::RW - 1) / XBlock<V, N>::RW) * a.ny;
        SG_X4LENS(SG_CASE)
#undef SG_CASE
    default:
        return 0;
    }
}

bool launch_xrow4_press(const Step2Args<double2> &a, const double2 *wx, hipStream_t s)
{
    switch (a.nxp / 2) {
#define SG_CASE(N)                                 \
    case N:                                        \
        xrow4<double2, N, true>(a, wx, s);         \
        return true;
        SG_X4LENS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

template <typename V>
bool launch_xclos(const Step2Args<V> &a, const V *wx, const ClosArgs &c, hipStream_t s)
{
    switch (a.nxp / 2) {
#define SG_CASE(N)                      \
    case N:                             \
        return xclos<V, N>(a, wx, c, s);
        SG_X4LENS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

template bool launch_xclos<double2>(const Step2Args<double2> &, const double2 *, const ClosArgs &,
                                    hipStream_t);
template bool launch_xclos<float2>(const Step2Args<float2> &, const float2 *, const ClosArgs &,
                                   hipStream_t);
template bool launch_xrow4<double2>(const Step2Args<double2> &, const double2 *, hipStream_t);
template bool launch_xrow4<float2>(const Step2Args<float2> &, const float2 *, hipStream_t);
template int xrow4_blocks<double2>(const Step2Args<double2> &);
template int xrow4_blocks<float2>(const Step2Args<float2> &);

}  // namespace sg
