// Step 3 of the bla substep (bla/linearbl.f, tau method, ibc = 0, 100 or 101) with one
// wavefront per (kx,kz) column.
//
// Lane L of the wave holds, for parity class c (k = c mod 2), the E class
// elements M = L*E + e, i.e. Chebyshev index k = 2M + c; both classes and real
// and imaginary parts live in registers.  Every sequential recurrence of the
// reference becomes a wave-level scan:
//   * dcheb (dcheb.f:24-36): b_k is a suffix sum over the opposite class,
//     b_2M = sum_{M'>=M} 2(2M'+1) a_{2M'+1}, b_2M+1 = sum_{M'>M} 4M' a_{2M'};
//   * the tau solves (setmacrhs.f + trid.f): with the per-column factors
//     gamma_k, 1/pivot_k precomputed for each lam (sg_linear.hip), the
//     bottom-up sweep rho_k = (R_k - U_k rho_{k+2})/piv_k is a suffix scan
//     of affine maps, the forward substitution a_k = rho_k - gamma_k a_{k-2}
//     a prefix scan of affine maps, and the boundary row a sum;
//   * parity sums (dv/dy at the wall, the mass flux) are wave reductions.
// The scans run on DPP lane moves within 16-lane rows.  Across the four rows
// of a 64-lane segment, prefix scans use DPP row broadcasts; suffix scans
// and sums take ds_bpermute lane reads, which go through the LDS pipe and so
// leave the VALU, this kernel's bound, to the arithmetic (sg_wave.h).
#include <cstdlib>

#include "sg_internal.h"
#include "sg_tau.h"
#include "sg_wave.h"

namespace sg {

namespace {

constexpr int WV = wv::kWave;

// inclusive suffix sums over the class index of NV channels
template <int SW, int NV, int E>
__device__ __forceinline__ void suffix_sum(double (&x)[NV][E], int l)
{
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int e = E - 2; e >= 0; e--)
            x[v][e] += x[v][e + 1];
    double ex[NV];
#pragma unroll
    for (int v = 0; v < NV; v++)
        ex[v] = wv::from_next<SW>(x[v][0], 0.0, l);
    double one = 1.0;
    wv::scan_suffix<SW, NV, false>(ex, one, l);
#pragma unroll
    for (int v = 0; v < NV; v++)
#pragma unroll
        for (int e = 0; e < E; e++)
            x[v][e] += ex[v];
}

// sum over the segment, identical on every lane of it
template <int SW, int NV>
__device__ __forceinline__ void wave_sum(double (&x)[NV], int l)
{
    wv::sum_all<SW, NV>(x, l);
}

// d/dy in Chebyshev space of NC channels (dcheb.f:24-36)
template <int SW, int NC, int E>
__device__ __forceinline__ void dcheb_w(const double (&a)[NC][2][E], double (&b)[NC][2][E], int l)
{
    const int L = l & (SW - 1);
    double s[2 * NC][E];
    double nxt[NC];
#pragma unroll
    for (int q = 0; q < NC; q++)
        nxt[q] = wv::from_next<SW>(a[q][0][0], 0.0, l);   // next lane's first even element
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int M = L * E + e;
        const double wo = 2.0 * (double)(2 * M + 1);
        const double we = 4.0 * (double)(M + 1);
#pragma unroll
        for (int q = 0; q < NC; q++) {
            s[q][e] = wo * a[q][1][e];
            s[NC + q][e] = we * ((e + 1 < E) ? a[q][0][e + 1] : nxt[q]);
        }
    }
    suffix_sum<SW, 2 * NC, E>(s, l);
#pragma unroll
    for (int e = 0; e < E; e++)
#pragma unroll
        for (int q = 0; q < NC; q++) {
            b[q][0][e] = s[q][e];
            b[q][1][e] = s[NC + q][e];
        }
    if (L == 0)
#pragma unroll
        for (int q = 0; q < NC; q++)
            b[q][0][0] *= 0.5;
}

// Row weights of the tau system (they depend on k only), one table per
// block in LDS: wm = c_{k-2}/(4k(k-1)), w0 = [k<=n-3]/(2(k^2-1)),
// wp = [k<=n-5]/(4k(k+1)), so that R_k = wm f_{k-2} - w0 f_k + wp f_{k+2}
// and U_k = -lam wp (sg_tau.h); zero outside the rows k = 2..n-1.
template <int KT>
struct TauTab {
    double wm[KT], w0[KT], wp[KT];
    __device__ __forceinline__ void fill(int n)
    {
        for (int k = threadIdx.x; k < KT; k += blockDim.x) {
            const double kk = (double)k;
            const bool row = k >= 2 && k < n;
            wm[k] = row ? ((k == 2) ? 2.0 : 1.0) / (4.0 * kk * (kk - 1.0)) : 0.0;
            w0[k] = (row && k <= n - 3) ? 1.0 / (2.0 * (kk * kk - 1.0)) : 0.0;
            wp[k] = (row && k <= n - 5) ? 1.0 / (4.0 * kk * (kk + 1.0)) : 0.0;
        }
    }
};

// Tau solve of u'' - lam u = f, in place, for NC real channels sharing lam
// (one column), boundary rows bc[q][class]; gp holds (gamma, 1/pivot).
template <int SW, int NC, int E, int KT>
__device__ __forceinline__ void solve_w(int n, double lam, const double2 (&gp)[2][E],
                                        const TauTab<KT> &tw, double (&f)[NC][2][E],
                                        const double (&bc)[NC][2], int l)
{
    const int L = l & (SW - 1);
    double (&u)[NC][2][E] = f;
    double al[NC][2][E], be[2][E], bq[2][E];
    double fprev[NC][2], fnext[NC][2];
#pragma unroll
    for (int q = 0; q < NC; q++)
#pragma unroll
        for (int c = 0; c < 2; c++) {
            fprev[q][c] = wv::from_prev<SW>(f[q][c][E - 1], 0.0, l);
            fnext[q][c] = wv::from_next<SW>(f[q][c][0], 0.0, l);
        }
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int M = L * E + e, k = 2 * M + c;
            const bool row = (M >= 1) && (k < n);
            const double ip = gp[c][e].y;
            const double wm = tw.wm[k], w0 = tw.w0[k], wp = tw.wp[k];
            // -U_k / piv_k and R_k / piv_k with the row weights
            be[c][e] = row ? lam * wp * ip : 0.0;
            bq[c][e] = (M == 0) ? 1.0 : (k < n ? -gp[c][e].x : 0.0);
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const double fm = e > 0 ? f[q][c][e - 1] : fprev[q][c];
                const double fp = e + 1 < E ? f[q][c][e + 1] : fnext[q][c];
                const double R = wm * fm - w0 * f[q][c][e] + wp * fp;
                al[q][c][e] = row ? R * ip : 0.0;
            }
        }
    // rho_M = al_M + be_M rho_{M+1}: local composition then suffix scan
    double A[NC][2][E], B[2][E];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        double accB = 1.0, accA[NC];
#pragma unroll
        for (int q = 0; q < NC; q++)
            accA[q] = 0.0;
#pragma unroll
        for (int e = E - 1; e >= 0; e--) {
#pragma unroll
            for (int q = 0; q < NC; q++) {
                accA[q] = al[q][c][e] + be[c][e] * accA[q];
                A[q][c][e] = accA[q];
            }
            accB = be[c][e] * accB;
            B[c][e] = accB;
        }
    }
    double X[NC][2], Y[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Y[c] = wv::from_next<SW>(B[c][0], 1.0, l);
#pragma unroll
        for (int q = 0; q < NC; q++)
            X[q][c] = wv::from_next<SW>(A[q][c][0], 0.0, l);
    }
    {
        double Xc[2][NC];
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < NC; q++)
                Xc[c][q] = X[q][c];
        wv::scan_suffix2<SW, NC>(Xc, Y, l);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int q = 0; q < NC; q++)
                X[q][c] = Xc[c][q];
    }
    double rho[NC][2][E];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++)
#pragma unroll
            for (int q = 0; q < NC; q++)
                rho[q][c][e] = A[q][c][e] + B[c][e] * X[q][c];
    // P_M = rho_M + bq_M P_{M-1}, Q_M = bq_M Q_{M-1}: prefix scan
    double Ap[NC][2][E], Bp[2][E];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        double accB = 1.0, accA[NC];
#pragma unroll
        for (int q = 0; q < NC; q++)
            accA[q] = 0.0;
#pragma unroll
        for (int e = 0; e < E; e++) {
#pragma unroll
            for (int q = 0; q < NC; q++) {
                accA[q] = rho[q][c][e] + bq[c][e] * accA[q];
                Ap[q][c][e] = accA[q];
            }
            accB = bq[c][e] * accB;
            Bp[c][e] = accB;
        }
    }
    double Xp[NC][2], Yp[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
        Yp[c] = wv::from_prev<SW>(Bp[c][E - 1], 1.0, l);
        double Xc[NC];
#pragma unroll
        for (int q = 0; q < NC; q++)
            Xc[q] = wv::from_prev<SW>(Ap[q][c][E - 1], 0.0, l);
        wv::scan_prefix<SW, NC, true>(Xc, Yp[c]);
#pragma unroll
        for (int q = 0; q < NC; q++)
            Xp[q][c] = Xc[q];
    }
    // P_M = A'_M + B'_M X', Q_M = B'_M Y' (B' the lane-local products), then
    // the boundary row of each class: sums [q*2+c] of P and [2*NC+c] of Q
    double sums[(NC + 1) * 2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
        for (int q = 0; q < NC; q++) {
            double sp = 0.0;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const double P = Ap[q][c][e] + Bp[c][e] * Xp[q][c];
                Ap[q][c][e] = P;
                sp += P;
            }
            sums[q * 2 + c] = sp;
        }
        double sq = 0.0;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const double Q = Bp[c][e] * Yp[c];
            sq += Q;
            Bp[c][e] = Q;
        }
        sums[2 * NC + c] = sq;
    }
    wave_sum<SW, (NC + 1) * 2>(sums, l);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        const double rq = 1.0 / sums[2 * NC + c];
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double a0 = (bc[q][c] - sums[q * 2 + c]) * rq;
#pragma unroll
            for (int e = 0; e < E; e++)
                u[q][c][e] = Ap[q][c][e] + Bp[c][e] * a0;
        }
    }
}

// Chebyshev-column storage: FP64 (C) or FP32 (C32, the mixed form, whose
// transforms run in FP32); the solve itself is FP64 either way.
template <typename CV>
__device__ __forceinline__ CV *cols(const ColArgs &a);
template <>
__device__ __forceinline__ double2 *cols<double2>(const ColArgs &a)
{
    return a.C;
}
template <>
__device__ __forceinline__ float2 *cols<float2>(const ColArgs &a)
{
    return a.C32;
}
__device__ __forceinline__ double2 widen(double2 v) { return v; }
__device__ __forceinline__ double2 widen(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void put(double2 *p, double x, double y) { *p = make_double2(x, y); }
__device__ __forceinline__ void put(float2 *p, double x, double y)
{
    *p = make_float2((float)x, (float)y);
}

__device__ __forceinline__ double2 widen(double v) { return make_double2(v, 0.); }

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Column I/O of one wave in the lane order of the scans (lane L of a
// segment holds elements k = 2(L E + e) + c).  With 16-lane segments the
// four columns of a wave, t0..t0+3, are contiguous in memory; they pass
// through a per-wave LDS buffer so that memory is read and written with
// consecutive lanes on consecutive elements (reading the blocked order
// straight from memory would touch a new cache line on every lane).  The
// buffer holds segment s at s*SEGP, element k at k + k/(2E) (one pad slot
// per lane block against bank conflicts).  Wider segments read directly.
template <int SW, int E>
struct ColIO {
    static constexpr int SEGP = 16 * (2 * E + 1);
    static constexpr int NBUF = SW == 16 ? 4 * SEGP : 1;
    double2 *buf;
    size_t t0n, tn;      // element offsets of column t0 (first of the wave) and t
    int n, l, L, nseg;   // nseg: segments with a column (t0 + s < ncol)
    unsigned ownmask;    // segments whose column is stored
    bool own;

    template <typename T>
    __device__ __forceinline__ void load(const T *f, double2 (&v)[2][E]) const
    {
        if constexpr (SW == 16) {
            for (int s = 0; s < nseg; s++)
                for (int k = l; k < n; k += 64)
                    buf[s * SEGP + k + k / (2 * E)] = widen(f[t0n + (size_t)s * n + k]);
            wave_sync();
            const int s = l >> 4;
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int k = 2 * (L * E + e) + c;
                    v[c][e] = (k < n && s < nseg) ? buf[s * SEGP + L * (2 * E + 1) + 2 * e + c]
                                                  : make_double2(0., 0.);
                }
            wave_sync();
        } else {
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int k = 2 * (L * E + e) + c;
                    v[c][e] = k < n ? widen(f[tn + k]) : make_double2(0., 0.);
                }
        }
    }

    template <typename T>
    __device__ __forceinline__ void store(T *f, const double2 (&v)[2][E]) const
    {
        if constexpr (SW == 16) {
            const int s = l >> 4;
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int e = 0; e < E; e++)
                    buf[s * SEGP + L * (2 * E + 1) + 2 * e + c] = v[c][e];
            wave_sync();
            for (int s2 = 0; s2 < nseg; s2++)
                if (ownmask & (1u << s2))
                    for (int k = l; k < n; k += 64) {
                        const double2 x = buf[s2 * SEGP + k + k / (2 * E)];
                        put(f + t0n + (size_t)s2 * n + k, x.x, x.y);
                    }
            wave_sync();
        } else {
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const int k = 2 * (L * E + e) + c;
                    if (own && k < n)
                        put(f + tn + k, v[c][e].x, v[c][e].y);
                }
        }
    }
};

// One column per SW-lane segment (64 / SW columns per wave).  A segment
// without a column of its own (the z oddball plane, linearbl.f:227, or past
// the end) computes on a valid column and stores nothing, so that every
// lane of the wave stays active through the lane moves.
//
// Register use is kept down by loading each input just before it is needed.
// With 16-lane segments (four elements per class and lane) the -c2 h_v,
// -c2 h_omega parts of the next partial right-hand sides also go to their
// output slots early and are read back and completed at the end by the same
// lanes that wrote them; wider segments keep them in registers, which saves
// that extra memory traffic.
template <int SW, int E>
constexpr int tau_waves_per_eu()
{
    return E == 1 ? 3 : 2;
}

// FT: the upper boundary y = +1: 0 a wall, 1 the free stream with the base
// flow's derivatives (ibc = 101), 2 with its values (ibc = 100)
template <int SW, int E, typename CV, int FT>
__global__ void __launch_bounds__(256, (tau_waves_per_eu<SW, E>()))
    k_tau_w(ColArgs a, LamFactor lf, LamFactor kf, double an, double bn, double anp1, double bnp1)
{
    constexpr int KT = 2 * SW * E;
    using IO = ColIO<SW, E>;
    __shared__ TauTab<KT> tw;
    __shared__ double2 sbuf[4][IO::NBUF];
    tw.fill(a.ny);
    __syncthreads();
    const int l = threadIdx.x & (WV - 1);
    const int L = l & (SW - 1);
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int t0 = wave * (WV / SW);
    const int tw_ = t0 + l / SW;
    const int ncol = a.nz * a.nxh;
    auto owns = [&](int tt) { return tt < ncol && tt / a.nxh != a.nz / 2; };
    const bool own = owns(tw_);
    const int t = SW == 16 ? min(tw_, ncol - 1) : (own ? tw_ : 1);
    const int z = t / a.nxh, x = t - z * a.nxh;
    const int n = a.ny;
    IO io;
    io.buf = sbuf[threadIdx.x >> 6];
    io.t0n = (size_t)t0 * n;
    io.tn = (size_t)t * n;
    io.n = n;
    io.l = l;
    io.L = L;
    io.nseg = min(WV / SW, ncol - t0);
    io.ownmask = 0;
    for (int s = 0; s < WV / SW; s++)
        if (owns(t0 + s))
            io.ownmask |= 1u << s;
    io.own = own;
    const double re = a.re;
    const double alfa = a.alfa[x], beta = a.beta[z];
    const double k2 = alfa * alfa + beta * beta;
    const double lam2 = k2 + 2. * re / (an + bn);
    const bool k0 = (t == 0);
    const double k2i = k0 ? 0. : 1. / k2;
    const double c1 = 2. * re / (an + bn) * an;
    const double c1n = 2. * re / (anp1 + bnp1);
    const double c2 = 2. * re * bnp1 / (anp1 + bnp1);
    const size_t fs = (size_t)ncol * n;
    const CV *C0 = cols<CV>(a);
    double2 *P0 = a.Pp;
    auto kidx = [&](int c, int e) { return 2 * (L * E + e) + c; };
    double2 tmp[2][E];
    // localized volume force (locf.f, loctyp 3) on H_i of the kz = 0 plane:
    // its Chebyshev coefficients times sin(tomeg tc)
    const bool with_force = a.lforce && z == 0;
    auto add_force = [&](int comp) {
        if (!with_force)
            return;
        const double2 *P = a.lforce + ((size_t)comp * a.nxh + x) * n;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = kidx(c, e);
                if (k < n) {
                    const double2 p = P[k];
                    tmp[c][e].x += a.lscale * p.x;
                    tmp[c][e].y += a.lscale * p.y;
                }
            }
    };

    // h_omega and g = i alpha H1 + i beta H3 (linearbl.f:325-343)
    double hw[2][2][E], dg[2][2][E];
    {
        double g[2][2][E];
        io.load(C0, tmp);                           // H1
        add_force(0);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                hw[0][c][e] = -beta * tmp[c][e].y;
                hw[1][c][e] = beta * tmp[c][e].x;
                g[0][c][e] = -alfa * tmp[c][e].y;
                g[1][c][e] = alfa * tmp[c][e].x;
            }
        io.load(C0 + 2 * fs, tmp);                  // H3
        add_force(2);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                hw[0][c][e] += alfa * tmp[c][e].y;
                hw[1][c][e] -= alfa * tmp[c][e].x;
                g[0][c][e] -= beta * tmp[c][e].y;
                g[1][c][e] += beta * tmp[c][e].x;
            }
        dcheb_w<SW, 2, E>(g, dg, l);
    }
    // h_v = -(dg/dy + k^2 H2), f_v = p_v - c1 h_v, f_omega = p_omega - c1 h_omega
    // (linearbl.f:345-361); channels 0, 1 of f are f_v, channels 2, 3 f_omega.
    constexpr bool early = SW == 16;
    double f[4][2][E], hv[2][2][E];
    io.load(C0 + fs, tmp);                          // H2
    add_force(1);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            dg[0][c][e] = -(dg[0][c][e] + k2 * tmp[c][e].x);   // h_v from here on
            dg[1][c][e] = -(dg[1][c][e] + k2 * tmp[c][e].y);
        }
    io.load(P0, tmp);                               // p_v
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            f[0][c][e] = k0 ? 0. : tmp[c][e].x - c1 * dg[0][c][e];
            f[1][c][e] = k0 ? 0. : tmp[c][e].y - c1 * dg[1][c][e];
            tmp[c][e] = make_double2(-c2 * dg[0][c][e], -c2 * dg[1][c][e]);
            hv[0][c][e] = dg[0][c][e];
            hv[1][c][e] = dg[1][c][e];
        }
    if (early)
        io.store(P0, tmp);
    io.load(P0 + fs, tmp);                          // p_omega
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            f[2][c][e] = k0 ? 0. : tmp[c][e].x - c1 * hw[0][c][e];
            f[3][c][e] = k0 ? 0. : tmp[c][e].y - c1 * hw[1][c][e];
            tmp[c][e] = make_double2(-c2 * hw[0][c][e], -c2 * hw[1][c][e]);
        }
    if (early)
        io.store(P0 + fs, tmp);
    // ibc = 100 (evalbc.f:142-159 with the values bu1jr(:, i, 1)): u_i(+1) =
    // U_i(+1) of the base flow on the kz = 0 plane (a wall elsewhere), so
    // v(+1) = V, and by continuity and the definition of omega_y dv/dy(+1)
    // = -i (alpha U + beta W), omega_y(+1) = i (beta U - alpha W)
    double2 vup = make_double2(0., 0.), dvp = vup, omp = vup;
    if constexpr (FT == 2) {
        if (z == 0 && a.bj && !k0) {
            const double2 ju = a.bj[x], jw = a.bj[2 * a.nxh + x];
            vup = a.bj[a.nxh + x];
            dvp = make_double2(alfa * ju.y + beta * jw.y, -(alfa * ju.x + beta * jw.x));
            omp = make_double2(-(beta * ju.y - alfa * jw.y), beta * ju.x - alfa * jw.x);
        }
    }
    // boundary rows per class: half the value at +1 in each (zero below)
    const double bc0[4][2] = {{0., 0.}, {0., 0.}, {.5 * omp.x, .5 * omp.x},
                              {.5 * omp.y, .5 * omp.y}};
    // phi and omega: Helmholtz solves with lam^2 (linearbl.f:625-626)
    io.load(lf.gp, tmp);
    solve_w<SW, 4, E, KT>(n, lam2, tmp, tw, f, bc0, l);
    double v[2][2][E], om[2][2][E];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            v[0][c][e] = f[0][c][e];
            v[1][c][e] = f[1][c][e];
            om[0][c][e] = f[2][c][e];
            om[1][c][e] = f[3][c][e];
        }
    // v from phi with k^2 (linearbl.f:1000-1001)
    {
        io.load(kf.gp, tmp);
        // v(+1) = 0 and v(-1) = the imposed wall velocity: (w/2, -w/2) for
        // the even and odd classes (linearbl.f:458-469, 944-945)
        double2 wv = make_double2(0., 0.);
        if (a.vwall && !k0) {
            wv = a.vwall[t];
            wv.x *= .5 * a.vscale;
            wv.y *= .5 * a.vscale;
        }
        const double bcv[2][2] = {{wv.x + .5 * vup.x, -wv.x + .5 * vup.x},
                                  {wv.y + .5 * vup.y, -wv.y + .5 * vup.y}};
        solve_w<SW, 2, E, KT>(n, k2, tmp, tw, v, bcv, l);
    }
    // superpose v_h per parity so that dv/dy = 0 at the walls
    // (linearbl.f:1023-1070), or dv/dy(+1) = dvp with ibc = 100 (half of it
    // in each class): class-c wall slope sum_k k^2 v_k
    {
        double s[4] = {0., 0., 0., 0.};
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const double kk = (double)kidx(c, e);
                s[0 + c] += kk * kk * v[0][c][e];
                s[2 + c] += kk * kk * v[1][c][e];
            }
        wave_sum<SW, 4>(s, l);
        const double2 sh = lf.sh[t];
        const double shc[2] = {sh.x, sh.y};
        io.load(lf.vh, tmp);
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const double cr = (.5 * dvp.x - s[c]) * (1. / shc[c]);
            const double ci = (.5 * dvp.y - s[2 + c]) * (1. / shc[c]);
#pragma unroll
            for (int e = 0; e < E; e++) {
                v[0][c][e] = v[0][c][e] + cr * tmp[c][e].x;
                v[1][c][e] = v[1][c][e] + ci * tmp[c][e].x;
            }
        }
        if constexpr (FT == 1) {
            // free stream at y = +1 (ibc = 101; evald.f, evalbc.f:142-159,
            // solvbc.f, linearbl.f:946-1150): du_i/dy = dU_i/dy of the base
            // flow there (kz = 0 plane, zero elsewhere).  By continuity and
            // the definition of omega_y that fixes d2v/dy2 = -i (alpha U' +
            // beta W') and d om/dy = i (beta U' - alpha W') at +1, and dv/dy
            // = V'.  With v_h1 = v_h scaled per class to unit slope at +1,
            // v_h2 and phi_h = omega_h (A1, A2, B their derivatives there):
            //   v -= c1 v_h1 + c2 v_h2,  om -= c3 phi_h,
            //   c1 = -V'/2, c2 = (d2v_p - c1 A1 - d2v*)/A2,
            //   c3 = (dom_p - dom*)/B.
            // The mean column (k0) keeps its own condition below.
            double q[4] = {0., 0., 0., 0.};
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const double kk = (double)kidx(c, e) * (double)kidx(c, e);
                    const double w2 = kk * (kk - 1.) * (1. / 3.);
                    q[0] += w2 * v[0][c][e];
                    q[1] += w2 * v[1][c][e];
                    q[2] += kk * om[0][c][e];
                    q[3] += kk * om[1][c][e];
                }
            wave_sum<SW, 4>(q, l);
            const double A1 = lf.hc[4 * t], A2 = lf.hc[4 * t + 1], B = lf.hc[4 * t + 2];
            double2 ju = make_double2(0., 0.), jv = ju, jw = ju;
            if (z == 0 && a.bj) {
                ju = a.bj[x];
                jv = a.bj[a.nxh + x];
                jw = a.bj[2 * a.nxh + x];
            }
            const double sr = alfa * ju.x + beta * jw.x, si = alfa * ju.y + beta * jw.y;
            const double tr = beta * ju.x - alfa * jw.x, ti = beta * ju.y - alfa * jw.y;
            const double c1r = k0 ? 0. : -.5 * jv.x, c1i = k0 ? 0. : -.5 * jv.y;
            const double c2r = k0 ? 0. : (q[0] - c1r * A1 - si) / A2;
            const double c2i = k0 ? 0. : (q[1] - c1i * A1 + sr) / A2;
            const double c3r = k0 ? 0. : (q[2] + ti) / B;
            const double c3i = k0 ? 0. : (q[3] - tr) / B;
            double2 hb[2][E];
            io.load(lf.hb, hb);                     // (phi_h, v_h2)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const double r1 = 1. / shc[c];
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const double h1 = tmp[c][e].x * r1, h2 = hb[c][e].y, ph = hb[c][e].x;
                    v[0][c][e] -= c1r * h1 + c2r * h2;
                    v[1][c][e] -= c1i * h1 + c2i * h2;
                    om[0][c][e] -= c3r * ph;
                    om[1][c][e] -= c3i * ph;
                }
            }
        }
    }
    // dv, phi = D^2 v - k^2 v (linearbl.f:1151-1166)
    double dv[2][2][E], ph[2][2][E];
    dcheb_w<SW, 2, E>(v, dv, l);
    dcheb_w<SW, 2, E>(dv, ph, l);
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                ph[q][c][e] = ph[q][c][e] - k2 * v[q][c][e];
    CV *Cw = cols<CV>(a);
    // u, v, w (linearbl.f:1190-1212) in Chebyshev space; the k = 0 column
    // stores nothing here.  Its segment (segment 0 of the wave with t0 = 0)
    // is masked off for the whole wave, since every lane of the wave takes
    // part in the staged stores of every segment.
    const unsigned ownmask = io.ownmask;
    if (t0 == 0)
        io.ownmask &= ~1u;
    io.own = own && !k0;
    {
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                tmp[c][e] = make_double2((-alfa * dv[1][c][e] + beta * om[1][c][e]) * k2i,
                                         (alfa * dv[0][c][e] - beta * om[0][c][e]) * k2i);
        io.store(Cw, tmp);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                tmp[c][e] = make_double2(v[0][c][e], v[1][c][e]);
        io.store(Cw + fs, tmp);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                tmp[c][e] = make_double2((-alfa * om[1][c][e] - beta * dv[1][c][e]) * k2i,
                                         (alfa * om[0][c][e] + beta * dv[0][c][e]) * k2i);
        io.store(Cw + 2 * fs, tmp);
    }
    // om_x, om_z
    double dom[2][2][E];
    dcheb_w<SW, 2, E>(om, dom, l);
    {
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                tmp[c][e] = make_double2((-alfa * dom[1][c][e] - beta * ph[1][c][e]) * k2i,
                                         (alfa * dom[0][c][e] + beta * ph[0][c][e]) * k2i);
        io.store(Cw + 3 * fs, tmp);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                tmp[c][e] = make_double2((-beta * dom[1][c][e] + alfa * ph[1][c][e]) * k2i,
                                         (beta * dom[0][c][e] - alfa * ph[0][c][e]) * k2i);
        io.store(Cw + 4 * fs, tmp);
    }
    // next partial right-hand sides (linearbl.f:1761-1784, tau branch),
    // (with 16-lane segments completing the -c2 h parts stored above)
    io.ownmask = ownmask;
    io.own = own;
    const double k2m = k2 - c1n;
    {
        double d1[2][2][E], d2[2][2][E];
        dcheb_w<SW, 2, E>(ph, d1, l);
        dcheb_w<SW, 2, E>(d1, d2, l);
        if (early)
            io.load(P0, tmp);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (!early)
                    tmp[c][e] = make_double2(-c2 * hv[0][c][e], -c2 * hv[1][c][e]);
                tmp[c][e] = make_double2(tmp[c][e].x - d2[0][c][e] + k2m * ph[0][c][e],
                                         tmp[c][e].y - d2[1][c][e] + k2m * ph[1][c][e]);
            }
        io.store(P0, tmp);
    }
    {
        double d1[2][2][E];
        dcheb_w<SW, 2, E>(dom, d1, l);
        if (early)
            io.load(P0 + fs, tmp);
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                if (!early)
                    tmp[c][e] = make_double2(-c2 * hw[0][c][e], -c2 * hw[1][c][e]);
                tmp[c][e] = make_double2(tmp[c][e].x - d1[0][c][e] + k2m * om[0][c][e],
                                         tmp[c][e].y - d1[1][c][e] + k2m * om[1][c][e]);
            }
        io.store(P0 + fs, tmp);
    }
    if (!k0)
        return;
    // k = 0: mean flow and constant-flux pressure gradient
    // (linearbl.f:303-319, 1216-1417, 1734-1746; tau branch, ibc = 0); the
    // real parts of H1, H3 are read again (slots 0-2 are not yet rewritten)
    const CV *Cc = C0 + (size_t)t * n;
    double h1r[2][E], h3r[2][E], uw[2][2][E];
    double2 gl[2][E];
    const double2 *glp = lf.gp + (size_t)t * n;
    double px = *a.px;
    const double2 z2 = make_double2(0., 0.);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = kidx(c, e);
            const bool ok = k < n;
            h1r[c][e] = ok ? widen(Cc[k]).x : 0.;
            h3r[c][e] = ok ? widen(Cc[2 * fs + k]).x : 0.;
            if (a.mforce && ok) {
                // the parallel-flow forcing of the mean mode (nonlinbl.f:
                // 771-795), already in Chebyshev space
                h1r[c][e] += a.mforce[k];
                h3r[c][e] += a.mforce[n + k];
            }
            if (a.lforce && ok) {
                h1r[c][e] += a.lscale * a.lforce[k].x;
                h3r[c][e] += a.lscale * a.lforce[(size_t)2 * a.nxh * n + k].x;
            }
            gl[c][e] = ok ? glp[k] : z2;
        }
    auto rhs = [&](double pxs) {
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = kidx(c, e);
                const bool ok = k < n;
                uw[0][c][e] = ok ? a.puw[k] - c1 * h1r[c][e] : 0.;
                uw[1][c][e] = ok ? a.puw[n + k] - c1 * h3r[c][e] : 0.;
            }
        if (L == 0)
            uw[0][0][0] += pxs;
    };
    const double bcu[2][2] = {{.5 * (a.u0upp + a.u0low), .5 * (a.u0upp - a.u0low)},
                              {.5 * (a.w0upp + a.w0low), .5 * (a.w0upp - a.w0low)}};
    rhs(2. * re * px);
    solve_w<SW, 2, E, KT>(n, lam2, gl, tw, uw, bcu, l);
    if (a.cflux) {
        auto flux = [&]() {
            // integral over [-1,1] of sum_k u_k T_k (linearbl.f:1367-1370)
            double m[1] = {0.};
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = 2 * (L * E + e);
                if (k < n)
                    m[0] -= uw[0][0][e] * 2. / (double)(k * k - 1);
            }
            wave_sum<SW, 1>(m, l);
            return m[0];
        };
        const double m1 = flux();
        rhs(4. * re * px);
        solve_w<SW, 2, E, KT>(n, lam2, gl, tw, uw, bcu, l);
        const double m2 = flux();
        px = px * (1. + (a.mflux - m1) / (m2 - m1));
        rhs(2. * re * px);
        solve_w<SW, 2, E, KT>(n, lam2, gl, tw, uw, bcu, l);
        if (L == 0)
            *a.px = px;
    }
    if constexpr (FT == 1) {
        // free stream above (linearbl.f:1500-1531): add phi_h (the k = 0
        // lam^2 solution with 1 per class) so that du0/dy(+1) = du0upp and
        // dw0/dy(+1) = 0; B = dphi_h/dy at +1
        double q[2] = {0., 0.};
        double ph[2][E];
        const double2 *hbp = lf.hb + (size_t)t * n;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = kidx(c, e);
                const double kk = (double)k * (double)k;
                ph[c][e] = k < n ? hbp[k].x : 0.;
                q[0] += kk * uw[0][c][e];
                q[1] += kk * uw[1][c][e];
            }
        wave_sum<SW, 2>(q, l);
        const double B = lf.hc[4 * t + 2];
        const double cu = (q[0] - a.du0upp) / B, cw = q[1] / B;
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                uw[0][c][e] -= cu * ph[c][e];
                uw[1][c][e] -= cw * ph[c][e];
            }
    }
    double duw[2][2][E], d2uw[2][2][E];
    dcheb_w<SW, 2, E>(uw, duw, l);
    dcheb_w<SW, 2, E>(duw, d2uw, l);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = kidx(c, e);
            if (k >= n)
                continue;
            a.puw[k] = -d2uw[0][c][e] - c1n * uw[0][c][e] - c2 * h1r[c][e];
            a.puw[n + k] = -d2uw[1][c][e] - c1n * uw[1][c][e] - c2 * h3r[c][e];
            CV *Ck = Cw + (size_t)t * n;
            put(Ck + k, uw[0][c][e], 0.);
            put(Ck + fs + k, k == 0 ? a.v0 : 0., 0.);   // mean v: the mean wall v
            put(Ck + 2 * fs + k, uw[1][c][e], 0.);
            put(Ck + 3 * fs + k, duw[1][c][e], 0.);
            put(Ck + 4 * fs + k, -duw[0][c][e], 0.);
            if constexpr (sizeof(CV) == sizeof(float2)) {
                // the mixed form also keeps the mean-flow column in FP64, so
                // that downloads and the pressure see the FP64 mean profiles
                double2 *C64 = a.C + (size_t)t * n;
                C64[k] = make_double2(uw[0][c][e], 0.);
                C64[fs + k] = make_double2(k == 0 ? a.v0 : 0., 0.);
                C64[2 * fs + k] = make_double2(uw[1][c][e], 0.);
                C64[3 * fs + k] = make_double2(duw[1][c][e], 0.);
                C64[4 * fs + k] = make_double2(-duw[0][c][e], 0.);
            }
        }
}

}  // namespace

template <int SW, int E, typename CV>
void launch_seg(const ColArgs &a, LamFactor lf, LamFactor kf, double an, double bn,
                double anp1, double bnp1, hipStream_t s)
{
    const int ncol = a.nz * a.nxh;
    constexpr int per_block = 4 * (WV / SW);
    dim3 grid((ncol + per_block - 1) / per_block);
    if (a.ibc == 101)
        hipLaunchKernelGGL((k_tau_w<SW, E, CV, 1>), grid, dim3(256), 0, s, a, lf, kf, an, bn,
                           anp1, bnp1);
    else if (a.ibc == 100)
        hipLaunchKernelGGL((k_tau_w<SW, E, CV, 2>), grid, dim3(256), 0, s, a, lf, kf, an, bn,
                           anp1, bnp1);
    else
        hipLaunchKernelGGL((k_tau_w<SW, E, CV, 0>), grid, dim3(256), 0, s, a, lf, kf, an, bn,
                           anp1, bnp1);
}

// Segment width: 64 lanes (one column per wave).  SIMSON_TAU_SEG=16 or 32
// selects narrower segments where a class fits (4 or 2 columns per wave,
// fewer lane moves per column).  At 128x97x128 they measured slower, since
// 8128 columns then make only 2 (16 lanes) or 4 (32 lanes) waves per SIMD
// and the scans' dependent chains are no longer hidden: tau 82 us (16) and
// 91 us (32, before the staged I/O) against 62 us (64) per FP64 substep.
int tau_segment(int ny)
{
    static int forced = -1;
    if (forced < 0) {
        const char *e = getenv("SIMSON_TAU_SEG");
        forced = e ? atoi(e) : 0;
    }
    const int ce = (ny + 1) / 2;              // elements of the even class
    if (forced == 32 && (ce + 31) / 32 <= 3)
        return 32;
    if (forced == 16 && (ce + 15) / 16 <= 4)
        return 16;
    return 64;
}

template <typename CV>
int launch_tau_t(const ColArgs &a, LamFactor lf, LamFactor kf, double an, double bn,
                 double anp1, double bnp1, hipStream_t s)
{
    const int ce = (a.ny + 1) / 2;
    const int SW = tau_segment(a.ny);
    const int E = (ce + SW - 1) / SW;
    if (SW == 16) {
        switch (E) {
        case 1: launch_seg<16, 1, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        case 2: launch_seg<16, 2, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        case 3: launch_seg<16, 3, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        case 4: launch_seg<16, 4, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        default: return SIMSON_E_ARG;
        }
    }
    if (SW == 32) {
        switch (E) {
        case 1: launch_seg<32, 1, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        case 2: launch_seg<32, 2, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        case 3: launch_seg<32, 3, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
        default: return SIMSON_E_ARG;
        }
    }
    switch (E) {
    case 1: launch_seg<64, 1, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
    case 2: launch_seg<64, 2, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
    case 3: launch_seg<64, 3, CV>(a, lf, kf, an, bn, anp1, bnp1, s); return 0;
    default: return SIMSON_E_ARG;
    }
}

int launch_tau(const ColArgs &a, LamFactor lf, LamFactor kf, double an, double bn, double anp1,
               double bnp1, bool fp32_cols, hipStream_t s)
{
    return fp32_cols ? launch_tau_t<float2>(a, lf, kf, an, bn, anp1, bnp1, s)
                     : launch_tau_t<double2>(a, lf, kf, an, bn, anp1, bnp1, s);
}

}  // namespace sg
