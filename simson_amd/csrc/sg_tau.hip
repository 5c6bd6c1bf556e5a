// Step 3 of the bla substep (bla/linearbl.f, tau method, ibc = 0) with one
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
// The scans run on DPP lane moves within 16-lane rows and row broadcasts or
// lane reads across rows (sg_wave.h), so no step waits on an LDS permute.
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

// Row weights of the tau system for the elements of this lane (they depend
// on k only): wm = c_{k-2}/(4k(k-1)), w0 = [k<=n-3]/(2(k^2-1)),
// wp = [k<=n-5]/(4k(k+1)), so that R_k = wm f_{k-2} - w0 f_k + wp f_{k+2}
// and U_k = -lam wp (sg_tau.h); zero outside the rows k = 2..n-1.
template <int E>
struct TauW {
    double wm[2][E], w0[2][E], wp[2][E];
    __device__ __forceinline__ TauW(int n, int L)
    {
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int M = L * E + e, k = 2 * M + c;
                const double kk = (double)k;
                const bool row = (M >= 1) && (k < n);
                wm[c][e] = row ? ((k == 2) ? 2.0 : 1.0) / (4.0 * kk * (kk - 1.0)) : 0.0;
                w0[c][e] = (row && k <= n - 3) ? 1.0 / (2.0 * (kk * kk - 1.0)) : 0.0;
                wp[c][e] = (row && k <= n - 5) ? 1.0 / (4.0 * kk * (kk + 1.0)) : 0.0;
            }
    }
};

// Tau solve of u'' - lam u = f for NC right-hand sides sharing lam (one
// column), boundary rows bc[q][class]; gp holds (gamma, 1/pivot).
template <int SW, int NC, int E>
__device__ __forceinline__ void solve_w(int n, double lam, const double2 (&gp)[2][E],
                                        const TauW<E> &tw, const double (&f)[NC][2][E],
                                        const double (&bc)[NC][2], double (&u)[NC][2][E], int l)
{
    const int L = l & (SW - 1);
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
            // -U_k / piv_k and R_k / piv_k with the row weights of TauW
            be[c][e] = row ? lam * tw.wp[c][e] * ip : 0.0;
            bq[c][e] = (M == 0) ? 1.0 : (k < n ? -gp[c][e].x : 0.0);
#pragma unroll
            for (int q = 0; q < NC; q++) {
                const double fm = e > 0 ? f[q][c][e - 1] : fprev[q][c];
                const double fp = e + 1 < E ? f[q][c][e + 1] : fnext[q][c];
                const double R = tw.wm[c][e] * fm - tw.w0[c][e] * f[q][c][e] + tw.wp[c][e] * fp;
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
#pragma unroll
    for (int c = 0; c < 2; c++) {
        double Xc[NC];
#pragma unroll
        for (int q = 0; q < NC; q++)
            Xc[q] = X[q][c];
        wv::scan_suffix<SW, NC, true>(Xc, Y[c], l);
#pragma unroll
        for (int q = 0; q < NC; q++)
            X[q][c] = Xc[q];
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
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int q = 0; q < NC; q++) {
            const double a0 = (bc[q][c] - sums[q * 2 + c]) / sums[2 * NC + c];
#pragma unroll
            for (int e = 0; e < E; e++)
                u[q][c][e] = Ap[q][c][e] + Bp[c][e] * a0;
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

// One column per SW-lane segment (64 / SW columns per wave).  A segment
// without a column of its own (past the end, or the z oddball plane,
// linearbl.f:227) computes on column 1 and stores nothing, so that every
// lane of the wave stays active through the lane moves.
template <int SW, int E, typename CV>
__global__ void __launch_bounds__(256, (SW == 64 && E == 1) ? 3 : 1)
    k_tau_w(ColArgs a, LamFactor lf, LamFactor kf, double an, double bn, double anp1, double bnp1)
{
    const int l = threadIdx.x & (WV - 1);
    const int L = l & (SW - 1);
    const int tw_ = (blockIdx.x * 4 + (threadIdx.x >> 6)) * (WV / SW) + l / SW;
    const int ncol = a.nz * a.nxh;
    const bool own = tw_ < ncol && tw_ / a.nxh != a.nz / 2;
    const int t = own ? tw_ : 1;
    const int z = t / a.nxh, x = t - z * a.nxh;
    const int n = a.ny;
    const TauW<E> tw(n, L);
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
    const CV *Cc = cols<CV>(a) + (size_t)t * n;
    const double2 *Pc = a.Pp + (size_t)t * n;
    const double2 *glp = lf.gp + (size_t)t * n;
    const double2 *gkp = kf.gp + (size_t)t * n;
    const double *vhp = lf.vh + (size_t)t * n;

    double H1[2][2][E], H2[2][2][E], H3[2][2][E], pv[2][2][E], pw[2][2][E], vh[2][E];
    double2 gl[2][E], gk[2][E];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = 2 * (L * E + e) + c;
            const bool ok = k < n;
            const double2 z2 = make_double2(0., 0.);
            const double2 h1 = ok ? widen(Cc[k]) : z2, h2 = ok ? widen(Cc[fs + k]) : z2;
            const double2 h3 = ok ? widen(Cc[2 * fs + k]) : z2;
            const double2 p0 = ok ? Pc[k] : z2, p1 = ok ? Pc[fs + k] : z2;
            H1[0][c][e] = h1.x; H1[1][c][e] = h1.y;
            H2[0][c][e] = h2.x; H2[1][c][e] = h2.y;
            H3[0][c][e] = h3.x; H3[1][c][e] = h3.y;
            pv[0][c][e] = p0.x; pv[1][c][e] = p0.y;
            pw[0][c][e] = p1.x; pw[1][c][e] = p1.y;
            gl[c][e] = ok ? glp[k] : z2;
            gk[c][e] = ok ? gkp[k] : z2;
            vh[c][e] = ok ? vhp[k] : 0.;
        }
    // h_omega, f_omega, g = i alpha H1 + i beta H3 (linearbl.f:325-343)
    double hw[2][2][E], fw[2][2][E], g[2][2][E], dg[2][2][E];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            hw[0][c][e] = -beta * H1[1][c][e] + alfa * H3[1][c][e];
            hw[1][c][e] = beta * H1[0][c][e] - alfa * H3[0][c][e];
            fw[0][c][e] = pw[0][c][e] - c1 * hw[0][c][e];
            fw[1][c][e] = pw[1][c][e] - c1 * hw[1][c][e];
            g[0][c][e] = -alfa * H1[1][c][e] - beta * H3[1][c][e];
            g[1][c][e] = alfa * H1[0][c][e] + beta * H3[0][c][e];
        }
    dcheb_w<SW, 2, E>(g, dg, l);
    // h_v = -(dg/dy + k^2 H2), f_v = p_v - c1 h_v (linearbl.f:345-361)
    double hv[2][2][E], fv[2][2][E];
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                hv[q][c][e] = -(dg[q][c][e] + k2 * H2[q][c][e]);
                fv[q][c][e] = pv[q][c][e] - c1 * hv[q][c][e];
                if (k0) {
                    fv[q][c][e] = 0.;
                    fw[q][c][e] = 0.;
                }
            }
    const double bc0[2][2] = {{0., 0.}, {0., 0.}};
    // Helmholtz solves with lam^2 (linearbl.f:625-626) and v with k^2
    // (linearbl.f:1000-1001); the homogeneous pair phi_h, v_h is precomputed
    double ph[2][2][E], om[2][2][E], v[2][2][E];
    solve_w<SW, 2, E>(n, lam2, gl, tw, fv, bc0, ph, l);
    solve_w<SW, 2, E>(n, lam2, gl, tw, fw, bc0, om, l);
    solve_w<SW, 2, E>(n, k2, gk, tw, ph, bc0, v, l);
    // superpose v_h per parity so that dv/dy = 0 at the walls
    // (linearbl.f:1023-1070): class-c wall slope sum_k k^2 v_k
    {
        double s[4] = {0., 0., 0., 0.};
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const double kk = (double)(2 * (L * E + e) + c);
                s[0 + c] += kk * kk * v[0][c][e];
                s[2 + c] += kk * kk * v[1][c][e];
            }
        wave_sum<SW, 4>(s, l);
        const double2 sh = lf.sh[t];
        const double shc[2] = {sh.x, sh.y};
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const double cr = (0. - s[c]) * (1. / shc[c]);
            const double ci = (0. - s[2 + c]) * (1. / shc[c]);
#pragma unroll
            for (int e = 0; e < E; e++) {
                v[0][c][e] = v[0][c][e] + cr * vh[c][e];
                v[1][c][e] = v[1][c][e] + ci * vh[c][e];
            }
        }
    }
    // dv, phi = D^2 v - k^2 v, d omega (linearbl.f:1151-1166)
    double dv[2][2][E], d2v[2][2][E], dom[2][2][E];
    dcheb_w<SW, 2, E>(v, dv, l);
    dcheb_w<SW, 2, E>(dv, d2v, l);
    dcheb_w<SW, 2, E>(om, dom, l);
#pragma unroll
    for (int q = 0; q < 2; q++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++)
                ph[q][c][e] = d2v[q][c][e] - k2 * v[q][c][e];
    // next partial right-hand sides (linearbl.f:1761-1784, tau branch)
    double d1[2][2][E], d2[2][2][E];
    dcheb_w<SW, 2, E>(ph, d1, l);
    dcheb_w<SW, 2, E>(d1, d2, l);
    const double k2m = k2 - c1n;
    CV *Cw = cols<CV>(a) + (size_t)t * n;
    double2 *Pw = a.Pp + (size_t)t * n;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = 2 * (L * E + e) + c;
            if (own && k < n)
                Pw[k] = make_double2(-d2[0][c][e] + k2m * ph[0][c][e] - c2 * hv[0][c][e],
                                     -d2[1][c][e] + k2m * ph[1][c][e] - c2 * hv[1][c][e]);
        }
    dcheb_w<SW, 2, E>(dom, d1, l);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = 2 * (L * E + e) + c;
            if (own && k < n)
                Pw[fs + k] = make_double2(-d1[0][c][e] + k2m * om[0][c][e] - c2 * hw[0][c][e],
                                          -d1[1][c][e] + k2m * om[1][c][e] - c2 * hw[1][c][e]);
        }
    if (!k0) {
        // u, w, om_x, om_z (linearbl.f:1190-1212) in Chebyshev space
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = 2 * (L * E + e) + c;
                if (!own || k >= n)
                    continue;
                const double dvr = dv[0][c][e], dvi = dv[1][c][e];
                const double omr = om[0][c][e], omi = om[1][c][e];
                const double dor = dom[0][c][e], doi = dom[1][c][e];
                const double phr = ph[0][c][e], phi = ph[1][c][e];
                put(Cw + k, (-alfa * dvi + beta * omi) * k2i, (alfa * dvr - beta * omr) * k2i);
                put(Cw + fs + k, v[0][c][e], v[1][c][e]);
                put(Cw + 2 * fs + k, (-alfa * omi - beta * dvi) * k2i,
                    (alfa * omr + beta * dvr) * k2i);
                put(Cw + 3 * fs + k, (-alfa * doi - beta * phi) * k2i,
                    (alfa * dor + beta * phr) * k2i);
                put(Cw + 4 * fs + k, (-beta * doi + alfa * phi) * k2i,
                    (beta * dor - alfa * phr) * k2i);
            }
        return;
    }
    // k = 0: mean flow and constant-flux pressure gradient
    // (linearbl.f:303-319, 1216-1417, 1734-1746; tau branch, ibc = 0)
    double fuw[2][2][E], uw[2][2][E];
    double px = *a.px;
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = 2 * (L * E + e) + c;
            const bool ok = k < n;
            fuw[0][c][e] = ok ? a.puw[k] - c1 * H1[0][c][e] : 0.;
            fuw[1][c][e] = ok ? a.puw[n + k] - c1 * H3[0][c][e] : 0.;
        }
    if (L == 0)
        fuw[0][0][0] += 2. * re * px;
    const double bcu[2][2] = {{.5 * (a.u0upp + a.u0low), .5 * (a.u0upp - a.u0low)},
                              {.5 * (a.w0upp + a.w0low), .5 * (a.w0upp - a.w0low)}};
    solve_w<SW, 2, E>(n, lam2, gl, tw, fuw, bcu, uw, l);
    if (a.cflux) {
        auto flux = [&](const double (&u)[2][2][E]) {
            // integral over [-1,1] of sum_k u_k T_k (linearbl.f:1367-1370)
            double m[1] = {0.};
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int k = 2 * (L * E + e);
                if (k < n)
                    m[0] -= u[0][0][e] * 2. / (double)(k * k - 1);
            }
            wave_sum<SW, 1>(m, l);
            return m[0];
        };
        const double m1 = flux(uw);
        if (L == 0)
            fuw[0][0][0] = fuw[0][0][0] + 2. * re * px;
        solve_w<SW, 2, E>(n, lam2, gl, tw, fuw, bcu, uw, l);
        const double m2 = flux(uw);
        if (L == 0)
            fuw[0][0][0] = fuw[0][0][0] - 4. * re * px;
        px = px * (1. + (a.mflux - m1) / (m2 - m1));
        if (L == 0)
            fuw[0][0][0] = fuw[0][0][0] + 2. * re * px;
        solve_w<SW, 2, E>(n, lam2, gl, tw, fuw, bcu, uw, l);
        if (L == 0)
            *a.px = px;
    }
    double duw[2][2][E], d2uw[2][2][E];
    dcheb_w<SW, 2, E>(uw, duw, l);
    dcheb_w<SW, 2, E>(duw, d2uw, l);
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int k = 2 * (L * E + e) + c;
            if (k >= n)
                continue;
            a.puw[k] = -d2uw[0][c][e] - c1n * uw[0][c][e] - c2 * H1[0][c][e];
            a.puw[n + k] = -d2uw[1][c][e] - c1n * uw[1][c][e] - c2 * H3[0][c][e];
            put(Cw + k, uw[0][c][e], 0.);
            put(Cw + fs + k, 0., 0.);
            put(Cw + 2 * fs + k, uw[1][c][e], 0.);
            put(Cw + 3 * fs + k, duw[1][c][e], 0.);
            put(Cw + 4 * fs + k, -duw[0][c][e], 0.);
            if constexpr (sizeof(CV) == sizeof(float2)) {
                // the mixed form also keeps the mean-flow column in FP64, so
                // that downloads and the pressure see the FP64 mean profiles
                double2 *C64 = a.C + (size_t)t * n;
                C64[k] = make_double2(uw[0][c][e], 0.);
                C64[fs + k] = make_double2(0., 0.);
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
    hipLaunchKernelGGL((k_tau_w<SW, E, CV>), grid, dim3(256), 0, s, a, lf, kf, an, bn, anp1, bnp1);
}

// Segment width: 64 lanes (one column per wave, three waves per SIMD at
// ny <= 127), unless SIMSON_TAU_SEG=32 asks for two columns per wave (fewer
// lane moves per column but one wave per SIMD: slower at 128x97x128,
// 91 against 62 us per substep).
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
    return 64;
}

template <typename CV>
int launch_tau_t(const ColArgs &a, LamFactor lf, LamFactor kf, double an, double bn,
                 double anp1, double bnp1, hipStream_t s)
{
    const int ce = (a.ny + 1) / 2;
    const int SW = tau_segment(a.ny);
    const int E = (ce + SW - 1) / SW;
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
