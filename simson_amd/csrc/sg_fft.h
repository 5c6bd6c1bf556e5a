// Batched mixed-radix Stockham FFT in LDS (device side).
#pragma once

#include <hip/hip_runtime.h>

#include "sg_internal.h"

namespace sg {

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 conjd(double2 a) { return make_double2(a.x, -a.y); }
// (i * s) * a for s = +-1
__device__ __forceinline__ double2 mul_is(double2 a, double s) { return make_double2(-s * a.y, s * a.x); }

// In-register R-point DFT, y_q = sum_r v_r exp(sign 2 pi i r q / R).
__device__ __forceinline__ void dft_small(double2 *v, int R, double sign)
{
    if (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        double2 a = v[0], b = v[1], c = v[2], d = v[3];
        double2 s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = csub(b, d);
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        double2 t = mul_is(d1, sign);
        v[1] = cadd(d0, t);
        v[3] = csub(d0, t);
    } else if (R == 3) {
        const double h = 0.86602540378443864676;   // sin(2 pi / 3)
        double2 a = v[0], b = v[1], c = v[2];
        double2 t1 = cadd(b, c);
        double2 t2 = make_double2(a.x - 0.5 * t1.x, a.y - 0.5 * t1.y);
        double2 t3 = mul_is(csub(b, c), sign * h);
        v[0] = cadd(a, t1);
        v[1] = cadd(t2, t3);
        v[2] = csub(t2, t3);
    } else {  // R == 5
        const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
        double2 a = v[0];
        double2 b = cadd(v[1], v[4]), bm = csub(v[1], v[4]);
        double2 c = cadd(v[2], v[3]), cm = csub(v[2], v[3]);
        double2 r1 = make_double2(a.x + c1 * b.x + c2 * c.x, a.y + c1 * b.y + c2 * c.y);
        double2 r2 = make_double2(a.x + c2 * b.x + c1 * c.x, a.y + c2 * b.y + c1 * c.y);
        double2 i1 = make_double2(s1 * bm.x + s2 * cm.x, s1 * bm.y + s2 * cm.y);
        double2 i2 = make_double2(s2 * bm.x - s1 * cm.x, s2 * bm.y - s1 * cm.y);
        v[0] = cadd(a, cadd(b, c));
        v[1] = cadd(r1, mul_is(i1, sign));
        v[4] = csub(r1, mul_is(i1, sign));
        v[2] = cadd(r2, mul_is(i2, sign));
        v[3] = csub(r2, mul_is(i2, sign));
    }
}

// ---- compile-time radix schedule: radix-4 passes, then 2, 3, 5 ----
constexpr int rad_at(int n, int p)
{
    const int order[4] = {4, 2, 3, 5};
    for (int i = 0; i < 4; i++) {
        const int r = order[i];
        while (n % r == 0) {
            if (p == 0)
                return r;
            --p;
            n /= r;
        }
    }
    return 0;
}

constexpr int nrad_of(int n)
{
    int c = 0;
    while (rad_at(n, c) != 0)
        c++;
    return c;
}

constexpr bool fft_len_ok(int n)
{
    int m = n;
    for (int r : {4, 2, 3, 5})
        while (m % r == 0)
            m /= r;
    return m == 1;
}

template <int R, int SIGN>
__device__ __forceinline__ void dft_small_t(double2 *v)
{
    dft_small(v, R, (double)SIGN);
}

// One Stockham pass of a length-N transform over nb transforms whose start
// points are SB apart; twiddles exp(+2 pi i t/TABN) from tw (LDS).
template <int N, int TABN, int SB, int P, int NS, int SIGN>
__device__ __forceinline__ void fft_pass_t(const double2 *src, double2 *dst, int nb,
                                           const double2 *tw)
{
    constexpr int R = rad_at(N, P);
    constexpr int NBF = N / R;
    constexpr int TSTEP = TABN / (NS * R);
    for (int idx = threadIdx.x; idx < nb * NBF; idx += blockDim.x) {
        const int b = idx / NBF;
        const int j = idx - b * NBF;
        const int k = j % NS;
        const double2 *s = src + b * SB;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            v[r] = s[j + r * NBF];
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 w = tw[r * k * TSTEP];
                if (SIGN < 0)
                    w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small_t<R, SIGN>(v);
        double2 *d = dst + b * SB;
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; r++)
            d[base + r * NS] = v[r];
    }
    __syncthreads();
}

// Full length-N transform (all passes); returns the buffer with the result.
template <int N, int TABN, int SB, int SIGN, int P = 0, int NS = 1>
__device__ __forceinline__ double2 *fft_t(double2 *src, double2 *dst, int nb, const double2 *tw)
{
    if constexpr (P == nrad_of(N)) {
        return src;
    } else {
        fft_pass_t<N, TABN, SB, P, NS, SIGN>(src, dst, nb, tw);
        return fft_t<N, TABN, SB, SIGN, P + 1, NS * rad_at(N, P)>(dst, src, nb, tw);
    }
}

// One Stockham pass of radix R (a compile-time constant, so that the R
// values stay in registers) over nb transforms of length n ([b][n]);
// twiddles exp(+2 pi i t/tabn) from pl.tw, t = r k tstep < tabn.
template <int R>
__device__ __forceinline__ void lds_pass(const double2 *src, double2 *dst, int n, int nb,
                                         int Ns, const FftPlan &pl, double sign)
{
    const int nbf = n / R;
    const int tstep = pl.tabn / (Ns * R);
    for (int idx = threadIdx.x; idx < nb * nbf; idx += blockDim.x) {
        const int b = idx / nbf;
        const int j = idx - b * nbf;
        const int k = j % Ns;
        const double2 *s = src + b * n;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            v[r] = s[j + r * nbf];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 w = pl.tw[r * k * tstep];
                w.y *= sign;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small(v, R, sign);
        double2 *d = dst + b * n;
        const int base = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; r++)
            d[base + r * Ns] = v[r];
    }
    __syncthreads();
}

// nb transforms of length n stored contiguously ([b][n]) in src.  Ping-pongs
// between src and dst; returns the buffer holding the result.  All threads of
// the block must call it; it ends with a barrier.
__device__ inline double2 *lds_fft(double2 *src, double2 *dst, int n, int nb,
                                   const FftPlan &pl, double sign)
{
    int Ns = 1;
    for (int p = 0; p < pl.nrad; p++) {
        const int R = pl.rad[p];
        switch (R) {
        case 2: lds_pass<2>(src, dst, n, nb, Ns, pl, sign); break;
        case 3: lds_pass<3>(src, dst, n, nb, Ns, pl, sign); break;
        case 4: lds_pass<4>(src, dst, n, nb, Ns, pl, sign); break;
        default: lds_pass<5>(src, dst, n, nb, Ns, pl, sign); break;
        }
        double2 *t = src;
        src = dst;
        dst = t;
        Ns *= R;
    }
    return src;
}

}  // namespace sg
