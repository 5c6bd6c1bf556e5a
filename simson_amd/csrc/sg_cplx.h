// Complex arithmetic and in-register DFTs for FP64 (double2) and FP32
// (float2) step-2 kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace sg {

template <typename R>
struct Cx;

template <>
struct Cx<double> {
    typedef double2 V;
    static __device__ __forceinline__ V mk(double a, double b) { return make_double2(a, b); }
};

template <>
struct Cx<float> {
    typedef float2 V;
    static __device__ __forceinline__ V mk(float a, float b) { return make_float2(a, b); }
};

template <typename V>
struct ScalarOf;
template <>
struct ScalarOf<double2> {
    typedef double T;
};
template <>
struct ScalarOf<float2> {
    typedef float T;
};

template <typename V>
__device__ __forceinline__ V vmk(typename ScalarOf<V>::T a, typename ScalarOf<V>::T b)
{
    V r;
    r.x = a;
    r.y = b;
    return r;
}

template <typename V>
__device__ __forceinline__ V vmul(V a, V b)
{
    return vmk<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ __forceinline__ V vadd(V a, V b)
{
    return vmk<V>(a.x + b.x, a.y + b.y);
}
template <typename V>
__device__ __forceinline__ V vsub(V a, V b)
{
    return vmk<V>(a.x - b.x, a.y - b.y);
}
template <typename V>
__device__ __forceinline__ V vconj(V a)
{
    return vmk<V>(a.x, -a.y);
}
// (i * s) * a, s = +-1 or a real factor
template <typename V>
__device__ __forceinline__ V vmul_is(V a, typename ScalarOf<V>::T s)
{
    return vmk<V>(-s * a.y, s * a.x);
}
template <typename V>
__device__ __forceinline__ V vscale(V a, typename ScalarOf<V>::T s)
{
    return vmk<V>(s * a.x, s * a.y);
}

// In-register R-point DFT, y_q = sum_r v_r exp(SIGN 2 pi i r q / R), R <= 5.
template <typename V, int R, int SIGN>
__device__ __forceinline__ void dft_r(V *v)
{
    typedef typename ScalarOf<V>::T T;
    constexpr T sg = (T)SIGN;
    if constexpr (R == 2) {
        const V a = v[0], b = v[1];
        v[0] = vadd(a, b);
        v[1] = vsub(a, b);
    } else if constexpr (R == 4) {
        const V a = v[0], b = v[1], c = v[2], d = v[3];
        const V s0 = vadd(a, c), d0 = vsub(a, c), s1 = vadd(b, d), d1 = vsub(b, d);
        v[0] = vadd(s0, s1);
        v[2] = vsub(s0, s1);
        const V t = vmul_is(d1, sg);
        v[1] = vadd(d0, t);
        v[3] = vsub(d0, t);
    } else if constexpr (R == 3) {
        const T h = (T)0.86602540378443864676;        // sin(2 pi / 3)
        const V a = v[0], b = v[1], c = v[2];
        const V t1 = vadd(b, c);
        const V t2 = vmk<V>(a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y);
        const V t3 = vmul_is(vsub(b, c), sg * h);
        v[0] = vadd(a, t1);
        v[1] = vadd(t2, t3);
        v[2] = vsub(t2, t3);
    } else if constexpr (R == 5) {
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        const V a = v[0];
        const V b = vadd(v[1], v[4]), bm = vsub(v[1], v[4]);
        const V c = vadd(v[2], v[3]), cm = vsub(v[2], v[3]);
        const V r1 = vmk<V>(a.x + c1 * b.x + c2 * c.x, a.y + c1 * b.y + c2 * c.y);
        const V r2 = vmk<V>(a.x + c2 * b.x + c1 * c.x, a.y + c2 * b.y + c1 * c.y);
        const V i1 = vmk<V>(s1 * bm.x + s2 * cm.x, s1 * bm.y + s2 * cm.y);
        const V i2 = vmk<V>(s2 * bm.x - s1 * cm.x, s2 * bm.y - s1 * cm.y);
        v[0] = vadd(a, vadd(b, c));
        v[1] = vadd(r1, vmul_is(i1, sg));
        v[4] = vsub(r1, vmul_is(i1, sg));
        v[2] = vadd(r2, vmul_is(i2, sg));
        v[3] = vsub(r2, vmul_is(i2, sg));
    }
}

template <int N>
constexpr int first_radix()
{
    return N % 4 == 0 ? 4 : N % 2 == 0 ? 2 : N % 3 == 0 ? 3 : N % 5 == 0 ? 5 : 0;
}

constexpr bool radix235(int n)
{
    while (n % 2 == 0)
        n /= 2;
    while (n % 3 == 0)
        n /= 3;
    while (n % 5 == 0)
        n /= 5;
    return n == 1;
}

// Full N-point DFT in registers, y_k = sum_n v_n exp(SIGN 2 pi i n k / N),
// by decimation in time with the twiddles exp(+2 pi i t / TABN) of a table
// (N divides TABN; the indices are compile-time constants).
template <typename V, int N, int SIGN, int TABN>
__device__ __forceinline__ void reg_dft(V *v, const V *tw)
{
    static_assert(radix235(N), "length must factor into 2, 3 and 5");
    if constexpr (N == 1) {
        return;
    } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5) {
        dft_r<V, N, SIGN>(v);
    } else {
        constexpr int R = first_radix<N>(), NN = N / R;
        V sub[R][NN];
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int m = 0; m < NN; m++)
                sub[r][m] = v[r + R * m];
#pragma unroll
        for (int r = 0; r < R; r++)
            reg_dft<V, NN, SIGN, TABN>(sub[r], tw);
#pragma unroll
        for (int k = 0; k < NN; k++) {
            V t[R];
            t[0] = sub[0][k];
#pragma unroll
            for (int r = 1; r < R; r++) {
                if (k == 0) {
                    t[r] = sub[r][0];
                } else {
                    V w = tw[r * k * (TABN / N)];
                    if (SIGN < 0)
                        w.y = -w.y;
                    t[r] = vmul(sub[r][k], w);
                }
            }
            dft_r<V, R, SIGN>(t);
#pragma unroll
            for (int q = 0; q < R; q++)
                v[k + NN * q] = t[q];
        }
    }
}

}  // namespace sg
