// Wavefront-level data movement for 64-lane waves on gfx950, built on DPP
// (data-parallel primitives: a VALU operand modifier that reads another
// lane's register in the same instruction) instead of LDS permutes.
//
// DPP controls used (GFX9 encoding):
//   row_shr:s  0x110+s  lane i reads lane i-s within its 16-lane row
//   row_shl:s  0x100+s  lane i reads lane i+s within its row
//   wave_shl1  0x130    lane i reads lane i+1 across the wave
//   wave_shr1  0x138    lane i reads lane i-1 across the wave
//   row_bcast15 0x142   lane 15 of row r is read by row r+1
//   row_bcast31 0x143   lane 31 is read by rows 2 and 3
// A lane whose source does not exist, or whose row is masked off, keeps
// the 'old' operand (bound_ctrl off), which the helpers set to the identity
// of the combining operation.
#pragma once

#include <hip/hip_runtime.h>

namespace sg {

namespace wv {

constexpr int kWave = 64;

template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp(double old, double x)
{
    const long long o = __double_as_longlong(old), v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, ROWMASK,
                                               0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double readlane(double x, int lane)
{
    const long long v = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)v, lane);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// value of lane i+1 (lane 63 gets fill)
__device__ __forceinline__ double from_next(double x, double fill) { return dpp<0x130>(fill, x); }
// value of lane i-1 (lane 0 gets fill)
__device__ __forceinline__ double from_prev(double x, double fill) { return dpp<0x138>(fill, x); }

// Inclusive scans of affine maps r -> X[q] + Y r over the lanes.  Each lane
// holds the map of its segment; after the call lane i holds the composition
// of lanes 0..i (prefix: earlier lanes applied first, i.e.
// X_i <- X_i + Y_i X_{i-1}) or of lanes i..63 (suffix: X_i <- X_i + Y_i X_{i+1}).
// With HASY false, Y is 1 and the scan is a plain sum.
template <int NX, bool HASY>
__device__ __forceinline__ void combine(double (&X)[NX], double &Y, const double (&Xo)[NX],
                                        double Yo)
{
#pragma unroll
    for (int q = 0; q < NX; q++)
        X[q] = HASY ? fma(Y, Xo[q], X[q]) : X[q] + Xo[q];
    if (HASY)
        Y = Y * Yo;
}

template <int S, int NX, bool HASY, bool SUFFIX>
__device__ __forceinline__ void row_step(double (&X)[NX], double &Y)
{
    constexpr int ctrl = SUFFIX ? 0x100 + S : 0x110 + S;
    double Xo[NX];
#pragma unroll
    for (int q = 0; q < NX; q++)
        Xo[q] = dpp<ctrl>(0.0, X[q]);
    const double Yo = HASY ? dpp<ctrl>(1.0, Y) : 1.0;
    combine<NX, HASY>(X, Y, Xo, Yo);
}

template <int NX, bool HASY>
__device__ __forceinline__ void scan_prefix(double (&X)[NX], double &Y)
{
    row_step<1, NX, HASY, false>(X, Y);
    row_step<2, NX, HASY, false>(X, Y);
    row_step<4, NX, HASY, false>(X, Y);
    row_step<8, NX, HASY, false>(X, Y);
    {
        double Xo[NX];
#pragma unroll
        for (int q = 0; q < NX; q++)
            Xo[q] = dpp<0x142, 0xa>(0.0, X[q]);
        const double Yo = HASY ? dpp<0x142, 0xa>(1.0, Y) : 1.0;
        combine<NX, HASY>(X, Y, Xo, Yo);
    }
    {
        double Xo[NX];
#pragma unroll
        for (int q = 0; q < NX; q++)
            Xo[q] = dpp<0x143, 0xc>(0.0, X[q]);
        const double Yo = HASY ? dpp<0x143, 0xc>(1.0, Y) : 1.0;
        combine<NX, HASY>(X, Y, Xo, Yo);
    }
}

template <int NX, bool HASY>
__device__ __forceinline__ void scan_suffix(double (&X)[NX], double &Y, int lane)
{
    row_step<1, NX, HASY, true>(X, Y);
    row_step<2, NX, HASY, true>(X, Y);
    row_step<4, NX, HASY, true>(X, Y);
    row_step<8, NX, HASY, true>(X, Y);
    // rows 2, 1, 0 in turn take the total of the rows above them, which
    // the head lane (16 r) of the next row holds
#pragma unroll
    for (int r = 2; r >= 0; r--) {
        double Xo[NX];
#pragma unroll
        for (int q = 0; q < NX; q++)
            Xo[q] = readlane(X[q], 16 * (r + 1));
        const double Yo = HASY ? readlane(Y, 16 * (r + 1)) : 1.0;
        if ((lane >> 4) == r)
            combine<NX, HASY>(X, Y, Xo, Yo);
    }
}

// sum over the wave of NX channels, the same value on every lane
template <int NX>
__device__ __forceinline__ void sum_all(double (&X)[NX])
{
    double one = 1.0;
    scan_prefix<NX, false>(X, one);
#pragma unroll
    for (int q = 0; q < NX; q++)
        X[q] = readlane(X[q], kWave - 1);
}

}  // namespace wv

}  // namespace sg
