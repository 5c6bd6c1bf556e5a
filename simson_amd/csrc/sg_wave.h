// Lane-segment data movement for 64-lane waves on gfx950, built on DPP
// (data-parallel primitives: a VALU operand modifier that reads another
// lane's register in the same instruction) within 16-lane rows and, for
// prefix scans, the row broadcasts.  Suffix scans and sums across the rows
// of a 64-lane segment read other lanes with ds_bpermute, an LDS-pipe
// instruction, rather than with v_readlane, which occupies the VALU like an
// arithmetic instruction.
//
// A wave is split into segments of SW = 16, 32 or 64 lanes (one column of
// the column solver per segment); every scan, shift and sum below stays
// inside its segment.
//
// DPP controls used (GFX9 encoding):
//   row_shr:s  0x110+s  lane i reads lane i-s within its 16-lane row
//   row_shl:s  0x100+s  lane i reads lane i+s within its row
//   wave_shl1  0x130    lane i reads lane i+1 across the wave
//   wave_shr1  0x138    lane i reads lane i-1 across the wave
//   row_bcast15 0x142   lane 15 of row r is read by row r+1
//   row_bcast31 0x143   lane 31 is read by rows 2 and 3
// With bound_ctrl set a lane whose source does not exist reads 0; without
// it, such a lane (or one in a row masked off) keeps the 'old' operand,
// which the helpers set to the identity of the combining operation.
#pragma once

#include <hip/hip_runtime.h>

namespace sg {

namespace wv {

constexpr int kWave = 64;

__device__ __forceinline__ long long bits(double x) { return __double_as_longlong(x); }
__device__ __forceinline__ double dbl(int lo, int hi)
{
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// lanes without a source keep 'old'
template <int CTRL, int ROWMASK = 0xf>
__device__ __forceinline__ double dpp(double old, double x)
{
    const long long o = bits(old), v = bits(x);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, ROWMASK,
                                               0xf, false);
    return dbl(lo, hi);
}

// lanes without a source read 0
template <int CTRL>
__device__ __forceinline__ double dpp0(double x)
{
    const long long v = bits(x);
    const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xf, 0xf, true);
    return dbl(lo, hi);
}

// value of lane src (any lane of the wave) through the LDS crossbar
// (ds_bpermute, byte address 4 src): an LDS instruction, not a VALU one.
// Every lane of the wave must be active.
__device__ __forceinline__ double bperm(double x, int src)
{
    const long long v = bits(x);
    const int lo = __builtin_amdgcn_ds_bpermute(4 * src, (int)v);
    const int hi = __builtin_amdgcn_ds_bpermute(4 * src, (int)(v >> 32));
    return dbl(lo, hi);
}

__device__ __forceinline__ double readlane(double x, int lane)
{
    const long long v = bits(x);
    const int lo = __builtin_amdgcn_readlane((int)v, lane);
    const int hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
    return dbl(lo, hi);
}

// value of segment lane i+1 (the last lane of the segment gets fill);
// l is the lane index in the wave
template <int SW>
__device__ __forceinline__ double from_next(double x, double fill, int l)
{
    if (SW == 16)              // row_shl:1 stays in the row; lane 15 keeps fill
        return (fill == 0.0) ? dpp0<0x101>(x) : dpp<0x101>(fill, x);
    double r = (fill == 0.0) ? dpp0<0x130>(x) : dpp<0x130>(fill, x);
    if (SW < kWave && (l & (SW - 1)) == SW - 1)
        r = fill;
    return r;
}

// value of segment lane i-1 (the first lane of the segment gets fill)
template <int SW>
__device__ __forceinline__ double from_prev(double x, double fill, int l)
{
    if (SW == 16)              // row_shr:1; lane 0 of the row keeps fill
        return (fill == 0.0) ? dpp0<0x111>(x) : dpp<0x111>(fill, x);
    double r = (fill == 0.0) ? dpp0<0x138>(x) : dpp<0x138>(fill, x);
    if (SW < kWave && (l & (SW - 1)) == 0)
        r = fill;
    return r;
}

// Inclusive scans of affine maps r -> X[q] + Y r over the lanes of each
// segment.  Each lane holds the map of its part; afterwards lane i holds the
// composition of segment lanes up to i (prefix: X_i <- X_i + Y_i X_{i-1}) or
// from i on (suffix: X_i <- X_i + Y_i X_{i+1}).  With HASY false, Y is 1 and
// the scan is a plain sum.
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

template <int CTRL, int NX, bool HASY>
__device__ __forceinline__ void dpp_step(double (&X)[NX], double &Y)
{
    double Xo[NX];
#pragma unroll
    for (int q = 0; q < NX; q++)
        Xo[q] = dpp0<CTRL>(X[q]);
    const double Yo = HASY ? dpp<CTRL>(1.0, Y) : 1.0;
    combine<NX, HASY>(X, Y, Xo, Yo);
}

template <int CTRL, int ROWMASK, int NX, bool HASY>
__device__ __forceinline__ void bcast_step(double (&X)[NX], double &Y)
{
    double Xo[NX];
#pragma unroll
    for (int q = 0; q < NX; q++)
        Xo[q] = dpp<CTRL, ROWMASK>(0.0, X[q]);
    const double Yo = HASY ? dpp<CTRL, ROWMASK>(1.0, Y) : 1.0;
    combine<NX, HASY>(X, Y, Xo, Yo);
}

template <int SW, int NX, bool HASY>
__device__ __forceinline__ void scan_prefix(double (&X)[NX], double &Y)
{
    dpp_step<0x111, NX, HASY>(X, Y);
    dpp_step<0x112, NX, HASY>(X, Y);
    dpp_step<0x114, NX, HASY>(X, Y);
    dpp_step<0x118, NX, HASY>(X, Y);
    if (SW >= 32)
        bcast_step<0x142, 0xa, NX, HASY>(X, Y);   // rows 1, 3 take rows 0, 2
    if (SW >= 64)
        bcast_step<0x143, 0xc, NX, HASY>(X, Y);   // rows 2, 3 take rows 0-1
}

template <int SW, int NX, bool HASY>
__device__ __forceinline__ void scan_suffix(double (&X)[NX], double &Y, int l)
{
    dpp_step<0x101, NX, HASY>(X, Y);
    dpp_step<0x102, NX, HASY>(X, Y);
    dpp_step<0x104, NX, HASY>(X, Y);
    dpp_step<0x108, NX, HASY>(X, Y);
    const int row = l >> 4;
    if (SW == 32) {
        // rows 0 and 2 take the totals of rows 1 and 3 (head lanes 16, 48)
        double Xa[NX], Xb[NX];
#pragma unroll
        for (int q = 0; q < NX; q++) {
            Xa[q] = readlane(X[q], 16);
            Xb[q] = readlane(X[q], 48);
        }
        const double Ya = HASY ? readlane(Y, 16) : 1.0, Yb = HASY ? readlane(Y, 48) : 1.0;
        if ((row & 1) == 0) {
            double Xo[NX];
#pragma unroll
            for (int q = 0; q < NX; q++)
                Xo[q] = row == 0 ? Xa[q] : Xb[q];
            combine<NX, HASY>(X, Y, Xo, row == 0 ? Ya : Yb);
        }
    } else if (SW == 64) {
        // rows 0-2 take the head of the next row (lane 16(r+1)), then rows
        // 0-1 the head two rows on, which by then composes rows r+2..3
#pragma unroll
        for (int d = 1; d <= 2; d *= 2) {
            const int src = 16 * ((row + d) & 3);
            double Xo[NX];
#pragma unroll
            for (int q = 0; q < NX; q++)
                Xo[q] = bperm(X[q], src);
            const double Yo = HASY ? bperm(Y, src) : 1.0;
            if (row + d < 4)
                combine<NX, HASY>(X, Y, Xo, Yo);
        }
    }
}

// two independent affine suffix scans (classes c = 0, 1 with their own Y)
// in one pass, so that the lane reads of both issue before either is used;
// each class gets exactly the operations of scan_suffix
template <int SW, int NX>
__device__ __forceinline__ void scan_suffix2(double (&X)[2][NX], double (&Y)[2], int l)
{
    if (SW != kWave) {
        scan_suffix<SW, NX, true>(X[0], Y[0], l);
        scan_suffix<SW, NX, true>(X[1], Y[1], l);
        return;
    }
    dpp_step<0x101, NX, true>(X[0], Y[0]);
    dpp_step<0x101, NX, true>(X[1], Y[1]);
    dpp_step<0x102, NX, true>(X[0], Y[0]);
    dpp_step<0x102, NX, true>(X[1], Y[1]);
    dpp_step<0x104, NX, true>(X[0], Y[0]);
    dpp_step<0x104, NX, true>(X[1], Y[1]);
    dpp_step<0x108, NX, true>(X[0], Y[0]);
    dpp_step<0x108, NX, true>(X[1], Y[1]);
    const int row = l >> 4;
#pragma unroll
    for (int d = 1; d <= 2; d *= 2) {
        const int src = 16 * ((row + d) & 3);
        double Xo[2][NX], Yo[2];
#pragma unroll
        for (int c = 0; c < 2; c++) {
#pragma unroll
            for (int q = 0; q < NX; q++)
                Xo[c][q] = bperm(X[c][q], src);
            Yo[c] = bperm(Y[c], src);
        }
        if (row + d < 4) {
            combine<NX, true>(X[0], Y[0], Xo[0], Yo[0]);
            combine<NX, true>(X[1], Y[1], Xo[1], Yo[1]);
        }
    }
}

// sum over each segment of NX channels, the same value on every lane of
// the segment
template <int SW, int NX>
__device__ __forceinline__ void sum_all(double (&X)[NX], int l)
{
    if (SW == 16) {
        // butterfly within the row: swap pairs, swap pair halves of each
        // quad (quad_perm 1,0,3,2 and 2,3,0,1), mirror each half row, mirror
        // the row.  Every lane adds the same operands (in commuted order),
        // so all 16 lanes hold the identical total.
#pragma unroll
        for (int q = 0; q < NX; q++) {
            X[q] += dpp0<0xb1>(X[q]);
            X[q] += dpp0<0x4e>(X[q]);
            X[q] += dpp0<0x141>(X[q]);
            X[q] += dpp0<0x140>(X[q]);
        }
        return;
    }
    double one = 1.0;
    dpp_step<0x111, NX, false>(X, one);
    dpp_step<0x112, NX, false>(X, one);
    dpp_step<0x114, NX, false>(X, one);
    dpp_step<0x118, NX, false>(X, one);
    // lane 16 r + 15 holds the total of row r
    const int row = l >> 4;
    if (SW == 64) {
        // every lane fetches its row's total, adds the partner row's (lane
        // xor 16), then the other half's (lane xor 32): (r0 + r1) + (r2 + r3)
        // on every lane, the operands of each addition only commuted.  Each
        // round reads all channels before any is used.
        double t[NX];
#pragma unroll
        for (int q = 0; q < NX; q++)
            t[q] = bperm(X[q], 16 * row + 15);
#pragma unroll
        for (int q = 0; q < NX; q++)
            X[q] = bperm(t[q], l ^ 16);
#pragma unroll
        for (int q = 0; q < NX; q++)
            t[q] = t[q] + X[q];
#pragma unroll
        for (int q = 0; q < NX; q++)
            X[q] = bperm(t[q], l ^ 32);
#pragma unroll
        for (int q = 0; q < NX; q++)
            X[q] = t[q] + X[q];
        return;
    }
#pragma unroll
    for (int q = 0; q < NX; q++) {
        const double r0 = readlane(X[q], 15), r1 = readlane(X[q], 31);
        const double r2 = readlane(X[q], 47), r3 = readlane(X[q], 63);
        if (SW == 64) {
            X[q] = (r0 + r1) + (r2 + r3);
        } else if (SW == 32) {
            X[q] = row < 2 ? r0 + r1 : r2 + r3;
        } else {
            X[q] = row == 0 ? r0 : (row == 1 ? r1 : (row == 2 ? r2 : r3));
        }
    }
}

}  // namespace wv

}  // namespace sg
