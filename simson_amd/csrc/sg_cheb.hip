// Batched y Chebyshev transforms on the matrix cores (MFMA 16x16x4).
//
// Each transform is a dense product Y = A X with the (ny x ny) matrix A of
// vchbf + scaling (analysis, linearbl.f:262-273, prhs.f:302) or vchbb
// (synthesis, linearbl.f:1830-1846), applied to every plane of every field.
// X and Y are either physical planes [y][x] (row layout: main storage) or
// Chebyshev columns [x][k] (column layout: one contiguous (kx,kz) column per
// x for the column solver), complex values as (re, im) pairs, so a plane is
// an ny x 2 nxh real matrix in either layout.
//
// A workgroup of four waves owns 32 real columns of one plane at a time.
// Wave w computes the output rows of M-blocks 2w and 2w+1 (16 rows each)
// and keeps the A fragments of those rows for all k in registers for the
// whole kernel, loaded once; the workgroup walks over column tiles
// persistently.  The input tile is staged in LDS (coalesced along x for
// rows, along k for columns) and read as the B operand; column-layout
// output is staged through the same LDS so that stores run along k.
//
// Operand layout of v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32 (lane l):
// A[i = l%16][k = l/16], B[k = l/16][j = l%16]; the result register r of
// lane l is D[l/16 + 4 r][l%16] for f64 and D[4 (l/16) + r][l%16] for f32.
#include "sg_internal.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace sg {

namespace {

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
};

template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 op(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
};

constexpr int kNB = 32;              // real columns per tile
constexpr int kLdb = kNB + 2;        // LDS row stride

template <typename T, int KS>
constexpr int lds_rows()
{
    return 4 * KS > 128 ? 4 * KS : 128;
}

template <typename T, int KS, bool IN_COL, bool OUT_COL>
__global__ void __launch_bounds__(256, (sizeof(T) == 8 ? 2 : 3)) k_cheb_mfma(const T *__restrict__ A, const T *__restrict__ in,
                                                    T *__restrict__ out, GemmIO io, int ntile_x)
{
    typedef typename Mfma<T>::v4 v4;
    __shared__ T Bs[lds_rows<T, KS>() * kLdb];
    // w is uniform over the wave; readfirstlane lets the block tests below
    // be scalar branches
    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
    const int ny = io.ny, nc = 2 * io.nxh, nxh = io.nxh;
    const int li = l & 15, lk = l >> 4;

    // A fragments of this wave's two M-blocks, all k
    T a[2][KS];
    bool act[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int row = 16 * (2 * w + j) + li;
        act[j] = 16 * (2 * w + j) < ny;
#pragma unroll
        for (int s = 0; s < KS; s++) {
            const int k = 4 * s + lk;
            a[j][s] = (row < ny && k < ny) ? A[(size_t)row * ny + k] : (T)0;
        }
    }

    const int nplane = io.nf * io.nz;
    const int ntile = ntile_x * nplane;
    const size_t plane = (size_t)ny * nc;
    const int len = 2 * ny;                      // reals per Chebyshev column

    // The LDS rows k >= ny are never loaded: A is zero there, and the rows
    // only need finite values (zero now; later, column-layout output).
    for (int e = tid; e < (lds_rows<T, KS>() - ny) * kLdb; e += 256)
        Bs[ny * kLdb + e] = (T)0;

    // Tile input is fetched into registers one tile ahead (the loads of the
    // next tile are in flight during the products of this one), then staged
    // into LDS.  Row layout: element e = k * kNB + n; column layout: the 16
    // columns of a tile are contiguous, element e = xl * len + rem.
    constexpr int NE = (4 * KS * kNB + 255) / 256;
    T buf[NE];
    auto fetch = [&](int tile) {
        const int pl = tile / ntile_x, tx = tile - pl * ntile_x;
        const int f = pl / io.nz, z = pl - f * io.nz;
        const T *X = in + ((size_t)io.in_slot[f] * io.nz + z) * plane;
        const int n0 = tx * kNB;
        if (!IN_COL) {
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const int e = tid + 256 * i;
                const int k = min(e / kNB, ny - 1), n = min(n0 + (e % kNB), nc - 1);
                buf[i] = X[(size_t)k * nc + n];
            }
        } else {
            const int x0 = n0 >> 1;
            const int last = min(kNB / 2, nxh - x0) * len - 1;
            const T *Xt = X + (size_t)x0 * len;
#pragma unroll
            for (int i = 0; i < NE; i++)
                buf[i] = Xt[min(tid + 256 * i, last)];
        }
    };
    auto stage = [&](int tile) {
        const int tx = tile % ntile_x;
        const int n0 = tx * kNB;
        if (!IN_COL) {
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const int e = tid + 256 * i;
                const int k = e / kNB, n = e % kNB;
                if (k < ny)
                    Bs[k * kLdb + n] = n0 + n < nc ? buf[i] : (T)0;
            }
        } else {
            const int xn = min(kNB / 2, nxh - (n0 >> 1));
            int xl = tid / len, rem = tid - xl * len;
#pragma unroll
            for (int i = 0; i < NE; i++) {
                if (xl < kNB / 2)
                    Bs[(rem >> 1) * kLdb + 2 * xl + (rem & 1)] = xl < xn ? buf[i] : (T)0;
                rem += 256;
                while (rem >= len) {
                    rem -= len;
                    xl++;
                }
            }
        }
    };

    if (blockIdx.x < ntile)
        fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int pl = tile / ntile_x, tx = tile - pl * ntile_x;
        const int f = pl / io.nz, z = pl - f * io.nz;
        T *Y = out + ((size_t)io.out_slot[f] * io.nz + z) * plane;
        const int n0 = tx * kNB, x0 = n0 >> 1;

        __syncthreads();                  // previous tile's LDS reads are done
        stage(tile);
        __syncthreads();
        if (tile + (int)gridDim.x < ntile)
            fetch(tile + gridDim.x);
        if (z == io.skip_z)
            continue;

        v4 acc[2][2];
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int nb = 0; nb < 2; nb++)
                acc[j][nb] = v4{0, 0, 0, 0};
        if (act[1]) {
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const T *brow = Bs + (4 * s + lk) * kLdb + li;
                const T b0 = brow[0], b1 = brow[16];
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    acc[j][0] = Mfma<T>::op(a[j][s], b0, acc[j][0]);
                    acc[j][1] = Mfma<T>::op(a[j][s], b1, acc[j][1]);
                }
            }
        } else if (act[0]) {
#pragma unroll
            for (int s = 0; s < KS; s++) {
                const T *brow = Bs + (4 * s + lk) * kLdb + li;
                const T b0 = brow[0], b1 = brow[16];
                acc[0][0] = Mfma<T>::op(a[0][s], b0, acc[0][0]);
                acc[0][1] = Mfma<T>::op(a[0][s], b1, acc[0][1]);
            }
        }

        if (!OUT_COL) {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                if (!act[j])
                    continue;
#pragma unroll
                for (int nb = 0; nb < 2; nb++) {
                    const int col = n0 + 16 * nb + li;
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int row = 16 * (2 * w + j) + Mfma<T>::row(lk, r);
                        if (row < ny && col < nc)
                            Y[(size_t)row * nc + col] = acc[j][nb][r];
                    }
                }
            }
            continue;
        }
        __syncthreads();                  // all B reads of this tile are done
#pragma unroll
        for (int j = 0; j < 2; j++) {
            if (!act[j])
                continue;
#pragma unroll
            for (int nb = 0; nb < 2; nb++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    Bs[(16 * (2 * w + j) + Mfma<T>::row(lk, r)) * kLdb + 16 * nb + li] =
                        acc[j][nb][r];
        }
        __syncthreads();
        T *Yt = Y + (size_t)x0 * len;
        const int xn = min(kNB / 2, nxh - x0);
        for (int rem = tid; rem < len; rem += 256) {
            const int k = rem >> 1, r = rem & 1;
            const T *src = Bs + k * kLdb + r;
#pragma unroll 4
            for (int xl = 0; xl < kNB / 2; xl++)
                if (xl < xn)
                    Yt[(size_t)xl * len + rem] = src[2 * xl];
        }
    }
}

// persistent grid: as many workgroups as can be resident at once
template <typename K>
int resident_blocks(K kernel)
{
    static int n = 0;
    if (n == 0) {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess ||
            per < 1)
            return 256;
        n = per * cus;
    }
    return n;
}

template <typename T, int KS, bool IN_COL, bool OUT_COL>
void launch_one(const T *A, const T *in, T *out, const GemmIO &io, hipStream_t s)
{
    const int ntile_x = (2 * io.nxh + kNB - 1) / kNB;
    const int ntile = ntile_x * io.nf * io.nz;
    auto kern = k_cheb_mfma<T, KS, IN_COL, OUT_COL>;
    int grid = resident_blocks(kern);
    if (grid > ntile)
        grid = ntile;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, s, A, in, out, io, ntile_x);
}

template <typename T, int KS>
void launch_ks(const T *A, const T *in, T *out, const GemmIO &io, hipStream_t s)
{
    if (!io.in_col && io.out_col)
        launch_one<T, KS, false, true>(A, in, out, io, s);
    else if (io.in_col && !io.out_col)
        launch_one<T, KS, true, false>(A, in, out, io, s);
    else if (io.in_col)
        launch_one<T, KS, true, true>(A, in, out, io, s);
    else
        launch_one<T, KS, false, false>(A, in, out, io, s);
}

}  // namespace

// ---------------------------------------------------------------------------
// Parity-split transforms.  The ny = 2H + 1 Gauss-Lobatto points are
// symmetric about y = 0 (y_{ny-1-j} = -y_j) and T_k(-y) = (-1)^k T_k(y), so
// every y transform here maps the even and the odd parts of a column
// separately.  The input is split into part 0 (H + 1 values: the folded sums
// s_j = x_j + x_{ny-1-j}, s_H = x_H, or the even Chebyshev coefficients) and
// part 1 (H values: the differences d_j = x_j - x_{ny-1-j}, or the odd
// coefficients).  The even output (H + 1 rows: the symmetric half
// (y_i + y_{ny-1-i})/2 with y_H, or the even coefficients) depends on one
// part only, the odd output (H rows) on the other; a derivative (FLIP)
// swaps them.  Two half-size products replace the dense one, which halves
// the matrix-core work, and the output is recombined as
// y_i = E_i + O_i, y_{ny-1-i} = E_i - O_i, or interleaved as c_2i = E_i,
// c_2i+1 = O_i.
//
// ny = 8 (K - 1) + 1, so H = 4 (K - 1) and the even block takes K k-steps
// of 4 (H + 1 inputs) or K - 1 (FLIP).  A workgroup of NW = ceil(H/16)
// waves owns a tile of 32 real columns; wave w computes rows 16w..16w+15 of
// both output blocks with their A fragments resident in registers, so the
// two results of one output row meet in the same lane and register.  Row H
// of the even block (the highest even coefficient, or y = 0) is a single
// dot product per column, summed on the VALU by all waves.
//
// LDS holds part 0 at rows [0, 4K) and part 1 at rows [4K, 8K) of the tile.
// The packed matrix M holds the even block [H + 1][4K] and then the odd
// block [H][4K] (zero beyond each block's inputs).

constexpr int kPNB = 32;             // real columns per tile

template <int K>
constexpr int par_waves()
{
    return (K - 1 + 3) / 4;
}

template <typename T, int K, bool FLIP, bool IN_COL, bool OUT_COL>
__global__ void __launch_bounds__(64 * par_waves<K>())
    k_cheb_par(const T *__restrict__ M, const T *__restrict__ in, T *__restrict__ out, GemmIO io,
               int ntile_x)
{
    typedef typename Mfma<T>::v4 v4;
    constexpr int NW = par_waves<K>();
    constexpr int NT = 64 * NW;
    constexpr int H = 4 * (K - 1);
    constexpr int NY = 2 * H + 1;
    constexpr int KE = FLIP ? K - 1 : K;      // k-steps of the even block
    constexpr int KO = FLIP ? K : K - 1;
    constexpr int PE = FLIP ? 1 : 0;          // input part of the even block
    constexpr int LDM = 4 * K;                // row stride of M
    constexpr int NQ = NT / kPNB;             // partial sums per column of row H
    constexpr int CH = (4 * KE + NQ - 1) / NQ;
    __shared__ T Bs[8 * K * kLdb];
    __shared__ T mh[4 * K];
    __shared__ T red[NQ * kPNB];

    const int tid = threadIdx.x, w = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63;
    const int nc = 2 * io.nxh, nxh = io.nxh;
    const int li = l & 15, lk = l >> 4;
    const int row = 16 * w + li;

    T ae[KE], ao[KO];
#pragma unroll
    for (int s = 0; s < KE; s++)
        ae[s] = row < H ? M[(size_t)row * LDM + 4 * s + lk] : (T)0;
#pragma unroll
    for (int s = 0; s < KO; s++)
        ao[s] = row < H ? M[(size_t)(H + 1 + row) * LDM + 4 * s + lk] : (T)0;
    for (int e = tid; e < 4 * K; e += NT)
        mh[e] = M[(size_t)H * LDM + e];
    // rows never staged (padding of both parts) only need finite values
    for (int e = tid; e < 8 * K * kLdb; e += NT)
        Bs[e] = (T)0;

    const int nplane = io.nf * io.nz;
    const int ntile = ntile_x * nplane;
    const size_t plane = (size_t)NY * nc;
    constexpr int len = 2 * NY;                 // reals per Chebyshev column

    // Row input: item e = (j, c), j <= H, loads x_j and x_{ny-1-j} of
    // column c.  Column input: the 16 columns of a tile are contiguous,
    // element e = xl * len + rem.
    constexpr int NE = IN_COL ? (kPNB * NY + NT - 1) / NT : ((H + 1) * kPNB + NT - 1) / NT;
    constexpr int NB = IN_COL ? NE : 2 * NE;
    T buf[NB];
    auto fetch = [&](int tile) {
        const int pl = tile / ntile_x, tx = tile - pl * ntile_x;
        const int f = pl / io.nz, z = pl - f * io.nz;
        const T *X = in + ((size_t)io.in_slot[f] * io.nz + z) * plane;
        const int n0 = tx * kPNB;
        if (!IN_COL) {
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const int e = tid + NT * i;
                const int j = min(e / kPNB, H), n = min(n0 + (e % kPNB), nc - 1);
                buf[2 * i] = X[(size_t)j * nc + n];
                buf[2 * i + 1] = X[(size_t)(NY - 1 - j) * nc + n];
            }
        } else {
            const int x0 = n0 >> 1;
            const int last = min(kPNB / 2, nxh - x0) * len - 1;
            const T *Xt = X + (size_t)x0 * len;
#pragma unroll
            for (int i = 0; i < NE; i++)
                buf[i] = Xt[min(tid + NT * i, last)];
        }
    };
    auto stage = [&](int tile) {
        const int tx = tile % ntile_x;
        const int n0 = tx * kPNB;
        if (!IN_COL) {
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const int e = tid + NT * i;
                const int j = e / kPNB, n = e % kPNB;
                const bool ok = n0 + n < nc;
                const T a = buf[2 * i], b = buf[2 * i + 1];
                if (j < H) {
                    Bs[j * kLdb + n] = ok ? a + b : (T)0;
                    Bs[(4 * K + j) * kLdb + n] = ok ? a - b : (T)0;
                } else if (j == H) {
                    Bs[H * kLdb + n] = ok ? a : (T)0;
                }
            }
        } else {
            const int xn = min(kPNB / 2, nxh - (n0 >> 1));
#pragma unroll
            for (int i = 0; i < NE; i++) {
                const int e = tid + NT * i;
                const int xl = e / len, rem = e - xl * len;
                const int k = rem >> 1;
                const int r = (k & 1) ? 4 * K + (k >> 1) : (k >> 1);
                if (xl < kPNB / 2)
                    Bs[r * kLdb + 2 * xl + (rem & 1)] = xl < xn ? buf[i] : (T)0;
            }
        }
    };

    if (blockIdx.x < ntile)
        fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int pl = tile / ntile_x, tx = tile - pl * ntile_x;
        const int f = pl / io.nz, z = pl - f * io.nz;
        T *Y = out + ((size_t)io.out_slot[f] * io.nz + z) * plane;
        const int n0 = tx * kPNB, x0 = n0 >> 1;

        __syncthreads();                  // previous tile's LDS reads are done
        stage(tile);
        __syncthreads();
        if (tile + (int)gridDim.x < ntile)
            fetch(tile + gridDim.x);
        if (z == io.skip_z)
            continue;

        v4 ce[2], co[2];
#pragma unroll
        for (int nb = 0; nb < 2; nb++) {
            ce[nb] = v4{0, 0, 0, 0};
            co[nb] = v4{0, 0, 0, 0};
        }
        {
            // k-steps in chunks of CK, the B values of the next chunk loaded
            // while the products of this one issue; the scheduling barriers
            // keep the compiler from hoisting every LDS read of the tile
            // (which would hold 4 K registers)
            const T *be = Bs + (PE * 4 * K + lk) * kLdb + li;
            const T *bo = Bs + ((1 - PE) * 4 * K + lk) * kLdb + li;
            constexpr int CK = 2, NCH = (K + CK - 1) / CK;
            T bb[2][CK][4];
            auto load = [&](int c, T(&d)[CK][4]) {
#pragma unroll
                for (int t = 0; t < CK; t++) {
                    const int s = c * CK + t;
                    if (s < KE) {
                        d[t][0] = be[4 * s * kLdb];
                        d[t][1] = be[4 * s * kLdb + 16];
                    }
                    if (s < KO) {
                        d[t][2] = bo[4 * s * kLdb];
                        d[t][3] = bo[4 * s * kLdb + 16];
                    }
                }
            };
            load(0, bb[0]);
#pragma unroll
            for (int c = 0; c < NCH; c++) {
                if (c + 1 < NCH)
                    load(c + 1, bb[(c + 1) & 1]);
#pragma unroll
                for (int t = 0; t < CK; t++) {
                    const int s = c * CK + t;
                    if (s < KE) {
                        ce[0] = Mfma<T>::op(ae[s], bb[c & 1][t][0], ce[0]);
                        ce[1] = Mfma<T>::op(ae[s], bb[c & 1][t][1], ce[1]);
                    }
                    if (s < KO) {
                        co[0] = Mfma<T>::op(ao[s], bb[c & 1][t][2], co[0]);
                        co[1] = Mfma<T>::op(ao[s], bb[c & 1][t][3], co[1]);
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // row H of the even block: partial dot products over CH inputs
        {
            const int c = tid % kPNB, q = tid / kPNB;
            const T *b = Bs + PE * 4 * K * kLdb + c;
            T acc = (T)0;
#pragma unroll
            for (int t = 0; t < CH; t++) {
                const int kk = q * CH + t;
                if (kk < 4 * KE)
                    acc += mh[kk] * b[kk * kLdb];
            }
            red[q * kPNB + c] = acc;
        }

        if (!OUT_COL) {
#pragma unroll
            for (int nb = 0; nb < 2; nb++) {
                const int col = n0 + 16 * nb + li;
                if (col < nc) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int i = 16 * w + Mfma<T>::row(lk, r);
                        if (i < H) {
                            const T e = ce[nb][r], o = co[nb][r];
                            Y[(size_t)i * nc + col] = e + o;
                            Y[(size_t)(NY - 1 - i) * nc + col] = e - o;
                        }
                    }
                }
            }
            __syncthreads();              // partial sums of row H are in LDS
            if (tid < kPNB && n0 + tid < nc) {
                T sum = (T)0;
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    sum += red[q * kPNB + tid];
                Y[(size_t)H * nc + n0 + tid] = sum;
            }
            continue;
        }
        __syncthreads();                  // all B reads of this tile are done
#pragma unroll
        for (int nb = 0; nb < 2; nb++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * w + Mfma<T>::row(lk, r);
                if (i < H) {
                    Bs[(2 * i) * kLdb + 16 * nb + li] = ce[nb][r];
                    Bs[(2 * i + 1) * kLdb + 16 * nb + li] = co[nb][r];
                }
            }
        if (tid < kPNB) {
            T sum = (T)0;
#pragma unroll
            for (int q = 0; q < NQ; q++)
                sum += red[q * kPNB + tid];
            Bs[(2 * H) * kLdb + tid] = sum;
        }
        __syncthreads();
        T *Yt = Y + (size_t)x0 * len;
        const int xn = min(kPNB / 2, nxh - x0);
        for (int rem = tid; rem < len; rem += NT) {
            const int k = rem >> 1, r = rem & 1;
            const T *src = Bs + k * kLdb + r;
#pragma unroll 4
            for (int xl = 0; xl < kPNB / 2; xl++)
                if (xl < xn)
                    Yt[(size_t)xl * len + rem] = src[2 * xl];
        }
    }
}

template <typename T, int K, bool FLIP, bool IN_COL, bool OUT_COL>
void launch_par_one(const T *M, const T *in, T *out, const GemmIO &io, hipStream_t s)
{
    const int ntile_x = (2 * io.nxh + kPNB - 1) / kPNB;
    const int ntile = ntile_x * io.nf * io.nz;
    auto kern = k_cheb_par<T, K, FLIP, IN_COL, OUT_COL>;
    static int nres = 0;
    if (nres == 0) {
        int dev = 0, cus = 0, per = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) ==
                hipSuccess &&
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64 * par_waves<K>(), 0) ==
                hipSuccess &&
            per >= 1)
            nres = per * cus;
        else
            nres = 256;
    }
    const int grid = nres < ntile ? nres : ntile;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * par_waves<K>()), 0, s, M, in, out, io,
                       ntile_x);
}

// the four (input, output) layouts in use: rows -> columns (analysis),
// columns -> rows (synthesis), rows -> rows (a physical-space operator:
// parity preserving, or swapping for a derivative)
template <typename T, int K>
bool launch_par_k(const T *M, int flip, const T *in, T *out, const GemmIO &io, hipStream_t s)
{
    if (!io.in_col && io.out_col && !flip)
        launch_par_one<T, K, false, false, true>(M, in, out, io, s);
    else if (io.in_col && !io.out_col && !flip)
        launch_par_one<T, K, false, true, false>(M, in, out, io, s);
    else if (!io.in_col && !io.out_col)
        flip ? launch_par_one<T, K, true, false, false>(M, in, out, io, s)
             : launch_par_one<T, K, false, false, false>(M, in, out, io, s);
    else
        return false;
    return true;
}

// K values with a parity-split kernel: ny = 9, 17, 25, 33, 49, 65, 97, 129,
// 193, 257
#define SG_PAR_KS(X) X(2) X(3) X(4) X(5) X(7) X(9) X(13) X(17) X(25) X(33)

template <typename T>
bool launch_par(const T *M, int K, int flip, const T *in, T *out, const GemmIO &io, hipStream_t s)
{
    switch (K) {
#define SG_CASE(KK) \
    case KK:        \
        return launch_par_k<T, KK>(M, flip, in, out, io, s);
        SG_PAR_KS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

// k-step counts (of 4) with a specialised kernel; ny <= 128
#define SG_CHEB_KS(X) X(4) X(8) X(12) X(16) X(20) X(25) X(28) X(32)

int cheb_mfma_ks(int ny)
{
    const int need = (ny + 3) / 4;
#define SG_PICK(K)   \
    if (need <= K)   \
        return K;
    SG_CHEB_KS(SG_PICK)
#undef SG_PICK
    return 0;
}

bool launch_cheb_mfma(const double *A, const GemmIO &io, hipStream_t s)
{
    switch (cheb_mfma_ks(io.ny)) {
#define SG_CASE(K)                                   \
    case K:                                          \
        launch_ks<double, K>(A, io.in, io.out, io, s); \
        return true;
        SG_CHEB_KS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

// Parity split of a transform A [ny][ny] for the layouts in_col/out_col
// (column layout = Chebyshev coefficients, row layout = physical points):
// the packed blocks of k_cheb_par in M and the kernel size K, or 0 when ny
// has no split kernel or A lacks the parity structure (then *flip is -1).
int cheb_split(const double *A, int ny, int in_col, int out_col, std::vector<double> &M,
               int *flip)
{
    *flip = -1;
    if (ny < 9 || ny % 8 != 1)
        return 0;
    const int H = (ny - 1) / 2, K = H / 4 + 1;
    bool have = false;
#define SG_HAVE(KK) have = have || K == KK;
    SG_PAR_KS(SG_HAVE)
#undef SG_HAVE
    if (!have)
        return 0;
    auto a = [&](int i, int j) { return A[(size_t)i * ny + j]; };
    // A seen from input part p (0: s or even k, H + 1 values; 1: d or odd k)
    auto P = [&](int p, int i, int kk) {
        if (in_col)
            return a(i, 2 * kk + p);
        if (p == 0)
            return kk == H ? a(i, H) : 0.5 * (a(i, kk) + a(i, ny - 1 - kk));
        return 0.5 * (a(i, kk) - a(i, ny - 1 - kk));
    };
    // even output rows i <= H, odd output rows i < H
    auto E = [&](int p, int i, int kk) {
        if (out_col)
            return P(p, 2 * i, kk);
        return i == H ? P(p, H, kk) : 0.5 * (P(p, i, kk) + P(p, ny - 1 - i, kk));
    };
    auto O = [&](int p, int i, int kk) {
        return out_col ? P(p, 2 * i + 1, kk) : 0.5 * (P(p, i, kk) - P(p, ny - 1 - i, kk));
    };
    double amax = 0.;
    for (size_t i = 0; i < (size_t)ny * ny; i++)
        amax = std::max(amax, std::fabs(A[i]));
    // the cross terms that vanish: even <- part 1 and odd <- part 0 (parity
    // preserving), or even <- part 0 and odd <- part 1 (parity swapping)
    double x_keep = 0., x_swap = 0.;
    for (int i = 0; i <= H; i++)
        for (int kk = 0; kk <= H; kk++) {
            if (kk < H)
                x_keep = std::max(x_keep, std::fabs(E(1, i, kk)));
            x_swap = std::max(x_swap, std::fabs(E(0, i, kk)));
            if (i < H) {
                x_keep = std::max(x_keep, std::fabs(O(0, i, kk)));
                if (kk < H)
                    x_swap = std::max(x_swap, std::fabs(O(1, i, kk)));
            }
        }
    const double tol = 1e-13 * amax;
    int f;
    if (x_keep <= tol)
        f = 0;
    else if (x_swap <= tol)
        f = 1;
    else
        return 0;
    const int ld = 4 * K;
    M.assign((size_t)(2 * H + 1) * ld, 0.);
    // inputs of the even block's part and of the odd block's part
    const int ne = f ? H : H + 1, no = f ? H + 1 : H;
    for (int i = 0; i <= H; i++)
        for (int kk = 0; kk < ne; kk++)
            M[(size_t)i * ld + kk] = E(f, i, kk);
    for (int i = 0; i < H; i++)
        for (int kk = 0; kk < no; kk++)
            M[(size_t)(H + 1 + i) * ld + kk] = O(1 - f, i, kk);
    *flip = f;
    return K;
}

bool launch_cheb_par(const double *M, int K, int flip, const GemmIO &io, hipStream_t s)
{
    return launch_par<double>(M, K, flip, io.in, io.out, io, s);
}

bool launch_cheb_par_f32(const float *M, int K, int flip, const float *in, float *out,
                         const GemmIO &io, hipStream_t s)
{
    return launch_par<float>(M, K, flip, in, out, io, s);
}

bool launch_cheb_mfma_f32(const float *A, const float *in, float *out, const GemmIO &io,
                          hipStream_t s)
{
    switch (cheb_mfma_ks(io.ny)) {
#define SG_CASE(K)                                \
    case K:                                       \
        launch_ks<float, K>(A, in, out, io, s);   \
        return true;
        SG_CHEB_KS(SG_CASE)
#undef SG_CASE
    default:
        return false;
    }
}

bool launch_cheb_f32(const ChebOp &op, const float *in, float *out, const GemmIO &io,
                     hipStream_t s)
{
    if (op.S32 && launch_cheb_par_f32(op.S32, op.K, op.flip, in, out, io, s))
        return true;
    return op.A32 && launch_cheb_mfma_f32(op.A32, in, out, io, s);
}

bool cheb_f32_ok(int ny)
{
    if (cheb_mfma_ks(ny) != 0)
        return true;
    if (ny % 8 != 1)
        return false;
    const int K = (ny - 1) / 8 + 1;
    bool have = false;
#define SG_HAVE(KK) have = have || K == KK;
    SG_PAR_KS(SG_HAVE)
#undef SG_HAVE
    return have;
}

namespace {

__global__ void k_narrow(const double2 *__restrict__ src, float2 *__restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = make_float2((float)src[i].x, (float)src[i].y);
}

__global__ void k_widen(const float2 *__restrict__ src, double2 *__restrict__ dst, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = make_double2(src[i].x, src[i].y);
}

unsigned copy_blocks(size_t n)
{
    const size_t b = (n + 255) / 256;
    return (unsigned)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

}  // namespace

namespace {

// u(y) of the kx = kz = 0 column of five slots from its Chebyshev
// coefficients (Tb is the synthesis matrix, symmetric)
__global__ void k_mean_synth(const double *__restrict__ Tb, const double2 *__restrict__ C,
                             double2 *__restrict__ U, int ny, int nz, int nxh)
{
    const int s = blockIdx.x;
    const double2 *c = C + (size_t)s * nz * nxh * ny;          // column (z 0, x 0)
    for (int y = threadIdx.x; y < ny; y += blockDim.x) {
        double re = 0., im = 0.;
        for (int k = 0; k < ny; k++) {
            const double t = Tb[(size_t)y * ny + k];
            re = fma(t, c[k].x, re);
            im = fma(t, c[k].y, im);
        }
        U[((size_t)s * nz * ny + y) * nxh] = make_double2(re, im);
    }
}

}  // namespace

void launch_mean_synth(const double *Tb, const double2 *C, double2 *U, int ny, int nz, int nxh,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_mean_synth, dim3(5), dim3(128), 0, s, Tb, C, U, ny, nz, nxh);
}

void launch_narrow(const double2 *src, float2 *dst, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_narrow, dim3(copy_blocks(n)), dim3(256), 0, s, src, dst, n);
}

void launch_widen(const float2 *src, double2 *dst, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_widen, dim3(copy_blocks(n)), dim3(256), 0, s, src, dst, n);
}

}  // namespace sg
