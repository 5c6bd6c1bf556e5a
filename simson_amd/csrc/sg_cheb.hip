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
