// Learned subgrid closure: a small MLP (10 -> 64 -> 64 -> 6, GELU) from the
// filtered velocity-gradient tensor and y+ to the six subgrid stresses,
// trained and evaluated by hand-written kernels on the BF16 matrix cores
// (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation).
//
// Parameters are kept in FP32 (Adam state too); every product rounds its
// operands to BF16 and accumulates in FP32.  Layer widths are padded to
// multiples of 16: the input to 16 (features 10-15 zero), the output to 16
// (stresses 6-15 dropped from the loss), and the padded rows of W1 and
// columns of W3 stay zero because their gradients are masked.
//
// One wavefront carries a tile of 16 points through the forward pass, the
// loss gradient and the backward pass.  The tile's matrices live in the
// wave's share of LDS, row-major [point][unit] in FP32, and every MFMA
// operand is gathered from there in the lane layout of the instruction
// (lane l: A[l%16][4(l/16)+j], B[4(l/16)+j][l%16], result D[4(l/16)+r][l%16]),
// so transposed operands are index arithmetic, not data movement.  Weight
// gradients accumulate in registers over all tiles of the wave and are
// written once as per-wave partial sums; a second kernel reduces them and
// applies Adam.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "simson_gpu.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int NIN = 10, NOUT = 6, KI = 16, NH = 64, KO = 16;
// parameter offsets (FP32): W1 [KI][NH], b1 [NH], W2 [NH][NH], b2 [NH],
// W3 [NH][KO], b3 [KO]
constexpr int OW1 = 0, OB1 = OW1 + KI * NH, OW2 = OB1 + NH, OB2 = OW2 + NH * NH,
              OW3 = OB2 + NH, OB3 = OW3 + NH * KO, NPAR = OB3 + KO;
constexpr int TP = 16;              // points per wave tile
constexpr int WPB = 4;              // waves per block

__device__ __forceinline__ short bf16(float x)
{
    // round to nearest even (finite inputs)
    uint32_t u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z)
{
    return 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_d(float z)
{
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) +
           z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Operand gathers from row-major FP32 LDS matrices.
// A block (rows i0.., k0..) of matrix M with element (i, k) at M[i*ld + k]
// (trans: at M[k*ld + i]).
template <bool TRANS>
__device__ __forceinline__ s4 opA(const float *M, int ld, int i0, int k0, int l)
{
    const int i = i0 + (l & 15), k = k0 + 4 * (l >> 4);
    s4 r;
#pragma unroll
    for (int j = 0; j < 4; j++)
        r[j] = bf16(TRANS ? M[(k + j) * ld + i] : M[i * ld + k + j]);
    return r;
}

// B block (k0.., n0..) of matrix M with element (k, n) at M[k*ld + n]
// (trans: at M[n*ld + k]).
template <bool TRANS>
__device__ __forceinline__ s4 opB(const float *M, int ld, int k0, int n0, int l)
{
    const int n = n0 + (l & 15), k = k0 + 4 * (l >> 4);
    s4 r;
#pragma unroll
    for (int j = 0; j < 4; j++)
        r[j] = bf16(TRANS ? M[n * ld + k + j] : M[(k + j) * ld + n]);
    return r;
}

// the lanes of a wave exchange tile data through LDS between phases
__device__ __forceinline__ void wsync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ f4 mfma(s4 a, s4 b, f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// result block D at rows i0.., cols n0.. of row-major M
__device__ __forceinline__ void putD(float *M, int ld, int i0, int n0, f4 d, int l)
{
#pragma unroll
    for (int r = 0; r < 4; r++)
        M[(i0 + 4 * (l >> 4) + r) * ld + n0 + (l & 15)] = d[r];
}

struct WaveLds {
    float X[TP * KI], Y[TP * KO];                 // inputs; outputs, then dY
    float Z1[TP * NH], H1[TP * NH], Z2[TP * NH], H2[TP * NH];
    float G2[TP * NH], G1[TP * NH];               // dZ2, dZ1
};

struct BlockLds {
    float W1[KI * NH], W2[NH * NH], W3[NH * KO], b1[NH], b2[NH], b3[KO];
    WaveLds w[WPB];
};

// forward pass of the tile in L (X filled); Y = outputs
__device__ __forceinline__ void forward(const BlockLds &S, WaveLds &L, int l)
{
#pragma unroll
    for (int nb = 0; nb < NH / 16; nb++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
        acc = mfma(opA<false>(L.X, KI, 0, 0, l), opB<false>(S.W1, NH, 0, 16 * nb, l), acc);
        putD(L.Z1, NH, 0, 16 * nb, acc, l);
    }
    wsync();
    for (int e = l; e < TP * NH; e += 64)
        L.H1[e] = gelu(L.Z1[e] += S.b1[e % NH]);
    wsync();
#pragma unroll
    for (int nb = 0; nb < NH / 16; nb++) {
        f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < NH / 16; kb++)
            acc = mfma(opA<false>(L.H1, NH, 0, 16 * kb, l), opB<false>(S.W2, NH, 16 * kb, 16 * nb, l),
                       acc);
        putD(L.Z2, NH, 0, 16 * nb, acc, l);
    }
    wsync();
    for (int e = l; e < TP * NH; e += 64)
        L.H2[e] = gelu(L.Z2[e] += S.b2[e % NH]);
    wsync();
    f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < NH / 16; kb++)
        acc = mfma(opA<false>(L.H2, NH, 0, 16 * kb, l), opB<false>(S.W3, KO, 16 * kb, 0, l), acc);
    putD(L.Y, KO, 0, 0, acc, l);
    wsync();
    for (int e = l; e < TP * KO; e += 64)
        L.Y[e] += S.b3[e % KO];
    wsync();
}

__device__ __forceinline__ void load_params(BlockLds &S, const float *P)
{
    for (int e = threadIdx.x; e < KI * NH; e += blockDim.x)
        S.W1[e] = P[OW1 + e];
    for (int e = threadIdx.x; e < NH * NH; e += blockDim.x)
        S.W2[e] = P[OW2 + e];
    for (int e = threadIdx.x; e < NH * KO; e += blockDim.x)
        S.W3[e] = P[OW3 + e];
    for (int e = threadIdx.x; e < NH; e += blockDim.x) {
        S.b1[e] = P[OB1 + e];
        S.b2[e] = P[OB2 + e];
    }
    for (int e = threadIdx.x; e < KO; e += blockDim.x)
        S.b3[e] = P[OB3 + e];
}

// tile t of points [t*16, t*16+16) of X [n][NIN] into L.X (padded, zero past n)
__device__ __forceinline__ void load_tile(WaveLds &L, const float *X, int n, int t, int l)
{
    for (int e = l; e < TP * KI; e += 64) {
        const int p = e / KI, k = e % KI, q = t * TP + p;
        L.X[e] = (k < NIN && q < n) ? X[(size_t)q * NIN + k] : 0.f;
    }
    wsync();
}

// Forward + backward over the tiles [t0, t0 + ntile) of the points [0, n);
// per-wave partial gradients to part[wave][NPAR], partial loss to
// loss[wave] (sum of squared errors).
__global__ void __launch_bounds__(64 * WPB)
    k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
               const float *__restrict__ T, int n, int t0, int ntile, float scale,
               float *__restrict__ part, float *__restrict__ loss)
{
    extern __shared__ __align__(16) unsigned char smem[];
    BlockLds &S = *reinterpret_cast<BlockLds *>(smem);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    WaveLds &L = S.w[wv];
    load_params(S, P);
    __syncthreads();
    const int gw = blockIdx.x * WPB + wv, nw = gridDim.x * WPB;
    f4 g1[NH / 16], g2[NH / 16][NH / 16], g3[NH / 16];
#pragma unroll
    for (int a = 0; a < NH / 16; a++) {
        g1[a] = f4{0.f, 0.f, 0.f, 0.f};
        g3[a] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int b = 0; b < NH / 16; b++)
            g2[a][b] = f4{0.f, 0.f, 0.f, 0.f};
    }
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, sse = 0.f;
    for (int t = t0 + gw; t < t0 + ntile; t += nw) {
        load_tile(L, X, n, t, l);
        forward(S, L, l);
        // dY = scale (Y - T) on the real outputs of real points
        for (int e = l; e < TP * KO; e += 64) {
            const int p = e / KO, o = e % KO, q = t * TP + p;
            float d = 0.f;
            if (o < NOUT && q < n) {
                d = L.Y[e] - T[(size_t)q * NOUT + o];
                sse += d * d;
                d *= scale;
            }
            L.Y[e] = d;
        }
        wsync();
        // dH2 = dY W3^T, dZ2 = dH2 gelu'(Z2)
#pragma unroll
        for (int nb = 0; nb < NH / 16; nb++) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
            acc = mfma(opA<false>(L.Y, KO, 0, 0, l), opB<true>(S.W3, KO, 0, 16 * nb, l), acc);
            putD(L.G2, NH, 0, 16 * nb, acc, l);
        }
        wsync();
        for (int e = l; e < TP * NH; e += 64)
            L.G2[e] *= gelu_d(L.Z2[e]);
        wsync();
        // dH1 = dZ2 W2^T, dZ1 = dH1 gelu'(Z1)
#pragma unroll
        for (int nb = 0; nb < NH / 16; nb++) {
            f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kb = 0; kb < NH / 16; kb++)
                acc = mfma(opA<false>(L.G2, NH, 0, 16 * kb, l),
                           opB<true>(S.W2, NH, 16 * kb, 16 * nb, l), acc);
            putD(L.G1, NH, 0, 16 * nb, acc, l);
        }
        wsync();
        for (int e = l; e < TP * NH; e += 64)
            L.G1[e] *= gelu_d(L.Z1[e]);
        wsync();
        // weight gradients, contracted over the 16 points of the tile:
        // dW3 = H2^T dY, dW2 = H1^T dZ2, dW1 = X^T dZ1
#pragma unroll
        for (int mb = 0; mb < NH / 16; mb++) {
            g3[mb] = mfma(opA<true>(L.H2, NH, 16 * mb, 0, l), opB<false>(L.Y, KO, 0, 0, l), g3[mb]);
            g1[mb] = mfma(opA<true>(L.X, KI, 0, 0, l), opB<false>(L.G1, NH, 0, 16 * mb, l), g1[mb]);
#pragma unroll
            for (int nb = 0; nb < NH / 16; nb++)
                g2[mb][nb] = mfma(opA<true>(L.H1, NH, 16 * mb, 0, l),
                                  opB<false>(L.G2, NH, 0, 16 * nb, l), g2[mb][nb]);
        }
        // bias gradients: column sums (lane l owns unit l)
        for (int p = 0; p < TP; p++) {
            gb1 += L.G1[p * NH + l];
            gb2 += L.G2[p * NH + l];
            if (l < KO)
                gb3 += L.Y[p * KO + l];
        }
        wsync();                          // the next tile overwrites L
    }
    // per-wave partial sums, in the parameter layout
    float *g = part + (size_t)gw * NPAR;
    const int rr = 4 * (l >> 4), cc = l & 15;
#pragma unroll
    for (int mb = 0; mb < NH / 16; mb++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            // dW1 [KI][NH]: rows are input features (mask the padded ones)
            g[OW1 + (rr + r) * NH + 16 * mb + cc] = (rr + r < NIN) ? g1[mb][r] : 0.f;
            // dW3 [NH][KO]: columns are outputs (mask the padded ones)
            g[OW3 + (16 * mb + rr + r) * KO + cc] = (cc < NOUT) ? g3[mb][r] : 0.f;
#pragma unroll
            for (int nb = 0; nb < NH / 16; nb++)
                g[OW2 + (16 * mb + rr + r) * NH + 16 * nb + cc] = g2[mb][nb][r];
        }
    g[OB1 + l] = gb1;
    g[OB2 + l] = gb2;
    if (l < KO)
        g[OB3 + l] = l < NOUT ? gb3 : 0.f;
    // loss partial: wave sum of the per-lane sums
    float s = sse;
    for (int off = 32; off > 0; off >>= 1)
        s += __shfl_xor(s, off);
    if (l == 0)
        loss[gw] = s;
}

// sum of the per-wave partials; with adam, one Adam step (Kingma & Ba 2015)
// on the FP32 parameters
__global__ void k_mlp_reduce(const float *__restrict__ part, const float *__restrict__ lossp,
                             int nw, float *__restrict__ grad, float *__restrict__ loss, int adam,
                             float *__restrict__ P, float *__restrict__ m, float *__restrict__ v,
                             float lr, float b1t, float b2t)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == NPAR) {
        float s = 0.f;
        for (int w = 0; w < nw; w++)
            s += lossp[w];
        *loss = s;
    }
    if (i >= NPAR)
        return;
    float s = 0.f;
    for (int w = 0; w < nw; w++)
        s += part[(size_t)w * NPAR + i];
    grad[i] = s;
    if (adam) {
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float mi = b1 * m[i] + (1.f - b1) * s, vi = b2 * v[i] + (1.f - b2) * s * s;
        m[i] = mi;
        v[i] = vi;
        P[i] -= lr * (mi / (1.f - b1t)) / (sqrtf(vi / (1.f - b2t)) + eps);
    }
}

__global__ void __launch_bounds__(64 * WPB)
    k_mlp_infer(const float *__restrict__ P, const float *__restrict__ X, int n,
                float *__restrict__ Y)
{
    extern __shared__ __align__(16) unsigned char smem[];
    BlockLds &S = *reinterpret_cast<BlockLds *>(smem);
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    WaveLds &L = S.w[wv];
    load_params(S, P);
    __syncthreads();
    const int ntile = (n + TP - 1) / TP;
    for (int t = blockIdx.x * WPB + wv; t < ntile; t += gridDim.x * WPB) {
        load_tile(L, X, n, t, l);
        forward(S, L, l);
        for (int e = l; e < TP * KO; e += 64) {
            const int p = e / KO, o = e % KO, q = t * TP + p;
            if (o < NOUT && q < n)
                Y[(size_t)q * NOUT + o] = L.Y[e];
        }
        wsync();
    }
}

}  // namespace

struct simson_mlp {
    hipStream_t stream = nullptr;
    float *P = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
    float *part = nullptr, *lossp = nullptr;
    float *X = nullptr, *T = nullptr;     // resident training set
    int n = 0, step = 0, maxwaves = 0;
    std::string err;
};

namespace {

constexpr int kMaxBlocks = 256;

int fail(simson_mlp_t *h, hipError_t e, const char *what)
{
    h->err = std::string(what) + ": " + hipGetErrorString(e);
    return SIMSON_E_HIP;
}

#define MK(expr)                                  \
    do {                                          \
        hipError_t e_ = (expr);                   \
        if (e_ != hipSuccess)                     \
            return fail(h, e_, #expr);            \
    } while (0)

// deterministic initial weights: uniform(-a, a), a = sqrt(6/(fan_in+fan_out))
// (Glorot), from a 64-bit linear congruential stream; biases zero
void init_params(std::vector<float> &p, uint64_t seed)
{
    uint64_t s = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto uni = [&]() {
        s = s * 6364136223846793005ull + 1442695040888963407ull;
        return (float)((double)(s >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0);
    };
    p.assign(NPAR, 0.f);
    const float a1 = sqrtf(6.f / (NIN + NH)), a2 = sqrtf(6.f / (NH + NH)),
                a3 = sqrtf(6.f / (NH + NOUT));
    for (int k = 0; k < NIN; k++)
        for (int h = 0; h < NH; h++)
            p[OW1 + k * NH + h] = a1 * uni();
    for (int k = 0; k < NH * NH; k++)
        p[OW2 + k] = a2 * uni();
    for (int h = 0; h < NH; h++)
        for (int o = 0; o < NOUT; o++)
            p[OW3 + h * KO + o] = a3 * uni();
}

size_t lds_bytes() { return sizeof(BlockLds); }

int grad_pass(simson_mlp_t *h, const float *X, const float *T, int n, int t0, int ntile,
              float scale, int adam, float lr)
{
    int blocks = (ntile + WPB - 1) / WPB;
    if (blocks > kMaxBlocks)
        blocks = kMaxBlocks;
    const int nw = blocks * WPB;
    hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(64 * WPB), lds_bytes(), h->stream, h->P, X,
                       T, n, t0, ntile, scale, h->part, h->lossp);
    const float b1t = adam ? powf(0.9f, (float)h->step) : 0.f;
    const float b2t = adam ? powf(0.999f, (float)h->step) : 0.f;
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPAR + 1 + 255) / 256), dim3(256), 0, h->stream,
                       h->part, h->lossp, nw, h->grad, h->loss, adam, h->P, h->m, h->v, lr, b1t,
                       b2t);
    MK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int simson_mlp_create(unsigned long long seed, simson_mlp_t **out)
{
    if (!out)
        return SIMSON_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return SIMSON_E_NODEV;
    auto *h = new simson_mlp_t();
    std::vector<float> p;
    init_params(p, seed);
    MK(hipStreamCreate(&h->stream));
    MK(hipMalloc(&h->P, NPAR * sizeof(float)));
    MK(hipMalloc(&h->m, NPAR * sizeof(float)));
    MK(hipMalloc(&h->v, NPAR * sizeof(float)));
    MK(hipMalloc(&h->grad, NPAR * sizeof(float)));
    MK(hipMalloc(&h->loss, sizeof(float)));
    MK(hipMalloc(&h->part, (size_t)kMaxBlocks * WPB * NPAR * sizeof(float)));
    MK(hipMalloc(&h->lossp, (size_t)kMaxBlocks * WPB * sizeof(float)));
    MK(hipMemcpy(h->P, p.data(), NPAR * sizeof(float), hipMemcpyHostToDevice));
    MK(hipMemset(h->m, 0, NPAR * sizeof(float)));
    MK(hipMemset(h->v, 0, NPAR * sizeof(float)));
    MK(hipFuncSetAttribute((const void *)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize,
                           (int)lds_bytes()));
    MK(hipFuncSetAttribute((const void *)k_mlp_infer, hipFuncAttributeMaxDynamicSharedMemorySize,
                           (int)lds_bytes()));
    *out = h;
    return 0;
}

int simson_mlp_nparams(void) { return NPAR; }

int simson_mlp_params(simson_mlp_t *h, float *get, const float *set)
{
    if (!h)
        return SIMSON_E_ARG;
    if (set) {
        MK(hipMemcpy(h->P, set, NPAR * sizeof(float), hipMemcpyHostToDevice));
        MK(hipMemset(h->m, 0, NPAR * sizeof(float)));
        MK(hipMemset(h->v, 0, NPAR * sizeof(float)));
        h->step = 0;
    }
    if (get)
        MK(hipMemcpy(get, h->P, NPAR * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int simson_mlp_set_data(simson_mlp_t *h, const float *X, const float *T, int n)
{
    if (!h || !X || !T || n < 1)
        return SIMSON_E_ARG;
    if (h->X)
        MK(hipFree(h->X));
    if (h->T)
        MK(hipFree(h->T));
    h->X = h->T = nullptr;
    MK(hipMalloc(&h->X, (size_t)n * NIN * sizeof(float)));
    MK(hipMalloc(&h->T, (size_t)n * NOUT * sizeof(float)));
    MK(hipMemcpy(h->X, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
    MK(hipMemcpy(h->T, T, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice));
    h->n = n;
    return 0;
}

int simson_mlp_grad(simson_mlp_t *h, const float *X, const float *T, int n, float *grad,
                    float *loss)
{
    if (!h || !X || !T || n < 1 || !grad)
        return SIMSON_E_ARG;
    float *dX = nullptr, *dT = nullptr;
    MK(hipMalloc(&dX, (size_t)n * NIN * sizeof(float)));
    MK(hipMalloc(&dT, (size_t)n * NOUT * sizeof(float)));
    MK(hipMemcpy(dX, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
    MK(hipMemcpy(dT, T, (size_t)n * NOUT * sizeof(float), hipMemcpyHostToDevice));
    const int ntile = (n + TP - 1) / TP;
    int st = grad_pass(h, dX, dT, n, 0, ntile, 2.f / ((float)n * NOUT), 0, 0.f);
    if (st == 0) {
        MK(hipMemcpyAsync(grad, h->grad, NPAR * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        float l = 0.f;
        MK(hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
        MK(hipStreamSynchronize(h->stream));
        if (loss)
            *loss = l / ((float)n * NOUT);
    }
    MK(hipFree(dX));
    MK(hipFree(dT));
    return st;
}

int simson_mlp_train(simson_mlp_t *h, int steps, int batch, float lr, float *losses)
{
    if (!h || !h->X || steps < 0 || batch < TP || batch % TP != 0)
        return SIMSON_E_ARG;
    const int ntile_all = h->n / TP;
    if (ntile_all < 1)
        return SIMSON_E_ARG;
    const int bt = batch / TP < ntile_all ? batch / TP : ntile_all;
    int cursor = 0;
    for (int s = 0; s < steps; s++) {
        if (cursor + bt > ntile_all)
            cursor = 0;
        h->step++;
        int st = grad_pass(h, h->X, h->T, h->n, cursor, bt, 2.f / ((float)bt * TP * NOUT), 1, lr);
        if (st)
            return st;
        cursor += bt;
        if (losses) {
            float l = 0.f;
            MK(hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
            MK(hipStreamSynchronize(h->stream));
            losses[s] = l / ((float)bt * TP * NOUT);
        }
    }
    MK(hipStreamSynchronize(h->stream));
    return 0;
}

int simson_mlp_infer(simson_mlp_t *h, const float *X, int n, float *Y)
{
    if (!h || !X || !Y || n < 1)
        return SIMSON_E_ARG;
    float *dX = nullptr, *dY = nullptr;
    MK(hipMalloc(&dX, (size_t)n * NIN * sizeof(float)));
    MK(hipMalloc(&dY, (size_t)n * NOUT * sizeof(float)));
    MK(hipMemcpy(dX, X, (size_t)n * NIN * sizeof(float), hipMemcpyHostToDevice));
    const int ntile = (n + TP - 1) / TP;
    int blocks = (ntile + WPB - 1) / WPB;
    if (blocks > 1024)
        blocks = 1024;
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(64 * WPB), lds_bytes(), h->stream, h->P,
                       dX, n, dY);
    MK(hipGetLastError());
    MK(hipMemcpyAsync(Y, dY, (size_t)n * NOUT * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    MK(hipStreamSynchronize(h->stream));
    MK(hipFree(dX));
    MK(hipFree(dY));
    return 0;
}

const char *simson_mlp_last_error(const simson_mlp_t *h)
{
    return h ? h->err.c_str() : "no handle";
}

void simson_mlp_destroy(simson_mlp_t *h)
{
    if (!h)
        return;
    void *ptrs[] = {h->P, h->m, h->v, h->grad, h->loss, h->part, h->lossp, h->X, h->T};
    for (void *p : ptrs)
        if (p)
            (void)hipFree(p);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
