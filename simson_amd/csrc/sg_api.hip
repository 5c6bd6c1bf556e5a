// C ABI of the MI355X Simson step library (declared in include/simson_gpu.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sg_internal.h"

using namespace sg;

// Plane statistics of simson_gpu_amp_noalloc: the per-y sums [20][ny] go to
// this file-scope array instead of the handle's buffer, for callers that
// cannot allocate device memory.  Each device holds its own copy until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (amp_scratch_address).
constexpr int kAmpScratchNy = 257;
static __device__ __align__(256) double scratch[20 * kAmpScratchNy];

namespace {

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double *scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on device dev, which must be current; looked
// up once per device
hipError_t amp_scratch_address(int dev, double **out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess)
            return e;
        scratch_ptr[dev] = (double *)p;
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

const char *const kFamilies[] = {"zsyn", "xrow", "zana", "cfl_reduce", "cheb_fwd",
                                 "tau", "cheb_bwd", "prhs_cheb_fwd", "prhs",
                                 "prhs_cheb_bwd", "tau_factor", "les_filter",
                                 "closure_zsyn", "closure_xrow", "closure_zana",
                                 "closure_dy", "smag"};
enum Fam { F_ZSYN, F_XROW, F_ZANA, F_CFL, F_CHF, F_TAU, F_CHB, F_PCHF, F_PRHS, F_PCHB,
           F_FAC, F_LES, F_CZS, F_CXR, F_CZA, F_CDY, F_SMAG, F_N };

struct TimedLaunch {
    int fam;
    hipEvent_t a, b;
};

}  // namespace

struct simson_gpu {
    simson_params_t p{};
    int nx = 0, ny = 0, nz = 0, nxh = 0, nxp = 0, nzp = 0;
    hipStream_t stream = nullptr;
    // U: slots 0-4 at physical y, [5][nz][ny][nx/2]; C: Chebyshev columns
    // [5][nz][nx/2][ny]; Pp: partial right-hand sides p_v, p_omega as
    // Chebyshev columns [2][nz][nx/2][ny]; P: step-2 planes.
    double2 *U = nullptr, *P = nullptr, *C = nullptr, *Pp = nullptr;
    double2 *twz = nullptr, *twx = nullptr;
    double *W = nullptr, *puw = nullptr, *px = nullptr;
    double *cfl_blk = nullptr, *cfl = nullptr;
    double *amp = nullptr;                // [20][ny] plane statistics
    double *amp_scratch = nullptr;        // this device's `scratch` (amp_noalloc)
    // pressure work (allocated at the first pressure call): Q rows
    // [5][nz][ny][nxh] (H1, H2, H3, W, v), Qc the same as Chebyshev columns,
    // PR the pressure at physical y, pscr per-column scratch
    double2 *Q = nullptr, *Qc = nullptr, *PR = nullptr;
    double *pscr = nullptr;
    // tau factorisations: k^2 once, lam^2 for up to three values of a+b
    LamFactor kfac;
    double *kfT = nullptr;            // the k^2 factors as [2][ny][ncol] planes
    LamFactor lam[3];
    double lam_apb[3] = {0., 0., 0.};
    long lam_used[3] = {0, 0, 0};
    long clock = 0;
    double *Tf = nullptr, *Tp = nullptr, *Tb = nullptr;
    double *alfa = nullptr, *beta = nullptr, *deta = nullptr;
    // mixed form: FP32 copies of U, P, C and of the tables; U is refreshed
    // from U32 before a download (u64_stale) and U32 from U after an upload
    // or prhs (u32_stale)
    bool mixed = false, u32_stale = false, u64_stale = false;
    bool mean64 = false;              // C holds the FP64 mean-flow column of U32's state
    float2 *U32 = nullptr, *P32 = nullptr, *C32 = nullptr, *twz32 = nullptr, *twx32 = nullptr;
    float *Tf32 = nullptr, *Tb32 = nullptr;
    // ADM-RT model (simson_gpu_les): the wall-normal low-pass L (FP64 and
    // FP32), the x-z transfer function lpf [nz][nxh] and the filtered
    // velocities Wf = L u of slots 0-2
    bool les = false;
    double les_chi = 0.;
    double *Lm = nullptr, *lpf = nullptr;
    float *Lm32 = nullptr;
    double2 *Wf = nullptr;
    float2 *Wf32 = nullptr;
    // learned closure (simson_gpu_closure): network parameters and feature
    // maps, the nine gradient fields and six stresses on the padded x-z grid
    // (PG [9][ny][nzp][nxh]), the stress spectra TAU [6][nz][ny][nxh], their
    // y derivatives DT [3][...] and the y differentiation matrix Dm [ny][ny]
    bool clos = false;
    ClosArgs cl{};
    float *clW = nullptr, *clY = nullptr;
    double2 *PG = nullptr, *TAU = nullptr, *DT = nullptr;
    float2 *PG32 = nullptr, *TAU32 = nullptr, *DT32 = nullptr;
    double *Dm = nullptr;
    float *Dm32 = nullptr;
    FftPlan pz, pm;
    // x-y statistics (allocated at the first simson_gpu_xys call): unpadded
    // transforms of lengths nz and nx/2, their twiddles and the sums
    // [15][ny][nx]; the step-2 planes P serve as work space
    FftPlan pzs, pms;
    double2 *twzs = nullptr, *twxs = nullptr;
    double *xysd = nullptr;
    // eddy-viscosity LES (simson_gpu_les with iles = 2): settings and work
    // space of sg_smag.hip (SQ [12][ny][nz][nxh], SS [7][...], SW
    // [21][ny][nz][nx/4] for the dynamic procedure), the filter width per y
    bool smag = false;
    SmagArgs sm{};
    double2 *SQ = nullptr, *SS = nullptr, *SW = nullptr;
    int sq_nf = 0;                    // fields SQ holds (12, or 21 for the dynamic HPF model)
    // HPF model (iles 3): L u of slots 0-4 and the secondary x-z transfer
    // function (Lm and lpf are the ADM-RT ones above)
    double2 *LW = nullptr;
    double *lpf2 = nullptr;
    double *dlt2 = nullptr, *SP = nullptr;
    size_t smag_lds = 0;
    // fringe of spatial simulations (simson_gpu_fringe): lambda [nxp/2] and
    // the base flow [3][ny][nxp/2] packed as the x rows, FP64 and FP32
    int fringe = 0;
    double2 *flam = nullptr, *fbase = nullptr;
    float2 *flam32 = nullptr, *fbase32 = nullptr;
    // lower-wall blowing and suction (simson_gpu_wall): the spectra of
    // max(P, P damp) and min(P, P damp), their (0,0) values, and the time
    // at the start of the next substep (simson_gpu_set_time, advanced by run)
    // free stream above (ibc = 101, simson_gpu_free_stream): the base
    // flow's dU_i/dy at y = +1 per kx of the kz = 0 plane [3][nxh] (bj_set
    // once given) and its mean streamwise value
    double2 *bj = nullptr;
    bool bj_set = false;
    double du0upp = 0.;
    // mean forcing (simson_gpu_mean_force): Chebyshev coefficients [2][ny]
    double *mforce = nullptr;
    bool mforce_on = false;
    // localized volume force (simson_gpu_local_force): Chebyshev
    // coefficients [3][nxh][ny] of the kz = 0 plane and its frequency
    double2 *lforce = nullptr;
    bool lforce_on = false;
    double lforce_omega = 0.;
    int wall = 0;
    double wall_tomeg = 0., wall_p0 = 0., wall_m0 = 0., tnow = 0.;
    double2 *wallP = nullptr, *wallM = nullptr;
    double mflux = 0., u0upp = 0., u0low = 0., w0upp = 0., w0low = 0.;
    // the y transforms with their parity-split forms: analysis (linear step
    // and pressure, prhs), synthesis, the y derivative of the stresses and
    // the ADM-RT low-pass; split blocks live in 'extra'
    ChebOp oTf, oTp, oTb, oDm, oLm;
    std::vector<void *> extra;
    // CFL read-back of simson_gpu_run: the latest value of each evaluation
    // copied to pinned host memory (two slots) behind an event
    double *cfl_host = nullptr;
    hipEvent_t cfl_ev[2] = {nullptr, nullptr};
    bool timing = false;
    std::vector<TimedLaunch> launches;
    std::vector<hipEvent_t> pool;
    double tot_ms[F_N] = {};
    long cnt[F_N] = {};
    std::string err;
};

namespace {

#define CK(expr)                                                                  \
    do {                                                                          \
        hipError_t e_ = (expr);                                                   \
        if (e_ != hipSuccess) {                                                   \
            g->err = std::string(#expr) + ": " + hipGetErrorString(e_);         \
            return SIMSON_E_HIP;                                                  \
        }                                                                         \
    } while (0)

hipEvent_t get_event(simson_gpu_t *g)
{
    if (!g->pool.empty()) {
        hipEvent_t e = g->pool.back();
        g->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess)
        return nullptr;
    return e;
}

template <class F>
void timed(simson_gpu_t *g, int fam, F &&launch)
{
    if (!g->timing) {
        launch();
        return;
    }
    TimedLaunch t{fam, get_event(g), get_event(g)};
    (void)hipEventRecord(t.a, g->stream);
    launch();
    (void)hipEventRecord(t.b, g->stream);
    g->launches.push_back(t);
}

// exp(+2 pi i t/n) for t < n with the angle reduced to the first octant
std::vector<double2> roots(int n)
{
    std::vector<double2> w(n);
    for (int t = 0; t < n; t++) {
        long num = 4L * t, q = num / n, r = num % n;
        double a = 2.0 * M_PI * (double)r / (4.0 * (double)n);
        double c = std::cos(a), s = std::sin(a);
        switch (q & 3) {
        case 0: w[t] = make_double2(c, s); break;
        case 1: w[t] = make_double2(-s, c); break;
        case 2: w[t] = make_double2(-c, -s); break;
        default: w[t] = make_double2(s, -c); break;
        }
    }
    return w;
}

bool make_radices(int n, FftPlan &pl)
{
    pl.n = n;
    pl.nrad = 0;
    int m = n;
    const int order[4] = {4, 2, 3, 5};
    for (int r : order) {
        while (m % r == 0) {
            if (pl.nrad >= kMaxRad)
                return false;
            pl.rad[pl.nrad++] = r;
            m /= r;
        }
    }
    return m == 1;
}

void free_all(simson_gpu_t *g)
{
    void *ptrs[] = {g->U, g->P, g->C, g->Pp, g->twz, g->twx, g->W, g->puw, g->px,
                    g->cfl_blk, g->cfl, g->amp, g->Tf, g->Tp, g->Tb, g->alfa, g->beta, g->deta,
                    g->kfac.gp, g->kfT, g->lam[0].gp, g->lam[0].vh, g->lam[0].sh,
                    g->lam[1].gp, g->lam[1].vh, g->lam[1].sh,
                    g->lam[2].gp, g->lam[2].vh, g->lam[2].sh, g->U32, g->P32, g->C32,
                    g->twz32, g->twx32, g->Tf32, g->Tb32, g->Q, g->Qc, g->PR, g->pscr,
                    g->Lm, g->lpf, g->Lm32, g->Wf, g->Wf32, g->clW, g->clY, g->PG, g->TAU,
                    g->DT, g->PG32, g->TAU32, g->DT32, g->Dm, g->Dm32, g->twzs,
                    g->twxs, g->xysd, g->SQ, g->SS, g->SW, g->dlt2, g->SP, g->flam, g->fbase,
                    g->flam32, g->fbase32, g->wallP, g->wallM, g->bj,
                    g->lam[0].hb, g->lam[0].hc, g->lam[1].hb, g->lam[1].hc,
                    g->lam[2].hb, g->lam[2].hc, g->mforce, g->lforce, g->LW, g->lpf2};
    for (void *q : ptrs)
        if (q)
            (void)hipFree(q);
    for (void *q : g->extra)
        (void)hipFree(q);
    if (g->cfl_host)
        (void)hipHostFree(g->cfl_host);
    for (auto e : g->cfl_ev)
        if (e)
            (void)hipEventDestroy(e);
    for (auto &t : g->launches) {
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    for (auto e : g->pool)
        (void)hipEventDestroy(e);
    if (g->stream)
        (void)hipStreamDestroy(g->stream);
}

template <class T>
int dalloc(simson_gpu_t *g, T **p, size_t n)
{
    CK(hipMalloc((void **)p, n * sizeof(T)));
    CK(hipMemsetAsync(*p, 0, n * sizeof(T), g->stream));
    return 0;
}

template <class T>
int h2d(simson_gpu_t *g, T *d, const T *h, size_t n)
{
    CK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, g->stream));
    return 0;
}

// A transform operator: the dense device matrices A (FP64) and A32 (FP32 or
// null) of the host matrix Ah, and its parity-split blocks when ny has a
// split kernel (FP32 blocks when A32 is given).  SIMSON_CHEB=dense keeps the
// dense kernels (for comparisons).
int make_op(simson_gpu_t *g, ChebOp &op, const std::vector<double> &Ah, const double *A,
            const float *A32, int in_col, int out_col)
{
    op = ChebOp{};
    op.A = A;
    op.A32 = A32;
    const char *env = getenv("SIMSON_CHEB");
    if (env && strcmp(env, "dense") == 0)
        return 0;
    std::vector<double> M;
    int flip = 0;
    const int K = cheb_split(Ah.data(), g->ny, in_col, out_col, M, &flip);
    if (K == 0)
        return 0;
    double *S = nullptr;
    CK(hipMalloc((void **)&S, M.size() * sizeof(double)));
    g->extra.push_back(S);
    if (h2d(g, S, M.data(), M.size()))
        return SIMSON_E_HIP;
    if (A32) {
        std::vector<float> M32(M.begin(), M.end());
        float *S32 = nullptr;
        CK(hipMalloc((void **)&S32, M32.size() * sizeof(float)));
        g->extra.push_back(S32);
        if (h2d(g, S32, M32.data(), M32.size()))
            return SIMSON_E_HIP;
        // the copy reads host memory that goes out of scope on return
        CK(hipStreamSynchronize(g->stream));
        op.S32 = S32;
    }
    CK(hipStreamSynchronize(g->stream));
    op.S = S;
    op.K = K;
    op.flip = flip;
    return 0;
}

NonlinArgs nonlin_args(simson_gpu_t *g, int do_cfl)
{
    NonlinArgs a;
    a.U = g->U;
    a.Uw = g->U;
    a.P = g->P;
    a.alfa = g->alfa;
    a.beta = g->beta;
    a.deta = g->deta;
    a.cfl_blk = g->cfl_blk;
    a.nx = g->nx;
    a.ny = g->ny;
    a.nz = g->nz;
    a.nxh = g->nxh;
    a.nxp = g->nxp;
    a.nzp = g->nzp;
    a.rot = g->p.rot;
    a.xl = g->p.xl;
    a.zl = g->p.zl;
    a.do_cfl = do_cfl;
    a.split_mean = 0;
    a.les = 0;
    a.Wf = nullptr;
    a.lpf = g->lpf;
    a.chi_n = g->les_chi * (double)g->nxp * (double)g->nzp;
    a.clos = 0;
    a.TAU = nullptr;
    a.DT = nullptr;
    a.fringe = g->fringe;
    a.flam = g->flam;
    a.fbase = g->fbase;
    a.nocap = g->wall;
    return a;
}

ColArgs col_args(simson_gpu_t *g)
{
    ColArgs a;
    a.U = g->U;
    a.C = g->C;
    a.C32 = g->C32;
    a.Pp = g->Pp;
    a.W = g->W;
    a.puw = g->puw;
    a.px = g->px;
    a.alfa = g->alfa;
    a.beta = g->beta;
    a.ny = g->ny;
    a.nz = g->nz;
    a.nxh = g->nxh;
    a.re = g->p.re;
    a.mflux = g->mflux;
    a.cflux = g->p.cflux;
    a.u0upp = g->u0upp;
    a.u0low = g->u0low;
    a.w0upp = g->w0upp;
    a.w0low = g->w0low;
    a.vwall = nullptr;
    a.vscale = 0.;
    a.v0 = 0.;
    a.ibc = g->p.ibc;
    a.bj = g->bj_set ? g->bj : nullptr;
    a.du0upp = g->du0upp;
    a.mforce = g->mforce_on ? g->mforce : nullptr;
    a.lforce = nullptr;
    a.lscale = 0.;
    if (g->p.ibc == 100 || g->p.ibc == 101) {
        // with a base flow the mean mode takes U = 1 above and no-slip
        // below (bla.f:1360-1365); with ibc = 101 the free-stream condition
        // then replaces the upper value
        a.u0upp = 1.;
        a.u0low = a.w0upp = a.w0low = 0.;
    }
    return a;
}

// in_col / out_col: the input / output planes are Chebyshev columns
GemmIO gemm_io(simson_gpu_t *g, const void *in, int in_col, void *out, int out_col,
               std::initializer_list<int> sin, std::initializer_list<int> sout, int skip_z)
{
    GemmIO io{};
    io.in = (const double *)in;
    io.out = (double *)out;
    int i = 0;
    for (int s : sin)
        io.in_slot[i++] = s;
    io.nf = i;
    i = 0;
    for (int s : sout)
        io.out_slot[i++] = s;
    io.nz = g->nz;
    io.ny = g->ny;
    io.nxh = g->nxh;
    io.in_col = in_col;
    io.out_col = out_col;
    io.skip_z = skip_z;
    return io;
}

// Factorisation slot for lam^2 = k^2 + 2 Re/apb; computed when apb is new
// (three values recur with a fixed dt).
int lam_slot(simson_gpu_t *g, double apb)
{
    g->clock++;
    for (int i = 0; i < 3; i++)
        if (g->lam_used[i] && g->lam_apb[i] == apb) {
            g->lam_used[i] = g->clock;
            return i;
        }
    int v = 0;
    for (int i = 1; i < 3; i++)
        if (g->lam_used[i] < g->lam_used[v])
            v = i;
    ColArgs ca = col_args(g);
    timed(g, F_FAC, [&] { launch_factor_lam(ca, apb, g->lam[v], g->kfT, g->stream); });
    g->lam_apb[v] = apb;
    g->lam_used[v] = g->clock;
    return v;
}

Step2Args<float2> nonlin_args32(simson_gpu_t *g, int do_cfl)
{
    const NonlinArgs d = nonlin_args(g, do_cfl);
    Step2Args<float2> a;
    a.U = g->U32;
    a.Uw = g->U32;
    a.P = g->P32;
    a.alfa = d.alfa;
    a.beta = d.beta;
    a.deta = d.deta;
    a.cfl_blk = d.cfl_blk;
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.nxh = d.nxh;
    a.nxp = d.nxp;
    a.nzp = d.nzp;
    a.rot = d.rot;
    a.xl = d.xl;
    a.zl = d.zl;
    a.do_cfl = d.do_cfl;
    a.split_mean = 1;
    a.les = 0;
    a.Wf = nullptr;
    a.lpf = d.lpf;
    a.chi_n = d.chi_n;
    a.clos = 0;
    a.TAU = nullptr;
    a.DT = nullptr;
    a.fringe = d.fringe;
    a.flam = g->flam32;
    a.fbase = g->fbase32;
    a.nocap = d.nocap;
    return a;
}

// Learned closure: stresses of the current velocities and their y
// derivatives (the tau_ij of les.f and the first half of sgs_stress.f)
template <typename V>
void closure_stresses(simson_gpu_t *g, Step2Args<V> c, V *PG, V *TAU, V *DT, const V *twz,
                      const V *twx, bool fp32)
{
    c.P = PG;
    c.Uw = TAU;
    c.split_mean = 0;
    timed(g, F_CZS, [&] { launch_gsyn4(c, twz, g->stream); });
    timed(g, F_CXR, [&] { launch_xclos(c, twx, g->cl, g->stream); });
    timed(g, F_CZA, [&] { launch_zana4(c, twz, g->stream, 6); });
    GemmIO io = gemm_io(g, TAU, 0, DT, 0, {1, 3, 4}, {0, 1, 2}, g->nz / 2);
    timed(g, F_CDY, [&] {
        if (fp32)
            launch_cheb_f32(g->oDm, (const float *)TAU, (float *)DT, io, g->stream);
        else
            launch_cheb_gemm(g->oDm, io, g->stream);
    });
}

size_t field_size(const simson_gpu_t *g) { return (size_t)g->nz * g->ny * g->nxh; }

// mixed form: FP64 state from the FP32 one before a download, the
// statistics, the pressure or prhs; after a linear step the mean-flow
// column comes from its FP64 Chebyshev coefficients instead
void refresh_u64(simson_gpu_t *g)
{
    if (g->mixed && g->u64_stale) {
        launch_widen(g->U32, g->U, kNPhys * field_size(g), g->stream);
        if (g->mean64)
            launch_mean_synth(g->Tb, g->C, g->U, g->ny, g->nz, g->nxh, g->stream);
        g->u64_stale = false;
    }
}

// mixed form: FP32 state from the FP64 one after an upload or prhs
void refresh_u32(simson_gpu_t *g)
{
    if (g->mixed && g->u32_stale) {
        launch_narrow(g->U, g->U32, kNPhys * field_size(g), g->stream);
        g->u32_stale = false;
    }
}

// the unpadded transforms of lengths nz and nx/2 (statistics, eddy
// viscosity): plans and twiddles, made once
int unpadded_plans(simson_gpu_t *g)
{
    if (g->twzs)
        return 0;
    if (!make_radices(g->nz, g->pzs) || !make_radices(g->nxh, g->pms)) {
        g->err = "unsupported nx or nz for the unpadded transforms";
        return SIMSON_E_ARG;
    }
    std::vector<double2> tz = roots(g->nz), tx = roots(g->nx);
    if (dalloc(g, &g->twzs, tz.size()) || dalloc(g, &g->twxs, tx.size()) ||
        h2d(g, g->twzs, tz.data(), tz.size()) || h2d(g, g->twxs, tx.data(), tx.size()))
        return SIMSON_E_HIP;
    g->pzs.tw = g->twzs;
    g->pzs.tabn = g->nz;
    g->pms.tw = g->twxs;          // length-nx/2 transform uses every other root
    g->pms.tabn = g->nx;
    return 0;
}

// stress spectra TAU [6][nz][ny][nxh], their y derivatives DT [3][...] and
// the y differentiation matrix on the Gauss-Lobatto points, the
// physical-space form of vchbf, dcheb, vchbb (sgs_stress.f:98-121); shared
// by the learned closure and the eddy-viscosity model
int alloc_stress(simson_gpu_t *g, bool fp64, bool fp32)
{
    const int ny = g->ny;
    const size_t nt = 6 * field_size(g), nd = 3 * field_size(g);
    if (!g->Dm) {
        std::vector<double> D((size_t)ny * ny), y(ny), c(ny);
        for (int j = 0; j < ny; j++) {
            y[j] = cos(M_PI * j / (ny - 1));
            c[j] = ((j == 0 || j == ny - 1) ? 2. : 1.) * ((j & 1) ? -1. : 1.);
        }
        for (int i = 0; i < ny; i++) {
            double sum = 0.;
            for (int j = 0; j < ny; j++)
                if (i != j) {
                    D[(size_t)i * ny + j] = c[i] / c[j] / (y[i] - y[j]);
                    sum += D[(size_t)i * ny + j];
                }
            D[(size_t)i * ny + i] = -sum;
        }
        if (dalloc(g, &g->Dm, D.size()) || h2d(g, g->Dm, D.data(), D.size()))
            return SIMSON_E_HIP;
        if (g->mixed) {
            std::vector<float> D32(D.begin(), D.end());
            if (dalloc(g, &g->Dm32, D32.size()) || h2d(g, g->Dm32, D32.data(), D32.size()))
                return SIMSON_E_HIP;
        }
        if (int st = make_op(g, g->oDm, D, g->Dm, g->Dm32, 0, 0))
            return st;
        if (g->mixed && !g->oDm.S32 && cheb_mfma_ks(ny) == 0) {
            g->err = "the y derivative has no matrix-core form for this ny";
            return SIMSON_E_ARG;
        }
    }
    if (fp64 && !g->TAU && (dalloc(g, &g->TAU, nt) || dalloc(g, &g->DT, nd)))
        return SIMSON_E_HIP;
    if (fp32 && !g->TAU32 && (dalloc(g, &g->TAU32, nt) || dalloc(g, &g->DT32, nd)))
        return SIMSON_E_HIP;
    return 0;
}

// eddy-viscosity stresses of the current FP64 state into TAU (and TAU32 in
// the mixed form) and their y derivatives into DT (DT32)
void smag_stresses(simson_gpu_t *g)
{
    SmagArgs a = g->sm;
    a.U = g->U;
    a.u0low = g->u0low;
    a.w0low = g->w0low;
    if (a.hpf) {
        // L u of the five stored fields (linearbl.f:1871-2066 and, for the
        // dynamic procedure, 2067-2256 filter the same L u in x-z)
        GemmIO io = gemm_io(g, g->U, 0, g->LW, 0, {0, 1, 2, 3, 4}, {0, 1, 2, 3, 4}, g->nz / 2);
        timed(g, F_LES, [&] { launch_cheb_gemm(g->oLm, io, g->stream); });
    }
    timed(g, F_SMAG, [&] { launch_smag(a, g->pzs, g->pms, g->smag_lds, g->stream); });
    if (g->mixed) {
        launch_narrow(g->TAU, g->TAU32, 6 * field_size(g), g->stream);
        GemmIO io = gemm_io(g, g->TAU32, 0, g->DT32, 0, {1, 3, 4}, {0, 1, 2}, g->nz / 2);
        timed(g, F_CDY, [&] {
            launch_cheb_f32(g->oDm, (const float *)g->TAU32, (float *)g->DT32, io, g->stream);
        });
    } else {
        GemmIO io = gemm_io(g, g->TAU, 0, g->DT, 0, {1, 3, 4}, {0, 1, 2}, g->nz / 2);
        timed(g, F_CDY, [&] { launch_cheb_gemm(g->oDm, io, g->stream); });
    }
}

int do_nonlin(simson_gpu_t *g, int do_cfl)
{
    const int odd = g->nz / 2;
    if (g->mixed) {
        refresh_u32(g);
        Step2Args<float2> a = nonlin_args32(g, do_cfl);
        if (g->les) {
            // Wf = L u of the velocities (linearbl.f:1871-1962), row layout
            GemmIO io = gemm_io(g, g->U32, 0, g->Wf32, 0, {0, 1, 2}, {0, 1, 2}, odd);
            timed(g, F_LES, [&] {
                launch_cheb_f32(g->oLm, (const float *)g->U32, (float *)g->Wf32, io, g->stream);
            });
            a.les = 1;
            a.Wf = g->Wf32;
        }
        if (g->clos) {
            closure_stresses<float2>(g, a, g->PG32, g->TAU32, g->DT32, g->twz32, g->twx32,
                                     true);
            a.clos = 1;
            a.TAU = g->TAU32;
            a.DT = g->DT32;
        }
        if (g->smag) {
            // the model runs in FP64 on the FP64 image of the state
            refresh_u64(g);
            smag_stresses(g);
            a.clos = 1;
            a.TAU = g->TAU32;
            a.DT = g->DT32;
        }
        timed(g, F_ZSYN, [&] { launch_zsyn4(a, g->twz32, g->stream); });
        timed(g, F_XROW, [&] { launch_xrow4(a, g->twx32, g->stream); });
        if (do_cfl)
            timed(g, F_CFL, [&] {
                launch_cfl_reduce(g->cfl_blk, xrow4_blocks(a), g->cfl, g->stream);
            });
        timed(g, F_ZANA, [&] { launch_zana4(a, g->twz32, g->stream); });
        g->u64_stale = true;
        g->mean64 = false;
        CK(hipGetLastError());
        return 0;
    }
    NonlinArgs a = nonlin_args(g, do_cfl);
    if (g->les) {
        GemmIO io = gemm_io(g, g->U, 0, g->Wf, 0, {0, 1, 2}, {0, 1, 2}, odd);
        timed(g, F_LES, [&] { launch_cheb_gemm(g->oLm, io, g->stream); });
        a.les = 1;
        a.Wf = g->Wf;
    }
    if (g->clos) {
        closure_stresses<double2>(g, a, g->PG, g->TAU, g->DT, (const double2 *)g->twz,
                                  (const double2 *)g->twx, false);
        a.clos = 1;
        a.TAU = g->TAU;
        a.DT = g->DT;
    }
    if (g->smag) {
        smag_stresses(g);
        a.clos = 1;
        a.TAU = g->TAU;
        a.DT = g->DT;
    }
    timed(g, F_ZSYN, [&] {
        if (!launch_zsyn4(a, (const double2 *)g->twz, g->stream))
            launch_zsyn(a, g->pz, g->stream);
    });
    bool four = false;
    timed(g, F_XROW, [&] {
        four = launch_xrow4(a, (const double2 *)g->twx, g->stream);
        if (!four)
            launch_xrow(a, g->pm, g->twx, g->stream);
    });
    const int nblk = four ? xrow4_blocks(a) : xrow_blocks(a);
    if (do_cfl)
        timed(g, F_CFL, [&] { launch_cfl_reduce(g->cfl_blk, nblk, g->cfl, g->stream); });
    timed(g, F_ZANA, [&] {
        if (!launch_zana4(a, (const double2 *)g->twz, g->stream))
            launch_zana(a, g->pz, g->stream);
    });
    CK(hipGetLastError());
    return 0;
}

int do_linear(simson_gpu_t *g, double an, double bn, double anp1, double bnp1)
{
    const int odd = g->nz / 2;
    const int slot = lam_slot(g, an + bn);
    if (g->mixed)
        refresh_u32(g);
    GemmIO fw = gemm_io(g, g->U, 0, g->C, 1, {0, 1, 2}, {0, 1, 2}, odd);
    timed(g, F_CHF, [&] {
        if (g->mixed)
            launch_cheb_f32(g->oTf, (const float *)g->U32, (float *)g->C32, fw, g->stream);
        else
            launch_cheb_gemm(g->oTf, fw, g->stream);
    });
    ColArgs ca = col_args(g);
    if (g->wall) {
        // the wall velocity at the start of the substep (bla.f:1797, 2405)
        const double sv = g->wall_tomeg == 0. ? 1. : std::sin(g->wall_tomeg * g->tnow);
        ca.vwall = sv >= 0. ? g->wallP : g->wallM;
        ca.vscale = sv;
        ca.v0 = sv * (sv >= 0. ? g->wall_p0 : g->wall_m0);
    }
    if (g->lforce_on) {
        // the force at the start of the substep (locf.f:393, tc of bla.f:1797)
        ca.lforce = g->lforce;
        ca.lscale = std::sin(g->lforce_omega * g->tnow);
    }
    int st = 0;
    timed(g, F_TAU, [&] {
        st = launch_tau(ca, g->lam[slot], g->kfac, an, bn, anp1, bnp1, g->mixed, g->stream);
    });
    if (st) {
        g->err = "unsupported ny for the column solver";
        return st;
    }
    GemmIO bw = gemm_io(g, g->C, 1, g->U, 0, {0, 1, 2, 3, 4}, {0, 1, 2, 3, 4}, odd);
    timed(g, F_CHB, [&] {
        if (g->mixed)
            launch_cheb_f32(g->oTb, (const float *)g->C32, (float *)g->U32, bw, g->stream);
        else
            launch_cheb_gemm(g->oTb, bw, g->stream);
    });
    if (g->mixed) {
        g->u64_stale = true;
        g->mean64 = true;
    }
    CK(hipGetLastError());
    return 0;
}

// The start-up partial RHS runs in FP64 in both forms, on the FP64 fields.
int do_prhs(simson_gpu_t *g, double an)
{
    refresh_u64(g);
    GemmIO fw = gemm_io(g, g->U, 0, g->C, 1, {0, 1, 2}, {0, 1, 2}, -1);
    timed(g, F_PCHF, [&] { launch_cheb_gemm(g->oTp, fw, g->stream); });
    ColArgs ca = col_args(g);
    timed(g, F_PRHS, [&] { launch_prhs(ca, an, g->stream); });
    GemmIO bw = gemm_io(g, g->C, 1, g->U, 0, {3, 4}, {3, 4}, -1);
    timed(g, F_PCHB, [&] { launch_cheb_gemm(g->oTb, bw, g->stream); });
    g->u32_stale = g->mixed;
    CK(hipGetLastError());
    return 0;
}

int valid_fft_len(int n)
{
    FftPlan pl;
    return make_radices(n, pl);
}

}  // namespace

extern "C" {

int simson_gpu_init(const simson_params_t *p, simson_gpu_t **out)
{
    if (!p || !out)
        return SIMSON_E_ARG;
    *out = nullptr;
    if ((p->precision != SIMSON_FP64 && p->precision != SIMSON_MIXED) ||
        (p->ibc != 0 && p->ibc != 100 && p->ibc != 101) || (p->ibc >= 100 && p->cflux) ||
        p->nx < 4 || p->nz < 4 ||
        p->nx % 4 != 0 || p->nz % 2 != 0 || p->ny < 9 || p->ny % 2 == 0 ||
        p->re <= 0 || p->xl <= 0 || p->zl <= 0)
        return SIMSON_E_ARG;
    const int nxp = p->nx + p->nx / 2, nzp = p->nz + p->nz / 2;
    if (!valid_fft_len(nzp) || !valid_fft_len(nxp / 2))
        return SIMSON_E_ARG;
    // one tau column per wavefront holds at most 3 x 128 Chebyshev
    // coefficients
    if (p->ny > 383)
        return SIMSON_E_ARG;
    // the FP32 path exists only as the specialised kernels
    if (p->precision == SIMSON_MIXED &&
        (!has_zfft4(nzp) || !has_xrow4(nxp / 2) || !cheb_f32_ok(p->ny)))
        return SIMSON_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return SIMSON_E_NODEV;
    {
        // lengths without a specialised kernel use the generic ones, which
        // keep two z lines, or twelve half x rows and their twiddles, in
        // LDS: up to the device's limit per block (160 KiB on gfx950), with
        // the kernels allowed more than the default 64 KiB where needed
        int dev = 0, lds_max = 65536;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
                hipSuccess)
            return SIMSON_E_HIP;
        const long zneed = has_zfft4(nzp) ? 0 : 2L * nzp * 16;
        const long xneed = has_xrow4(nxp / 2) ? 0 : 14L * (nxp / 2) * 16;
        if (zneed > lds_max || xneed > lds_max)
            return SIMSON_E_ARG;
        if ((zneed > 65536 || xneed > 65536) && nonlin_prepare(lds_max) != 0)
            return SIMSON_E_HIP;
    }
    simson_gpu_t *g = new simson_gpu_t;
    g->p = *p;
    g->nx = p->nx;
    g->ny = p->ny;
    g->nz = p->nz;
    g->nxh = p->nx / 2;
    g->nxp = nxp;
    g->nzp = nzp;
    auto fail = [&](int code) {
        fprintf(stderr, "simson_gpu_init: %s\n", g->err.c_str());
        free_all(g);
        delete g;
        return code;
    };
    if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
        g->err = "hipStreamCreate failed";
        return fail(SIMSON_E_HIP);
    }
    {
        // this device's copy of the statistics scratch (amp_noalloc)
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess)
            e = amp_scratch_address(dev, &g->amp_scratch);
        if (e != hipSuccess) {
            g->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return fail(SIMSON_E_HIP);
        }
    }
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = g->nxh;
    const size_t F = (size_t)nz * ny * nxh;
    const size_t ncol = (size_t)nz * nxh;
    const int xblk = nzp * ny;                  // >= xrow blocks for any RW >= 1
    if (dalloc(g, &g->U, kNPhys * F) || dalloc(g, &g->P, (size_t)6 * ny * nzp * nxh) ||
        dalloc(g, &g->C, 5 * F) || dalloc(g, &g->Pp, 2 * F) ||
        dalloc(g, &g->W, (size_t)kScratch * ny * ncol) ||
        dalloc(g, &g->kfac.gp, F) || dalloc(g, &g->kfT, 2 * F) || dalloc(g, &g->lam[0].gp, F) ||
        dalloc(g, &g->lam[1].gp, F) || dalloc(g, &g->lam[2].gp, F) ||
        dalloc(g, &g->lam[0].vh, F) || dalloc(g, &g->lam[1].vh, F) ||
        dalloc(g, &g->lam[2].vh, F) || dalloc(g, &g->lam[0].sh, ncol) ||
        dalloc(g, &g->lam[1].sh, ncol) || dalloc(g, &g->lam[2].sh, ncol) ||
        dalloc(g, &g->puw, 2 * (size_t)ny) || dalloc(g, &g->px, 1) ||
        dalloc(g, &g->cfl_blk, (size_t)xblk) ||
        dalloc(g, &g->cfl, 2) || dalloc(g, &g->amp, 20 * (size_t)ny) ||
        dalloc(g, &g->Tf, (size_t)ny * ny) ||
        dalloc(g, &g->Tp, (size_t)ny * ny) || dalloc(g, &g->Tb, (size_t)ny * ny) ||
        dalloc(g, &g->alfa, (size_t)nxh) || dalloc(g, &g->beta, (size_t)nz) ||
        dalloc(g, &g->deta, (size_t)ny) || dalloc(g, &g->twz, (size_t)nzp) ||
        dalloc(g, &g->twx, (size_t)nxp))
        return fail(SIMSON_E_HIP);
    if (p->ibc == 101)
        for (int i = 0; i < 3; i++)
            if (dalloc(g, &g->lam[i].hb, F) || dalloc(g, &g->lam[i].hc, 4 * ncol))
                return fail(SIMSON_E_HIP);
    if (p->ibc >= 100 && dalloc(g, &g->bj, 3 * (size_t)nxh))
        return fail(SIMSON_E_HIP);
    // wavenumbers and grid (preprbl.f:34-55)
    std::vector<double> alfa(nxh), beta(nz), eta(ny), deta(ny);
    const double pi = 3.14159265358979323846;
    for (int x = 0; x < nxh; x++)
        alfa[x] = 2. * pi / p->xl * (double)x;
    beta[0] = 0.;
    for (int z = 1; z <= nz / 2; z++) {
        beta[z] = 2. * pi / p->zl * (double)z;
        beta[nz - z] = -2. * pi / p->zl * (double)z;
    }
    std::vector<double2> rc = roots(2 * (ny - 1));   // exp(i pi m/(ny-1))
    for (int y = 0; y < ny; y++)
        eta[y] = rc[y].x;
    deta[0] = 1. - eta[1];
    deta[ny - 1] = deta[0];
    for (int y = 1; y < ny - 1; y++)
        deta[y] = (eta[y - 1] - eta[y + 1]) * .5;
    // Chebyshev matrices: analysis (vchbf with c_0 = c_{n-1} = 1/2) times the
    // linearbl scaling 2/(nxp (ny-1) nzp) or the prhs scaling 2/(ny-1);
    // synthesis (vchbb) cos(pi j k/(ny-1)).
    std::vector<double> Tf((size_t)ny * ny), Tp((size_t)ny * ny), Tb((size_t)ny * ny);
    const double cn = 2. / ((double)nxp * (double)(ny - 1) * (double)nzp);
    const double cp = 2. / (double)(ny - 1);
    const int M2 = 2 * (ny - 1);
    for (int k = 0; k < ny; k++) {
        for (int j = 0; j < ny; j++) {
            const double cs = rc[(long)k * j % M2].x;
            const double ck = (k == 0 || k == ny - 1) ? 0.5 : 1.0;
            const double cj = (j == 0 || j == ny - 1) ? 0.5 : 1.0;
            Tf[(size_t)k * ny + j] = ck * cj * cs * cn;
            Tp[(size_t)k * ny + j] = ck * cj * cs * cp;
            Tb[(size_t)k * ny + j] = cs;
        }
    }
    if (!make_radices(nzp, g->pz) || !make_radices(nxp / 2, g->pm)) {
        g->err = "unsupported FFT length";
        return fail(SIMSON_E_ARG);
    }
    std::vector<double2> twz = roots(nzp), twx = roots(nxp);
    g->pz.tw = g->twz;
    g->pz.tabn = nzp;
    g->pm.tw = g->twx;         // length-nxp/2 transform uses every other root
    g->pm.tabn = nxp;
    if (h2d(g, g->alfa, alfa.data(), alfa.size()) || h2d(g, g->beta, beta.data(), beta.size()) ||
        h2d(g, g->deta, deta.data(), deta.size()) || h2d(g, g->Tf, Tf.data(), Tf.size()) ||
        h2d(g, g->Tp, Tp.data(), Tp.size()) || h2d(g, g->Tb, Tb.data(), Tb.size()) ||
        h2d(g, g->twz, twz.data(), twz.size()) || h2d(g, g->twx, twx.data(), twx.size()))
        return fail(SIMSON_E_HIP);
    // FP32 state and tables of the mixed form (rounded from the FP64 ones)
    std::vector<float2> twz32(nzp), twx32(nxp);
    std::vector<float> Tf32((size_t)ny * ny), Tb32((size_t)ny * ny);
    g->mixed = p->precision == SIMSON_MIXED;
    if (g->mixed) {
        for (int i = 0; i < nzp; i++)
            twz32[i] = make_float2((float)twz[i].x, (float)twz[i].y);
        for (int i = 0; i < nxp; i++)
            twx32[i] = make_float2((float)twx[i].x, (float)twx[i].y);
        for (size_t i = 0; i < Tf32.size(); i++) {
            Tf32[i] = (float)Tf[i];
            Tb32[i] = (float)Tb[i];
        }
        if (dalloc(g, &g->U32, kNPhys * F) || dalloc(g, &g->P32, (size_t)6 * ny * nzp * nxh) ||
            dalloc(g, &g->C32, 5 * F) || dalloc(g, &g->twz32, (size_t)nzp) ||
            dalloc(g, &g->twx32, (size_t)nxp) || dalloc(g, &g->Tf32, Tf32.size()) ||
            dalloc(g, &g->Tb32, Tb32.size()) ||
            h2d(g, g->twz32, twz32.data(), twz32.size()) ||
            h2d(g, g->twx32, twx32.data(), twx32.size()) ||
            h2d(g, g->Tf32, Tf32.data(), Tf32.size()) || h2d(g, g->Tb32, Tb32.data(), Tb32.size()))
            return fail(SIMSON_E_HIP);
    }
    if (int st = make_op(g, g->oTf, Tf, g->Tf, g->Tf32, 0, 1))
        return fail(st);
    if (int st = make_op(g, g->oTp, Tp, g->Tp, nullptr, 0, 1))
        return fail(st);
    if (int st = make_op(g, g->oTb, Tb, g->Tb, g->Tb32, 1, 0))
        return fail(st);
    if (g->mixed && !g->oTf.S32 && cheb_mfma_ks(ny) == 0) {
        g->err = "no matrix-core transform for this ny";
        return fail(SIMSON_E_ARG);
    }
    // pressure gradient set-up (bla.f:1622-1662)
    double px;
    if (p->cflux) {
        g->mflux = 4. / 3.;
        px = -2. / p->re;
    } else {
        g->mflux = 0.;
        px = -(p->retau / p->re) * (p->retau / p->re);
    }
    if (h2d(g, g->px, &px, 1))
        return fail(SIMSON_E_HIP);
    // tau factorisation for lam = k^2 (fixed for the run)
    ColArgs ca = col_args(g);
    launch_factor_k2(ca, g->kfac, g->kfT, g->stream);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g->stream) != hipSuccess) {
        g->err = "k^2 factorisation failed";
        return fail(SIMSON_E_HIP);
    }
    (void)nx;
    *out = g;
    return 0;
}

int simson_gpu_upload(simson_gpu_t *g, const double *ur, const double *ui, int nslots,
                      const double *puw, double px)
{
    if (!g || !ur || !ui || (nslots != 3 && nslots != kNSlot) || (nslots == kNSlot && !puw))
        return SIMSON_E_ARG;
    const size_t F = (size_t)g->nz * g->ny * g->nxh;
    const int nphys = nslots < kNPhys ? nslots : kNPhys;
    std::vector<double2> h(nphys * F);
    for (size_t i = 0; i < nphys * F; i++)
        h[i] = make_double2(ur[i], ui[i]);
    if (h2d(g, g->U, h.data(), h.size()))
        return SIMSON_E_HIP;
    if (nslots == kNSlot) {
        // p_v, p_omega: main-storage [z][y][x] -> Chebyshev columns [z][x][y]
        std::vector<double2> hp(2 * F);
        const int nz = g->nz, ny = g->ny, nxh = g->nxh;
        for (int s = 0; s < 2; s++)
            for (int z = 0; z < nz; z++)
                for (int y = 0; y < ny; y++)
                    for (int x = 0; x < nxh; x++) {
                        const size_t src = (((size_t)(kNPhys + s) * nz + z) * ny + y) * nxh + x;
                        const size_t dst = (((size_t)s * nz + z) * nxh + x) * ny + y;
                        hp[dst] = make_double2(ur[src], ui[src]);
                    }
        if (h2d(g, g->Pp, hp.data(), hp.size()))
            return SIMSON_E_HIP;
        CK(hipStreamSynchronize(g->stream));
    }
    if (puw && h2d(g, g->puw, puw, 2 * (size_t)g->ny))
        return SIMSON_E_HIP;
    if (h2d(g, g->px, &px, 1))
        return SIMSON_E_HIP;
    // wall values of the mean mode (rdiscbl.f:319-323), zeroed below 1e-8
    // (bla.f:1041-1044)
    const size_t ylast = (size_t)(g->ny - 1) * g->nxh;
    double v[4] = {ur[0], ur[ylast], ur[2 * F], ur[2 * F + ylast]};
    for (double &q : v)
        if (std::fabs(q) < 1e-8)
            q = 0.;
    g->u0upp = v[0];
    g->u0low = v[1];
    g->w0upp = v[2];
    g->w0low = v[3];
    g->u32_stale = g->mixed;
    g->u64_stale = false;
    g->mean64 = false;
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

int simson_gpu_prhs(simson_gpu_t *g, double an)
{
    if (!g || an == 0.)
        return SIMSON_E_ARG;
    return do_prhs(g, an);
}

int simson_gpu_nonlin(simson_gpu_t *g, int do_cfl, double *cfl_out)
{
    if (!g)
        return SIMSON_E_ARG;
    int st = do_nonlin(g, do_cfl);
    if (st)
        return st;
    if (do_cfl && cfl_out) {
        double c[2];
        CK(hipMemcpyAsync(c, g->cfl, 2 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        CK(hipStreamSynchronize(g->stream));
        *cfl_out = c[0];
    }
    return 0;
}

int simson_gpu_cfl(simson_gpu_t *g, double *cfl)
{
    if (!g || !cfl)
        return SIMSON_E_ARG;
    // the z synthesis and the x-row pass with the CFL maximum; the z
    // analysis is left out, so the state (and its storage slots) is
    // unchanged
    if (g->mixed) {
        refresh_u32(g);
        Step2Args<float2> a = nonlin_args32(g, 1);
        launch_zsyn4(a, g->twz32, g->stream);
        launch_xrow4(a, g->twx32, g->stream);
        launch_cfl_reduce(g->cfl_blk, xrow4_blocks(a), g->cfl, g->stream);
    } else {
        NonlinArgs a = nonlin_args(g, 1);
        if (!launch_zsyn4(a, (const double2 *)g->twz, g->stream))
            launch_zsyn(a, g->pz, g->stream);
        const bool four = launch_xrow4(a, (const double2 *)g->twx, g->stream);
        if (!four)
            launch_xrow(a, g->pm, g->twx, g->stream);
        launch_cfl_reduce(g->cfl_blk, four ? xrow4_blocks(a) : xrow_blocks(a), g->cfl,
                          g->stream);
    }
    CK(hipGetLastError());
    CK(hipMemcpyAsync(cfl, g->cfl, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

int simson_gpu_linear(simson_gpu_t *g, double an, double bn, double anp1, double bnp1,
                      double *px_out)
{
    if (!g || an + bn == 0. || anp1 + bnp1 == 0.)
        return SIMSON_E_ARG;
    int st = do_linear(g, an, bn, anp1, bnp1);
    if (st)
        return st;
    if (px_out) {
        CK(hipMemcpyAsync(px_out, g->px, sizeof(double), hipMemcpyDeviceToHost, g->stream));
        CK(hipStreamSynchronize(g->stream));
    }
    return 0;
}

int simson_gpu_substep(simson_gpu_t *g, double an, double bn, double anp1, double bnp1,
                       int do_cfl, double *cfl_out, double *px_out)
{
    if (!g || an + bn == 0. || anp1 + bnp1 == 0.)
        return SIMSON_E_ARG;
    int st = do_nonlin(g, do_cfl);
    if (st)
        return st;
    st = do_linear(g, an, bn, anp1, bnp1);
    if (st)
        return st;
    if (do_cfl && cfl_out) {
        double c[2];
        CK(hipMemcpyAsync(c, g->cfl, 2 * sizeof(double), hipMemcpyDeviceToHost, g->stream));
        CK(hipStreamSynchronize(g->stream));
        *cfl_out = c[0];
    }
    if (px_out) {
        CK(hipMemcpyAsync(px_out, g->px, sizeof(double), hipMemcpyDeviceToHost, g->stream));
        CK(hipStreamSynchronize(g->stream));
    }
    return 0;
}

int simson_gpu_run(simson_gpu_t *g, int nsteps, double dt, int icfl, int first, double *cfl_out)
{
    if (!g || nsteps < 0 || !(dt > 0.))
        return SIMSON_E_ARG;
    static const double anrk[3] = {8. / 15., 5. / 12., 3. / 4.};
    static const double bnrk[3] = {0., -17. / 60., -5. / 12.};
    const int nst = 3;
    const double dtn = dt, dtnp1 = dt;
    double an = dtn * anrk[0], bn = 0., anp1 = an, bnp1 = 0.;
    if (!g->cfl_host) {
        CK(hipHostMalloc((void **)&g->cfl_host, 2 * sizeof(double), hipHostMallocDefault));
        CK(hipEventCreateWithFlags(&g->cfl_ev[0], hipEventDisableTiming));
        CK(hipEventCreateWithFlags(&g->cfl_ev[1], hipEventDisableTiming));
    }
    CK(hipMemsetAsync(g->cfl, 0, 2 * sizeof(double), g->stream));
    if (first) {
        int st = do_prhs(g, anp1);
        if (st)
            return st;
    }
    static const double cnrk[3] = {0., 8. / 15., 2. / 3.};
    const double pi = 3.14159265358979323846;
    // CFL*dt of wcflbl.f:87 from the maximum of boxcfl.f; NaN (a field that
    // has blown up) reaches here as +infinity
    auto cfl_dt = [&](double m) { return (m * pi + 2. * std::fabs(g->p.rot)) * dt; };
    const double tstep = g->tnow;
    // The CFL stop of wcflbl.f:93-98 is checked at every evaluation, one
    // evaluation late: before the substep of evaluation n + 1 is queued the
    // host waits for the value of evaluation n, so the GPU keeps the
    // substeps in between queued.  On a stop the run ends after the substep
    // just before the next evaluation (the remaining substeps of the step
    // for icfl = 3).
    int nev = 0, done = 0;
    double last = 0.;
    auto check = [&](int ev) -> int {
        CK(hipEventSynchronize(g->cfl_ev[ev & 1]));
        last = cfl_dt(g->cfl_host[ev & 1]);
        return !(last <= 10.);
    };
    int stop = 0;
    for (int it = 1; it <= nst * nsteps; it++) {
        const int do_cfl = icfl > 0 && ((it - 1) % icfl) == 0;
        if (do_cfl && nev > 0 && check(nev - 1)) {
            stop = 1;
            break;
        }
        int st = do_nonlin(g, do_cfl);
        if (st)
            return st;
        if (do_cfl) {
            CK(hipMemcpyAsync(g->cfl_host + (nev & 1), g->cfl, sizeof(double),
                              hipMemcpyDeviceToHost, g->stream));
            CK(hipEventRecord(g->cfl_ev[nev & 1], g->stream));
            nev++;
        }
        an = anp1;
        bn = bnp1;
        anp1 = dtnp1 * anrk[it % nst];
        bnp1 = dtnp1 * bnrk[it % nst];
        // substep time t + c dt (bla.f:1797)
        g->tnow = tstep + dt * (double)((it - 1) / nst) + dt * cnrk[(it - 1) % nst];
        st = do_linear(g, an, bn, anp1, bnp1);
        if (st)
            return st;
        done = it;
    }
    // time at the end of the last completed substep
    g->tnow = tstep + dt * (double)(done / nst) + dt * cnrk[done % nst];
    CK(hipStreamSynchronize(g->stream));
    if (!stop && nev > 0 && check(nev - 1))
        stop = 1;
    if (cfl_out)
        *cfl_out = last;
    if (icfl > 0 && stop)
        return 324234;                                                  // wcflbl.f:93-98
    return 0;
}

int simson_gpu_download(simson_gpu_t *g, double *ur, double *ui, int nslots, double *puw,
                        double *px)
{
    if (!g || nslots < 0 || nslots > kNSlot || (nslots > 0 && (!ur || !ui)))
        return SIMSON_E_ARG;
    const size_t F = (size_t)g->nz * g->ny * g->nxh;
    refresh_u64(g);
    CK(hipGetLastError());
    if (nslots > 0) {
        const int nphys = nslots < kNPhys ? nslots : kNPhys;
        std::vector<double2> h(nphys * F);
        CK(hipMemcpyAsync(h.data(), g->U, h.size() * sizeof(double2), hipMemcpyDeviceToHost,
                          g->stream));
        CK(hipStreamSynchronize(g->stream));
        for (size_t i = 0; i < nphys * F; i++) {
            ur[i] = h[i].x;
            ui[i] = h[i].y;
        }
    }
    if (nslots > kNPhys) {
        // Chebyshev columns [z][x][y] -> main-storage [z][y][x]
        std::vector<double2> hp(2 * F);
        CK(hipMemcpyAsync(hp.data(), g->Pp, hp.size() * sizeof(double2), hipMemcpyDeviceToHost,
                          g->stream));
        CK(hipStreamSynchronize(g->stream));
        const int nz = g->nz, ny = g->ny, nxh = g->nxh;
        for (int s = 0; s < nslots - kNPhys; s++)
            for (int z = 0; z < nz; z++)
                for (int y = 0; y < ny; y++)
                    for (int x = 0; x < nxh; x++) {
                        const size_t dst = (((size_t)(kNPhys + s) * nz + z) * ny + y) * nxh + x;
                        const size_t src = (((size_t)s * nz + z) * nxh + x) * ny + y;
                        ur[dst] = hp[src].x;
                        ui[dst] = hp[src].y;
                    }
    }
    if (puw)
        CK(hipMemcpyAsync(puw, g->puw, 2 * (size_t)g->ny * sizeof(double), hipMemcpyDeviceToHost,
                          g->stream));
    if (px)
        CK(hipMemcpyAsync(px, g->px, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

int simson_gpu_amp(simson_gpu_t *g, double *amp)
{
    if (!g || !amp)
        return SIMSON_E_ARG;
    refresh_u64(g);
    launch_amp(g->U, g->alfa, g->beta, g->nz, g->ny, g->nxh, g->amp, g->stream);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(amp, g->amp, 20 * (size_t)g->ny * sizeof(double), hipMemcpyDeviceToHost,
                      g->stream));
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

int simson_gpu_amp_noalloc(simson_gpu_t *g, double *amp)
{
    if (!g || !amp)
        return SIMSON_E_ARG;
    if (g->ny > kAmpScratchNy) {
        g->err = "amp_noalloc: ny above the scratch limit of 257";
        return SIMSON_E_ARG;
    }
    // every entry [20][ny] is written by the kernel, so nothing needs clearing
    refresh_u64(g);
    launch_amp(g->U, g->alfa, g->beta, g->nz, g->ny, g->nxh, g->amp_scratch, g->stream);
    CK(hipGetLastError());
    CK(hipMemcpyAsync(amp, g->amp_scratch, 20 * (size_t)g->ny * sizeof(double),
                      hipMemcpyDeviceToHost, g->stream));
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

int simson_gpu_xys(simson_gpu_t *g, double dtn, double xs, double zs, int nxys, double *xys)
{
    if (!g || !xys)
        return SIMSON_E_ARG;
    if (nxys < 12 || nxys > 15) {
        g->err = "xy statistics: 12 <= nxys <= 15 (indices 1-15 of boxxys.f)";
        return SIMSON_E_ARG;
    }
    const int ny = g->ny, nx = g->nx;
    if (!g->xysd) {
        if (!xys_fits(g->nz, g->nxh)) {
            g->err = "xy statistics: unsupported nx or nz";
            return SIMSON_E_ARG;
        }
        if (int st = unpadded_plans(g))
            return st;
        if (dalloc(g, &g->xysd, (size_t)15 * ny * nx))
            return SIMSON_E_HIP;
    }
    refresh_u64(g);
    XysArgs a{};
    a.U = g->U;
    a.Q = g->P;                   // [6][ny][nz][nxh] within [6][ny][nzp][nxh]
    a.out = g->xysd;
    a.alfa = g->alfa;
    a.beta = g->beta;
    a.ny = ny;
    a.nz = g->nz;
    a.nxh = g->nxh;
    a.c1 = dtn / (double)g->nz;
    a.u0low = g->u0low;
    a.w0low = g->w0low;
    a.xs = xs;
    a.zs = zs;
    a.shift = xs != 0. || zs != 0.;
    launch_xys(a, g->pzs, g->pms, g->twxs, g->stream);
    CK(hipGetLastError());
    const int nout = nxys == 15 ? 15 : 12;
    std::vector<double> h((size_t)nout * ny * nx);
    CK(hipMemcpyAsync(h.data(), g->xysd, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                      g->stream));
    CK(hipStreamSynchronize(g->stream));
    for (size_t i = 0; i < h.size(); i++)
        xys[i] += h[i];
    return 0;
}

int simson_fsc_table(double m, int n, double ymax, double eps, double *tab, double *dstar2)
{
    return sg::fsc_table(m, n, ymax, eps, tab, dstar2);
}

int simson_base_flow(const simson_baseflow_t *b, const double *tab, int ntab, double *base,
                     double *bjr, double *bji, double *du0upp)
{
    return sg::spatial_base_flow(b, tab, ntab, base, bjr, bji, du0upp);
}

int simson_gpu_free_stream(simson_gpu_t *g, const double *bjr, const double *bji, double du0upp)
{
    if (!g || (g->p.ibc != 100 && g->p.ibc != 101) || (!bjr) != (!bji) ||
        !std::isfinite(du0upp))
        return SIMSON_E_ARG;
    g->du0upp = du0upp;
    g->bj_set = false;
    if (bjr) {
        const size_t m = 3 * (size_t)g->nxh;
        std::vector<double2> h(m);
        for (size_t i = 0; i < m; i++)
            h[i] = make_double2(bjr[i], bji[i]);
        if (h2d(g, g->bj, h.data(), m))
            return SIMSON_E_HIP;
        CK(hipStreamSynchronize(g->stream));
        g->bj_set = true;
    }
    return 0;
}

int simson_parallel_force(const simson_baseflow_t *b, const double *tab, int ntab, double cdev,
                          double t, double *force)
{
    return sg::parallel_force(b, tab, ntab, cdev, t, force);
}

int simson_gpu_mean_force(simson_gpu_t *g, const double *force)
{
    if (!g)
        return SIMSON_E_ARG;
    if (!force) {
        g->mforce_on = false;
        return 0;
    }
    const int ny = g->ny;
    for (int i = 0; i < 3 * ny; i++)
        if (!std::isfinite(force[i]))
            return SIMSON_E_ARG;
    if (!g->mforce && dalloc(g, &g->mforce, 2 * (size_t)ny))
        return SIMSON_E_HIP;
    // Chebyshev coefficients of f_u and f_w (the scaled forward transform
    // of linearbl.f:262-273 applied to the forcing times nxp nzp)
    std::vector<double> h(2 * (size_t)ny, 0.);
    const double pi = 3.14159265358979323846;
    for (int c = 0; c < 2; c++) {
        const double *f = force + (c == 0 ? 0 : 2 * ny);
        for (int k = 0; k < ny; k++) {
            double s = 0.;
            for (int j = 0; j < ny; j++) {
                const double w = (j == 0 || j == ny - 1) ? .5 : 1.;
                s += w * f[j] * std::cos(pi * (double)(((long)j * k) % (2 * (ny - 1))) /
                                         (double)(ny - 1));
            }
            const double ck = (k == 0 || k == ny - 1) ? .5 : 1.;
            h[(size_t)c * ny + k] = s * ck * 2. / (double)(ny - 1);
        }
    }
    if (h2d(g, g->mforce, h.data(), h.size()))
        return SIMSON_E_HIP;
    CK(hipStreamSynchronize(g->stream));
    g->mforce_on = true;
    return 0;
}

int simson_local_force_pattern(int nx, int ny, double xl, const simson_locf_t *lf, double *pr,
                               double *pi)
{
    if (!pr || !pi)
        return SIMSON_E_ARG;
    std::vector<double2> h(3 * (size_t)ny * (nx / 2));
    const int st = sg::local_force_pattern(nx, ny, xl, lf, h.data());
    if (st != 0)
        return st;
    for (size_t i = 0; i < h.size(); i++) {
        pr[i] = h[i].x;
        pi[i] = h[i].y;
    }
    return 0;
}

int simson_gpu_local_force(simson_gpu_t *g, const simson_locf_t *lf)
{
    if (!g)
        return SIMSON_E_ARG;
    if (!lf) {
        g->lforce_on = false;
        return 0;
    }
    const int ny = g->ny, nxh = g->nxh;
    std::vector<double2> phys(3 * (size_t)ny * nxh);
    const int st = sg::local_force_pattern(g->nx, ny, g->p.xl, lf, phys.data());
    if (st != 0)
        return st;
    if (!g->lforce && dalloc(g, &g->lforce, 3 * (size_t)ny * nxh))
        return SIMSON_E_HIP;
    // Chebyshev coefficients in y of each (component, kx) column: the
    // scaled forward transform of linearbl.f:262-273 applied to the force
    // times nxp nzp
    std::vector<double2> h(phys.size());
    const double pi = 3.14159265358979323846;
    std::vector<double> cs(2 * (size_t)(ny - 1));
    for (size_t q = 0; q < cs.size(); q++)
        cs[q] = std::cos(pi * (double)q / (double)(ny - 1));
    for (int c = 0; c < 3; c++)
        for (int x = 0; x < nxh; x++)
            for (int k = 0; k < ny; k++) {
                double sr = 0., si = 0.;
                for (int y = 0; y < ny; y++) {
                    const double w = (y == 0 || y == ny - 1) ? .5 : 1.;
                    const double cv = cs[(size_t)(((long)y * k) % (2 * (ny - 1)))];
                    const double2 v = phys[((size_t)c * ny + y) * nxh + x];
                    sr += w * v.x * cv;
                    si += w * v.y * cv;
                }
                const double ck = ((k == 0 || k == ny - 1) ? .5 : 1.) * 2. / (double)(ny - 1);
                h[((size_t)c * nxh + x) * ny + k] = make_double2(sr * ck, si * ck);
            }
    if (h2d(g, g->lforce, h.data(), h.size()))
        return SIMSON_E_HIP;
    CK(hipStreamSynchronize(g->stream));
    g->lforce_omega = lf->tomeg;
    g->lforce_on = true;
    return 0;
}

int simson_gpu_set_time(simson_gpu_t *g, double t)
{
    if (!g)
        return SIMSON_E_ARG;
    g->tnow = t;
    return 0;
}

int simson_wall_velocity(int nx, int nz, double xl, double zl, const simson_wall_t *w,
                         double tc, double *vr, double *vi)
{
    if (!w || !vr || !vi)
        return SIMSON_E_ARG;
    const size_t nc = (size_t)nz * (nx / 2);
    std::vector<double2> pm(nc), mn(nc);
    if (wall_spectra(nx, nz, xl, zl, w, pm.data(), mn.data()) != 0)
        return SIMSON_E_ARG;
    const double sv = w->tomeg == 0. ? 1. : std::sin(w->tomeg * tc);
    const std::vector<double2> &q = sv >= 0. ? pm : mn;
    for (size_t i = 0; i < nc; i++) {
        vr[i] = sv * q[i].x;
        vi[i] = sv * q[i].y;
    }
    return 0;
}

int simson_gpu_wall(simson_gpu_t *g, const simson_wall_t *w)
{
    if (!g)
        return SIMSON_E_ARG;
    if (!w) {
        g->wall = 0;
        return 0;
    }
    const size_t nc = (size_t)g->nz * g->nxh;
    std::vector<double2> pm(nc), mn(nc);
    if (wall_spectra(g->nx, g->nz, g->p.xl, g->p.zl, w, pm.data(), mn.data()) != 0) {
        g->err = "wall: xrise, xfall > 0 needed";
        return SIMSON_E_ARG;
    }
    if (!g->wallP && (dalloc(g, &g->wallP, nc) || dalloc(g, &g->wallM, nc)))
        return SIMSON_E_HIP;
    if (h2d(g, g->wallP, pm.data(), nc) || h2d(g, g->wallM, mn.data(), nc))
        return SIMSON_E_HIP;
    CK(hipStreamSynchronize(g->stream));
    g->wall_tomeg = w->tomeg;
    g->wall_p0 = pm[0].x;
    g->wall_m0 = mn[0].x;
    g->wall = 1;
    return 0;
}

int simson_fringe_function(int nx, double xl, const simson_fringe_t *f, double *lam)
{
    if (!f || !lam)
        return SIMSON_E_ARG;
    return fringe_function(nx, xl, f, lam);
}

int simson_gpu_fringe(simson_gpu_t *g, const simson_fringe_t *f, const double *base)
{
    if (!g)
        return SIMSON_E_ARG;
    if (!f) {
        g->fringe = 0;
        return 0;
    }
    if (!base)
        return SIMSON_E_ARG;
    if (!has_xrow4(g->nxp / 2)) {
        g->err = "fringe: nxp/2 needs a specialised x-row length";
        return SIMSON_E_ARG;
    }
    const int nxp = g->nxp, M = nxp / 2, ny = g->ny;
    std::vector<double> lam(nxp);
    if (fringe_function(g->nx, g->p.xl, f, lam.data()) != 0) {
        g->err = "fringe: fmax, frise, ffall > 0 and fstart < fend within -xl/2 to xl/2";
        return SIMSON_E_ARG;
    }
    std::vector<double2> hl(M), hb((size_t)3 * ny * M);
    for (int n = 0; n < M; n++)
        hl[n] = make_double2(lam[2 * n], lam[2 * n + 1]);
    for (int c = 0; c < 3; c++)
        for (int y = 0; y < ny; y++)
            for (int n = 0; n < M; n++) {
                const double *b = base + ((size_t)c * ny + y) * nxp;
                hb[((size_t)c * ny + y) * M + n] = make_double2(b[2 * n], b[2 * n + 1]);
            }
    if (!g->flam && (dalloc(g, &g->flam, hl.size()) || dalloc(g, &g->fbase, hb.size())))
        return SIMSON_E_HIP;
    if (h2d(g, g->flam, hl.data(), hl.size()) || h2d(g, g->fbase, hb.data(), hb.size()))
        return SIMSON_E_HIP;
    if (g->mixed) {
        std::vector<float2> l32(M), b32(hb.size());
        for (int n = 0; n < M; n++)
            l32[n] = make_float2((float)hl[n].x, (float)hl[n].y);
        for (size_t i = 0; i < hb.size(); i++)
            b32[i] = make_float2((float)hb[i].x, (float)hb[i].y);
        if (!g->flam32 &&
            (dalloc(g, &g->flam32, l32.size()) || dalloc(g, &g->fbase32, b32.size())))
            return SIMSON_E_HIP;
        if (h2d(g, g->flam32, l32.data(), l32.size()) ||
            h2d(g, g->fbase32, b32.data(), b32.size()))
            return SIMSON_E_HIP;
    }
    CK(hipStreamSynchronize(g->stream));
    g->fringe = 1;
    return 0;
}

int simson_les_filter_secondary(int nx, int nz, double cutoff, int iord, double *lpf2)
{
    if (!lpf2)
        return SIMSON_E_ARG;
    return les_filter_secondary(nx, nz, cutoff, iord, lpf2);
}

int simson_les_filter(int nx, int ny, int nz, double cutoff, int iord, double *lpf, double *L)
{
    if (!lpf || !L)
        return SIMSON_E_ARG;
    return les_filter(nx, ny, nz, cutoff, iord, lpf, L);
}

// the ADM-RT filters (init_filter.f, bla.f:1489-1518) on the device: L
// (FP64, and FP32 in the mixed form when fp32) and lpf; and, when lpf2 is
// given, the secondary x-z transfer function of bla.f:1520-1543
int setup_filters(simson_gpu_t *g, const simson_les_t *p, bool fp32, bool secondary)
{
    const int ny = g->ny;
    std::vector<double> lpf((size_t)g->nz * g->nxh), L((size_t)ny * ny), lpf2;
    if (les_filter(g->nx, ny, g->nz, p->cutoff, p->iord, lpf.data(), L.data()) != 0) {
        g->err = "invalid LES filter parameters (0 < cutoff < 1, 0 <= iord <= 5)";
        return SIMSON_E_ARG;
    }
    if (secondary) {
        lpf2.resize(lpf.size());
        if (les_filter_secondary(g->nx, g->nz, p->cutoff, p->iord, lpf2.data()) != 0) {
            g->err = "invalid LES filter parameters (0 < cutoff < 1, 0 <= iord <= 5)";
            return SIMSON_E_ARG;
        }
    }
    if (!g->Lm && (dalloc(g, &g->Lm, L.size()) || dalloc(g, &g->lpf, lpf.size())))
        return SIMSON_E_HIP;
    if (fp32 && !g->Lm32 && dalloc(g, &g->Lm32, L.size()))
        return SIMSON_E_HIP;
    if (secondary && !g->lpf2 && dalloc(g, &g->lpf2, lpf2.size()))
        return SIMSON_E_HIP;
    if (h2d(g, g->Lm, L.data(), L.size()) || h2d(g, g->lpf, lpf.data(), lpf.size()))
        return SIMSON_E_HIP;
    if (secondary && h2d(g, g->lpf2, lpf2.data(), lpf2.size()))
        return SIMSON_E_HIP;
    std::vector<float> L32;
    if (fp32) {
        L32.assign(L.begin(), L.end());
        if (h2d(g, g->Lm32, L32.data(), L32.size()))
            return SIMSON_E_HIP;
    }
    if (int st = make_op(g, g->oLm, L, g->Lm, fp32 ? g->Lm32 : nullptr, 0, 0))
        return st;
    // the copies read host vectors that go out of scope on return
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

// sgs.i iles = 2 (rparamles.f:72-81, 117-165): fixed cs or, with cs = 0,
// the dynamic procedure with clipping ineg; Delta^2 of preprbl.f:60-71.
// iles = 3 (rparamles.f:82-91): the same on the high-pass filtered fields.
int setup_smag(simson_gpu_t *g, const simson_les_t *p)
{
    const bool hpf = p->iles == 3;
    if (p->ineg < 0 || p->ineg > 2 || (p->cs != 0. && p->ineg != 1)) {
        g->err = "Smagorinsky: 0 <= ineg <= 2, and ineg = 1 with a fixed cs";
        return SIMSON_E_ARG;
    }
    int dev = 0, lmax = 0;
    CK(hipGetDevice(&dev));
    CK(hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    // the dynamic LDS of the x-row kernels, the static arrays of k_smag_xc
    // (12 KiB) aside
    const size_t lds = (size_t)std::max(65536, lmax) - 12288;
    if (!smag_fits(g->nz, g->nxh, lds) || g->nx % 4 != 0 || !has_zfft4(g->nzp) ||
        (g->mixed && !cheb_f32_ok(g->ny))) {
        g->err = "Smagorinsky: unsupported grid size";
        return SIMSON_E_ARG;
    }
    if (int st = unpadded_plans(g))
        return st;
    if (int st = alloc_stress(g, true, g->mixed))
        return st;
    const int ny = g->ny, nz = g->nz, nq = g->nx / 4;
    const size_t F = field_size(g);
    const bool dyn = p->cs == 0.;
    const int nqf = hpf && dyn ? 21 : 12;
    if (hpf) {
        if (int st = setup_filters(g, p, false, dyn))
            return st;
        if (!g->LW && dalloc(g, &g->LW, 5 * F))
            return SIMSON_E_HIP;
    }
    if (g->SQ && g->sq_nf < nqf) {
        (void)hipFree(g->SQ);
        g->SQ = nullptr;
        if (dalloc(g, &g->SQ, (size_t)nqf * F))
            return SIMSON_E_HIP;
        g->sq_nf = nqf;
    }
    if (!g->SQ) {
        if (dalloc(g, &g->SQ, (size_t)nqf * F) || dalloc(g, &g->SS, 7 * F) ||
            dalloc(g, &g->dlt2, (size_t)ny))
            return SIMSON_E_HIP;
        g->sq_nf = nqf;
        if (smag_prepare(lds) != 0) {
            g->err = "Smagorinsky: cannot raise the LDS limit of the x-row kernels";
            return SIMSON_E_HIP;
        }
        std::vector<double> d2(ny), eta(ny);
        for (int j = 0; j < ny; j++)
            eta[j] = cos(M_PI * j / (ny - 1));
        for (int j = 0; j < ny; j++) {
            const double dy = (j == 0 || j == ny - 1)
                                  ? 1. - eta[1]
                                  : sqrt((eta[j + 1] - eta[j]) * (eta[j] - eta[j - 1]));
            d2[j] = pow(g->p.xl / g->nx * g->p.zl / nz * dy, 2. / 3.);
        }
        if (h2d(g, g->dlt2, d2.data(), d2.size()))
            return SIMSON_E_HIP;
    }
    if (dyn && !g->SW &&
        (dalloc(g, &g->SW, (size_t)21 * ny * nz * nq) ||
         dalloc(g, &g->SP, (size_t)ny * smag_zblocks(nz, g->nxh, lds) * 2 * g->nx)))
        return SIMSON_E_HIP;
    g->smag_lds = lds;
    SmagArgs &a = g->sm;
    a = SmagArgs{};
    a.Q = g->SQ;
    a.SS = g->SS;
    a.W = g->SW;
    a.part = g->SP;
    a.TAU = g->TAU;
    a.alfa = g->alfa;
    a.beta = g->beta;
    a.dlt2 = g->dlt2;
    a.nx = g->nx;
    a.ny = ny;
    a.nz = nz;
    a.nxh = g->nxh;
    a.nq = nq;
    a.dynamic = dyn;
    a.hpf = hpf;
    a.nqf = nqf;
    a.LW = g->LW;
    a.lpf = g->lpf;
    a.lpf2 = g->lpf2;
    a.ineg = p->ineg;
    a.cs2 = p->cs * p->cs;
    a.tw2 = pow(4., 2. / 3.);
    a.re = g->p.re;
    a.hscale = (double)g->nxp * (double)g->nzp;
    g->smag = true;
    g->les = false;
    g->clos = false;
    return 0;
}

int simson_gpu_les(simson_gpu_t *g, const simson_les_t *p)
{
    if (!g || !p)
        return SIMSON_E_ARG;
    if (p->iles == 0) {
        g->les = false;
        g->smag = false;
        return 0;
    }
    if (p->iles == 2 || p->iles == 3)
        return setup_smag(g, p);
    if (p->iles != 1) {
        g->err = "iles must be 0, 1 (ADM-RT), 2 (Smagorinsky) or 3 (HPF Smagorinsky)";
        return SIMSON_E_ARG;
    }
    if (!has_zfft4(g->nzp) || (g->mixed && !cheb_f32_ok(g->ny))) {
        g->err = "ADM-RT needs nzp among the specialised z transform lengths";
        return SIMSON_E_ARG;
    }
    const int ny = g->ny;
    if (int st = setup_filters(g, p, g->mixed, false))
        return st;
    const size_t n3 = 3 * field_size(g);
    if (g->mixed ? !g->Wf32 && dalloc(g, &g->Wf32, n3) : !g->Wf && dalloc(g, &g->Wf, n3))
        return SIMSON_E_HIP;
    if (g->mixed && !g->oLm.S32 && cheb_mfma_ks(ny) == 0) {
        g->err = "the ADM-RT low-pass has no matrix-core form for this ny";
        return SIMSON_E_ARG;
    }
    g->les_chi = p->chi;
    g->les = true;
    g->clos = false;
    g->smag = false;
    return 0;
}

int simson_gpu_closure(simson_gpu_t *g, const float *params, const double *in_scale,
                       const double *in_shift, const double *yfeat, const double *out_scale,
                       const double *out_shift)
{
    if (!g)
        return SIMSON_E_ARG;
    if (!params) {
        g->clos = false;
        return 0;
    }
    if (!in_scale || !in_shift || !yfeat || !out_scale || !out_shift)
        return SIMSON_E_ARG;
    if (!has_zfft4(g->nzp) || !has_xrow4(g->nxp / 2) || !cheb_f32_ok(g->ny)) {
        g->err = "the learned closure needs the specialised FFT lengths and an ny with "
                 "matrix-core transforms";
        return SIMSON_E_ARG;
    }
    const int npar = simson_mlp_nparams(), ny = g->ny;
    if (!g->clW) {
        const size_t npg = 9 * (size_t)ny * g->nzp * g->nxh;
        if (dalloc(g, &g->clW, (size_t)npar) || dalloc(g, &g->clY, (size_t)ny))
            return SIMSON_E_HIP;
        if (g->mixed ? dalloc(g, &g->PG32, npg) : dalloc(g, &g->PG, npg))
            return SIMSON_E_HIP;
    }
    if (int st = alloc_stress(g, !g->mixed, g->mixed))
        return st;
    std::vector<float> yf(yfeat, yfeat + ny);
    if (h2d(g, g->clW, params, (size_t)npar) || h2d(g, g->clY, yf.data(), yf.size()))
        return SIMSON_E_HIP;
    ClosArgs &c = g->cl;
    c.W = g->clW;
    c.yfeat = g->clY;
    c.ob1 = 16 * 64;
    c.ow2 = c.ob1 + 64;
    c.ob2 = c.ow2 + 64 * 64;
    c.ow3 = c.ob2 + 64;
    c.ob3 = c.ow3 + 64 * 16;
    for (int k = 0; k < 9; k++) {
        c.in_a[k] = (float)in_scale[k];
        c.in_b[k] = (float)in_shift[k];
    }
    for (int o = 0; o < 6; o++) {
        c.out_a[o] = (float)out_scale[o];
        c.out_b[o] = (float)out_shift[o];
    }
    g->clos = true;
    g->les = false;
    g->smag = false;
    return 0;
}

int simson_gpu_pressure(simson_gpu_t *g, double *pr, double *pi)
{
    if (!g || !pr || !pi)
        return SIMSON_E_ARG;
    if (g->p.ibc != 0) {
        g->err = "pressure: only with walls at both boundaries (ibc = 0)";
        return SIMSON_E_ARG;
    }
    if (!has_zfft4(g->nzp) || !has_xrow4(g->nxp / 2) || g->ny > pressure_max_ny()) {
        g->err = "pressure: nzp and nxp/2 need a specialised FFT length";
        return SIMSON_E_ARG;
    }
    const size_t F = field_size(g);
    const int odd = g->nz / 2;
    if (!g->Q) {
        const size_t ncol = (size_t)g->nz * g->nxh;
        if (dalloc(g, &g->Q, 5 * F) || dalloc(g, &g->Qc, 5 * F) || dalloc(g, &g->PR, F) ||
            dalloc(g, &g->pscr, 6 * (size_t)g->ny * ncol))
            return SIMSON_E_HIP;
    }
    refresh_u64(g);
    // nonlinp.f: H and W = u.u of the current state into Q slots 0-3
    NonlinArgs a = nonlin_args(g, 0);
    a.Uw = g->Q;
    launch_zsyn4(a, (const double2 *)g->twz, g->stream);
    launch_xrow4_press(a, (const double2 *)g->twx, g->stream);
    launch_zana4(a, (const double2 *)g->twz, g->stream, 4);
    CK(hipMemcpyAsync(g->Q + 4 * F, g->U + F, F * sizeof(double2), hipMemcpyDeviceToDevice,
                      g->stream));
    // Chebyshev analysis, linearp.f per column, synthesis of p
    GemmIO fw = gemm_io(g, g->Q, 0, g->Qc, 1, {0, 1, 2, 3, 4}, {0, 1, 2, 3, 4}, odd);
    launch_cheb_gemm(g->oTf, fw, g->stream);
    launch_pressure_cols(g->Qc, g->kfac.gp, g->alfa, g->beta, g->nz, g->nxh, g->ny, g->p.re,
                         (double)g->nxp * (double)g->nzp, g->pscr, g->Qc, g->stream);
    GemmIO bw = gemm_io(g, g->Qc, 1, g->PR, 0, {0}, {0}, odd);
    launch_cheb_gemm(g->oTb, bw, g->stream);
    // the z-oddball plane keeps W, as slot 8 does after nonlinp.f
    const size_t plane = (size_t)g->ny * g->nxh;
    CK(hipMemcpyAsync(g->PR + (size_t)odd * plane, g->Q + 3 * F + (size_t)odd * plane,
                      plane * sizeof(double2), hipMemcpyDeviceToDevice, g->stream));
    CK(hipGetLastError());
    std::vector<double2> h(F);
    CK(hipMemcpyAsync(h.data(), g->PR, F * sizeof(double2), hipMemcpyDeviceToHost, g->stream));
    CK(hipStreamSynchronize(g->stream));
    for (size_t i = 0; i < F; i++) {
        pr[i] = h[i].x;
        pi[i] = h[i].y;
    }
    return 0;
}

int simson_gpu_sync(simson_gpu_t *g)
{
    if (!g)
        return SIMSON_E_ARG;
    CK(hipStreamSynchronize(g->stream));
    return 0;
}

int simson_gpu_timing_enable(simson_gpu_t *g, int on)
{
    if (!g)
        return SIMSON_E_ARG;
    g->timing = on != 0;
    return 0;
}

int simson_gpu_timing_read(simson_gpu_t *g, const char **names, double *total_ms, long *count,
                           int cap)
{
    if (!g)
        return SIMSON_E_ARG;
    CK(hipStreamSynchronize(g->stream));
    for (auto &t : g->launches) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
            g->tot_ms[t.fam] += ms;
            g->cnt[t.fam] += 1;
        }
        g->pool.push_back(t.a);
        g->pool.push_back(t.b);
    }
    g->launches.clear();
    int n = 0;
    for (int f = 0; f < F_N && n < cap; f++) {
        if (names)
            names[n] = kFamilies[f];
        if (total_ms)
            total_ms[n] = g->tot_ms[f];
        if (count)
            count[n] = g->cnt[f];
        n++;
    }
    for (int f = 0; f < F_N; f++) {
        g->tot_ms[f] = 0.;
        g->cnt[f] = 0;
    }
    return n;
}

const char *simson_gpu_last_error(const simson_gpu_t *g)
{
    return g ? g->err.c_str() : "null handle";
}

void simson_gpu_finalize(simson_gpu_t *g)
{
    if (!g)
        return;
    (void)hipStreamSynchronize(g->stream);
    free_all(g);
    delete g;
}

}  // extern "C"
