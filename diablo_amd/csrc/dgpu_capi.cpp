// dgpu_capi.cpp -- host side of libdiablo_gpu.so: the extern "C" entry points
// declared in include/diablo_gpu.h, device memory, FFT plans and the
// Crank-Nicolson solve tables.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/diablo_gpu.h"
#include "dgpu_kernels.hpp"
#include "dgpu_plan.hpp"

using dgpu::cx;

namespace {

constexpr int kKinds = 5;  // timed kernel kinds: zinv, xfused, zfwd_nl, cn, remdiv

struct TimedLaunch {
    int kind;
    hipEvent_t a, b;
};

}  // namespace

extern "C" {
struct dgpu_ctx {
    dgpu_params p;
    dgpu::Geo g;
    int jstart, jend, jstart_th[DGPU_MAX_TH], jend_th[DGPU_MAX_TH];
    std::vector<double> gy, gyf, dy, dyf, kx, kz, kx2, kz2;
    double *d_vec = nullptr;  // dy, dyf, kx, kz, kx2, kz2 (device)
    double *d_dy = nullptr, *d_dyf = nullptr, *d_kx = nullptr, *d_kz = nullptr, *d_kx2 = nullptr, *d_kz2 = nullptr;
    dgpu::Fft<double> fx, fz;
    cx<double> *d_tw = nullptr;
    // in-place plans of the substep's transforms and their twiddle tables
    // (x table, then z table; FP32 copies for the mixed path)
    dgpu::IpLen ipx, ipz;
    cx<double> *d_iptw = nullptr;
    cx<float> *d_iptwf = nullptr;
    int ld = 0;      // row length of the intermediates (>= mk, multiple of 8)
    int xchunk = 64; // planes per block of the fused x kernel (set at creation)
    int colnw = 8;   // waves per block of the segmented column kernels; 0: whole-column sweeps
    int colfuse = -1;     // fused column kernel: -1 where NY >= 160 (FP64: at 8 waves), 0 never, 1 where it fits
    // mixed paths: FP32 transforms, intermediates and nonlinear terms
    int mixed = 0;
    // DGPU_MIXED: FP32 working state [slot][j][K][I] (same slots as d_state's
    // first nstate), FP64 side columns of the low modes [slot][j][l] (state
    // slots, then the solves' scratch), the low-mode map and FP32 Poisson
    // factors.  work_dirty: the working copy is newer than d_state;
    // master_dirty: d_state is newer than the working copy.
    bool f32state = false;
    int nstate = 0;
    cx<float> *d_state32 = nullptr;
    cx<double> *d_side = nullptr;
    int *d_map = nullptr;
    int nside = 0;
    float *d_pmultf = nullptr, *d_pinvf = nullptr;
    cx<float> *d_gs32 = nullptr;                     // FP32 solves' scratch (nin fields)
    hipStream_t side_stream = nullptr;               // the low-mode columns run beside the FP32 solves
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool work_dirty = false, master_dirty = true;
    cx<float> *d_ztf = nullptr, *d_wtf = nullptr, *d_nlf = nullptr;
    cx<float> *nlf[3 + DGPU_MAX_TH] = {};
    cx<float> **d_nlptrf = nullptr;
    // Fourier state, one allocation: [slot][j][K][I]
    cx<double> *d_state = nullptr;
    long fs = 0;  // complex elements per field
    int nslot = 0;
    cx<double> *cu[3] = {nullptr, nullptr, nullptr};
    cx<double> *cp = nullptr;
    cx<double> *cth[DGPU_MAX_TH] = {};
    cx<double> *cf[3] = {nullptr, nullptr, nullptr};
    cx<double> *cfth[DGPU_MAX_TH] = {};
    cx<double> *nl[3 + DGPU_MAX_TH] = {};
    cx<double> **d_nlptr = nullptr;
    // transform intermediates
    cx<double> *d_zt = nullptr, *d_wt = nullptr;
    long zfs = 0;  // complex elements per intermediate field: (NY+2)*NZ*ld
    double *d_fort = nullptr;  // Fortran-layout staging for host entry points
    double *d_fa = nullptr, *d_fb = nullptr;  // POISSON_P_CHAN work planes
    double *d_gyf = nullptr;    // GYF (0:NY+1), part of d_vec
    double *d_stats = nullptr;  // dgpu_stats_chan output
    double *d_noalloc = nullptr;  // this GPU's fixed array of dgpu_stats_chan_noalloc
    // COURANT: 0 idle, 1 requested for the next RK_STEP 3, 2 maxima recorded
    int courant_state = 0;
    double *d_vmax_part = nullptr, *d_vmax = nullptr;  // per-block, per-plane velocity maxima
    double *d_file = nullptr;  // one restart-file array, NX*NY*NZ
    // z transforms as split-BF16 DFT-as-GEMM (mixed precision, NZ <= 256)
    bool zdft = false;
    uint16_t *d_zdft = nullptr;
    dgpu::ZdftTab zinv_tab{}, zfwd_tab{};
    double *d_les = nullptr;   // LES work planes, NU_T, SGS force, filter widths
    bool les = false, les_model = false;  // input.dat's LES; les_chan called in each substep
    // CN tables: [stage 0..2][sys][NY+2][kRow]
    double *d_cn = nullptr;
    std::vector<double> h_cn;
    double cn_key[3][4] = {{NAN}, {NAN}, {NAN}};
    // REM_DIV: per-mode Poisson factors [NY+2][ps] and per-row constants
    double *d_pmult = nullptr, *d_pinv = nullptr, *d_rd = nullptr;
    int *d_codes = nullptr;
    double *d_c1 = nullptr;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    hipStream_t ens_stream = nullptr;  // an ensemble member: the ensemble's stream also reads d_cn
    bool use_graphs = true;
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    double gkey[3][4] = {};
    std::string err;
    bool timing = false;
    std::vector<TimedLaunch> pending;
    std::vector<hipEvent_t> event_pool;
    double t_ms[kKinds] = {0, 0, 0, 0, 0};
    long long t_n[kKinds] = {0, 0, 0, 0, 0};
};

// Members sharing NX, NZ and N_TH (hence transform plans, block shapes and
// the mode count): one launch per kernel kind covers all of them.
struct dgpu_ens_group {
    std::vector<int> idx;
    int max_planes = 0, max_span = 0, max_zf = 0;  // tallest member: NY+2, x-sweep planes, z-forward planes
    void *d_zinv = nullptr, *d_xf = nullptr, *d_zf = nullptr, *d_scat = nullptr;
    void *d_col[3] = {nullptr, nullptr, nullptr};   // column descriptors per RK stage
    void *d_cols[3] = {nullptr, nullptr, nullptr};  // low-mode columns (FP32 state)
    std::vector<char> col_host[3], cols_host[3];
    hipStream_t stream = nullptr;  // the group's own stream (several groups)
    hipEvent_t done = nullptr;
};

struct dgpu_ensemble {
    std::vector<dgpu_ctx *> m;
    hipStream_t main = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
    std::vector<double> gkey[3];
    std::vector<int> gshape[3];  // the members' launch shapes each graph was captured with
    bool use_graphs = true;
    // batched launches (all members at one precision)
    bool batched = false;
    std::vector<dgpu_ens_group> groups;
    hipStream_t side = nullptr;
    hipEvent_t ev_f = nullptr, ev_j = nullptr;
    // several groups: each runs its launches on its own stream, forked from
    // and joined back into the ensemble stream (DGPU_ENS_STREAMS=0: one stream)
    bool group_streams = false;
    hipEvent_t gfork = nullptr;
    std::string err;
};
}

namespace {

int fail(dgpu_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(c, expr)                                                                             \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return fail((c), DGPU_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
    } while (0)

bool make_len(int n, dgpu::FftLen &f)
{
    f = dgpu::fft_len_of(n);
    return f.nstages > 0;
}

void twiddles(int n, std::vector<cx<double>> &out)
{
    const long double twopi = 6.283185307179586476925286766559005768L;
    out.resize(n);
    for (int j = 0; j < n; ++j) {
        long double a = -twopi * (long double)j / (long double)n;
        out[j].re = (double)cosl(a);
        out[j].im = (double)sinl(a);
    }
}

// Matrices of one CN system exactly as RK_CHAN_1 + APPLY_BC_*_MPI build them
// (for/channel.f:771-1029, 1660-2282), THOMAS_FORWARD_REAL_MPI's elimination of
// the matrix part (for/mpi.f:673-699), and the other per-row constants of the
// Fourier-space solve (see the T_* layout in dgpu_kernels.hpp).
void build_cn(const dgpu_ctx *c, int sys, double temp1, double temp4, double *tab)
{
    using namespace dgpu;
    const int NY = c->g.ny;
    const double NUV = c->p.nu_v_scale;
    const double *DY = c->dy.data(), *DYF = c->dyf.data();
    std::vector<double> L(NY + 2, 0.0), D(NY + 2, 1.0), U(NY + 2, 0.0), bc00(NY + 2, 0.0), isbc(NY + 2, 1.0);
    std::vector<double> wp(NY + 2, 0.0), wm(NY + 2, 0.0);
    auto setrow = [&](int j, double l, double d, double u, double v) {
        L[j] = l; D[j] = d; U[j] = u; bc00[j] = v; isbc[j] = 1.0;
    };
    int cflo, cfhi, lo, hi;
    if (sys == SYS_U1 || sys == SYS_U3) {
        for (int j = c->jstart; j <= c->jend; ++j) {
            L[j] = -temp1 * NUV / (DY[j] * DYF[j]);
            D[j] = 1. - temp1 * NUV * (-1. / (DY[j + 1] * DYF[j]) - 1. / (DY[j] * DYF[j]));
            U[j] = -temp1 * NUV / (DY[j + 1] * DYF[j]);
            isbc[j] = 0.0;
            // R += TEMP1*NU_V_SCALE*((u(j+1)-u(j))/DY(j+1) - (u(j)-u(j-1))/DY(j))/DYF(j)
            wp[j] = temp1 * NUV / (DY[j + 1] * DYF[j]);
            wm[j] = temp1 * NUV / (DY[j] * DYF[j]);
        }
        const bool isu = (sys == SYS_U1);
        const int cmin = isu ? c->p.u_bc_ymin : c->p.w_bc_ymin;
        const int cmax = isu ? c->p.u_bc_ymax : c->p.w_bc_ymax;
        const double vmin = isu ? c->p.u_bc_ymin_c1 : c->p.w_bc_ymin_c1;
        const double vmax = isu ? c->p.u_bc_ymax_c1 : c->p.w_bc_ymax_c1;
        setrow(0, 0., 1., 0., 0.);
        if (cmin == 0) setrow(1, 0., 1., 0., vmin);
        else setrow(1, 0., -1., 1., DY[2] * vmin);
        if (cmax == 0) {
            setrow(NY + 1, 0., 1., 0., 0.);
            setrow(NY, 0., 1., 0., vmax);
        } else {
            setrow(NY, -1., 1., 0., DY[NY] * vmax);
            setrow(NY + 1, 0., 1., 0., 0.);
        }
        cflo = c->jstart; cfhi = c->jend;
        lo = c->jstart - 1; hi = c->jend + 1;
    } else if (sys == SYS_U2) {
        for (int j = 2; j <= NY; ++j) {
            L[j] = -temp1 * NUV / (DYF[j - 1] * DY[j]);
            D[j] = 1. + temp1 * NUV / (DYF[j] * DY[j]) + temp1 * NUV / (DYF[j - 1] * DY[j]);
            U[j] = -temp1 * NUV / (DYF[j] * DY[j]);
            isbc[j] = 0.0;
            wp[j] = temp1 * NUV / (DYF[j] * DY[j]);
            wm[j] = temp1 * NUV / (DYF[j - 1] * DY[j]);
        }
        if (c->p.v_bc_ymin == 0) {
            setrow(1, 0., 1., 0., c->p.v_bc_ymin_c1);
            setrow(2, 0., 1., 0., c->p.v_bc_ymin_c1);
        } else if (c->p.v_bc_ymin == 1) {
            setrow(1, 0., -1., 1., DYF[1] * c->p.v_bc_ymin_c1);
        }
        setrow(0, 0., 1., 0., 0.);
        if (c->p.v_bc_ymax == 0) {
            setrow(NY + 1, 0., 1., 0., c->p.v_bc_ymax_c1);
            setrow(NY, 0., 1., 0., c->p.v_bc_ymax_c1);
        } else if (c->p.v_bc_ymax == 1) {
            setrow(NY + 1, -1., 1., 0., DYF[NY] * c->p.v_bc_ymax_c1);
        }
        cflo = 2; cfhi = NY;
        lo = 1; hi = NY + 1;
    } else {
        const int n = sys - SYS_TH0;
        const double cc = temp1 / c->p.pr[n] * NUV;
        for (int j = c->jstart_th[n]; j <= c->jend_th[n]; ++j) {
            L[j] = -(cc) / (DY[j] * DYF[j]);
            D[j] = 1. + (cc) / (DY[j + 1] * DYF[j]) + (cc) / (DY[j] * DYF[j]);
            U[j] = -(cc) / (DY[j + 1] * DYF[j]);
            isbc[j] = 0.0;
            wp[j] = cc / (DY[j + 1] * DYF[j]);
            wm[j] = cc / (DY[j] * DYF[j]);
        }
        const double vmin = c->p.th_bc_ymin_c1[n], vmax = c->p.th_bc_ymax_c1[n];
        if (c->p.th_bc_ymin[n] == 0) {
            setrow(0, 0., 1., 0., 0.);
            setrow(1, 0., 1., 0., vmin);
        } else {
            setrow(1, 0., -1., 1., DY[2] * vmin);
            setrow(0, 0., 1., 0., 0.);
        }
        if (c->p.th_bc_ymax[n] == 0) {
            setrow(NY + 1, 0., 1., 0., 0.);
            setrow(NY, 0., 1., 0., vmax);
        } else {
            setrow(NY, -1., 1., 0., DY[NY] * vmax);
            setrow(NY + 1, 0., 1., 0., 0.);
        }
        cflo = c->jstart_th[n]; cfhi = c->jend_th[n];
        lo = c->jstart_th[n] - 1; hi = c->jend_th[n] + 1;
    }
    // forward elimination of the matrix (rows 1..NY+1)
    std::vector<double> A = L, B = D;
    for (int j = 1; j <= NY + 1; ++j) {
        A[j] = -A[j] / B[j - 1];
        B[j] = B[j] + A[j] * U[j - 1];
    }
    const double rgy = (c->g.nth > 0) ? c->p.ri[0] * c->p.grav_y : 0.0;
    for (int j = 0; j <= NY + 1; ++j) {
        double *t = tab + (long)j * kRow;
        for (int e = 0; e < kRow; ++e) t[e] = 0.0;
        t[T_MULT] = (j == 0) ? 0.0 : A[j];
        t[T_INVPIV] = (B[j] != 0.0) ? 1.0 / B[j] : 0.0;
        t[T_UP] = U[j];
        t[T_BC00] = bc00[j];
        t[T_ISBC] = isbc[j];
        if (isbc[j] == 0.0) {
            t[T_WP] = wp[j];
            t[T_WM] = wm[j];
            t[T_W0] = -(wp[j] + wm[j]);
        }
        if (sys == SYS_U2 && j >= 2 && j <= NY) {
            t[T_CPW] = temp4 / DY[j];
            // RI*GRAV_Y*(CTH(j)*DYF(j-1) + CTH(j-1)*DYF(j))/(2*DY(j))
            t[T_B0] = rgy * DYF[j - 1] / (2.0 * DY[j]);
            t[T_BM] = rgy * DYF[j] / (2.0 * DY[j]);
        }
        t[T_CFR] = (j >= cflo && j <= cfhi) ? 1.0 : 0.0;
        t[T_CPY] = (j >= lo && j <= hi) ? 1.0 : 0.0;
    }
}

int nsys_of(const dgpu_ctx *c) { return 3 + c->g.nth; }

long cn_block(const dgpu_ctx *c) { return (long)dgpu::kRow * (c->g.ny + 2); }

int ensure_cn(dgpu_ctx *c, int stage, double temp1, double temp4)
{
    const double key[4] = {temp1, temp4, 0, 0};
    if (c->cn_key[stage][0] == key[0] && c->cn_key[stage][1] == key[1]) return 0;
    const int nsys = nsys_of(c);
    const long blk = cn_block(c);
    double *base = c->h_cn.data() + (long)stage * nsys * blk;
    for (int q = 0; q < nsys; ++q) build_cn(c, q, temp1, temp4, base + q * blk);
    // every launch that may still read this stage's table must finish first:
    // the context's streams and, for an ensemble member, the ensemble's stream
    // (its batched launches and graphs run there)
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (c->side_stream) HIPCHK(c, hipStreamSynchronize(c->side_stream));
    if (c->ens_stream) HIPCHK(c, hipStreamSynchronize(c->ens_stream));
    HIPCHK(c, hipMemcpy(c->d_cn + (long)stage * nsys * blk, base, sizeof(double) * nsys * blk, hipMemcpyHostToDevice));
    c->cn_key[stage][0] = key[0];
    c->cn_key[stage][1] = key[1];
    return 0;
}

hipEvent_t get_event(dgpu_ctx *c)
{
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct TimerScope {
    dgpu_ctx *c;
    int kind;
    hipEvent_t a = nullptr;
    TimerScope(dgpu_ctx *c_, int k) : c(c_), kind(k)
    {
        if (c->timing) {
            a = get_event(c);
            if (a) (void)hipEventRecord(a, c->stream);
        }
    }
    ~TimerScope()
    {
        if (c->timing && a) {
            hipEvent_t b = get_event(c);
            if (b) {
                (void)hipEventRecord(b, c->stream);
                c->pending.push_back({kind, a, b});
            }
        }
    }
};

int check_launch(dgpu_ctx *c, const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, DGPU_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return 0;
}

int ensure_fort(dgpu_ctx *c)
{
    if (c->d_fort) return 0;
    const size_t n = (size_t)(c->g.nx + 2) * (c->g.nz + 2) * (c->g.ny + 2);
    HIPCHK(c, hipMalloc(&c->d_fort, n * sizeof(double)));
    return 0;
}

cx<double> *field_ptr(dgpu_ctx *c, int field)
{
    if (field >= DGPU_CU1 && field <= DGPU_CU3) return c->cu[field - DGPU_CU1];
    if (field == DGPU_CP) return c->cp;
    if (field >= DGPU_CF1 && field <= DGPU_CF3) return c->cf[field - DGPU_CF1];
    if (field >= DGPU_CTH0 && field < DGPU_CTH0 + c->g.nth) return c->cth[field - DGPU_CTH0];
    if (field >= DGPU_CFTH0 && field < DGPU_CFTH0 + c->g.nth) return c->cfth[field - DGPU_CFTH0];
    return nullptr;
}

__global__ void k_pack_fortran(const cx<double> *__restrict__ kept, double *__restrict__ fort, int nxp, int nz,
                               int mk, int mz)
{
    const int j = blockIdx.y;
    const long plane = (long)(nxp + 1) * (nz + 2);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < plane; e += (long)gridDim.x * blockDim.x) {
        const int K = (int)(e / (nxp + 1)), i = (int)(e - (long)K * (nxp + 1));
        cx<double> v;
        v.re = 0.0;
        v.im = 0.0;
        if (i < mk && K < mz) v = kept[(long)j * mk * mz + (long)K * mk + i];
        double *p = fort + ((long)j * plane + e) * 2;
        p[0] = v.re;
        p[1] = v.im;
    }
}

}  // namespace

extern "C" {

int32_t dgpu_abi_version(void) { return DGPU_ABI_VERSION; }

const char *dgpu_build_info(void)
{
    return "libdiablo_gpu: DIABLO channel time-stepper, HIP kernels for gfx950 (FP64 path)";
}

const char *dgpu_last_error(const dgpu_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int32_t dgpu_create(const dgpu_params *p, dgpu_ctx **out)
{
    if (!p || !out) return DGPU_E_ARG;
    *out = nullptr;
    if (p->nx < 4 || p->nz < 4 || p->ny < 4 || (p->nx & 1) || (p->nz & 1)) return DGPU_E_ARG;
    if (p->n_th < 0 || p->n_th > DGPU_MAX_TH) return DGPU_E_ARG;  // the fused kernels carry one scalar
    if (p->precision != DGPU_FP64 && p->precision != DGPU_MIXED && p->precision != DGPU_MIXED_F64STATE)
        return DGPU_E_ARG;
    if (!p->gy) return DGPU_E_ARG;
    for (int code : {p->u_bc_ymin, p->w_bc_ymin, p->u_bc_ymax, p->w_bc_ymax})
        if (code != 0 && code != 1) return DGPU_E_ARG;
    for (int code : {p->v_bc_ymin, p->v_bc_ymax})
        if (code < 0 || code > 2) return DGPU_E_ARG;
    for (int n = 0; n < p->n_th; ++n)
        if ((p->th_bc_ymin[n] != 0 && p->th_bc_ymin[n] != 1) || (p->th_bc_ymax[n] != 0 && p->th_bc_ymax[n] != 1))
            return DGPU_E_ARG;
    dgpu_ctx *c = new (std::nothrow) dgpu_ctx();
    if (!c) return DGPU_E_NOMEM;
    c->p = *p;
    c->p.gy = c->p.gyf = c->p.dy = c->p.dyf = c->p.kx = c->p.kz = c->p.kx2 = c->p.kz2 = nullptr;
    c->mixed = (p->precision != DGPU_FP64);
    c->f32state = (p->precision == DGPU_MIXED);
    dgpu::Geo &g = c->g;
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz; g.nth = p->n_th;
    g.nkx = p->nx / 3; g.nkz = p->nz / 3; g.tnkz = 2 * g.nkz;
    g.mk = g.nkx + 1; g.mz = g.tnkz + 1;
    g.ps = (long)g.mk * g.mz;
    const int NY = g.ny, NXP = g.nx / 2;
    // INIT_CHAN_MPI at NPROCY = 1 (for/mpi.f:1080-1100)
    c->jstart = 2;
    c->jend = NY;
    for (int n = 0; n < g.nth; ++n) {
        c->jend_th[n] = NY;
        c->jstart_th[n] = (p->th_bc_ymin[n] == 1) ? 1 : 2;
    }
    // grid (for/hdf5.f:676-703, for/channel.f:1646-1652)
    c->gy.assign(NY + 2, 0.0);
    c->gyf.assign(NY + 2, 0.0);
    c->dy.assign(NY + 2, 0.0);
    c->dyf.assign(NY + 2, 0.0);
    for (int j = 1; j <= NY + 1; ++j) c->gy[j] = p->gy[j - 1];
    for (int j = 1; j <= NY; ++j) c->gyf[j] = 0.5 * (c->gy[j] + c->gy[j + 1]);
    c->gyf[0] = 2.0 * c->gyf[1] - c->gyf[2];
    c->gyf[NY + 1] = 2.0 * c->gyf[NY] - c->gyf[NY - 1];
    for (int j = 1; j <= NY + 1; ++j) c->dy[j] = c->gyf[j] - c->gyf[j - 1];
    for (int j = 1; j <= NY; ++j) c->dyf[j] = c->gy[j + 1] - c->gy[j];
    c->dyf[NY + 1] = c->dyf[NY];
    if (p->gyf) c->gyf.assign(p->gyf, p->gyf + NY + 2);
    if (p->dy) c->dy.assign(p->dy, p->dy + NY + 2);
    if (p->dyf) c->dyf.assign(p->dyf, p->dyf + NY + 2);
    // wavenumbers (for/fft.f:72-105)
    const double PI = 4.0 * std::atan(1.0);
    c->kx.assign(NXP + 1, 0.0);
    c->kz.assign(g.tnkz + 1, 0.0);
    for (int i = 0; i <= NXP - 1; ++i) c->kx[i] = ((double)i * (2.0 * PI)) / p->lx;
    for (int k = 0; k <= g.nkz; ++k) c->kz[k] = ((double)k * (2.0 * PI)) / p->lz;
    for (int k = 1; k <= g.nkz; ++k) c->kz[g.tnkz + 1 - k] = ((double)(-k) * (2.0 * PI)) / p->lz;
    if (p->kx) c->kx.assign(p->kx, p->kx + NXP + 1);
    if (p->kz) c->kz.assign(p->kz, p->kz + g.tnkz + 1);
    c->kx2.resize(NXP + 1);
    c->kz2.resize(g.tnkz + 1);
    for (int i = 0; i <= NXP; ++i) c->kx2[i] = c->kx[i] * c->kx[i];
    for (int k = 0; k <= g.tnkz; ++k) c->kz2[k] = c->kz[k] * c->kz[k];
    if (p->kx2) c->kx2.assign(p->kx2, p->kx2 + NXP + 1);
    if (p->kz2) c->kz2.assign(p->kz2, p->kz2 + g.tnkz + 1);

    if (!make_len(g.nx, c->fx.len) || !make_len(g.nz, c->fz.len)) {
        delete c;
        return DGPU_E_ARG;
    }
    {
        // in-place plans; DGPU_FFT_EX / DGPU_FFT_EZ ask for a number of values
        // per thread, DGPU_XCHUNK for the planes per block of the x kernel
        const char *ex = std::getenv("DGPU_FFT_EX"), *ez = std::getenv("DGPU_FFT_EZ");
        const char *xc = std::getenv("DGPU_XCHUNK");
        // FP64 kernels spill above 12 values per thread; the FP32 z kernels
        // are faster at 12 too (NZ = 384: four radix stages at twice the
        // resident columns beat three stages of 24 values)
        const int emax = 12;
        if (!dgpu::ip_plan_x(g.nx, ex ? std::atoi(ex) : 0, c->mixed ? 16 : 12, 64, c->ipx) ||
            !dgpu::ip_plan_z(g.nz, ez ? std::atoi(ez) : 0, emax, c->ipz)) {
            delete c;
            return DGPU_E_ARG;
        }
        // planes per block: what dgpu_autotune measures best at 384x257x384
        // (64 in FP64, 32 with the FP32 transforms: three full rounds of
        // resident blocks either way, at two and four blocks per CU)
        c->xchunk = c->mixed ? 32 : 64;
        if (xc && std::atoi(xc) > 0) c->xchunk = std::atoi(xc);
        // column stage: DGPU_COLSEG=0 selects the whole-column sweeps,
        // DGPU_COLSEG_NW the waves per block of the segmented kernels
        const char *cs = std::getenv("DGPU_COLSEG"), *cw = std::getenv("DGPU_COLSEG_NW");
        if (cw && (std::atoi(cw) == 4 || std::atoi(cw) == 8 || std::atoi(cw) == 16)) c->colnw = std::atoi(cw);
        if (cs && std::atoi(cs) == 0) c->colnw = 0;
        const char *cfz = std::getenv("DGPU_COLFUSE");
        if (cfz && (std::atoi(cfz) == 0 || std::atoi(cfz) == 1)) c->colfuse = std::atoi(cfz);
        const char *sw = std::getenv("DGPU_XSWZ");  // 0 turns the row swizzle off
        // FP32 transforms (mixed contexts) hold 8-byte values in the rows
        if (sw && std::atoi(sw) == 0) {
            c->ipx.swz = 0;
            c->ipx.swzs = 3;
        } else {
            dgpu::ip_pick_swizzle(c->ipx, c->mixed ? 2 : 4);
        }
        // the z kernels' column batches: W = the largest power of two with W*T <= 256
        int zw = 1;
        while (2 * zw * c->ipz.t <= 256) zw *= 2;
        const char *zs = std::getenv("DGPU_ZSWZ");  // 0 turns the column-batch swizzle off
        c->ipz.cswz = (zs && std::atoi(zs) == 0) ? 0 : dgpu::ip_pick_cswz(c->ipz, zw);
        c->ld = (g.mk + 7) / 8 * 8;
    }
    int rc = 0;
    auto bail = [&](int code) { rc = code; dgpu_destroy(c); return code; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(DGPU_E_HIP);
    // device vectors: dy, dyf (NY+2), kx, kx2 (NXP+1), kz, kz2 (TNKZ+1), gyf (NY+2)
    const long nv = 3L * (NY + 2) + 2L * (NXP + 1) + 2L * (g.tnkz + 1);
    if (hipMalloc(&c->d_vec, sizeof(double) * nv) != hipSuccess) return bail(DGPU_E_NOMEM);
    {
        std::vector<double> h;
        h.insert(h.end(), c->dy.begin(), c->dy.end());
        h.insert(h.end(), c->dyf.begin(), c->dyf.end());
        h.insert(h.end(), c->kx.begin(), c->kx.end());
        h.insert(h.end(), c->kx2.begin(), c->kx2.end());
        h.insert(h.end(), c->kz.begin(), c->kz.end());
        h.insert(h.end(), c->kz2.begin(), c->kz2.end());
        h.insert(h.end(), c->gyf.begin(), c->gyf.end());
        if (hipMemcpy(c->d_vec, h.data(), sizeof(double) * nv, hipMemcpyHostToDevice) != hipSuccess)
            return bail(DGPU_E_HIP);
        c->d_dy = c->d_vec;
        c->d_dyf = c->d_dy + (NY + 2);
        c->d_kx = c->d_dyf + (NY + 2);
        c->d_kx2 = c->d_kx + (NXP + 1);
        c->d_kz = c->d_kx2 + (NXP + 1);
        c->d_kz2 = c->d_kz + (g.tnkz + 1);
        c->d_gyf = c->d_kz2 + (g.tnkz + 1);
    }
    // the fixed array of dgpu_stats_chan_noalloc on this (the current) device;
    // on failure only that entry point is unavailable
    if (dgpu::noalloc_stats_scratch(&c->d_noalloc) != hipSuccess) {
        c->d_noalloc = nullptr;
        (void)hipGetLastError();
    }
    // twiddles
    {
        std::vector<cx<double>> twx, twz;
        twiddles(g.nx, twx);
        twiddles(g.nz, twz);
        if (hipMalloc(&c->d_tw, sizeof(cx<double>) * (g.nx + g.nz)) != hipSuccess) return bail(DGPU_E_NOMEM);
        if (hipMemcpy(c->d_tw, twx.data(), sizeof(cx<double>) * g.nx, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_tw + g.nx, twz.data(), sizeof(cx<double>) * g.nz, hipMemcpyHostToDevice) != hipSuccess)
            return bail(DGPU_E_HIP);
        c->fx.tw = c->d_tw;
        c->fz.tw = c->d_tw + g.nx;
        std::vector<cx<double>> ipx, ipz;
        dgpu::ip_twiddles(c->ipx, ipx);
        dgpu::ip_twiddles(c->ipz, ipz);
        ipx.insert(ipx.end(), ipz.begin(), ipz.end());
        if (hipMalloc(&c->d_iptw, sizeof(cx<double>) * ipx.size()) != hipSuccess) return bail(DGPU_E_NOMEM);
        if (hipMemcpy(c->d_iptw, ipx.data(), sizeof(cx<double>) * ipx.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(DGPU_E_HIP);
        if (c->mixed) {
            std::vector<cx<float>> tf(ipx.size());
            for (size_t j = 0; j < ipx.size(); ++j) { tf[j].re = (float)ipx[j].re; tf[j].im = (float)ipx[j].im; }
            if (hipMalloc(&c->d_iptwf, sizeof(cx<float>) * tf.size()) != hipSuccess) return bail(DGPU_E_NOMEM);
            if (hipMemcpy(c->d_iptwf, tf.data(), sizeof(cx<float>) * tf.size(), hipMemcpyHostToDevice) != hipSuccess)
                return bail(DGPU_E_HIP);
        }
    }
    // state: cu1 cu2 cu3 cth[n] | cp cf1 cf2 cf3 cfth[n] | nl[3+n]
    const int nin = 3 + g.nth;
    c->fs = (long)(NY + 2) * g.ps;
    c->nslot = nin + 4 + g.nth + nin;
    if (hipMalloc(&c->d_state, sizeof(cx<double>) * c->fs * c->nslot) != hipSuccess) return bail(DGPU_E_NOMEM);
    if (hipMemset(c->d_state, 0, sizeof(cx<double>) * c->fs * c->nslot) != hipSuccess) return bail(DGPU_E_HIP);
    {
        int s = 0;
        for (int f = 0; f < 3; ++f) c->cu[f] = c->d_state + (s++) * c->fs;
        for (int n = 0; n < g.nth; ++n) c->cth[n] = c->d_state + (s++) * c->fs;
        c->cp = c->d_state + (s++) * c->fs;
        for (int f = 0; f < 3; ++f) c->cf[f] = c->d_state + (s++) * c->fs;
        for (int n = 0; n < g.nth; ++n) c->cfth[n] = c->d_state + (s++) * c->fs;
        for (int f = 0; f < nin; ++f) c->nl[f] = c->d_state + (s++) * c->fs;
    }
    // REM_DIV: per-mode Poisson factorisation (time independent) and row constants
    if (hipMalloc(&c->d_pmult, sizeof(double) * c->fs) != hipSuccess ||
        hipMalloc(&c->d_pinv, sizeof(double) * c->fs) != hipSuccess ||
        hipMalloc(&c->d_rd, sizeof(double) * 4 * (NY + 2)) != hipSuccess)
        return bail(DGPU_E_NOMEM);
    {
        std::vector<double> rd(4 * (NY + 2), 0.0);
        for (int j = 1; j <= NY; ++j) {
            if (j <= NY - 1) rd[4 * j + 0] = 1. / (c->dy[j + 1] * c->dyf[j]);
            rd[4 * j + 1] = 1. / c->dyf[j];
            rd[4 * j + 2] = 1. / c->dy[j];
        }
        rd[4 * (NY + 1) + 2] = 1. / c->dy[NY + 1];
        if (hipMemcpy(c->d_rd, rd.data(), sizeof(double) * rd.size(), hipMemcpyHostToDevice) != hipSuccess)
            return bail(DGPU_E_HIP);
        dgpu::launch_pois_factor(c->stream, g, c->d_dy, c->d_dyf, c->d_kx2, c->d_kz2, c->d_pmult, c->d_pinv);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) return bail(DGPU_E_HIP);
    }
    if (hipMalloc(&c->d_nlptr, sizeof(cx<double> *) * (3 + DGPU_MAX_TH)) != hipSuccess) return bail(DGPU_E_NOMEM);
    if (hipMemcpy(c->d_nlptr, c->nl, sizeof(cx<double> *) * (3 + DGPU_MAX_TH), hipMemcpyHostToDevice) != hipSuccess)
        return bail(DGPU_E_HIP);
    // intermediates
    c->zfs = (long)(NY + 2) * g.nz * c->ld;
    if (hipMalloc(&c->d_zt, sizeof(cx<double>) * c->zfs * nin) != hipSuccess) return bail(DGPU_E_NOMEM);
    if (hipMalloc(&c->d_wt, sizeof(cx<double>) * c->zfs * 2 * nin) != hipSuccess) return bail(DGPU_E_NOMEM);
    if (c->mixed) {
        if (hipMalloc(&c->d_ztf, sizeof(cx<float>) * c->zfs * nin) != hipSuccess ||
            hipMalloc(&c->d_wtf, sizeof(cx<float>) * c->zfs * 2 * nin) != hipSuccess ||
            hipMalloc(&c->d_nlf, sizeof(cx<float>) * c->fs * nin) != hipSuccess ||
            hipMalloc(&c->d_nlptrf, sizeof(cx<float> *) * (3 + DGPU_MAX_TH)) != hipSuccess)
            return bail(DGPU_E_NOMEM);
        if (hipMemset(c->d_nlf, 0, sizeof(cx<float>) * c->fs * nin) != hipSuccess) return bail(DGPU_E_HIP);
        for (int f = 0; f < nin; ++f) c->nlf[f] = c->d_nlf + f * c->fs;
        if (hipMemcpy(c->d_nlptrf, c->nlf, sizeof(cx<float> *) * (3 + DGPU_MAX_TH), hipMemcpyHostToDevice) !=
            hipSuccess)
            return bail(DGPU_E_HIP);
    }
    if (c->f32state) {
        // FP32 working state, FP64 side columns of the low modes, FP32 factors
        c->nstate = 2 * nin + 1;
        std::vector<int> map;
        const int kl = 2;
        for (int K = 0; K <= g.tnkz; ++K) {
            const int kz = K <= g.nkz ? K : K - (g.tnkz + 1);
            if (std::abs(kz) > kl) continue;
            for (int I = 0; I <= std::min(kl, g.nkx); ++I) map.push_back(K * g.mk + I);
        }
        c->nside = (int)map.size();
        const long side_fs = (long)c->nside * (NY + 2);
        if (hipMalloc(&c->d_state32, sizeof(cx<float>) * c->fs * c->nstate) != hipSuccess ||
            hipMalloc(&c->d_side, sizeof(cx<double>) * side_fs * (c->nstate + nin)) != hipSuccess ||
            hipMalloc(&c->d_map, sizeof(int) * map.size()) != hipSuccess ||
            hipMalloc(&c->d_pmultf, sizeof(float) * c->fs) != hipSuccess ||
            hipMalloc(&c->d_pinvf, sizeof(float) * c->fs) != hipSuccess ||
            hipMalloc(&c->d_gs32, sizeof(cx<float>) * c->fs * nin) != hipSuccess)
            return bail(DGPU_E_NOMEM);
        if (hipStreamCreateWithFlags(&c->side_stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess)
            return bail(DGPU_E_HIP);
        if (hipMemcpy(c->d_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemset(c->d_side, 0, sizeof(cx<double>) * side_fs * (c->nstate + nin)) != hipSuccess)
            return bail(DGPU_E_HIP);
        dgpu::launch_convert_real(c->stream, c->d_pmult, c->d_pmultf, c->fs);
        dgpu::launch_convert_real(c->stream, c->d_pinv, c->d_pinvf, c->fs);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) return bail(DGPU_E_HIP);
    }
    // CN tables
    c->h_cn.assign(3L * nsys_of(c) * cn_block(c), 0.0);
    if (hipMalloc(&c->d_cn, sizeof(double) * c->h_cn.size()) != hipSuccess) return bail(DGPU_E_NOMEM);
    // BC codes for APPLY_BC_VEL
    {
        int codes[6] = {p->u_bc_ymin, p->v_bc_ymin, p->w_bc_ymin, p->u_bc_ymax, p->v_bc_ymax, p->w_bc_ymax};
        double c1[6] = {p->u_bc_ymin_c1, p->v_bc_ymin_c1, p->w_bc_ymin_c1,
                        p->u_bc_ymax_c1, p->v_bc_ymax_c1, p->w_bc_ymax_c1};
        if (hipMalloc(&c->d_codes, sizeof(codes)) != hipSuccess || hipMalloc(&c->d_c1, sizeof(c1)) != hipSuccess)
            return bail(DGPU_E_NOMEM);
        if (hipMemcpy(c->d_codes, codes, sizeof(codes), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(c->d_c1, c1, sizeof(c1), hipMemcpyHostToDevice) != hipSuccess)
            return bail(DGPU_E_HIP);
    }
    (void)rc;
    *out = c;
    return DGPU_OK;
}

// DGPU_MIXED: bring the FP32 working copy (and the FP64 low-mode columns)
// up to date with d_state, or d_state up to date with them.
static int to_work(dgpu_ctx *c)
{
    if (!c->f32state || !c->master_dirty) return DGPU_OK;
    const dgpu::Geo &g = c->g;
    dgpu::launch_convert<float, double>(c->stream, c->d_state, c->d_state32, c->fs * c->nstate);
    dgpu::launch_side_gather<double, double>(c->stream, c->d_state, c->fs, g.ps, c->d_side,
                                             (long)c->nside * (g.ny + 2), c->d_map, c->nside, g.ny + 2, c->nstate);
    if (int rc = check_launch(c, "to_work")) return rc;
    c->master_dirty = false;
    return DGPU_OK;
}

static int to_master(dgpu_ctx *c)
{
    if (!c->f32state || !c->work_dirty) return DGPU_OK;
    const dgpu::Geo &g = c->g;
    dgpu::launch_convert<double, float>(c->stream, c->d_state32, c->d_state, c->fs * c->nstate);
    dgpu::launch_side_scatter<double, double>(c->stream, c->d_side, (long)c->nside * (g.ny + 2), c->d_state, c->fs,
                                              g.ps, c->d_map, c->nside, g.ny + 2, c->nstate);
    if (int rc = check_launch(c, "to_master")) return rc;
    c->work_dirty = false;
    return DGPU_OK;
}

int32_t dgpu_destroy(dgpu_ctx *c)
{
    if (!c) return DGPU_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &t : c->pending) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    for (auto &ge : c->gexec)
        if (ge) (void)hipGraphExecDestroy(ge);
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    void *bufs[] = {c->d_vec, c->d_tw, c->d_state, c->d_pmult, c->d_pinv, c->d_rd, c->d_nlptr, c->d_zt, c->d_wt,
                    c->d_fa, c->d_fb, c->d_iptw, c->d_iptwf, c->d_ztf, c->d_wtf, c->d_nlf, c->d_nlptrf,
                    c->d_fort, c->d_cn, c->d_codes, c->d_c1, c->d_state32, c->d_side, c->d_map,
                    c->d_pmultf, c->d_pinvf, c->d_gs32, c->d_stats, c->d_vmax_part, c->d_vmax, c->d_file, c->d_les, c->d_zdft};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (c->side_stream) {
        (void)hipStreamSynchronize(c->side_stream);
        (void)hipStreamDestroy(c->side_stream);
    }
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DGPU_OK;
}

// The statistics kernel into the device buffer buf (DGPU_NSTAT * NY doubles),
// copied to the caller's profiles; UBULK on the host.  Allocates nothing.
static int32_t stats_through(dgpu_ctx *c, double *buf, double *profiles, double *ubulk)
{
    const dgpu::Geo &g = c->g;
    dgpu::launch_stats(c->stream, g, c->cu[0], c->cu[1], c->cu[2], c->cp, g.nth ? c->cth[0] : nullptr, c->d_dy,
                       c->d_dyf, c->d_gyf, c->d_kx, c->d_kz, buf);
    if (int rc = check_launch(c, "stats")) return rc;
    HIPCHK(c, hipMemcpyAsync(profiles, buf, sizeof(double) * DGPU_NSTAT * g.ny, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // INTEGRATE_Y_VAR (for/mpi.f:1676-1698) of ume
    const double *ume = profiles + (long)DGPU_ST_UME * g.ny;  // ume(J) at index J-1
    double res = 0.;
    for (int j = 2; j <= g.ny; j++) res = res + 0.5 * (ume[j - 1] + ume[j - 2]) * c->dy[j];
    *ubulk = res / c->p.ly;
    return DGPU_OK;
}

int32_t dgpu_stats_chan(dgpu_ctx *c, double *profiles, double *ubulk)
{
    if (!c || !profiles || !ubulk) return DGPU_E_ARG;
    if (int rc = to_master(c)) return rc;
    if (!c->d_stats) HIPCHK(c, hipMalloc(&c->d_stats, sizeof(double) * DGPU_NSTAT * c->g.ny));
    return stats_through(c, c->d_stats, profiles, ubulk);
}

int32_t dgpu_stats_chan_noalloc(dgpu_ctx *c, double *profiles, double *ubulk)
{
    if (!c || !profiles || !ubulk) return DGPU_E_ARG;
    if (c->g.ny > DGPU_NOALLOC_MAX_NY)
        return fail(c, DGPU_E_ARG, "dgpu_stats_chan_noalloc: NY = " + std::to_string(c->g.ny) +
                                       " exceeds DGPU_NOALLOC_MAX_NY = " + std::to_string(DGPU_NOALLOC_MAX_NY) +
                                       " (use dgpu_stats_chan)");
    if (!c->d_noalloc)
        return fail(c, DGPU_E_STATE, "dgpu_stats_chan_noalloc: the fixed device array has no address on this GPU");
    if (int rc = to_master(c)) return rc;
    return stats_through(c, c->d_noalloc, profiles, ubulk);
}


static int fort_to_physical(dgpu_ctx *c, double *buf, int j0, int j1);
static int fort_to_fourier(dgpu_ctx *c, double *buf, int j0, int j1);

// WriteHDF5 / ReadHDF5's per-field conversion (for/hdf5.f:172-268, 470-534).
static cx<double> *flow_field(dgpu_ctx *c, int32_t field)
{
    if (field >= DGPU_CU1 && field <= DGPU_CU3) return c->cu[field - DGPU_CU1];
    if (field == DGPU_CP) return c->cp;
    if (field >= DGPU_CTH0 && field < DGPU_CTH0 + c->g.nth) return c->cth[field - DGPU_CTH0];
    return nullptr;
}

static int ensure_file(dgpu_ctx *c)
{
    if (!c->d_file) HIPCHK(c, hipMalloc(&c->d_file, sizeof(double) * c->g.nx * c->g.ny * c->g.nz));
    return ensure_fort(c);
}

int32_t dgpu_get_flow_physical(dgpu_ctx *c, int32_t field, double *out)
{
    if (!c || !out) return DGPU_E_ARG;
    cx<double> *src = flow_field(c, field);
    if (!src) return fail(c, DGPU_E_ARG, "field must be CU1..CU3, CP or CTH");
    if (int rc = ensure_file(c)) return rc;
    if (int rc = to_master(c)) return rc;
    const dgpu::Geo &g = c->g;
    const long plane = (long)(g.nx / 2 + 1) * (g.nz + 2);
    hipLaunchKernelGGL(k_pack_fortran, dim3((unsigned)((plane + 255) / 256), g.ny + 2), dim3(256), 0, c->stream, src,
                       c->d_fort, g.nx / 2, g.nz, g.mk, g.mz);
    if (int rc = check_launch(c, "pack")) return rc;
    if (int rc = fort_to_physical(c, c->d_fort, 0, g.ny + 1)) return rc;
    dgpu::launch_to_file(c->stream, g, c->d_fort, c->d_file, field == DGPU_CU2);
    if (int rc = check_launch(c, "to_file")) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->d_file, sizeof(double) * g.nx * g.ny * g.nz, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_set_flow_physical(dgpu_ctx *c, int32_t field, const double *in)
{
    if (!c || !in) return DGPU_E_ARG;
    cx<double> *dst = flow_field(c, field);
    if (!dst) return fail(c, DGPU_E_ARG, "field must be CU1..CU3, CP or CTH");
    if (int rc = ensure_file(c)) return rc;
    if (int rc = to_master(c)) return rc;
    c->master_dirty = true;
    const dgpu::Geo &g = c->g;
    HIPCHK(c, hipMemcpyAsync(c->d_file, in, sizeof(double) * g.nx * g.ny * g.nz, hipMemcpyHostToDevice, c->stream));
    dgpu::launch_from_file(c->stream, g, c->d_file, c->d_fort, field == DGPU_CU2);
    if (int rc = check_launch(c, "from_file")) return rc;
    if (int rc = fort_to_fourier(c, c->d_fort, 0, g.ny + 1)) return rc;
    dgpu::launch_unpack_fortran(c->stream, g, c->d_fort, dst, 0, g.ny + 2);
    if (int rc = check_launch(c, "unpack")) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_courant_request(dgpu_ctx *c)
{
    if (!c) return DGPU_E_ARG;
    if (dgpu::vmax_rows_per_block(c->ipx) < 1) return fail(c, DGPU_E_ARG, "x transform too long for COURANT");
    if (!c->d_vmax) {
        HIPCHK(c, hipMalloc(&c->d_vmax_part, sizeof(double) * dgpu::vmax_part_count(c->g, c->ipx)));
        HIPCHK(c, hipMalloc(&c->d_vmax, sizeof(double) * 3 * c->g.ny));
    }
    c->courant_state = 1;
    return DGPU_OK;
}

// COURANT (for/courant.f:14-62) from the recorded maxima.  For each plane,
// cfl * DY(J) / max|U2| is the smallest of the reference's per-point
// quotients (division rounds monotonically), and likewise for U1 and U3 with
// DX = LX/NX and DZ = LZ/NZ (for/channel.f:1574-1583); a zero maximum gives
// an infinite quotient, which the minimum ignores, as the reference's does.
int32_t dgpu_courant(dgpu_ctx *c, double cfl, double *delta_t, double *h_bar)
{
    if (!c || !delta_t || !h_bar) return DGPU_E_ARG;
    if (c->courant_state != 2) return fail(c, DGPU_E_STATE, "no velocity maxima recorded since dgpu_courant_request");
    const int NY = c->g.ny;
    std::vector<double> m(3 * (size_t)NY);
    HIPCHK(c, hipMemcpyAsync(m.data(), c->d_vmax, sizeof(double) * m.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->courant_state = 0;
    const dgpu_params &p = c->p;
    const double NU = p.nu, PI = 4.0 * std::atan(1.0);
    const double dx = p.lx / c->g.nx, dz = p.lz / c->g.nz;
    double dt = 999.0;
    dt = std::fmin(dt, 0.5 * std::pow(std::fmin(dx, c->dy[1]), 2.0) / NU);
    for (int n = 0; n < c->g.nth; ++n) dt = std::fmin(dt, dt * NU / (NU / p.pr[n]));
    if (p.i_ro != 0.0) dt = std::fmin(dt, 2.0 * PI / std::fabs(p.i_ro) / 20.0);
    for (int n = 0; n < c->g.nth; ++n) {
        if (p.ri[n] != 0.0) {
            const double Nmax = std::sqrt(std::fabs(p.ri[n] * p.th_bc_ymin_c1[n]));
            dt = std::fmin(dt, (double)0.1f * 2.0 * PI / Nmax);  // the literal 0.1 is a default REAL
        }
    }
    for (int j = 1; j <= NY; ++j) {
        const double dt_x = cfl * dx / m[j - 1];
        const double dt_y = cfl * c->dy[j] / m[NY + j - 1];
        const double dt_z = cfl * dz / m[2 * NY + j - 1];
        dt = std::fmin(dt, std::fmin(dt_x, std::fmin(dt_y, dt_z)));
    }
    const double DELTA_T = dt <= 0.0 ? 0.0001 : (dt >= 999.0 ? 999.0 : dt);
    *delta_t = DELTA_T;
    h_bar[0] = DELTA_T * (8.0 / 15.0);
    h_bar[1] = DELTA_T * (2.0 / 15.0);
    h_bar[2] = DELTA_T * (5.0 / 15.0);
    return DGPU_OK;
}

int32_t dgpu_tke_chan(dgpu_ctx *c, double *epsilon)
{
    if (!c || !epsilon) return DGPU_E_ARG;
    if (int rc = to_master(c)) return rc;
    const dgpu::Geo &g = c->g;
    const int NY = g.ny;
    if (!c->d_stats) HIPCHK(c, hipMalloc(&c->d_stats, sizeof(double) * DGPU_NSTAT * NY));
    static_assert(DGPU_NSTAT >= 10, "the statistics buffer holds the 10 epsilon sums");
    dgpu::launch_tke(c->stream, g, c->cu[0], c->cu[1], c->cu[2], c->d_kx, c->d_kz, c->d_stats);
    if (int rc = check_launch(c, "tke")) return rc;
    std::vector<double> o(10 * (size_t)NY);
    HIPCHK(c, hipMemcpyAsync(o.data(), c->d_stats, sizeof(double) * o.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    auto P = [&](int q, int j) { return (j >= 1 && j <= NY) ? o[(size_t)q * NY + (j - 1)] : 0.0; };
    const std::vector<double> &DY = c->dy, &DYF = c->dyf, &GY = c->gy;
    for (int j = 1; j <= NY; ++j) {
        // x and z derivatives of U1, U3 on GY: the DYF-weighted average of
        // planes j and j-1 (plane 0's term is left out: the reference reads
        // an unset scratch plane there)
        auto gyw = [&](int q) { return (DYF[j - 1] * P(q, j) + DYF[j] * (j > 1 ? P(q, j - 1) : 0.0)) / (2.0 * DY[j]); };
        double dvdy;
        if (j >= 2 && j <= NY - 1) {
            const double d = GY[j + 1] - GY[j - 1];
            const double mean = (P(9, j + 1) - P(9, j)) / d;  // vme(j), vme(j-1) beside U2(j+1), U2(j-1)
            dvdy = P(8, j) / (d * d) + mean * mean;
        } else {
            dvdy = P(3, j);  // the reference's scratch still holds du/dz there
        }
        const double e = gyw(0) + P(1, j) + P(6, j) / (DY[j] * DY[j]) + gyw(2) + gyw(3) + dvdy +
                         P(7, j) / (DY[j] * DY[j]) + P(4, j) + gyw(5);
        epsilon[j - 1] = c->p.nu * e;
    }
    return DGPU_OK;
}

// ---------------------------------------------------------------------------
// z transforms as DFT-as-GEMM on the matrix cores (mixed precision)
// ---------------------------------------------------------------------------
static uint16_t bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static double bf16_value(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Real weights A (rows x cols, row-major) -> fragment-major split-BF16 halves:
// [mt][ks][lane][t] = A[mt*32 + lane%32][ks*16 + (lane/32)*8 + t], zero-padded.
static void zdft_fragments(const std::vector<double> &A, int rows, int cols, int nmt, int nks,
                           std::vector<uint16_t> &out)
{
    const size_t n = (size_t)nmt * nks * 64 * 8;
    out.assign(2 * n, 0);
    for (int mt = 0; mt < nmt; ++mt)
        for (int ks = 0; ks < nks; ++ks)
            for (int l = 0; l < 64; ++l)
                for (int t = 0; t < 8; ++t) {
                    const int m = mt * 32 + l % 32, k = ks * 16 + (l / 32) * 8 + t;
                    const double v = (m < rows && k < cols) ? A[(size_t)m * cols + k] : 0.0;
                    const uint16_t h = bf16_rne((float)v);
                    const size_t at = (((size_t)mt * nks + ks) * 64 + l) * 8 + t;
                    out[at] = h;
                    out[n + at] = bf16_rne((float)(v - bf16_value(h)));
                }
}

// One complex weight w at (row q, column p) of a complex matrix as the real
// 2x2 block [[Re w, -Im w], [Im w, Re w]].
static void put_cx(std::vector<double> &A, int cols, int q, int p, double re, double im)
{
    A[(size_t)(2 * q) * cols + 2 * p] = re;
    A[(size_t)(2 * q) * cols + 2 * p + 1] = -im;
    A[(size_t)(2 * q + 1) * cols + 2 * p] = im;
    A[(size_t)(2 * q + 1) * cols + 2 * p + 1] = re;
}

static int build_zdft(dgpu_ctx *c)
{
    const dgpu::Geo &g = c->g;
    const int NZ = g.nz, mz = g.mz;
    const double PI = 4.0 * std::atan(1.0);
    auto kint = [&](int s) { return s <= g.nkz ? s : s - (g.tnkz + 1); };  // signed wavenumber of slot s
    auto angle = [&](long kz) { return 2.0 * PI * (double)(((kz % NZ) + NZ) % NZ) / NZ; };
    // backward: Z[z] = sum_s exp(+i 2 pi k(s) z / NZ) X[s]
    const int ri = 2 * NZ, ci = 2 * mz;
    const int nmt_i = (ri + 31) / 32, nks_i = (ci + 15) / 16;
    std::vector<double> Ai((size_t)ri * ci, 0.0);
    for (int z = 0; z < NZ; ++z)
        for (int s = 0; s < mz; ++s) {
            const double th = angle((long)kint(s) * z);
            put_cx(Ai, ci, z, s, std::cos(th), std::sin(th));
        }
    // forward: NL[s] = sum_z exp(-i 2 pi k(s) z / NZ) / NZ (G[z] + i KZ(s) H[z])
    const int rf = 2 * mz, cf = 4 * NZ;
    const int nmt_f = (rf + 31) / 32, nks_f = cf / 16;
    std::vector<double> Af((size_t)rf * cf, 0.0);
    for (int s = 0; s < mz; ++s)
        for (int z = 0; z < NZ; ++z) {
            const double th = angle((long)kint(s) * z);
            const double fr = std::cos(th) / NZ, fi = -std::sin(th) / NZ;
            const double kz = c->kz[s];
            put_cx(Af, cf, s, z, fr, fi);
            put_cx(Af, cf, s, NZ + z, -kz * fi, kz * fr);
        }
    std::vector<uint16_t> fi_, ff_;
    zdft_fragments(Ai, ri, ci, nmt_i, nks_i, fi_);
    zdft_fragments(Af, rf, cf, nmt_f, nks_f, ff_);
    HIPCHK(c, hipMalloc(&c->d_zdft, sizeof(uint16_t) * (fi_.size() + ff_.size())));
    HIPCHK(c, hipMemcpy(c->d_zdft, fi_.data(), sizeof(uint16_t) * fi_.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_zdft + fi_.size(), ff_.data(), sizeof(uint16_t) * ff_.size(), hipMemcpyHostToDevice));
    const uint4 *bi = reinterpret_cast<const uint4 *>(c->d_zdft);
    const uint4 *bf = reinterpret_cast<const uint4 *>(c->d_zdft + fi_.size());
    c->zinv_tab = {bi, bi + fi_.size() / 16, nmt_i, nks_i};
    c->zfwd_tab = {bf, bf + ff_.size() / 16, nmt_f, nks_f};
    return DGPU_OK;
}

int32_t dgpu_set_zdft(dgpu_ctx *c, int32_t on)
{
    if (!c) return DGPU_E_ARG;
    if (on) {
        if (!c->mixed) return fail(c, DGPU_E_ARG, "the MFMA z transforms run in mixed-precision contexts");
        if (c->g.nz > 256 || c->g.nz % 16 != 0 || c->g.nx > 256)
            return fail(c, DGPU_E_ARG, "the MFMA z transforms need NX <= 256 and NZ <= 256, a multiple of 16");
        if (!c->d_zdft)
            if (int rc = build_zdft(c)) return rc;
    }
    c->zdft = on != 0;
    for (auto &ge : c->gexec)
        if (ge) {
            HIPCHK(c, hipGraphExecDestroy(ge));
            ge = nullptr;
        }
    return DGPU_OK;
}

static int fused_waves(const dgpu_ctx *c, int ny);
// the fused column kernel runs with the FP32 state or in FP64 (not with the
// FP64 state under FP32 transforms)
static bool fused_fits(const dgpu_ctx *c, int ny, int waves)
{
    if (c->f32state) return dgpu::colfused_supported<float>(ny, waves);
    return !c->mixed && dgpu::colfused_supported<double>(ny, waves);
}

// a shape some segmented kernel of the context's precision runs with
static bool column_shape_fits(const dgpu_ctx *c, int waves)
{
    if (waves == 0) return true;
    return c->f32state
               ? (dgpu::colseg_supported<float>(c->g.ny, waves) || dgpu::colfused_supported<float>(c->g.ny, waves))
               : dgpu::colseg_supported<double>(c->g.ny, waves);
}

int32_t dgpu_set_column_shape(dgpu_ctx *c, int32_t waves, int32_t fused)
{
    if (!c) return DGPU_E_ARG;
    if (waves != 0 && waves != 4 && waves != 8 && waves != 16)
        return fail(c, DGPU_E_ARG, "column shape: waves per block must be 0, 4, 8 or 16");
    if (fused < -1 || fused > 1) return fail(c, DGPU_E_ARG, "column shape: fused must be -1, 0 or 1");
    if (!column_shape_fits(c, waves))
        return fail(c, DGPU_E_ARG, "column shape: the segmented column kernels do not fit NY = " +
                                       std::to_string(c->g.ny) + " with " + std::to_string(waves) + " waves");
    if (fused == 1 && !(waves && fused_fits(c, c->g.ny, waves)))
        return fail(c, DGPU_E_ARG, "column shape: the fused column kernel needs the FP32 state (DGPU_MIXED) or "
                                   "the FP64 path, 4, 8 or 16 waves (4 or 8 in FP64) and NY <= 268");
    if (c->colnw == waves && c->colfuse == fused) return DGPU_OK;
    c->colnw = waves;
    c->colfuse = fused;
    for (auto &ge : c->gexec)
        if (ge) {
            HIPCHK(c, hipGraphExecDestroy(ge));
            ge = nullptr;
        }
    return DGPU_OK;
}

int32_t dgpu_get_column_shape(const dgpu_ctx *c, int32_t *waves, int32_t *fused)
{
    if (!c || !waves) return DGPU_E_ARG;
    *waves = column_shape_fits(c, c->colnw) ? c->colnw : 0;
    if (fused) *fused = fused_waves(c, c->g.ny) ? 1 : 0;
    return DGPU_OK;
}

// ---------------------------------------------------------------------------
// LES (for/les.f, for/channel.f:303-386, 628-659, 748-1003)
// ---------------------------------------------------------------------------
// Work planes, Fortran layout (NX+2) x (NZ+2) x (NY+2) doubles each:
//   0-7   S_ij (3-7, 1) and U1, U3 (0, 2) for the model; in the substep
//         physical U1, U2, U3, R1, R2, R3 (0-5) and Thomas scratch (6, 7)
//   8-12  the stress parts to transform; in the substep CS1 (8), TH (9), RTH
//         (10), Thomas scratch (8, 11, 12, with 14-16 once the SGS force is
//         spent)
//   13    NU_T (kept between substeps; KAPPA_T equals it)
//   14-17 the SGS force for CF1, CF2, CF3, CFTH (kept modes)
//   18    rows 1 and NY of CU1, CU3 before les_chan's BCs (kept modes)
// then DELTA_YF, DELTA_Y.
static const int kLesBufs = 19;

static size_t les_n(const dgpu_ctx *c)
{
    return (size_t)(c->g.nx + 2) * (c->g.nz + 2) * (c->g.ny + 2);
}

static double *les_buf(dgpu_ctx *c, int q) { return c->d_les + (size_t)q * les_n(c); }

static int ensure_les(dgpu_ctx *c)
{
    if (c->d_les) return DGPU_OK;
    const dgpu::Geo &g = c->g;
    const int NY = g.ny;
    const size_t n = les_n(c);
    HIPCHK(c, hipMalloc(&c->d_les, (kLesBufs * n + 2 * (size_t)(NY + 2)) * sizeof(double)));
    // NU_T's padding and the unset rows of the work planes stay zero
    HIPCHK(c, hipMemsetAsync(c->d_les, 0, kLesBufs * n * sizeof(double), c->stream));
    // filter widths with -2 C_SMAG^2 absorbed (for/les.f:200-223):
    // C_SMAG = 0.13, beta_sgs = 1, DELTA = (DX DY 2 DZ)^(1/3)
    const double C_SMAG = 0.13, beta_sgs = 1.0;
    const double dx = c->p.lx / g.nx, dz = c->p.lz / g.nz;
    std::vector<double> dl(2 * (size_t)(NY + 2), 0.0);
    for (int j = 1; j <= NY + 1; ++j) {
        dl[j] = -2.0 * std::pow(C_SMAG, 2.0) * std::pow(dx * beta_sgs * c->dyf[j] * 2.0 * dz * beta_sgs, 2.0 / 3.0);
        dl[NY + 2 + j] =
            -2.0 * std::pow(C_SMAG, 2.0) * std::pow(dx * beta_sgs * c->dy[j] * 2.0 * dz * beta_sgs, 2.0 / 3.0);
    }
    HIPCHK(c, hipMemcpyAsync(c->d_les + kLesBufs * n, dl.data(), dl.size() * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

// les_chan with LES_MODEL_TYPE = 1 (for/les.f:1-331) at NPROCY = 1: the
// velocity BCs and APPLY_BC_LES on the state, S_ij and U1, U3 to physical
// space, |S| and NU_T, tau_ij back to Fourier space and its divergence
// subtracted from f1-f3 (the reference's CF1-3), then the explicit cross
// terms NU_T dU/dy added to f2.  fresh: f1-f3 start from zero (the SGS
// force alone) and need no clear beforehand.
static int les_model(dgpu_ctx *c, cx<double> *f1, cx<double> *f2, cx<double> *f3, bool fresh)
{
    const dgpu::Geo &g = c->g;
    const int NY = g.ny;
    const size_t n = les_n(c);
    const double *delta_yf = c->d_les + kLesBufs * n, *delta_y = delta_yf + (NY + 2);
    dgpu::LesPtrs S, O;
    // S_ij in planes 3-7 and 1, U1 and U3 in planes 0 and 2 (where the
    // substep keeps its physical U1, U3)
    const int splane[8] = {3, 4, 5, 6, 7, 1, 0, 2};
    for (int q = 0; q < 8; ++q) S.p[q] = les_buf(c, splane[q]);
    for (int q = 0; q < 6; ++q) O.p[q] = les_buf(c, 8 + q);
    O.p[6] = O.p[7] = nullptr;
    dgpu::launch_apply_bc_vel(c->stream, g, c->cu[0], c->cu[1], c->cu[2], c->d_codes, c->d_c1, c->d_dy, c->d_dyf);
    if (int rc = check_launch(c, "apply_bc_vel")) return rc;
    dgpu::launch_les_bc(c->stream, g, c->cu[0], c->cu[2], c->p.u_bc_ymin, c->p.u_bc_ymax, c->p.w_bc_ymin,
                        c->p.w_bc_ymax);
    if (int rc = check_launch(c, "les_bc")) return rc;
    dgpu::launch_les_strain(c->stream, g, c->cu[0], c->cu[1], c->cu[2], c->d_kx, c->d_kz, c->d_dy, c->d_dyf, S);
    if (int rc = check_launch(c, "les_strain")) return rc;
    for (int q = 0; q < 8; ++q)
        if (int rc = fort_to_physical(c, S.p[q], 0, NY + 1)) return rc;
    dgpu::launch_les_nut(c->stream, g, S, O, c->cu[0], c->cu[2], c->d_dy, c->d_dyf, delta_y, c->jstart, c->jend);
    if (int rc = check_launch(c, "les_nut")) return rc;
    for (int q = 0; q < 5; ++q)
        if (int rc = fort_to_fourier(c, O.p[q], 0, NY + 1)) return rc;
    // CF1, CF3 on J1 = JSTART .. J2 = NY-1 (the RANKY = NPROCY-1 bounds, for/les.f:46-50)
    dgpu::launch_les_cf(c->stream, g, O, f1, f2, f3, c->d_kx, c->d_kz, c->d_dyf, delta_y, delta_yf, c->jstart,
                        NY - 1, fresh);
    if (int rc = check_launch(c, "les_cf")) return rc;
    for (int pass = 0; pass < 2; ++pass) {
        dgpu::launch_les_cross_src(c->stream, g, O.p[5], S.p[6 + pass], c->d_dy, O.p[0]);
        if (int rc = check_launch(c, "les_cross_src")) return rc;
        if (int rc = fort_to_fourier(c, O.p[0], 2, NY)) return rc;
        dgpu::launch_les_cross_add(c->stream, g, O.p[0], f2, c->d_kx, c->d_kz, pass);
        if (int rc = check_launch(c, "les_cross_add")) return rc;
    }
    return DGPU_OK;
}

int32_t dgpu_les_chan(dgpu_ctx *c, double *nu_t)
{
    if (!c) return DGPU_E_ARG;
    if (int rc = to_master(c)) return rc;
    c->master_dirty = true;
    if (int rc = ensure_les(c)) return rc;
    if (int rc = les_model(c, c->cf[0], c->cf[1], c->cf[2], false)) return rc;
    if (nu_t) HIPCHK(c, hipMemcpyAsync(nu_t, les_buf(c, 13), les_n(c) * sizeof(double), hipMemcpyDeviceToHost,
                                       c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_tke_les(dgpu_ctx *c, double *eps_sgs)
{
    if (!c || !eps_sgs) return DGPU_E_ARG;
    if (!c->les) return fail(c, DGPU_E_ARG, "tkebudget_chan_les needs LES on");
    if (int rc = to_master(c)) return rc;
    const dgpu::Geo &g = c->g;
    const int NY = g.ny;
    double *U[3] = {les_buf(c, 0), les_buf(c, 1), les_buf(c, 2)};
    for (int q = 0; q < 3; ++q) {
        dgpu::launch_les_pack(c->stream, g, c->cu[q], U[q], c->d_kx, c->d_kz, 0, 0, NY + 1);
        if (int rc = check_launch(c, "les_pack")) return rc;
        if (int rc = fort_to_physical(c, U[q], 0, NY + 1)) return rc;
    }
    double *out = les_buf(c, 6);
    dgpu::launch_les_eps(c->stream, g, U[0], U[1], U[2], les_buf(c, 13), c->d_dy, c->d_dyf, out);
    if (int rc = check_launch(c, "les_eps")) return rc;
    HIPCHK(c, hipMemcpyAsync(eps_sgs, out, sizeof(double) * NY, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_set_les(dgpu_ctx *c, int32_t les, int32_t model)
{
    if (!c) return DGPU_E_ARG;
    if (les && c->mixed) return fail(c, DGPU_E_ARG, "LES runs in FP64 contexts only");
    c->les = les != 0;
    c->les_model = c->les && model != 0;
    if (c->les) return ensure_les(c);
    return DGPU_OK;
}

int32_t dgpu_synchronize(dgpu_ctx *c)
{
    if (!c) return DGPU_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_set_field(dgpu_ctx *c, int32_t field, const double *a)
{
    if (!c || !a) return DGPU_E_ARG;
    cx<double> *dst = field_ptr(c, field);
    if (!dst) return fail(c, DGPU_E_ARG, "unknown field id");
    if (int rc = ensure_fort(c)) return rc;
    if (int rc = to_master(c)) return rc;
    c->master_dirty = true;
    const size_t n = (size_t)(c->g.nx + 2) * (c->g.nz + 2) * (c->g.ny + 2);
    HIPCHK(c, hipMemcpyAsync(c->d_fort, a, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    dgpu::launch_unpack_fortran(c->stream, c->g, c->d_fort, dst, 0, c->g.ny + 2);
    if (int rc = check_launch(c, "unpack")) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_get_field(dgpu_ctx *c, int32_t field, double *a)
{
    if (!c || !a) return DGPU_E_ARG;
    cx<double> *src = field_ptr(c, field);
    if (!src) return fail(c, DGPU_E_ARG, "unknown field id");
    if (int rc = ensure_fort(c)) return rc;
    if (int rc = to_master(c)) return rc;
    const dgpu::Geo &g = c->g;
    const long plane = (long)(g.nx / 2 + 1) * (g.nz + 2);
    hipLaunchKernelGGL(k_pack_fortran, dim3((unsigned)((plane + 255) / 256), g.ny + 2), dim3(256), 0, c->stream, src,
                       c->d_fort, g.nx / 2, g.nz, g.mk, g.mz);
    if (int rc = check_launch(c, "pack")) return rc;
    const size_t n = (size_t)(g.nx + 2) * (g.nz + 2) * (g.ny + 2);
    HIPCHK(c, hipMemcpyAsync(a, c->d_fort, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

extern "C++" {
template <typename NLT>
static void fill_ptrs(dgpu_ctx *c, dgpu::CnPtrs<double, NLT> &P, cx<NLT> *const *nlsrc)
{
    std::memset(&P, 0, sizeof(P));
    P.cu1 = c->cu[0];
    P.cu2 = c->cu[1];
    P.cu3 = c->cu[2];
    P.cp = c->cp;
    P.cf1 = c->cf[0];
    P.cf2 = c->cf[1];
    P.cf3 = c->cf[2];
    P.cth = c->g.nth ? c->cth[0] : nullptr;
    P.cfth = c->g.nth ? c->cfth[0] : nullptr;
    // zfwd_nl writes the x-, y-, z-momentum and scalar terms to slots 0..3;
    // the FP64 slots c->nl also serve as the solves' scratch.
    const int map[4] = {0, 2, 1, 3};  // SYS_U1, SYS_U3, SYS_U2, SYS_TH0 -> slot
    for (int q = 0; q < 3 + c->g.nth; ++q) {
        P.nl[q] = nlsrc[map[q]];
        P.gs[q] = c->nl[map[q]];
    }
}

// DGPU_MIXED: the FP32 working state and the FP32 solves' own scratch (the
// low-mode columns read the nonlinear terms at the same time).
static void fill_ptrs32(dgpu_ctx *c, dgpu::CnPtrs<float, float> &P)
{
    std::memset(&P, 0, sizeof(P));
    cx<float> *w = c->d_state32;
    const long fs = c->fs;
    const int nin = 3 + c->g.nth;
    P.cu1 = w;
    P.cu2 = w + fs;
    P.cu3 = w + 2 * fs;
    P.cth = c->g.nth ? w + 3 * fs : nullptr;
    P.cp = w + nin * fs;
    P.cf1 = w + (nin + 1) * fs;
    P.cf2 = w + (nin + 2) * fs;
    P.cf3 = w + (nin + 3) * fs;
    P.cfth = c->g.nth ? w + (nin + 4) * fs : nullptr;
    const int map[4] = {0, 2, 1, 3};
    for (int q = 0; q < nin; ++q) {
        P.nl[q] = c->nlf[map[q]];
        P.gs[q] = c->d_gs32 + map[q] * fs;  // field order: U1, U2, U3, TH
    }
}

// DGPU_MIXED: the FP64 side columns of the low modes (same slot order), the
// FP32 nonlinear terms read at the modes' own indices.
static void fill_ptrs_side(dgpu_ctx *c, dgpu::CnPtrs<double, float> &P)
{
    std::memset(&P, 0, sizeof(P));
    cx<double> *w = c->d_side;
    const long fs = (long)c->nside * (c->g.ny + 2);
    const int nin = 3 + c->g.nth;
    P.cu1 = w;
    P.cu2 = w + fs;
    P.cu3 = w + 2 * fs;
    P.cth = c->g.nth ? w + 3 * fs : nullptr;
    P.cp = w + nin * fs;
    P.cf1 = w + (nin + 1) * fs;
    P.cf2 = w + (nin + 2) * fs;
    P.cf3 = w + (nin + 3) * fs;
    P.cfth = c->g.nth ? w + (nin + 4) * fs : nullptr;
    const int map[4] = {0, 2, 1, 3};
    for (int q = 0; q < nin; ++q) {
        P.nl[q] = c->nlf[map[q]];
        P.gs[q] = w + (c->nstate + q) * fs;
    }
}
}  // extern "C++"

static void fill_rdargs(dgpu_ctx *c, dgpu::RdArgs &R, int update_cp, double temp4)
{
    std::memset(&R, 0, sizeof(R));
    R.g = c->g;
    R.update_cp = update_cp;
    R.jstart = c->jstart;
    R.jend = c->jend;
    R.inv_temp4 = 1.0 / temp4;
    R.pmult = c->d_pmult;
    R.pinv = c->d_pinv;
    R.rd = c->d_rd;
    R.kx = c->d_kx;
    R.kz = c->d_kz;
    R.pmultf = c->d_pmultf;
    R.pinvf = c->d_pinvf;
    R.map = nullptr;
    R.fac_ps = c->g.ps;
}

// Per-stage constants of the CN kernel (for/channel.f:77-81 and the forcing).
static void make_cn_args(dgpu_ctx *c, int rk_step, double time, double temp3, double temp4, double temp5,
                         dgpu::CnArgs &A)
{
    const dgpu::Geo &g = c->g;
    const int k = rk_step - 1;
    std::memset(&A, 0, sizeof(A));
    A.g = g;
    A.rk_gt1 = rk_step > 1;
    A.f_type = c->p.f_type;
    A.temp3 = temp3;
    A.temp4 = temp4;
    A.temp5 = temp5;
    A.nu = c->p.nu;
    A.nupr = g.nth ? c->p.nu / c->p.pr[0] : 0.0;
    A.beta = c->p.beta;
    A.i_ro = c->p.i_ro;
    A.force00 = (c->p.f_type == 2) ? (c->p.px0 + c->p.amp_omega0 * std::cos(c->p.omega0 * time)) : c->p.px0;
    A.rgx = g.nth ? c->p.ri[0] * c->p.grav_x : 0.0;
    A.rgz = g.nth ? c->p.ri[0] * c->p.grav_z : 0.0;
    A.drhodx = g.nth ? c->p.drhodx[0] : 0.0;
    A.drhodz = g.nth ? c->p.drhodz[0] : 0.0;
    A.tab = c->d_cn + (long)k * nsys_of(c) * cn_block(c);
    A.kx = c->d_kx;
    A.kz = c->d_kz;
    A.kx2 = c->d_kx2;
    A.kz2 = c->d_kz2;
    A.map = nullptr;
    A.nl_ps = g.ps;
}

// Enqueues the five kernels of one substep on the context's stream.
// COURANT's velocity maxima (for/channel.f:1037-1040): the velocity after the
// implicit solves and before the projection, taken to physical space in the
// precision of the run's transforms; planes 1..NY of CU1..CU3.
// vel: the velocity after the implicit solves, three consecutive fields
// (U1, U2, U3) with the state's field stride; null: the state's own CU1..3.
static int enqueue_vmax(dgpu_ctx *c, const void *vel = nullptr)
{
    const dgpu::Geo &g = c->g;
    const long zps = (long)g.nz * c->ld;
    if (c->mixed) {
        const cx<float> *twx = c->d_iptwf, *twz = c->d_iptwf + c->ipx.twn;
        if (c->f32state)
            dgpu::launch_zinv_ip<float, float>(c->stream, g, c->ipz, twz,
                                               vel ? (const cx<float> *)vel : c->d_state32, c->fs, g.ps, c->d_ztf,
                                               c->zfs, zps, c->ld, 1, g.ny, 3);
        else
            dgpu::launch_zinv_ip<float, double>(c->stream, g, c->ipz, twz, vel ? (const cx<double> *)vel : c->cu[0],
                                                c->fs, g.ps, c->d_ztf, c->zfs, zps, c->ld, 1, g.ny, 3);
        dgpu::launch_vmax_ip<float>(c->stream, g, c->ipx, twx, c->ld, c->d_ztf, c->zfs, c->d_vmax_part, c->d_vmax);
    } else {
        const cx<double> *twx = c->d_iptw, *twz = c->d_iptw + c->ipx.twn;
        dgpu::launch_zinv_ip<double, double>(c->stream, g, c->ipz, twz, vel ? (const cx<double> *)vel : c->cu[0],
                                             c->fs, g.ps, c->d_zt, c->zfs, zps, c->ld, 1, g.ny, 3);
        dgpu::launch_vmax_ip<double>(c->stream, g, c->ipx, twx, c->ld, c->d_zt, c->zfs, c->d_vmax_part, c->d_vmax);
    }
    if (int rc = check_launch(c, "courant maxima")) return rc;
    c->courant_state = 2;
    return DGPU_OK;
}

// Waves per block of the segmented column kernels (dgpu_colseg.hip) for a
// column of NY rows in FP32 or FP64, or 0 for the whole-column sweeps
// (dgpu_columns.hip): the context's choice where the column fits it.
static int seg_waves(const dgpu_ctx *c, int ny, bool fp32)
{
    if (!c->colnw) return 0;
    const bool fits = fp32 ? dgpu::colseg_supported<float>(ny, c->colnw) : dgpu::colseg_supported<double>(ny, c->colnw);
    return fits ? c->colnw : 0;
}
// The fused column kernel (dgpu_colseg.hip): the context's shape where the
// column fits it, else 0.
// Left to itself (colfuse -1) the fused kernel runs from NY = 160 up (six or
// more rows per segment): on shorter columns its four solves in a row are
// mostly barriers, and the separate kernels are faster.  In FP64 only at 8
// waves: its 252 VGPRs and 131 KB of LDS (NY = 257) leave one block per CU,
// and at 4 waves that is too few wavefronts (1.73 ms against 1.22 at 8 and
// 1.24 for the two kernels, 384 x 257 x 384).
static int fused_waves(const dgpu_ctx *c, int ny)
{
    if (!c->colnw) return 0;
    // tall columns; with the FP32 state also the short ones the fused kernel
    // cuts into 16 segments (NY <= 142): the 64-member ensemble 17.13k ->
    // 17.65k Mpts*substeps/s against the two kernels
    const bool auto_on = (ny >= 160 && (c->f32state || c->colnw == 8)) || (c->f32state && ny + 2 <= 144);
    const bool want = c->colfuse == 1 || (c->colfuse < 0 && auto_on);
    return (want && fused_fits(c, ny, c->colnw)) ? c->colnw : 0;
}
// the low-mode side columns (a few modes per system): the smallest blocks
static int side_waves(const dgpu_ctx *c, int ny)
{
    return (c->colnw && dgpu::colseg_supported<double>(ny, 4)) ? 4 : 0;
}

static int rk_enqueue(dgpu_ctx *c, int rk_step, double time, double temp1, double temp3, double temp4, double temp5)
{
    (void)temp1;
    const dgpu::Geo &g = c->g;
    const bool courant = rk_step == 3 && c->courant_state == 1;
    const int nin = 3 + g.nth;
    int jlo = 2;
    for (int n = 0; n < g.nth; ++n) jlo = std::min(jlo, c->jstart_th[n]);
    const long zps = (long)g.nz * c->ld;
    if (c->mixed) {
        const cx<float> *twx = c->d_iptwf, *twz = c->d_iptwf + c->ipx.twn;
        {
            TimerScope t(c, 0);
            if (c->zdft && c->f32state)
                dgpu::launch_zdft_inv<float>(c->stream, g, c->zinv_tab, c->d_state32, c->fs, g.ps, c->d_ztf, c->zfs,
                                             zps, c->ld, 0, g.ny + 2, nin);
            else if (c->zdft)
                dgpu::launch_zdft_inv<double>(c->stream, g, c->zinv_tab, c->cu[0], c->fs, g.ps, c->d_ztf, c->zfs,
                                              zps, c->ld, 0, g.ny + 2, nin);
            else if (c->f32state)
                dgpu::launch_zinv_ip<float, float>(c->stream, g, c->ipz, twz, c->d_state32, c->fs, g.ps, c->d_ztf,
                                                   c->zfs, zps, c->ld, 0, g.ny + 2, nin);
            else
                dgpu::launch_zinv_ip<float, double>(c->stream, g, c->ipz, twz, c->cu[0], c->fs, g.ps, c->d_ztf,
                                                    c->zfs, zps, c->ld, 0, g.ny + 2, nin);
        }
        if (int rc = check_launch(c, "zinv")) return rc;
        {
            TimerScope t(c, 1);
            dgpu::launch_xfused_ip<float>(c->stream, g, c->ipx, twx, c->ld, c->d_ztf, c->zfs, c->d_wtf, c->zfs,
                                          c->d_dy, c->d_dyf, c->d_kx, jlo, g.ny, c->xchunk);
        }
        if (int rc = check_launch(c, "xfused")) return rc;
        {
            TimerScope t(c, 2);
            if (c->zdft)
                dgpu::launch_zdft_fwd(c->stream, g, c->zfwd_tab, c->d_wtf, c->zfs, c->ld, c->d_nlptrf, jlo, g.ny, nin);
            else
                dgpu::launch_zfwd_nl_ip<float, float>(c->stream, g, c->ipz, twz, c->d_wtf, c->zfs, c->ld,
                                                      c->d_nlptrf, c->d_kz, jlo, g.ny, nin);
        }
        if (int rc = check_launch(c, "zfwd_nl")) return rc;
    } else {
        const cx<double> *twx = c->d_iptw, *twz = c->d_iptw + c->ipx.twn;
        {
            TimerScope t(c, 0);
            dgpu::launch_zinv_ip<double, double>(c->stream, g, c->ipz, twz, c->cu[0], c->fs, g.ps, c->d_zt, c->zfs,
                                                 zps, c->ld, 0, g.ny + 2, nin);
        }
        if (int rc = check_launch(c, "zinv")) return rc;
        {
            TimerScope t(c, 1);
            dgpu::launch_xfused_ip<double>(c->stream, g, c->ipx, twx, c->ld, c->d_zt, c->zfs, c->d_wt, c->zfs,
                                           c->d_dy, c->d_dyf, c->d_kx, jlo, g.ny, c->xchunk);
        }
        if (int rc = check_launch(c, "xfused")) return rc;
        {
            TimerScope t(c, 2);
            dgpu::launch_zfwd_nl_ip<double, double>(c->stream, g, c->ipz, twz, c->d_wt, c->zfs, c->ld, c->d_nlptr,
                                                    c->d_kz, jlo, g.ny, nin);
        }
        if (int rc = check_launch(c, "zfwd_nl")) return rc;
    }
    dgpu::CnArgs A;
    make_cn_args(c, rk_step, time, temp3, temp4, temp5, A);
    dgpu::RdArgs R;
    fill_rdargs(c, R, 1, temp4);
    // the segmented column kernels where the column fits them
    const int seg64 = seg_waves(c, g.ny, false), seg32 = seg_waves(c, g.ny, true), segs = side_waves(c, g.ny);
    const int fused32 = fused_waves(c, g.ny);
    if (c->f32state) {
        // FP32 solves on every mode, then the low modes again in FP64 on their
        // side columns, copied into the working state for the transforms.
        dgpu::CnPtrs<float, float> P;
        fill_ptrs32(c, P);
        dgpu::CnPtrs<double, float> Ps;
        fill_ptrs_side(c, Ps);
        dgpu::CnArgs As = A;
        As.g.ps = c->nside;
        As.map = c->d_map;
        dgpu::RdArgs Rs = R;
        Rs.g.ps = c->nside;
        Rs.map = c->d_map;
        const long side_fs = (long)c->nside * (g.ny + 2);
        // The low-mode columns (a few dozen lanes whose sweeps are latency
        // bound) run on the side stream while the FP32 solves run; their
        // results then replace the FP32 ones in the working state.
        HIPCHK(c, hipEventRecord(c->ev_fork, c->stream));
        HIPCHK(c, hipStreamWaitEvent(c->side_stream, c->ev_fork, 0));
        if (segs) {
            dgpu::launch_cn_seg<double, float>(c->side_stream, segs, As, Ps);
            dgpu::launch_remdiv_seg<double, float>(c->side_stream, segs, Rs, Ps);
        } else {
            dgpu::launch_cn_side<double, float>(c->side_stream, As, Ps);
            dgpu::launch_remdiv_side<double, float>(c->side_stream, Rs, Ps);
        }
        if (int rc = check_launch(c, "low-mode columns")) return rc;
        HIPCHK(c, hipEventRecord(c->ev_join, c->side_stream));
        if (fused32 && !courant) {
            // the whole column stage in one kernel (timed as the cn kind)
            TimerScope t(c, 3);
            dgpu::launch_colfused<float, float>(c->stream, fused32, A, R, P);
        } else {
            {
                TimerScope t(c, 3);
                if (seg32) dgpu::launch_cn_seg<float, float>(c->stream, seg32, A, P);
                else dgpu::launch_cn<float, float>(c->stream, A, P);
            }
            if (int rc = check_launch(c, "cn")) return rc;
            if (courant) {
                if (int rc = enqueue_vmax(c, seg32 ? (const void *)P.gs[dgpu::SYS_U1] : nullptr)) return rc;
            }
            {
                TimerScope t(c, 4);
                if (seg32) dgpu::launch_remdiv_seg<float, float>(c->stream, seg32, R, P);
                else dgpu::launch_remdiv<float, float>(c->stream, R, P);
            }
        }
        if (int rc = check_launch(c, "remdiv")) return rc;
        HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_join, 0));
        // CU1-3, CTH and CP of the low modes
        dgpu::launch_side_scatter<double, float>(c->stream, c->d_side, side_fs, c->d_state32, c->fs, g.ps, c->d_map,
                                                 c->nside, g.ny + 2, nin + 1);
    } else if (c->mixed) {
        dgpu::CnPtrs<double, float> P;
        fill_ptrs<float>(c, P, c->nlf);
        {
            TimerScope t(c, 3);
            if (seg64) dgpu::launch_cn_seg<double, float>(c->stream, seg64, A, P);
            else dgpu::launch_cn<double, float>(c->stream, A, P);
        }
        if (int rc = check_launch(c, "cn")) return rc;
        if (courant) {
            if (int rc = enqueue_vmax(c, seg64 ? (const void *)P.gs[dgpu::SYS_U1] : nullptr)) return rc;
        }
        {
            TimerScope t(c, 4);
            if (seg64) dgpu::launch_remdiv_seg<double, float>(c->stream, seg64, R, P);
            else dgpu::launch_remdiv<double, float>(c->stream, R, P);
        }
    } else {
        dgpu::CnPtrs<double, double> P;
        fill_ptrs<double>(c, P, c->nl);
        if (const int nf = fused_waves(c, g.ny); nf && !courant) {
            TimerScope t(c, 3);
            dgpu::launch_colfused<double, double>(c->stream, nf, A, R, P);
            return check_launch(c, "colfused");
        }
        {
            TimerScope t(c, 3);
            if (seg64) dgpu::launch_cn_seg<double, double>(c->stream, seg64, A, P);
            else dgpu::launch_cn<double, double>(c->stream, A, P);
        }
        if (int rc = check_launch(c, "cn")) return rc;
        if (courant) {
            if (int rc = enqueue_vmax(c, seg64 ? (const void *)P.gs[dgpu::SYS_U1] : nullptr)) return rc;
        }
        {
            TimerScope t(c, 4);
            if (seg64) dgpu::launch_remdiv_seg<double, double>(c->stream, seg64, R, P);
            else dgpu::launch_remdiv<double, double>(c->stream, R, P);
        }
    }
    if (int rc = check_launch(c, "remdiv")) return rc;
    return DGPU_OK;
}

// RK_CHAN_1 with LES = .TRUE. (for/channel.f:62-1073): the explicit terms of
// the fused path (with the SGS force when the model is called) give the
// right-hand sides; the Crank-Nicolson solves run in physical space with the
// eddy viscosity, then REM_DIV_CHAN and the CP update as in the DNS path.
static int les_substep(dgpu_ctx *c, int rk_step, double time, double temp1, double temp3, double temp4,
                       double temp5)
{
    const dgpu::Geo &g = c->g;
    const int NY = g.ny;
    const int nin = 3 + g.nth;
    const bool courant = rk_step == 3 && c->courant_state == 1;
    const double temp2 = temp4 / 2.0;  // TEMP2 = H_BAR(RK_STEP) / 2
    double *U[3] = {les_buf(c, 0), les_buf(c, 1), les_buf(c, 2)};
    double *Rb[3] = {les_buf(c, 3), les_buf(c, 4), les_buf(c, 5)};
    double *bm = les_buf(c, 6), *gw = les_buf(c, 7), *cs = les_buf(c, 8), *thp = les_buf(c, 9),
           *rthp = les_buf(c, 10), *nut = les_buf(c, 13);
    cx<double> *sgs[4];
    for (int q = 0; q < 4; ++q) sgs[q] = reinterpret_cast<cx<double> *>(les_buf(c, 14 + q));
    cx<double> *pre = reinterpret_cast<cx<double> *>(les_buf(c, 18));
    const size_t fsb = sizeof(cx<double>) * (size_t)c->fs;
    if (c->les_model) {
        // the model writes the whole of sgs[0..2]; the scalar's part is
        // accumulated over its own rows
        if (g.nth) HIPCHK(c, hipMemsetAsync(sgs[3], 0, fsb, c->stream));
        for (int r = 0; r < 4; ++r)
            HIPCHK(c, hipMemcpyAsync(pre + (size_t)r * g.ps, c->cu[r < 2 ? 0 : 2] + (size_t)(r % 2 ? NY : 1) * g.ps,
                                     sizeof(cx<double>) * g.ps, hipMemcpyDeviceToDevice, c->stream));
        if (int rc = les_model(c, sgs[0], sgs[1], sgs[2], true)) return rc;
        // APPLY_BC_VEL_PHYS_MPI (for/channel.f:309-311)
        dgpu::LesBc b;
        b.u_lo = c->p.u_bc_ymin; b.u_hi = c->p.u_bc_ymax; b.w_lo = c->p.w_bc_ymin; b.w_hi = c->p.w_bc_ymax;
        b.v_lo = c->p.v_bc_ymin; b.v_hi = c->p.v_bc_ymax;
        b.u_lo_c = c->p.u_bc_ymin_c1; b.u_hi_c = c->p.u_bc_ymax_c1; b.w_lo_c = c->p.w_bc_ymin_c1;
        b.w_hi_c = c->p.w_bc_ymax_c1; b.v_lo_c = c->p.v_bc_ymin_c1; b.v_hi_c = c->p.v_bc_ymax_c1;
        dgpu::launch_les_bc_phys(c->stream, g, c->cu[0], c->cu[1], c->cu[2], b, c->d_dy, c->d_dyf);
        if (int rc = check_launch(c, "les_bc_phys")) return rc;
        // and on the model's physical U1, U3 (planes 0, 2), as the reference
        // applies it to the arrays les_chan left in physical space
        dgpu::launch_les_bc_phys_planes(c->stream, g, U[0], U[2], b, c->d_dy);
        if (int rc = check_launch(c, "les_bc_phys_planes")) return rc;
        // KAPPA_T = NU_T and the scalar's horizontal eddy diffusion, explicit
        // (for/channel.f:313-375)
        for (int nth = 0; nth < g.nth; ++nth)
            for (int dir = 1; dir <= 2; ++dir) {
                dgpu::launch_les_pack(c->stream, g, c->cth[nth], cs, c->d_kx, c->d_kz, dir, 1, NY + 1);
                if (int rc = check_launch(c, "les_pack")) return rc;
                if (int rc = fort_to_physical(c, cs, 0, NY + 1)) return rc;
                dgpu::launch_les_mul(c->stream, g, cs, nut, 1, NY + 1);
                if (int rc = check_launch(c, "les_mul")) return rc;
                if (int rc = fort_to_fourier(c, cs, 0, NY + 1)) return rc;
                dgpu::launch_les_acc(c->stream, g, sgs[3], cs, c->d_kx, c->d_kz, dir, c->jstart_th[nth],
                                     c->jend_th[nth]);
                if (int rc = check_launch(c, "les_acc")) return rc;
            }
    }
    // nonlinear terms (the fused transforms of the DNS path, FP64)
    int jlo = 2;
    for (int q = 0; q < g.nth; ++q) jlo = std::min(jlo, c->jstart_th[q]);
    const long zps = (long)g.nz * c->ld;
    const cx<double> *twx = c->d_iptw, *twz = c->d_iptw + c->ipx.twn;
    dgpu::launch_zinv_ip<double, double>(c->stream, g, c->ipz, twz, c->cu[0], c->fs, g.ps, c->d_zt, c->zfs, zps,
                                         c->ld, 0, g.ny + 2, nin);
    if (int rc = check_launch(c, "zinv")) return rc;
    dgpu::launch_xfused_ip<double>(c->stream, g, c->ipx, twx, c->ld, c->d_zt, c->zfs, c->d_wt, c->zfs, c->d_dy,
                                   c->d_dyf, c->d_kx, jlo, g.ny, c->xchunk);
    if (int rc = check_launch(c, "xfused")) return rc;
    dgpu::launch_zfwd_nl_ip<double, double>(c->stream, g, c->ipz, twz, c->d_wt, c->zfs, c->ld, c->d_nlptr, c->d_kz,
                                            jlo, g.ny, nin);
    if (int rc = check_launch(c, "zfwd_nl")) return rc;
    // the SGS force joins the nonlinear terms (slots: x, y, z momentum, scalar)
    if (c->les_model)
        for (int q = 0; q < nin; ++q) {
            dgpu::launch_les_add(c->stream, c->nl[q], sgs[q], c->fs);
            if (int rc = check_launch(c, "les_add")) return rc;
        }
    // new CF and the right-hand sides (sweep A of the column kernel)
    dgpu::CnArgs A;
    make_cn_args(c, rk_step, time, temp3, temp4, temp5, A);
    A.rhs_only = 1;
    dgpu::CnPtrs<double, double> P;
    fill_ptrs<double>(c, P, c->nl);
    dgpu::launch_cn<double, double>(c->stream, A, P);
    if (int rc = check_launch(c, "cn(rhs)")) return rc;
    if (c->les_model) {
        dgpu::LesCfFix f;
        f.nu = c->p.nu; f.beta = c->p.beta; f.i_ro = c->p.i_ro;
        f.drhodx = g.nth ? c->p.drhodx[0] : 0.0; f.drhodz = g.nth ? c->p.drhodz[0] : 0.0;
        f.jstart = c->jstart; f.jend = c->jend;
        f.jstart_th = g.nth ? c->jstart_th[0] : 1; f.jend_th = g.nth ? c->jend_th[0] : 0;
        dgpu::launch_les_cf_rows(c->stream, g, pre, c->cu[0], c->cu[2], c->cf[0], c->cf[2],
                                 g.nth ? c->cfth[0] : nullptr, c->d_kx2, c->d_kz2, f);
        if (int rc = check_launch(c, "les_cf_rows")) return rc;
    }
    // physical right-hand sides and fields
    for (int q = 0; q < 3; ++q) {
        dgpu::launch_les_pack(c->stream, g, c->nl[q], Rb[q], c->d_kx, c->d_kz, 0, 0, NY + 1);
        if (int rc = check_launch(c, "les_pack")) return rc;
        if (int rc = fort_to_physical(c, Rb[q], 0, NY + 1)) return rc;
        if (c->les_model && q != 1) continue;  // the model's U1, U3 with their BCs
        dgpu::launch_les_pack(c->stream, g, c->cu[q], U[q], c->d_kx, c->d_kz, 0, 0, NY + 1);
        if (int rc = check_launch(c, "les_pack")) return rc;
        if (int rc = fort_to_physical(c, U[q], 0, NY + 1)) return rc;
    }
    if (g.nth) {
        dgpu::launch_les_pack(c->stream, g, c->nl[3], rthp, c->d_kx, c->d_kz, 0, 0, NY + 1);
        dgpu::launch_les_pack(c->stream, g, c->cth[0], thp, c->d_kx, c->d_kz, 0, 0, NY + 1);
        if (int rc = check_launch(c, "les_pack")) return rc;
        if (int rc = fort_to_physical(c, rthp, 0, NY + 1)) return rc;
        if (int rc = fort_to_physical(c, thp, 0, NY + 1)) return rc;
    }
    // implicit solves: TH, U2, U1, U3 (for/channel.f:762-1029), independent,
    // in one launch, each right-hand side with its explicit eddy-viscosity
    // half (for/channel.f:628-659, 748-760) added on the fly; scratch planes: 6, 7 (work planes of the model), 8, 11,
    // 12, and 14-16 (the SGS force, already added to the nonlinear terms)
    const double nuv = c->p.nu_v_scale;
    double *scratch[8] = {bm, gw, les_buf(c, 8), les_buf(c, 11), les_buf(c, 12), les_buf(c, 14), les_buf(c, 15),
                          les_buf(c, 16)};
    dgpu::LesCnSet set{};
    auto add = [&](int sys, int jl, int jh, int lo, int hi, double clo, double chi, double cmol, const double *rhs,
                   double *out) {
        dgpu::LesCn &a = set.a[set.n];
        a.sys = sys; a.ny = NY; a.jlo = jl; a.jhi = jh; a.clo = jl - 1; a.chi = jh + 1;
        if (sys == dgpu::LES_SYS_U2) { a.clo = 1; a.chi = NY + 1; }
        a.lo = lo; a.hi = hi; a.cmol = cmol; a.temp2 = temp2; a.clo_c = clo; a.chi_c = chi;
        set.rhs[set.n] = rhs;
        set.out[set.n] = out;
        set.bm[set.n] = scratch[2 * set.n];
        set.gw[set.n] = scratch[2 * set.n + 1];
        ++set.n;
    };
    if (g.nth)
        add(dgpu::LES_SYS_TH, c->jstart_th[0], c->jend_th[0], c->p.th_bc_ymin[0], c->p.th_bc_ymax[0],
            c->p.th_bc_ymin_c1[0], c->p.th_bc_ymax_c1[0], temp1 / c->p.pr[0] * nuv, rthp, thp);
    add(dgpu::LES_SYS_U2, 2, NY, c->p.v_bc_ymin, c->p.v_bc_ymax, c->p.v_bc_ymin_c1, c->p.v_bc_ymax_c1, temp1 * nuv,
        Rb[1], U[1]);
    add(dgpu::LES_SYS_U13, c->jstart, c->jend, c->p.u_bc_ymin, c->p.u_bc_ymax, c->p.u_bc_ymin_c1, c->p.u_bc_ymax_c1,
        temp1 * nuv, Rb[0], U[0]);
    add(dgpu::LES_SYS_U13, c->jstart, c->jend, c->p.w_bc_ymin, c->p.w_bc_ymax, c->p.w_bc_ymin_c1, c->p.w_bc_ymax_c1,
        temp1 * nuv, Rb[2], U[2]);
    dgpu::launch_les_cn(c->stream, g, set, nut, c->d_dy, c->d_dyf);
    if (int rc = check_launch(c, "les_cn")) return rc;
    // COURANT on the physical velocity after the solves (for/channel.f:1037-1040)
    if (courant) {
        dgpu::launch_les_vmax(c->stream, g, U[0], U[1], U[2], c->d_vmax);
        if (int rc = check_launch(c, "les_vmax")) return rc;
        c->courant_state = 2;
    }
    // back to Fourier space (for/channel.f:1043-1050)
    for (int q = 0; q < 3 + g.nth; ++q) {
        double *f = q < 3 ? U[q] : thp;
        if (int rc = fort_to_fourier(c, f, 0, NY + 1)) return rc;
        dgpu::launch_unpack_fortran(c->stream, g, f, q < 3 ? c->cu[q] : c->cth[0], 0, NY + 2);
        if (int rc = check_launch(c, "unpack")) return rc;
    }
    dgpu::RdArgs R;
    fill_rdargs(c, R, 1, temp4);
    dgpu::launch_remdiv<double, double>(c->stream, R, P);
    return check_launch(c, "remdiv");
}

int32_t dgpu_rk_chan_1(dgpu_ctx *c, int32_t rk_step, double time, const double *h_bar, const double *beta_bar,
                       const double *zeta_bar)
{
    if (!c || !h_bar || !beta_bar || !zeta_bar || rk_step < 1 || rk_step > 3) return DGPU_E_ARG;
    const int k = rk_step - 1;
    // for/channel.f:77-81
    const double temp1 = c->p.nu * h_bar[k] / 2.0;
    const double temp3 = zeta_bar[k] * h_bar[k];
    const double temp4 = h_bar[k];
    const double temp5 = beta_bar[k] * h_bar[k];
    if (int rc = ensure_cn(c, k, temp1, temp4)) return rc;
    if (int rc = to_work(c)) return rc;
    c->work_dirty = c->f32state;
    if (c->les) return les_substep(c, rk_step, time, temp1, temp3, temp4, temp5);
    // One captured graph per RK stage.  Not used while kernels are timed one
    // by one, or when the forcing depends on TIME (F_TYPE = 2).
    const bool graphable = c->use_graphs && !c->timing && c->p.f_type != 2 &&
                           !(rk_step == 3 && c->courant_state == 1);
    if (!graphable) return rk_enqueue(c, rk_step, time, temp1, temp3, temp4, temp5);
    const double key[4] = {temp1, temp3, temp4, temp5};
    if (!(c->gexec[k] && std::memcmp(key, c->gkey[k], sizeof(key)) == 0)) {
        if (c->gexec[k]) {
            HIPCHK(c, hipGraphExecDestroy(c->gexec[k]));
            c->gexec[k] = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int rc = rk_enqueue(c, rk_step, time, temp1, temp3, temp4, temp5);
        const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        HIPCHK(c, ec);
        const hipError_t ei = hipGraphInstantiate(&c->gexec[k], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        HIPCHK(c, ei);
        std::memcpy(c->gkey[k], key, sizeof(key));
    }
    HIPCHK(c, hipGraphLaunch(c->gexec[k], c->stream));
    return DGPU_OK;
}

// Launch-shape autotuner: times `steps` RK3 steps (after one untimed step)
// of this context's own state (with time step delta_t) for each candidate -- planes per block of the
// fused x kernel, and for mixed contexts with NX, NZ <= 256 the FFT or the
// MFMA z transforms -- and keeps the fastest.  The state is copied aside
// before and put back after, so tuning does not advance the run.
int32_t dgpu_autotune(dgpu_ctx *c, int32_t steps, double delta_t, int32_t *xchunk_out, int32_t *zdft_out,
                      double *ms_per_step)
{
    if (!c || steps < 1 || !(delta_t > 0.0)) return DGPU_E_ARG;
    if (c->courant_state != 0) return fail(c, DGPU_E_ARG, "a COURANT request is pending");
    if (int rc = to_master(c)) return rc;
    const size_t bytes = sizeof(cx<double>) * c->fs * c->nslot;
    void *backup = nullptr;
    HIPCHK(c, hipMalloc(&backup, bytes));
    auto restore = [&]() -> int {
        HIPCHK(c, hipMemcpyAsync(c->d_state, backup, bytes, hipMemcpyDeviceToDevice, c->stream));
        c->master_dirty = true;
        c->work_dirty = false;
        return DGPU_OK;
    };
    int rc = DGPU_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMemcpyAsync(backup, c->d_state, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fail(c, DGPU_E_HIP, "autotune set-up");
    const double dt = delta_t;
    const double h_bar[3] = {dt * (8.0 / 15.0), dt * (2.0 / 15.0), dt * (5.0 / 15.0)};
    const double beta_bar[3] = {1.0, 25.0 / 8.0, 9.0 / 4.0}, zeta_bar[3] = {0.0, -17.0 / 8.0, -5.0 / 4.0};
    const int xchunk0 = c->xchunk, colnw0 = c->colnw, colfuse0 = c->colfuse;
    const bool zdft0 = c->zdft;
    const bool try_zdft = c->mixed && !c->les && c->g.nz <= 256 && c->g.nz % 16 == 0 && c->g.nx <= 256;
    // ms per step of the current shape, from the saved state
    auto time_shape = [&](double &ms_out) -> int {
        for (auto &ge : c->gexec)
            if (ge) {
                (void)hipGraphExecDestroy(ge);
                ge = nullptr;
            }
        if (int r = restore()) return r;
        for (int k = 1; k <= 3; ++k)
            if (int r = dgpu_rk_chan_1(c, k, 0.0, h_bar, beta_bar, zeta_bar)) return r;
        if (hipEventRecord(e0, c->stream) != hipSuccess) return fail(c, DGPU_E_HIP, "autotune event");
        for (int n = 0; n < steps; ++n)
            for (int k = 1; k <= 3; ++k)
                if (int r = dgpu_rk_chan_1(c, k, 0.0, h_bar, beta_bar, zeta_bar)) return r;
        float ms = 0.f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return fail(c, DGPU_E_HIP, "autotune timing");
        ms_out = ms / steps;
        return DGPU_OK;
    };
    // x-kernel planes per block and the z transforms, then the column shape
    // (LES runs keep the column stage of the LES substep)
    const int chunks[] = {4, 8, 16, 32, 64};
    int best_x = xchunk0, best_c = colnw0, best_f = colfuse0;
    bool best_z = zdft0;
    double best = 1e300;
    for (int z = 0; z <= (try_zdft ? 1 : 0) && rc == DGPU_OK; ++z)
        for (int xc : chunks) {
            if (rc != DGPU_OK) break;
            if ((rc = dgpu_set_zdft(c, z))) break;
            c->xchunk = xc;
            double ms = 0.0;
            if ((rc = time_shape(ms))) break;
            if (ms < best) {
                best = ms;
                best_x = xc;
                best_z = z != 0;
            }
        }
    if (rc == DGPU_OK && !c->les && colnw0 != 0) {
        c->xchunk = best_x;
        rc = dgpu_set_zdft(c, best_z);
        // with the FP32 state and in FP64 also the fused column kernel on and off
        const bool f0 = fused_waves(c, c->g.ny) != 0;
        for (int fz = 0; fz <= ((c->f32state || !c->mixed) ? 1 : 0) && rc == DGPU_OK; ++fz)
            for (int nw : {4, 8, 16}) {
                const bool fz_on = fz ? !f0 : f0;
                if (rc != DGPU_OK || !column_shape_fits(c, nw)) continue;
                if (fz_on && !fused_fits(c, c->g.ny, nw)) continue;
                if (nw == colnw0 && fz_on == f0) continue;  // timed above
                c->colnw = nw;
                c->colfuse = fz_on ? 1 : 0;
                double ms = 0.0;
                if ((rc = time_shape(ms))) break;
                if (ms < best) {
                    best = ms;
                    best_c = nw;
                    best_f = c->colfuse;
                }
            }
    }
    // the chosen shape, and the state as it was
    c->xchunk = rc == DGPU_OK ? best_x : xchunk0;
    c->colnw = rc == DGPU_OK ? best_c : colnw0;
    c->colfuse = rc == DGPU_OK ? best_f : colfuse0;
    const int rz = dgpu_set_zdft(c, rc == DGPU_OK ? best_z : zdft0);
    if (rc == DGPU_OK) rc = rz;
    const int rr = restore();
    if (rc == DGPU_OK) rc = rr;
    if (hipStreamSynchronize(c->stream) != hipSuccess && rc == DGPU_OK) rc = fail(c, DGPU_E_HIP, "autotune sync");
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(backup);
    if (rc) return rc;
    if (xchunk_out) *xchunk_out = c->xchunk;
    if (zdft_out) *zdft_out = c->zdft ? 1 : 0;
    if (ms_per_step) *ms_per_step = best;
    return DGPU_OK;
}

int32_t dgpu_set_graphs(dgpu_ctx *c, int32_t enable)
{
    if (!c) return DGPU_E_ARG;
    c->use_graphs = enable != 0;
    return DGPU_OK;
}

int32_t dgpu_rem_div_chan(dgpu_ctx *c)
{
    if (!c) return DGPU_E_ARG;
    if (int rc = to_master(c)) return rc;
    c->master_dirty = true;
    dgpu::CnPtrs<double, double> P;
    fill_ptrs<double>(c, P, c->nl);
    dgpu::RdArgs R;
    fill_rdargs(c, R, 0, 1.0);
    dgpu::launch_remdiv<double, double>(c->stream, R, P);
    return check_launch(c, "rem_div");
}

int32_t dgpu_apply_bc_vel(dgpu_ctx *c)
{
    if (!c) return DGPU_E_ARG;
    if (int rc = to_master(c)) return rc;
    c->master_dirty = true;
    dgpu::launch_apply_bc_vel(c->stream, c->g, c->cu[0], c->cu[1], c->cu[2], c->d_codes, c->d_c1, c->d_dy, c->d_dyf);
    return check_launch(c, "apply_bc_vel");
}

int32_t dgpu_fft_xz_to_fourier(dgpu_ctx *c, double *v, int32_t jmin, int32_t jmax)
{
    if (!c || !v) return DGPU_E_ARG;
    const dgpu::Geo &g = c->g;
    if (jmin < 0 || jmax > g.ny + 1) return fail(c, DGPU_E_ARG, "plane range outside 0..NY+1");
    if (jmax < jmin) return DGPU_OK;
    if (int rc = ensure_fort(c)) return rc;
    const size_t plane = (size_t)(g.nx + 2) * (g.nz + 2);
    const int np = jmax - jmin + 1;
    HIPCHK(c, hipMemcpyAsync(c->d_fort + plane * jmin, v + plane * jmin, plane * np * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    dgpu::launch_x_r2c<double>(c->stream, g, c->fx, c->d_fort, c->d_zt, jmin, np);
    if (int rc = check_launch(c, "x_r2c")) return rc;
    dgpu::launch_zfwd_pack<double>(c->stream, g, c->fz, c->d_zt, c->d_fort, jmin, np);
    if (int rc = check_launch(c, "zfwd_pack")) return rc;
    HIPCHK(c, hipMemcpyAsync(v + plane * jmin, c->d_fort + plane * jmin, plane * np * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

int32_t dgpu_fft_xz_to_physical(dgpu_ctx *c, double *vv, int32_t jmin, int32_t jmax)
{
    if (!c || !vv) return DGPU_E_ARG;
    const dgpu::Geo &g = c->g;
    if (jmin < 0 || jmax > g.ny + 1) return fail(c, DGPU_E_ARG, "plane range outside 0..NY+1");
    if (jmax < jmin) return DGPU_OK;
    if (int rc = ensure_fort(c)) return rc;
    const size_t plane = (size_t)(g.nx + 2) * (g.nz + 2);
    const int np = jmax - jmin + 1;
    const int nxp = g.nx / 2;
    HIPCHK(c, hipMemcpyAsync(c->d_fort + plane * jmin, vv + plane * jmin, plane * np * sizeof(double),
                             hipMemcpyHostToDevice, c->stream));
    const cx<double> *in = reinterpret_cast<const cx<double> *>(c->d_fort);
    const long in_ps = (long)(nxp + 1) * (g.nz + 2);
    dgpu::launch_zinv<double, double>(c->stream, g, c->fz, in, 0, in_ps, nxp + 1, nxp, c->d_wt, 0, (long)g.nz * nxp,
                                      nxp, jmin, np, 1);
    if (int rc = check_launch(c, "zinv(host)")) return rc;
    dgpu::launch_x_c2r<double>(c->stream, g, c->fx, c->d_wt, nxp, c->d_fort, jmin, np);
    if (int rc = check_launch(c, "x_c2r")) return rc;
    HIPCHK(c, hipMemcpyAsync(vv + plane * jmin, c->d_fort + plane * jmin, plane * np * sizeof(double),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

// Device-side Fortran-layout plane transforms used by POISSON_P_CHAN.
static int fort_to_physical(dgpu_ctx *c, double *buf, int j0, int j1)
{
    const dgpu::Geo &g = c->g;
    const int nxp = g.nx / 2, np = j1 - j0 + 1;
    const cx<double> *in = reinterpret_cast<const cx<double> *>(buf);
    dgpu::launch_zinv<double, double>(c->stream, g, c->fz, in, 0, (long)(nxp + 1) * (g.nz + 2), nxp + 1, nxp,
                                      c->d_wt, 0, (long)g.nz * nxp, nxp, j0, np, 1);
    if (int rc = check_launch(c, "zinv(planes)")) return rc;
    dgpu::launch_x_c2r<double>(c->stream, g, c->fx, c->d_wt, nxp, buf, j0, np);
    return check_launch(c, "x_c2r(planes)");
}

static int fort_to_fourier(dgpu_ctx *c, double *buf, int j0, int j1)
{
    const dgpu::Geo &g = c->g;
    const int np = j1 - j0 + 1;
    dgpu::launch_x_r2c<double>(c->stream, g, c->fx, buf, c->d_zt, j0, np);
    if (int rc = check_launch(c, "x_r2c(planes)")) return rc;
    dgpu::launch_zfwd_pack<double>(c->stream, g, c->fz, c->d_zt, buf, j0, np);
    return check_launch(c, "zfwd_pack(planes)");
}

int32_t dgpu_poisson_p_chan(dgpu_ctx *c)
{
    if (!c) return DGPU_E_ARG;
    if (int rc = to_master(c)) return rc;
    c->master_dirty = true;
    const dgpu::Geo &g = c->g;
    if (int rc = ensure_fort(c)) return rc;
    const size_t n = (size_t)(g.nx + 2) * (g.nz + 2) * (g.ny + 2);
    if (!c->d_fa) HIPCHK(c, hipMalloc(&c->d_fa, n * sizeof(double)));
    if (!c->d_fb) HIPCHK(c, hipMalloc(&c->d_fb, n * sizeof(double)));
    double *fa = c->d_fa, *fb = c->d_fb, *fc = c->d_fort;
    cx<double> *cs = c->nl[1], *scr = c->nl[0];
    const int j0 = 2, j1 = g.ny;
    for (int group = 0; group < 4; ++group) {
        dgpu::launch_pp_spec(c->stream, group, g, c->cu[0], c->cu[1], c->cu[2], c->d_dyf, c->d_kx, c->d_kz, fa, fb,
                             group == 0 ? fc : nullptr);
        if (int rc = check_launch(c, "pp_spec")) return rc;
        if (int rc = fort_to_physical(c, fa, j0, j1)) return rc;
        if (int rc = fort_to_physical(c, fb, j0, j1)) return rc;
        if (group == 0) {
            if (int rc = fort_to_physical(c, fc, j0, j1)) return rc;
            dgpu::launch_pp_phys(c->stream, 0, g, fa, fb, fb, fc);
        } else {
            dgpu::launch_pp_phys(c->stream, 1, g, fa, fb, fb, nullptr);
        }
        if (int rc = check_launch(c, "pp_phys")) return rc;
        if (int rc = fort_to_fourier(c, fa, j0, j1)) return rc;
        if (group == 0) {
            if (int rc = fort_to_fourier(c, fb, j0, j1)) return rc;
            if (int rc = fort_to_fourier(c, fc, j0, j1)) return rc;
            dgpu::launch_pp_accum(c->stream, g, cs, fa, fb, fc, 1);
        } else {
            dgpu::launch_pp_accum(c->stream, g, cs, fa, nullptr, nullptr, 0);
        }
        if (int rc = check_launch(c, "pp_accum")) return rc;
    }
    dgpu::launch_pp_solve(c->stream, g, cs, g.nth ? c->cth[0] : nullptr, g.nth ? c->p.ri[0] : 0.0, c->d_gyf,
                          c->d_pmult, c->d_pinv, c->d_rd, scr, c->cp);
    if (int rc = check_launch(c, "pp_solve")) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return DGPU_OK;
}

// ---------------------------------------------------------------------------
// Ensemble: independent runs (their own grids and parameters) advanced
// together.  One graph per RK stage holds every member's substep as a
// parallel branch: fork event on the ensemble stream, each member's kernels on
// its own stream, join events back.
// ---------------------------------------------------------------------------
extern "C++" {
static int member_jlo(const dgpu_ctx *c)
{
    int jlo = 2;
    for (int n = 0; n < c->g.nth; ++n) jlo = std::min(jlo, c->jstart_th[n]);
    return jlo;
}

template <typename D>
static int upload(void **dptr, const std::vector<D> &h)
{
    if (!*dptr && hipMalloc(dptr, sizeof(D) * h.size()) != hipSuccess) return DGPU_E_NOMEM;
    if (hipMemcpy(*dptr, h.data(), sizeof(D) * h.size(), hipMemcpyHostToDevice) != hipSuccess) return DGPU_E_HIP;
    return DGPU_OK;
}

// Groups the members by (NX, NZ, N_TH) and writes the transform descriptors
// (they do not change between steps).  Members of different precisions keep
// the per-member path.
static int ens_build_groups(dgpu_ensemble *e)
{
    const int n = (int)e->m.size();
    for (int i = 1; i < n; ++i)
        if (e->m[i]->p.precision != e->m[0]->p.precision) return DGPU_OK;
    for (int i = 0; i < n; ++i) {
        const dgpu_ctx *c = e->m[i];
        bool placed = false;
        for (auto &gr : e->groups) {
            const dgpu_ctx *c0 = e->m[gr.idx[0]];
            if (c0->g.nx == c->g.nx && c0->g.nz == c->g.nz && c0->g.nth == c->g.nth) {
                gr.idx.push_back(i);
                placed = true;
                break;
            }
        }
        if (!placed) {
            e->groups.emplace_back();
            e->groups.back().idx.push_back(i);
        }
    }
    const bool mixed = e->m[0]->mixed, f32 = e->m[0]->f32state;
    for (auto &gr : e->groups) {
        std::vector<dgpu::XfDesc<double>> xf64;
        std::vector<dgpu::XfDesc<float>> xf32;
        std::vector<dgpu::ZinvDesc<double, double>> zi64;
        std::vector<dgpu::ZinvDesc<float, double>> zi32d;
        std::vector<dgpu::ZinvDesc<float, float>> zi32;
        std::vector<dgpu::ZfDesc<double, double>> zf64;
        std::vector<dgpu::ZfDesc<float, float>> zf32;
        std::vector<dgpu::ScatDesc<double, float>> sc;
        for (int i : gr.idx) {
            dgpu_ctx *c = e->m[i];
            const int jlo = member_jlo(c), ny = c->g.ny;
            gr.max_planes = std::max(gr.max_planes, ny + 2);
            gr.max_span = std::max(gr.max_span, ny - jlo + 1);
            gr.max_zf = std::max(gr.max_zf, ny - jlo + 1);
            if (!mixed) {
                zi64.push_back({c->cu[0], c->d_zt, c->fs, c->zfs, ny + 2});
                xf64.push_back({c->d_zt, c->d_wt, c->zfs, c->zfs, c->d_dy, c->d_dyf, c->d_kx, jlo, ny});
                zf64.push_back({c->d_wt, c->zfs, c->d_nlptr, c->d_kz, jlo, ny - jlo + 1});
            } else {
                if (f32) zi32.push_back({c->d_state32, c->d_ztf, c->fs, c->zfs, ny + 2});
                else zi32d.push_back({c->cu[0], c->d_ztf, c->fs, c->zfs, ny + 2});
                xf32.push_back({c->d_ztf, c->d_wtf, c->zfs, c->zfs, c->d_dy, c->d_dyf, c->d_kx, jlo, ny});
                zf32.push_back({c->d_wtf, c->zfs, c->d_nlptrf, c->d_kz, jlo, ny - jlo + 1});
                if (f32)
                    sc.push_back({c->d_side, c->d_state32, (long)c->nside * (ny + 2), c->fs, c->d_map, ny + 2});
            }
        }
        int rc = DGPU_OK;
        if (!mixed) {
            if (!rc) rc = upload(&gr.d_zinv, zi64);
            if (!rc) rc = upload(&gr.d_xf, xf64);
            if (!rc) rc = upload(&gr.d_zf, zf64);
        } else {
            if (!rc) rc = f32 ? upload(&gr.d_zinv, zi32) : upload(&gr.d_zinv, zi32d);
            if (!rc) rc = upload(&gr.d_xf, xf32);
            if (!rc) rc = upload(&gr.d_zf, zf32);
            if (!rc && f32) rc = upload(&gr.d_scat, sc);
        }
        if (rc) return rc;
    }
    if (f32) {
        if (hipStreamCreateWithFlags(&e->side, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&e->ev_f, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&e->ev_j, hipEventDisableTiming) != hipSuccess)
            return DGPU_E_HIP;
    }
    const char *es = std::getenv("DGPU_ENS_STREAMS");
    if (e->groups.size() > 1 && !(es && std::atoi(es) == 0)) {
        if (hipEventCreateWithFlags(&e->gfork, hipEventDisableTiming) != hipSuccess) return DGPU_E_HIP;
        for (auto &gr : e->groups)
            if (hipStreamCreateWithFlags(&gr.stream, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&gr.done, hipEventDisableTiming) != hipSuccess)
                return DGPU_E_HIP;
        e->group_streams = true;
    }
    e->batched = true;
    return DGPU_OK;
}

// Writes a group's column descriptors of one RK stage if they changed (the
// ensemble stream is idle first: a launched graph may be reading them).
template <typename D>
static int refresh(dgpu_ensemble *e, void **dptr, std::vector<char> &last, const std::vector<D> &h)
{
    const char *b = reinterpret_cast<const char *>(h.data());
    if (*dptr && last.size() == sizeof(D) * h.size() && std::memcmp(last.data(), b, last.size()) == 0)
        return DGPU_OK;
    if (hipStreamSynchronize(e->main) != hipSuccess) return DGPU_E_HIP;
    if (int rc = upload(dptr, h)) return rc;
    last.assign(b, b + sizeof(D) * h.size());
    return DGPU_OK;
}

static int ens_refresh_columns(dgpu_ensemble *e, int rk_step, const std::vector<double> &key)
{
    const int k = rk_step - 1;
    const bool mixed = e->m[0]->mixed, f32 = e->m[0]->f32state;
    for (auto &gr : e->groups) {
        std::vector<dgpu::ColDesc<double, double>> c64;
        std::vector<dgpu::ColDesc<double, float>> c64f, side;
        std::vector<dgpu::ColDesc<float, float>> c32;
        for (int i : gr.idx) {
            dgpu_ctx *c = e->m[i];
            const double *q = &key[4 * i];
            dgpu::CnArgs A;
            make_cn_args(c, rk_step, 0.0, q[1], q[2], q[3], A);
            dgpu::RdArgs R;
            fill_rdargs(c, R, 1, q[2]);
            if (!mixed) {
                dgpu::ColDesc<double, double> d;
                std::memset(&d, 0, sizeof(d));
                d.A = A;
                d.R = R;
                fill_ptrs<double>(c, d.P, c->nl);
                c64.push_back(d);
            } else if (!f32) {
                dgpu::ColDesc<double, float> d;
                std::memset(&d, 0, sizeof(d));
                d.A = A;
                d.R = R;
                fill_ptrs<float>(c, d.P, c->nlf);
                c64f.push_back(d);
            } else {
                dgpu::ColDesc<float, float> d;
                std::memset(&d, 0, sizeof(d));
                d.A = A;
                d.R = R;
                fill_ptrs32(c, d.P);
                c32.push_back(d);
                dgpu::ColDesc<double, float> ds;
                std::memset(&ds, 0, sizeof(ds));
                ds.A = A;
                ds.A.g.ps = c->nside;
                ds.A.map = c->d_map;
                ds.R = R;
                ds.R.g.ps = c->nside;
                ds.R.map = c->d_map;
                fill_ptrs_side(c, ds.P);
                side.push_back(ds);
            }
        }
        int rc = DGPU_OK;
        if (!mixed) rc = refresh(e, &gr.d_col[k], gr.col_host[k], c64);
        else if (!f32) rc = refresh(e, &gr.d_col[k], gr.col_host[k], c64f);
        else {
            rc = refresh(e, &gr.d_col[k], gr.col_host[k], c32);
            if (!rc) rc = refresh(e, &gr.d_cols[k], gr.cols_host[k], side);
        }
        if (rc) return rc;
    }
    return DGPU_OK;
}

// One substep of every member: per group, the transform kernels and the
// column kernels each as one launch over the group's members.
static int ens_enqueue_batched(dgpu_ensemble *e, int rk_step)
{
    const int k = rk_step - 1;
    const bool mixed = e->m[0]->mixed, f32 = e->m[0]->f32state;
    const bool par = e->group_streams;
    // with several groups each group's chain runs on its own stream (the
    // groups share nothing), so the small groups' launches fill the GPU
    // beside the large ones'
    if (par) {
        if (hipEventRecord(e->gfork, e->main) != hipSuccess) return DGPU_E_HIP;
        for (auto &gr : e->groups)
            if (hipStreamWaitEvent(gr.stream, e->gfork, 0) != hipSuccess) return DGPU_E_HIP;
    }
    auto gs = [&](const dgpu_ens_group &gr) { return par ? gr.stream : e->main; };
    for (auto &gr : e->groups) {
        dgpu_ctx *c0 = e->m[gr.idx[0]];
        const dgpu::Geo &g = c0->g;
        const int nm = (int)gr.idx.size(), nin = 3 + g.nth;
        hipStream_t s = gs(gr);
        if (!mixed) {
            const cx<double> *twx = c0->d_iptw, *twz = c0->d_iptw + c0->ipx.twn;
            dgpu::launch_zinv_ip_batch<double, double>(s, g, c0->ipz, twz,
                                                       (const dgpu::ZinvDesc<double, double> *)gr.d_zinv, nm,
                                                       gr.max_planes, c0->ld, nin);
            dgpu::launch_xfused_ip_batch<double>(s, g, c0->ipx, twx, (const dgpu::XfDesc<double> *)gr.d_xf, nm,
                                                 gr.max_span, c0->ld, c0->xchunk);
            dgpu::launch_zfwd_nl_ip_batch<double, double>(s, g, c0->ipz, twz,
                                                          (const dgpu::ZfDesc<double, double> *)gr.d_zf, nm,
                                                          gr.max_zf, c0->ld, nin);
        } else {
            const cx<float> *twx = c0->d_iptwf, *twz = c0->d_iptwf + c0->ipx.twn;
            if (f32)
                dgpu::launch_zinv_ip_batch<float, float>(s, g, c0->ipz, twz,
                                                         (const dgpu::ZinvDesc<float, float> *)gr.d_zinv, nm,
                                                         gr.max_planes, c0->ld, nin);
            else
                dgpu::launch_zinv_ip_batch<float, double>(s, g, c0->ipz, twz,
                                                          (const dgpu::ZinvDesc<float, double> *)gr.d_zinv, nm,
                                                          gr.max_planes, c0->ld, nin);
            dgpu::launch_xfused_ip_batch<float>(s, g, c0->ipx, twx, (const dgpu::XfDesc<float> *)gr.d_xf, nm,
                                                gr.max_span, c0->ld, c0->xchunk);
            dgpu::launch_zfwd_nl_ip_batch<float, float>(s, g, c0->ipz, twz,
                                                        (const dgpu::ZfDesc<float, float> *)gr.d_zf, nm, gr.max_zf,
                                                        c0->ld, nin);
        }
        if (hipGetLastError() != hipSuccess) return DGPU_E_HIP;
    }
    if (f32) {
        // the low-mode FP64 columns after the nonlinear terms: on each
        // group's stream, or on the side stream once every group has them
        if (!par && (hipEventRecord(e->ev_f, e->main) != hipSuccess ||
                     hipStreamWaitEvent(e->side, e->ev_f, 0) != hipSuccess))
            return DGPU_E_HIP;
        for (auto &gr : e->groups) {
            hipStream_t ss = par ? gr.stream : e->side;
            dgpu::Geo gsd = e->m[gr.idx[0]]->g;
            gsd.ps = e->m[gr.idx[0]]->nside;
            const int nm = (int)gr.idx.size();
            const int max_ny = gr.max_planes - 2;
            const auto *d = (const dgpu::ColDesc<double, float> *)gr.d_cols[k];
            if (const int nw = side_waves(e->m[gr.idx[0]], max_ny)) {
                dgpu::launch_cn_seg_batch<double, float>(ss, nw, gsd, max_ny, d, nm);
                dgpu::launch_remdiv_seg_batch<double, float>(ss, nw, gsd, max_ny, d, nm);
            } else {
                dgpu::launch_cn_side_batch<double, float>(ss, gsd, d, nm);
                dgpu::launch_remdiv_side_batch<double, float>(ss, gsd, d, nm);
            }
        }
        if (hipGetLastError() != hipSuccess) return DGPU_E_HIP;
        if (!par && hipEventRecord(e->ev_j, e->side) != hipSuccess) return DGPU_E_HIP;
    }
    for (auto &gr : e->groups) {
        const dgpu::Geo &g = e->m[gr.idx[0]]->g;
        const int nm = (int)gr.idx.size();
        const int max_ny = gr.max_planes - 2;
        hipStream_t s = gs(gr);
        if (!mixed) {
            const auto *d = (const dgpu::ColDesc<double, double> *)gr.d_col[k];
            if (const int nf = fused_waves(e->m[gr.idx[0]], max_ny)) {
                dgpu::launch_colfused_batch<double, double>(s, nf, g, max_ny, d, nm);
            } else if (const int nw = seg_waves(e->m[gr.idx[0]], max_ny, false)) {
                dgpu::launch_cn_seg_batch<double, double>(s, nw, g, max_ny, d, nm);
                dgpu::launch_remdiv_seg_batch<double, double>(s, nw, g, max_ny, d, nm);
            } else {
                dgpu::launch_cn_batch<double, double>(s, g, d, nm);
                dgpu::launch_remdiv_batch<double, double>(s, g, d, nm);
            }
        } else if (!f32) {
            const auto *d = (const dgpu::ColDesc<double, float> *)gr.d_col[k];
            if (const int nw = seg_waves(e->m[gr.idx[0]], max_ny, false)) {
                dgpu::launch_cn_seg_batch<double, float>(s, nw, g, max_ny, d, nm);
                dgpu::launch_remdiv_seg_batch<double, float>(s, nw, g, max_ny, d, nm);
            } else {
                dgpu::launch_cn_batch<double, float>(s, g, d, nm);
                dgpu::launch_remdiv_batch<double, float>(s, g, d, nm);
            }
        } else {
            const auto *d = (const dgpu::ColDesc<float, float> *)gr.d_col[k];
            if (const int nf = fused_waves(e->m[gr.idx[0]], max_ny)) {
                dgpu::launch_colfused_batch<float, float>(s, nf, g, max_ny, d, nm);
            } else if (const int nw = seg_waves(e->m[gr.idx[0]], max_ny, true)) {
                dgpu::launch_cn_seg_batch<float, float>(s, nw, g, max_ny, d, nm);
                dgpu::launch_remdiv_seg_batch<float, float>(s, nw, g, max_ny, d, nm);
            } else {
                dgpu::launch_cn_batch<float, float>(s, g, d, nm);
                dgpu::launch_remdiv_batch<float, float>(s, g, d, nm);
            }
        }
        if (hipGetLastError() != hipSuccess) return DGPU_E_HIP;
    }
    if (f32) {
        // CU1-3, CTH and CP of the low modes into the working state, after
        // both column solves
        if (!par && hipStreamWaitEvent(e->main, e->ev_j, 0) != hipSuccess) return DGPU_E_HIP;
        for (auto &gr : e->groups) {
            const dgpu_ctx *c0 = e->m[gr.idx[0]];
            dgpu::launch_side_scatter_batch<double, float>(gs(gr), (const dgpu::ScatDesc<double, float> *)gr.d_scat,
                                                           (int)gr.idx.size(), c0->nside, c0->g.ps, gr.max_planes,
                                                           4 + c0->g.nth);
        }
        if (hipGetLastError() != hipSuccess) return DGPU_E_HIP;
    }
    if (par) {
        for (auto &gr : e->groups)
            if (hipEventRecord(gr.done, gr.stream) != hipSuccess || hipStreamWaitEvent(e->main, gr.done, 0) != hipSuccess)
                return DGPU_E_HIP;
    }
    return DGPU_OK;
}

}  // extern "C++"

int32_t dgpu_ensemble_destroy(dgpu_ensemble *e)
{
    if (!e) return DGPU_OK;
    if (e->main) (void)hipStreamSynchronize(e->main);
    if (e->side) {
        (void)hipStreamSynchronize(e->side);
        (void)hipStreamDestroy(e->side);
    }
    if (e->ev_f) (void)hipEventDestroy(e->ev_f);
    if (e->ev_j) (void)hipEventDestroy(e->ev_j);
    for (auto &gr : e->groups) {
        if (gr.stream) {
            (void)hipStreamSynchronize(gr.stream);
            (void)hipStreamDestroy(gr.stream);
        }
        if (gr.done) (void)hipEventDestroy(gr.done);
        void *bufs[] = {gr.d_zinv, gr.d_xf, gr.d_zf, gr.d_scat, gr.d_col[0], gr.d_col[1], gr.d_col[2],
                        gr.d_cols[0], gr.d_cols[1], gr.d_cols[2]};
        for (void *b : bufs)
            if (b) (void)hipFree(b);
    }
    for (auto &ge : e->gexec)
        if (ge) (void)hipGraphExecDestroy(ge);
    for (auto *c : e->m) dgpu_destroy(c);
    for (auto ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->gfork) (void)hipEventDestroy(e->gfork);
    if (e->main) (void)hipStreamDestroy(e->main);
    delete e;
    return DGPU_OK;
}

int32_t dgpu_ensemble_create(const dgpu_params *members, int32_t n, dgpu_ensemble **out)
{
    if (!members || n < 1 || !out) return DGPU_E_ARG;
    *out = nullptr;
    dgpu_ensemble *e = new (std::nothrow) dgpu_ensemble();
    if (!e) return DGPU_E_NOMEM;
    for (int i = 0; i < n; ++i) {
        dgpu_ctx *c = nullptr;
        const int rc = dgpu_create(&members[i], &c);
        if (rc) {
            dgpu_ensemble_destroy(e);
            return rc;
        }
        e->m.push_back(c);
    }
    if (hipStreamCreateWithFlags(&e->main, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
        dgpu_ensemble_destroy(e);
        return DGPU_E_HIP;
    }
    for (dgpu_ctx *c : e->m) c->ens_stream = e->main;
    e->join.assign(n, nullptr);
    for (int i = 0; i < n; ++i)
        if (hipEventCreateWithFlags(&e->join[i], hipEventDisableTiming) != hipSuccess) {
            dgpu_ensemble_destroy(e);
            return DGPU_E_HIP;
        }
    if (int rc = ens_build_groups(e)) {
        dgpu_ensemble_destroy(e);
        return rc;
    }
    *out = e;
    return DGPU_OK;
}

int32_t dgpu_ensemble_size(const dgpu_ensemble *e) { return e ? (int32_t)e->m.size() : 0; }

dgpu_ctx *dgpu_ensemble_member(dgpu_ensemble *e, int32_t i)
{
    if (!e || i < 0 || i >= (int32_t)e->m.size()) return nullptr;
    return e->m[i];
}

const char *dgpu_ensemble_last_error(const dgpu_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

int32_t dgpu_ensemble_set_graphs(dgpu_ensemble *e, int32_t enable)
{
    if (!e) return DGPU_E_ARG;
    e->use_graphs = enable != 0;
    return DGPU_OK;
}

int32_t dgpu_ensemble_rk_chan_1(dgpu_ensemble *e, int32_t rk_step, const double *h_bar, const double *beta_bar,
                                const double *zeta_bar)
{
    if (!e || !h_bar || !beta_bar || !zeta_bar || rk_step < 1 || rk_step > 3) return DGPU_E_ARG;
    const int k = rk_step - 1;
    const int n = (int)e->m.size();
    std::vector<double> key;
    bool graphable = e->use_graphs;
    // a member's COURANT request (dgpu_courant_request) is served at RK
    // stage 3 by the per-member launches, which record the velocity maxima;
    // that call runs without graphs and without the batched launches
    bool courant = false;
    for (int i = 0; i < n; ++i) courant = courant || (rk_step == 3 && e->m[i]->courant_state == 1);
    if (courant) graphable = false;
    const bool batched = e->batched && !courant;
    for (int i = 0; i < n; ++i) {
        dgpu_ctx *c = e->m[i];
        const double h = h_bar[3 * i + k];
        const double temp1 = c->p.nu * h / 2.0, temp3 = zeta_bar[3 * i + k] * h, temp4 = h;
        const double temp5 = beta_bar[3 * i + k] * h;
        if (int rc = ensure_cn(c, k, temp1, temp4)) {
            e->err = c->err;
            return rc;
        }
        key.insert(key.end(), {temp1, temp3, temp4, temp5});
        if (c->p.f_type == 2 || c->timing) graphable = false;
        if (c->f32state && c->master_dirty) {
            // the graph runs on the ensemble stream: finish the copy first
            if (int rc = to_work(c)) {
                e->err = c->err;
                return rc;
            }
            if (hipStreamSynchronize(c->stream) != hipSuccess) return DGPU_E_HIP;
        }
        c->work_dirty = c->f32state;
    }
    if (batched) {
        if (int rc = ens_refresh_columns(e, rk_step, key)) {
            e->err = "batched column descriptors";
            return rc;
        }
    } else if (e->batched) {
        // earlier batched launches on the ensemble stream finish before the
        // members' own streams take over
        if (hipStreamSynchronize(e->main) != hipSuccess) return DGPU_E_HIP;
    }
    auto enqueue_all = [&](bool joined) -> int {
        if (batched) {
            // one launch per kernel kind and member group on the ensemble stream
            if (!joined) {
                for (int i = 0; i < n; ++i)
                    if (hipStreamSynchronize(e->m[i]->stream) != hipSuccess) return DGPU_E_HIP;
            }
            const int rc = ens_enqueue_batched(e, rk_step);
            if (rc) e->err = "batched launch";
            if (!joined && !rc && hipStreamSynchronize(e->main) != hipSuccess) return DGPU_E_HIP;
            return rc;
        }
        for (int i = 0; i < n; ++i) {
            dgpu_ctx *c = e->m[i];
            if (joined && hipStreamWaitEvent(c->stream, e->fork, 0) != hipSuccess) return DGPU_E_HIP;
            const double *q = &key[4 * i];
            if (int rc = rk_enqueue(c, rk_step, 0.0, q[0], q[1], q[2], q[3])) {
                e->err = c->err;
                return rc;
            }
            if (joined) {
                if (hipEventRecord(e->join[i], c->stream) != hipSuccess) return DGPU_E_HIP;
                if (hipStreamWaitEvent(e->main, e->join[i], 0) != hipSuccess) return DGPU_E_HIP;
            }
        }
        return DGPU_OK;
    };
    if (!graphable) return enqueue_all(false);
    std::vector<int> shape;
    for (const dgpu_ctx *c : e->m) shape.insert(shape.end(), {c->xchunk, c->colnw, c->colfuse, c->zdft ? 1 : 0});
    if (!(e->gexec[k] && e->gkey[k] == key && e->gshape[k] == shape)) {
        if (e->gexec[k]) {
            (void)hipGraphExecDestroy(e->gexec[k]);
            e->gexec[k] = nullptr;
        }
        if (hipStreamBeginCapture(e->main, hipStreamCaptureModeThreadLocal) != hipSuccess) return DGPU_E_HIP;
        hipGraph_t graph = nullptr;
        int rc = DGPU_OK;
        if (hipEventRecord(e->fork, e->main) != hipSuccess) rc = DGPU_E_HIP;
        if (!rc) rc = enqueue_all(true);
        const hipError_t ec = hipStreamEndCapture(e->main, &graph);
        if (rc || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (!rc) e->err = std::string("graph capture: ") + hipGetErrorString(ec);
            return rc ? rc : DGPU_E_HIP;
        }
        const hipError_t ei = hipGraphInstantiate(&e->gexec[k], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            e->err = std::string("graph instantiate: ") + hipGetErrorString(ei);
            return DGPU_E_HIP;
        }
        e->gkey[k] = key;
        e->gshape[k] = shape;
    }
    if (hipGraphLaunch(e->gexec[k], e->main) != hipSuccess) return DGPU_E_HIP;
    return DGPU_OK;
}

// Ensemble launch shapes: every member's x-kernel planes per block (the
// batched x launch of a group takes its first member's) and column shape,
// timed over the ensemble's own states, which are put back afterwards.
int32_t dgpu_ensemble_autotune(dgpu_ensemble *e, int32_t steps, const double *h_bar, const double *beta_bar,
                               const double *zeta_bar, int32_t *xchunk_out, int32_t *waves_out, double *ms_per_step)
{
    if (!e || steps < 1 || !h_bar || !beta_bar || !zeta_bar) return DGPU_E_ARG;
    const int n = (int)e->m.size();
    for (const dgpu_ctx *c : e->m)
        if (c->courant_state != 0) {
            e->err = "ensemble autotune: a member has a COURANT request pending";
            return DGPU_E_ARG;
        }
    if (int rc = dgpu_ensemble_synchronize(e)) return rc;
    std::vector<void *> backup(n, nullptr);
    auto free_all = [&]() {
        for (void *b : backup)
            if (b) (void)hipFree(b);
    };
    for (int i = 0; i < n; ++i) {
        dgpu_ctx *c = e->m[i];
        const size_t bytes = sizeof(cx<double>) * c->fs * c->nslot;
        int rc = to_master(c);
        if (!rc && hipMalloc(&backup[i], bytes) != hipSuccess) rc = DGPU_E_NOMEM;
        if (!rc && (hipMemcpyAsync(backup[i], c->d_state, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
                    hipStreamSynchronize(c->stream) != hipSuccess))
            rc = DGPU_E_HIP;
        if (rc) {
            e->err = "ensemble autotune: state backup";
            free_all();
            return rc;
        }
    }
    auto restore = [&]() -> int {
        for (int i = 0; i < n; ++i) {
            dgpu_ctx *c = e->m[i];
            const size_t bytes = sizeof(cx<double>) * c->fs * c->nslot;
            if (hipMemcpyAsync(c->d_state, backup[i], bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess)
                return DGPU_E_HIP;
            c->master_dirty = true;
            c->work_dirty = false;
        }
        return DGPU_OK;
    };
    auto set_shape = [&](int xc, int nw) {
        for (dgpu_ctx *c : e->m) {
            c->xchunk = xc;
            c->colnw = nw;
        }
    };
    // wall time per step of the current shapes, from the saved states
    auto time_shape = [&](double &ms) -> int {
        if (int rc = restore()) return rc;
        for (int k = 1; k <= 3; ++k)
            if (int rc = dgpu_ensemble_rk_chan_1(e, k, h_bar, beta_bar, zeta_bar)) return rc;
        if (int rc = dgpu_ensemble_synchronize(e)) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s)
            for (int k = 1; k <= 3; ++k)
                if (int rc = dgpu_ensemble_rk_chan_1(e, k, h_bar, beta_bar, zeta_bar)) return rc;
        if (int rc = dgpu_ensemble_synchronize(e)) return rc;
        ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
        return DGPU_OK;
    };
    const int xc0 = e->m[0]->xchunk, nw0 = e->m[0]->colnw;
    int best_x = xc0, best_w = nw0;
    double best = 1e300;
    int rc = DGPU_OK;
    const int chunks[] = {8, 16, 32, 64};
    const int waves[] = {4, 8, 16};
    for (int xc : chunks) {
        for (int nw : waves) {
            if (nw0 == 0 && nw != waves[0]) continue;  // the sweeps were chosen: x chunks only
            bool fits = true;
            for (const dgpu_ctx *c : e->m) fits = fits && column_shape_fits(c, nw);
            const int w = (nw0 == 0 || !fits) ? nw0 : nw;
            set_shape(xc, w);
            double ms = 0.0;
            if ((rc = time_shape(ms))) break;
            if (ms < best) {
                best = ms;
                best_x = xc;
                best_w = w;
            }
        }
        if (rc) break;
    }
    set_shape(rc == DGPU_OK ? best_x : xc0, rc == DGPU_OK ? best_w : nw0);
    const int rr = restore();
    if (rc == DGPU_OK) rc = rr;
    free_all();
    if (rc) {
        if (e->err.empty()) e->err = "ensemble autotune";
        return rc;
    }
    if (xchunk_out) *xchunk_out = best_x;
    if (waves_out) *waves_out = best_w;
    if (ms_per_step) *ms_per_step = best;
    return DGPU_OK;
}

int32_t dgpu_ensemble_synchronize(dgpu_ensemble *e)
{
    if (!e) return DGPU_E_ARG;
    if (hipStreamSynchronize(e->main) != hipSuccess) return DGPU_E_HIP;
    for (auto *c : e->m)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return DGPU_E_HIP;
    return DGPU_OK;
}

int32_t dgpu_set_kernel_timing(dgpu_ctx *c, int32_t enable)
{
    if (!c) return DGPU_E_ARG;
    c->timing = enable != 0;
    return DGPU_OK;
}

int32_t dgpu_kernel_times(dgpu_ctx *c, double *ms, int64_t *count, int32_t n)
{
    if (!c) return DGPU_E_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (auto &t : c->pending) {
        float e = 0.f;
        if (hipEventElapsedTime(&e, t.a, t.b) == hipSuccess) {
            c->t_ms[t.kind] += e;
            c->t_n[t.kind] += 1;
        }
        c->event_pool.push_back(t.a);
        c->event_pool.push_back(t.b);
    }
    c->pending.clear();
    for (int k = 0; k < n && k < kKinds; ++k) {
        if (ms) ms[k] = c->t_ms[k];
        if (count) count[k] = c->t_n[k];
    }
    for (int k = 0; k < kKinds; ++k) { c->t_ms[k] = 0; c->t_n[k] = 0; }
    return DGPU_OK;
}

}  // extern "C"
