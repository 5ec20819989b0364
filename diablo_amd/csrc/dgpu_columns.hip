// dgpu_columns.hip -- per-(kx,kz)-mode y-column work of RK_CHAN_1.
//
// k_cn: explicit right-hand sides, new CF, and the four Crank-Nicolson solves
//   (TH, U2, U1, U3; for/channel.f:88-290, 581-626, 737-1029) in Fourier
//   space.  Block = 4 wavefronts x 64 modes; wavefront w solves system w for
//   the block's 64 modes.  The CN matrices do not depend on the mode, so their
//   elimination (multipliers, reciprocal pivots) and every other per-row
//   constant come from a table built on the host per RK stage and read with
//   uniform (scalar) loads.  Sweep A assembles each row's right-hand side and
//   forward-eliminates it; a block barrier separates it from sweep B (back
//   substitution), because sweep B overwrites CU/CTH that other wavefronts of
//   the block read as explicit terms in sweep A.
// k_remdiv: REM_DIV_CHAN (for/channel.f:1103-1214) per mode with the Poisson
//   factorisation precomputed once (it depends only on kx^2+kz^2 and the grid),
//   then the CP update (for/channel.f:1059-1065).
#include <algorithm>

#include "dgpu_kernels.hpp"

namespace dgpu {

template <typename T>
__device__ __forceinline__ cx<T> fma_c(T a, cx<T> x, cx<T> y) { return make_cx<T>(y.re + a * x.re, y.im + a * x.im); }

template <typename T>
__device__ __forceinline__ cx<T> neg_scale(T a, cx<T> x) { return make_cx<T>(-(a * x.re), -(a * x.im)); }

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Workgroups go to the eight XCDs in turn (block L to XCD L % 8).  For a grid
// padded to a multiple of 8, this gives each XCD one contiguous range of
// column blocks, in order: neighbouring blocks then share, in one L2, the
// cache line their rows straddle (planes of ps modes do not start on line
// boundaries).
__device__ __forceinline__ int xcd_contiguous(int L, int nb8) { return (L & 7) * (nb8 >> 3) + (L >> 3); }

// A sweep over planes j = j0, j0+dir, ... (n planes) whose per-plane inputs
// do not depend on the recurrence: the inputs of plane j are loaded D planes
// ahead into a ring of registers (ld(j) must be safe for any j in range, the
// caller clamps), so D planes of loads are in flight while the recurrence
// runs.  body(j, in) consumes them.  The loop body is unrolled D times so the
// ring is indexed statically.
template <int D, class Ld, class Body>
__device__ __forceinline__ void piped_sweep(int j0, int dir, int n, Ld &&ld, Body &&body)
{
    using In = decltype(ld(0));
    In ring[D];
#pragma unroll
    for (int d = 0; d < D; ++d) ring[d] = ld(j0 + dir * min(d, n - 1));
    for (int jb = 0; jb < n; jb += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int jj = jb + d;
            if (jj < n) {
                const In cur = ring[d];
                ring[d] = ld(j0 + dir * min(jj + D, n - 1));
                body(j0 + dir * jj, cur);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_cn
// ---------------------------------------------------------------------------
// 519 blocks of four wavefronts at 384x257x384: three waves per SIMD (<= 168
// VGPRs) keep them all resident at once instead of leaving a second round
// for a few CUs.
// AT / PT: the argument structs by value (a kernel's own arguments) or as
// references into a descriptor in global memory (batched launches); either
// way they stay out of private memory.
// DA, DB: planes in flight in sweeps A and B.  SIDE: the few low-mode
// columns of the mixed path (one block): deeper prefetch.  TPF: the table
// rows are fetched with the rest of a plane's inputs, DA/DB planes ahead,
// instead of read (scalar loads) when the plane is processed.
template <typename T, int NTH, typename NLT, typename AT, typename PT, int DA = (sizeof(T) == 4 ? 6 : 3),
          int DB = (sizeof(T) == 4 ? 16 : 8), bool SIDE = false, bool TPF = SIDE>
__device__ __forceinline__ void cn_body(AT A, PT P, int bx)
{
    const Geo &g = A.g;
    const int q = wave_uniform(threadIdx.x >> 6);  // system of this wavefront
    const long m = (long)bx * 64 + (threadIdx.x & 63);
    const bool live = m < g.ps && (q < 3 + NTH);
    const int NY = g.ny;
    const long ps = g.ps;
    const cx<T> zero = make_cx<T>(0, 0);
    const double *tab = A.tab + (long)q * (NY + 2) * kRow;
    int s = 0, i = 0;
    const long gm = (live && A.map) ? (long)A.map[m] : m;  // mode of this lane
    if (live) { s = (int)(gm / g.mk); i = (int)(gm - (long)s * g.mk); }
    const bool is00 = live && (i == 0 && s == 0);
    const long nps = A.nl_ps;
    const T kx = live ? T(A.kx[i]) : T(0), kz = live ? T(A.kz[s]) : T(0);
    const T k2 = live ? T(A.kx2[i]) + T(A.kz2[s]) : T(0);
    const T pw = (A.beta == 1.0) ? k2 : T(pow(double(k2), A.beta));
    const T temp3 = T(A.temp3), temp4 = T(A.temp4), temp5 = T(A.temp5);

    // pointers of this system: the field it solves and its accumulator
    cx<T> *fld = (q == SYS_U1) ? P.cu1 : (q == SYS_U3) ? P.cu3 : (q == SYS_U2) ? P.cu2 : P.cth;
    cx<T> *cf = (q == SYS_U1) ? P.cf1 : (q == SYS_U3) ? P.cf3 : (q == SYS_U2) ? P.cf2 : P.cfth;
    const cx<NLT> *nl = P.nl[q];
    cx<T> *gs = P.gs[q];
    const T visc = (q == SYS_TH0) ? T(A.nupr) * pw : T(A.nu) * pw;
    // Everything the sweeps' lambdas use, as plain locals: capturing the
    // argument structs themselves would give them an address (a private copy).
    const bool rk_gt1 = A.rk_gt1 != 0;
    const int f_type = A.f_type;
    const T i_ro = T(A.i_ro), rgx = T(A.rgx), rgz = T(A.rgz), drhodx = T(A.drhodx), drhodz = T(A.drhodz);
    const T force00 = T(A.force00);
    const cx<T> *const p_cp = P.cp;
    const cx<T> *const p_cu1 = P.cu1;
    const cx<T> *const p_cu3 = P.cu3;
    const cx<T> *const p_cth = P.cth;
    const bool rhs_only = A.rhs_only != 0;

    // ---------------- sweep A ----------------
    // Inputs of plane j are loaded four planes ahead (piped_sweep); every
    // load is unconditional at a valid plane and selected afterwards.
    if (live) {
        struct InA {
            cx<T> fp, nl, cf, aux1, aux2, cp;
            double tr[10];  // TPF: this row's table entries, fetched ahead
        };
        cx<T> fm = zero, f0 = fld[m];  // this system's field at j-1, j (j+1 comes in)
        cx<T> thm = zero;              // CTH at j-1 (U2 buoyancy)
        cx<T> cpm = zero;              // CP at j-1 (U2 pressure)
        cx<T> gprev = zero;
        const cx<T> *a1p = (q == SYS_U1) ? p_cu3 : p_cu1;   // Coriolis partner / TH advection
        const cx<T> *a2p = (q == SYS_TH0) ? p_cu3 : p_cth;  // buoyancy / TH advection
        const bool has1 = (q != SYS_U2);
        const bool has2 = (q == SYS_TH0) || NTH;
        const bool hascp = (q != SYS_TH0);
        piped_sweep<DA>(
            0, +1, NY + 2,
            [&](int j) {
                const long o = m + (long)j * ps;
                InA in;
                in.fp = fld[o + (j <= NY ? ps : 0)];
                const cx<NLT> nv = nl[gm + (long)j * nps];
                in.nl = make_cx<T>(T(nv.re), T(nv.im));
                in.cf = rk_gt1 ? cf[o] : zero;
                in.aux1 = has1 ? a1p[o] : zero;
                in.aux2 = has2 ? a2p[o] : zero;
                in.cp = hascp ? p_cp[o] : zero;
                if constexpr (TPF) {
                    const double *t = tab + (long)j * kRow;
                    in.tr[0] = t[T_MULT];
                    in.tr[1] = t[T_ISBC];
                    in.tr[2] = t[T_CFR];
                    in.tr[3] = t[T_B0];
                    in.tr[4] = t[T_BM];
                    in.tr[5] = t[T_BC00];
                    in.tr[6] = t[T_CPW];
                    in.tr[7] = t[T_WP];
                    in.tr[8] = t[T_W0];
                    in.tr[9] = t[T_WM];
                }
                return in;
            },
            [&](int j, const InA &in) {
                const long o = m + (long)j * ps;
                double trow[kRow];
                const double *t = tab + (long)j * kRow;
                if constexpr (TPF) {
                    trow[T_MULT] = in.tr[0];
                    trow[T_ISBC] = in.tr[1];
                    trow[T_CFR] = in.tr[2];
                    trow[T_B0] = in.tr[3];
                    trow[T_BM] = in.tr[4];
                    trow[T_BC00] = in.tr[5];
                    trow[T_CPW] = in.tr[6];
                    trow[T_WP] = in.tr[7];
                    trow[T_W0] = in.tr[8];
                    trow[T_WM] = in.tr[9];
                    t = trow;
                }
                const T mult = T(t[T_MULT]);
                const bool isbc = t[T_ISBC] != 0.0;
                const bool cfr = t[T_CFR] != 0.0;
                const cx<T> fp = (j <= NY) ? in.fp : zero;
                const bool use = cfr || !isbc;
                const cx<T> nlv = use ? in.nl : zero;
                const cx<T> cfo = use ? in.cf : zero;
                const cx<T> aux1 = in.aux1, aux2 = in.aux2, cp0 = in.cp;
                // new explicit term CF (for/channel.f:175-290 + nonlinear terms)
                cx<T> cfn = neg_scale(visc, f0);
                if (q == SYS_U1) {
                    cfn = fma_c(i_ro, aux1, cfn);
                    if (NTH) cfn = fma_c(rgx, aux2, cfn);
                } else if (q == SYS_U3) {
                    cfn = fma_c(-i_ro, aux1, cfn);
                    if (NTH) cfn = fma_c(rgz, aux2, cfn);
                } else if (q == SYS_U2) {
                    if (NTH && !is00) cfn = cfn + make_cx<T>(T(t[T_B0]) * aux2.re + T(t[T_BM]) * thm.re,
                                                              T(t[T_B0]) * aux2.im + T(t[T_BM]) * thm.im);
                } else {
                    cfn = cfn - make_cx<T>(aux1.re * drhodx + aux2.re * drhodz, aux1.im * drhodx + aux2.im * drhodz);
                }
                cfn = cfn + nlv;
                cx<T> vec;
                if (isbc) {
                    vec = is00 ? make_cx<T>(T(t[T_BC00]), 0) : zero;
                } else {
                    cx<T> cr = f0;
                    if (rk_gt1) cr = fma_c(temp3, cfo, cr);
                    if (q == SYS_U1) {
                        cr = cr - make_cx<T>(-temp4 * (kx * cp0.im), temp4 * (kx * cp0.re));
                        if (is00 && (f_type == 1 || f_type == 2)) cr.re = cr.re - temp4 * force00;
                    } else if (q == SYS_U3) {
                        cr = cr - make_cx<T>(-temp4 * (kz * cp0.im), temp4 * (kz * cp0.re));
                    } else if (q == SYS_U2) {
                        const T w = T(t[T_CPW]);
                        cr = cr - make_cx<T>(w * (cp0.re - cpm.re), w * (cp0.im - cpm.im));
                    }
                    cr = fma_c(temp5, cfn, cr);
                    const T wp = T(t[T_WP]), w0 = T(t[T_W0]), wm = T(t[T_WM]);
                    vec = make_cx<T>(cr.re + (wp * fp.re + w0 * f0.re + wm * fm.re),
                                     cr.im + (wp * fp.im + w0 * f0.im + wm * fm.im));
                }
                const cx<T> gj = (j == 0 || rhs_only) ? vec : fma_c(mult, gprev, vec);
                if (cfr) cf[o] = cfn;
                gs[o] = gj;
                gprev = gj;
                fm = f0;
                f0 = fp;
                thm = aux2;
                cpm = cp0;
            });
    }
    __syncthreads();
    // ---------------- sweep B ----------------
    if (live && !rhs_only) {
        struct InB {
            cx<T> g;
            double ip, u, cpy;  // TPF: fetched ahead
        };
        cx<T> x = zero;
        piped_sweep<DB>(
            NY + 1, -1, NY + 2,
            [&](int j) {
                InB in;
                in.g = gs[m + (long)j * ps];
                if constexpr (TPF) {
                    const double *t = tab + (long)j * kRow;
                    in.ip = t[T_INVPIV];
                    in.u = t[T_UP];
                    in.cpy = t[T_CPY];
                }
                return in;
            },
            [&](int j, const InB &in) {
                const cx<T> gj = in.g;
                const double *t = tab + (long)j * kRow;
                const double tip = TPF ? in.ip : t[T_INVPIV];
                const double tup = TPF ? in.u : t[T_UP];
                const double tcpy = TPF ? in.cpy : t[T_CPY];
                const T ip = T(tip);
                if (j == NY + 1) {
                    x = make_cx<T>(gj.re * ip, gj.im * ip);
                } else {
                    const T u = T(tup);
                    x = make_cx<T>((gj.re - u * x.re) * ip, (gj.im - u * x.im) * ip);
                }
                if (tcpy != 0.0) fld[m + (long)j * ps] = x;
            });
    }
}

// The table rows ride in the prefetch ring in FP32 (0.66 -> 0.60 ms at
// 384x257x384); in FP64 the ring's extra registers spill (1.2 -> 2.5 ms).
template <typename T, int NTH, typename NLT>
__global__ void __launch_bounds__(256, 3) k_cn(CnArgs A, CnPtrs<T, NLT> P)
{
    cn_body<T, NTH, NLT, const CnArgs, const CnPtrs<T, NLT>, (sizeof(T) == 4 ? 6 : 3), (sizeof(T) == 4 ? 16 : 8),
            false, sizeof(T) == 4>(A, P, xcd_contiguous(blockIdx.x, gridDim.x));
}

// The mixed path's low-mode columns: a block or two, so registers are free.
template <typename T, int NTH, typename NLT>
__global__ void __launch_bounds__(256, 1) k_cn_side(CnArgs A, CnPtrs<T, NLT> P)
{
    cn_body<T, NTH, NLT, const CnArgs, const CnPtrs<T, NLT>, 6, 16, true>(A, P, blockIdx.x);
}

// Ensemble members sharing NX, NZ (so ps): grid (blocks, members).
template <typename T, int NTH, typename NLT>
__global__ void __launch_bounds__(256, 3) k_cn_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    // references into the descriptor (uniform loads), not local copies that
    // would live in scratch
    const ColDesc<T, NLT> &D = d[blockIdx.y];
    cn_body<T, NTH, NLT, const CnArgs &, const CnPtrs<T, NLT> &, (sizeof(T) == 4 ? 6 : 3), (sizeof(T) == 4 ? 16 : 8),
            false, sizeof(T) == 4>(D.A, D.P, blockIdx.x);
}

// The ensemble's low-mode columns (same sweeps as k_cn_side).
template <typename T, int NTH, typename NLT>
__global__ void __launch_bounds__(256, 1) k_cn_side_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    const ColDesc<T, NLT> &D = d[blockIdx.y];
    cn_body<T, NTH, NLT, const CnArgs &, const CnPtrs<T, NLT> &, 6, 16, true>(D.A, D.P, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Poisson factorisation for REM_DIV_CHAN, once per context: the reference's
// THOMAS_FORWARD_COMPLEX_MPI elimination of the matrix (for/channel.f:1126-1134,
// for/mpi.f:1293-1332, for/mpi.f:741-749) per mode; stores the multipliers and
// the reciprocal pivots (0 where the reference's guard zeroes the solution).
// ---------------------------------------------------------------------------
__global__ void k_pois_factor(Geo g, const double *__restrict__ dy, const double *__restrict__ dyf,
                              const double *__restrict__ kx2, const double *__restrict__ kz2, double *pmult,
                              double *pinv)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.ps) return;
    const int s = (int)(m / g.mk), i = (int)(m - (long)s * g.mk);
    const bool is00 = (i == 0 && s == 0);
    const int NY = g.ny;
    const long ps = g.ps;
    double bprev = 1.0;  // row 1: D = 1
    pmult[m + ps] = 0.0;
    pinv[m + ps] = 1.0;
    for (int j = 2; j <= NY; ++j) {
        double L, D;
        const double Uprev = (j - 1 == 1) ? (is00 ? 0.0 : -1.0) : 1. / (dy[j] * dyf[j - 1]);
        if (j == NY) {
            L = 1.0;
            D = -1.0;
        } else {
            L = 1. / (dy[j] * dyf[j]);
            D = -kx2[i] - kz2[s] - 1. / (dy[j + 1] * dyf[j]) - 1. / (dy[j] * dyf[j]);
        }
        const double mult = (bprev != 0.0) ? -L / bprev : 0.0;
        const double b = D + mult * Uprev;
        pmult[m + (long)j * ps] = mult;
        pinv[m + (long)j * ps] = (b != 0.0) ? 1.0 / b : 0.0;
        bprev = b;
    }
}

// ---------------------------------------------------------------------------
// k_remdiv
// ---------------------------------------------------------------------------
template <typename T, typename NLT, int DC = (sizeof(T) == 4 ? 16 : 8), int DD = (sizeof(T) == 4 ? 12 : 6)>
__device__ __forceinline__ void remdiv_body(const RdArgs &A, const CnPtrs<T, NLT> &P, int bx)
{
    // Two lanes per mode: lane 2m+c takes component c (0 real, 1 imaginary)
    // of mode m.  The Poisson factors are real, so each component's sweeps
    // are independent; the divergence and the gradient update cross the
    // components (i kx, i kz), which the lane reads from memory (divergence)
    // or from its partner lane (update).  Twice the wavefronts of one lane
    // per mode, half the work each: these sweeps are latency bound.
    const Geo &g = A.g;
    const long lane = (long)bx * blockDim.x + threadIdx.x;
    const long m = lane >> 1;
    const int c = (int)(lane & 1);
    if (m >= g.ps) return;  // both lanes of a pair leave together
    const long gm = A.map ? (long)A.map[m] : m;  // mode of this lane
    const int s = (int)(gm / g.mk), i = (int)(gm - (long)s * g.mk);
    const bool is00 = (i == 0 && s == 0);
    const int NY = g.ny;
    const long ps = g.ps, fps = A.fac_ps;
    const T kx = T(A.kx[i]), kz = T(A.kz[s]);
    // i k x: real part -k x.im, imaginary part k x.re
    const T sgn = c ? T(1) : T(-1);
    auto pm = [&](int j) -> T {
        if constexpr (sizeof(T) == 4) return A.pmultf[gm + (long)j * fps];
        else return T(A.pmult[gm + (long)j * fps]);
    };
    auto pv = [&](int j) -> T {
        if constexpr (sizeof(T) == 4) return A.pinvf[gm + (long)j * fps];
        else return T(A.pinv[gm + (long)j * fps]);
    };
    // component views: element (plane offset o, component q) at 2 o + q
    const T *u1 = reinterpret_cast<const T *>(P.cu1);
    const T *u2 = reinterpret_cast<const T *>(P.cu2);
    const T *u3 = reinterpret_cast<const T *>(P.cu3);
    T *w1 = reinterpret_cast<T *>(P.cu1);
    T *w2 = reinterpret_cast<T *>(P.cu2);
    T *w3 = reinterpret_cast<T *>(P.cu3);
    T *wcp = reinterpret_cast<T *>(P.cp);
    T *scr = reinterpret_cast<T *>(P.gs[0]);  // forward-eliminated right-hand side
    const int co = c ^ 1;
    // sweep C: rows 2..NY (row 1 is phi(1) = 0 / phi(1) - phi(2) = 0, RHS 0)
    {
        struct InC {
            T u1o, u3o, c2p, mu, idf;
        };
        T gprev = T(0);
        T c2 = u2[2 * (m + 2 * ps) + c];
        scr[2 * (m + ps) + c] = T(0);
        piped_sweep<DC>(
            2, +1, NY - 1,
            [&](int j) {
                const long o = m + (long)j * ps;
                InC in;
                in.u1o = u1[2 * o + co];
                in.u3o = u3[2 * o + co];
                in.c2p = u2[2 * (o + ps) + c];
                in.mu = pm(j);
                in.idf = T(A.rd[(long)j * 4 + 1]);
                return in;
            },
            [&](int j, const InC &in) {
                const long o = m + (long)j * ps;
                T vec = T(0);
                if (j < NY) {
                    // (ikx CU1 + (CU2(j+1)-CU2(j))/DYF(j)) + ikz CU3
                    vec = (sgn * kx * in.u1o + (in.c2p - c2) * in.idf) + (sgn * kz * in.u3o);
                }
                const T gj = vec + in.mu * gprev;
                scr[2 * o + c] = gj;
                gprev = gj;
                c2 = in.c2p;
            });
    }
    // sweep D: back substitution (j = NY..1), projection, CP update.  Plane j
    // rewrites CU1(j), CU3(j), CU2(j+1) and CP(j); each is read (ahead) only
    // for that same plane, before its rewrite.
    {
        struct InD {
            T g, u1, u3, c2, cp, ip, U, idy;
        };
        const T u1row = is00 ? T(0) : T(-1);
        const T itemp4 = T(A.inv_temp4);
        const bool upd = A.update_cp;
        T phip = T(0);
        piped_sweep<DD>(
            NY, -1, NY,
            [&](int j) {
                const long o = m + (long)j * ps;
                InD in;
                in.g = scr[2 * o + c];
                in.ip = pv(j);
                in.U = (j == 1) ? u1row : T(A.rd[(long)j * 4 + 0]);
                in.u1 = u1[2 * o + c];
                in.u3 = u3[2 * o + c];
                in.c2 = u2[2 * (o + ps) + c];  // plane NY+1 exists (unused for j = NY)
                in.idy = T(A.rd[(long)(j + 1) * 4 + 2]);
                in.cp = upd ? wcp[2 * o + c] : T(0);
                return in;
            },
            [&](int j, const InD &in) {
                const long o = m + (long)j * ps;
                const T phi = (j == NY) ? in.g * in.ip : (in.g - in.U * phip) * in.ip;
                const T phio = __shfl_xor(phi, 1);  // the partner lane's component
                // CU1 - i kx phi: real part + kx phi.im, imaginary part - kx phi.re
                w1[2 * o + c] = in.u1 - sgn * kx * phio;
                w3[2 * o + c] = in.u3 - sgn * kz * phio;
                if (j + 1 <= NY) w2[2 * (o + ps) + c] = in.c2 - (phip - phi) * in.idy;
                if (upd && j >= A.jstart && j <= A.jend) wcp[2 * o + c] = in.cp + phi * itemp4;
                phip = phi;
            });
    }
}

template <typename T, typename NLT>
__global__ void __launch_bounds__(64) k_remdiv(RdArgs A, CnPtrs<T, NLT> P)
{
    remdiv_body<T, NLT>(A, P, xcd_contiguous(blockIdx.x, gridDim.x));
}

// The mixed path's low-mode columns: deeper prefetch (few wavefronts).
template <typename T, typename NLT>
__global__ void __launch_bounds__(64) k_remdiv_side(RdArgs A, CnPtrs<T, NLT> P)
{
    remdiv_body<T, NLT, 16, 16>(A, P, blockIdx.x);
}

template <typename T, typename NLT>
__global__ void __launch_bounds__(64) k_remdiv_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    const ColDesc<T, NLT> &D = d[blockIdx.y];
    remdiv_body<T, NLT>(D.R, D.P, blockIdx.x);
}

template <typename T, typename NLT>
__global__ void __launch_bounds__(64) k_remdiv_side_batch(const ColDesc<T, NLT> *__restrict__ d)
{
    const ColDesc<T, NLT> &D = d[blockIdx.y];
    remdiv_body<T, NLT, 16, 16>(D.R, D.P, blockIdx.x);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, typename NLT>
void launch_cn_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem)
{
    const long blocks = (g.ps + 63) / 64;
    if (g.nth == 0)
        hipLaunchKernelGGL((k_cn_batch<T, 0, NLT>), dim3(blocks, nmem), dim3(256), 0, s, d);
    else
        hipLaunchKernelGGL((k_cn_batch<T, 1, NLT>), dim3(blocks, nmem), dim3(256), 0, s, d);
}

template <typename T, typename NLT>
void launch_cn_side_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem)
{
    const long blocks = (g.ps + 63) / 64;
    if (g.nth == 0)
        hipLaunchKernelGGL((k_cn_side_batch<T, 0, NLT>), dim3(blocks, nmem), dim3(256), 0, s, d);
    else
        hipLaunchKernelGGL((k_cn_side_batch<T, 1, NLT>), dim3(blocks, nmem), dim3(256), 0, s, d);
}

template <typename T, typename NLT>
void launch_remdiv_side_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem)
{
    const long blocks = (2 * g.ps + 63) / 64;
    hipLaunchKernelGGL((k_remdiv_side_batch<T, NLT>), dim3(blocks, nmem), dim3(64), 0, s, d);
}

template <typename T, typename NLT>
void launch_remdiv_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem)
{
    const long blocks = (2 * g.ps + 63) / 64;  // two lanes per mode
    hipLaunchKernelGGL((k_remdiv_batch<T, NLT>), dim3(blocks, nmem), dim3(64), 0, s, d);
}

template void launch_cn_batch<double, double>(hipStream_t, const Geo &, const ColDesc<double, double> *, int);
template void launch_cn_batch<double, float>(hipStream_t, const Geo &, const ColDesc<double, float> *, int);
template void launch_cn_batch<float, float>(hipStream_t, const Geo &, const ColDesc<float, float> *, int);
template void launch_remdiv_batch<double, double>(hipStream_t, const Geo &, const ColDesc<double, double> *, int);
template void launch_remdiv_batch<double, float>(hipStream_t, const Geo &, const ColDesc<double, float> *, int);
template void launch_remdiv_batch<float, float>(hipStream_t, const Geo &, const ColDesc<float, float> *, int);
template void launch_cn_side_batch<double, float>(hipStream_t, const Geo &, const ColDesc<double, float> *, int);
template void launch_remdiv_side_batch<double, float>(hipStream_t, const Geo &, const ColDesc<double, float> *,
                                                       int);

template <typename T, typename NLT>
void launch_cn(hipStream_t s, const CnArgs &a, const CnPtrs<T, NLT> &p)
{
    const long blocks = ((a.g.ps + 63) / 64 + 7) / 8 * 8;  // a multiple of 8 (xcd_contiguous)
    if (a.g.nth == 0)
        hipLaunchKernelGGL((k_cn<T, 0, NLT>), dim3(blocks), dim3(256), 0, s, a, p);
    else
        hipLaunchKernelGGL((k_cn<T, 1, NLT>), dim3(blocks), dim3(256), 0, s, a, p);
}

template <typename T, typename NLT>
void launch_cn_side(hipStream_t s, const CnArgs &a, const CnPtrs<T, NLT> &p)
{
    const long blocks = (a.g.ps + 63) / 64;
    if (a.g.nth == 0)
        hipLaunchKernelGGL((k_cn_side<T, 0, NLT>), dim3(blocks), dim3(256), 0, s, a, p);
    else
        hipLaunchKernelGGL((k_cn_side<T, 1, NLT>), dim3(blocks), dim3(256), 0, s, a, p);
}

template <typename T, typename NLT>
void launch_remdiv_side(hipStream_t s, const RdArgs &a, const CnPtrs<T, NLT> &p)
{
    const long blocks = (2 * a.g.ps + 63) / 64;
    hipLaunchKernelGGL((k_remdiv_side<T, NLT>), dim3(blocks), dim3(64), 0, s, a, p);
}

template <typename T, typename NLT>
void launch_remdiv(hipStream_t s, const RdArgs &a, const CnPtrs<T, NLT> &p)
{
    const long blocks = ((2 * a.g.ps + 63) / 64 + 7) / 8 * 8;  // two lanes per mode; a multiple of 8
    hipLaunchKernelGGL((k_remdiv<T, NLT>), dim3(blocks), dim3(64), 0, s, a, p);
}

void launch_pois_factor(hipStream_t s, const Geo &g, const double *dy, const double *dyf, const double *kx2,
                        const double *kz2, double *pmult, double *pinv)
{
    const long blocks = (g.ps + 255) / 256;
    hipLaunchKernelGGL(k_pois_factor, dim3(blocks), dim3(256), 0, s, g, dy, dyf, kx2, kz2, pmult, pinv);
}

// ---------------------------------------------------------------------------
// state copies of the mixed path
// ---------------------------------------------------------------------------
template <typename TO, typename FROM>
__global__ void k_convert(const cx<FROM> *__restrict__ src, cx<TO> *__restrict__ dst, long n)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        dst[k] = make_cx<TO>(TO(src[k].re), TO(src[k].im));
}

__global__ void k_convert_real(const double *__restrict__ src, float *__restrict__ dst, long n)
{
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long)gridDim.x * blockDim.x)
        dst[k] = float(src[k]);
}

template <typename TS, typename TF, bool GATHER>
__global__ void k_side(const cx<TF> *full_in, cx<TF> *full_out, long full_fs, long full_ps, const cx<TS> *side_in,
                       cx<TS> *side_out, long side_fs, const int *__restrict__ map, int nside, int planes)
{
    const int f = blockIdx.y;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)nside * planes;
         k += (long)gridDim.x * blockDim.x) {
        const int j = (int)(k / nside), l = (int)(k - (long)j * nside);
        const long fo = f * full_fs + (long)j * full_ps + map[l];
        const long so = f * side_fs + (long)j * nside + l;
        if (GATHER) {
            const cx<TF> v = full_in[fo];
            side_out[so] = make_cx<TS>(TS(v.re), TS(v.im));
        } else {
            const cx<TS> v = side_in[so];
            full_out[fo] = make_cx<TF>(TF(v.re), TF(v.im));
        }
    }
}

static inline unsigned grid1(long n) { return (unsigned)std::min<long>((n + 255) / 256, 65536); }

template <typename TO, typename FROM>
void launch_convert(hipStream_t s, const cx<FROM> *src, cx<TO> *dst, long n)
{
    hipLaunchKernelGGL((k_convert<TO, FROM>), dim3(grid1(n)), dim3(256), 0, s, src, dst, n);
}

void launch_convert_real(hipStream_t s, const double *src, float *dst, long n)
{
    hipLaunchKernelGGL(k_convert_real, dim3(grid1(n)), dim3(256), 0, s, src, dst, n);
}

template <typename TS, typename TF>
void launch_side_gather(hipStream_t s, const cx<TF> *full, long full_fs, long full_ps, cx<TS> *side, long side_fs,
                        const int *map, int nside, int planes, int nf)
{
    const long n = (long)nside * planes;
    hipLaunchKernelGGL((k_side<TS, TF, true>), dim3(grid1(n), nf), dim3(256), 0, s, full, nullptr, full_fs, full_ps,
                       nullptr, side, side_fs, map, nside, planes);
}

template <typename TS, typename TF>
void launch_side_scatter(hipStream_t s, const cx<TS> *side, long side_fs, cx<TF> *full, long full_fs, long full_ps,
                         const int *map, int nside, int planes, int nf)
{
    const long n = (long)nside * planes;
    hipLaunchKernelGGL((k_side<TS, TF, false>), dim3(grid1(n), nf), dim3(256), 0, s, nullptr, full, full_fs, full_ps,
                       side, nullptr, side_fs, map, nside, planes);
}

template <typename TS, typename TF>
__global__ void k_side_scatter_batch(const ScatDesc<TS, TF> *__restrict__ d, int nside, long ps)
{
    const ScatDesc<TS, TF> &D = d[blockIdx.z];
    const int f = blockIdx.y;
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < (long)nside * D.planes;
         k += (long)gridDim.x * blockDim.x) {
        const int j = (int)(k / nside), l = (int)(k - (long)j * nside);
        const cx<TS> v = D.side[f * D.side_fs + (long)j * nside + l];
        D.full[f * D.full_fs + (long)j * ps + D.map[l]] = make_cx<TF>(TF(v.re), TF(v.im));
    }
}

template <typename TS, typename TF>
void launch_side_scatter_batch(hipStream_t s, const ScatDesc<TS, TF> *d, int nmem, int nside, long ps,
                               int max_planes, int nf)
{
    const long n = (long)nside * max_planes;
    hipLaunchKernelGGL((k_side_scatter_batch<TS, TF>), dim3(grid1(n), nf, nmem), dim3(256), 0, s, d, nside, ps);
}

template void launch_side_scatter_batch<double, float>(hipStream_t, const ScatDesc<double, float> *, int, int, long,
                                                       int, int);

template void launch_convert<float, double>(hipStream_t, const cx<double> *, cx<float> *, long);
template void launch_convert<double, float>(hipStream_t, const cx<float> *, cx<double> *, long);
template void launch_side_gather<double, double>(hipStream_t, const cx<double> *, long, long, cx<double> *, long,
                                                 const int *, int, int, int);
template void launch_side_scatter<double, double>(hipStream_t, const cx<double> *, long, cx<double> *, long, long,
                                                  const int *, int, int, int);
template void launch_side_scatter<double, float>(hipStream_t, const cx<double> *, long, cx<float> *, long, long,
                                                 const int *, int, int, int);

template void launch_cn<double, double>(hipStream_t, const CnArgs &, const CnPtrs<double, double> &);
template void launch_cn<float, float>(hipStream_t, const CnArgs &, const CnPtrs<float, float> &);
template void launch_remdiv<float, float>(hipStream_t, const RdArgs &, const CnPtrs<float, float> &);
template void launch_cn<double, float>(hipStream_t, const CnArgs &, const CnPtrs<double, float> &);
template void launch_remdiv<double, double>(hipStream_t, const RdArgs &, const CnPtrs<double, double> &);
template void launch_remdiv<double, float>(hipStream_t, const RdArgs &, const CnPtrs<double, float> &);
template void launch_cn_side<double, float>(hipStream_t, const CnArgs &, const CnPtrs<double, float> &);
template void launch_remdiv_side<double, float>(hipStream_t, const RdArgs &, const CnPtrs<double, float> &);

}  // namespace dgpu
