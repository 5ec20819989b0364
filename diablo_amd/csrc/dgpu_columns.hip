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
#include "dgpu_kernels.hpp"

namespace dgpu {

template <typename T>
__device__ __forceinline__ cx<T> fma_c(T a, cx<T> x, cx<T> y) { return make_cx<T>(y.re + a * x.re, y.im + a * x.im); }

template <typename T>
__device__ __forceinline__ cx<T> neg_scale(T a, cx<T> x) { return make_cx<T>(-(a * x.re), -(a * x.im)); }

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---------------------------------------------------------------------------
// k_cn
// ---------------------------------------------------------------------------
template <typename T, int NTH, typename NLT>
__global__ void __launch_bounds__(256) k_cn(CnArgs A, CnPtrs<T, NLT> P)
{
    const Geo &g = A.g;
    const int q = wave_uniform(threadIdx.x >> 6);  // system of this wavefront
    const long m = (long)blockIdx.x * 64 + (threadIdx.x & 63);
    const bool live = m < g.ps && (q < 3 + NTH);
    const int NY = g.ny;
    const long ps = g.ps;
    const cx<T> zero = make_cx<T>(0, 0);
    const double *tab = A.tab + (long)q * (NY + 2) * kRow;
    int s = 0, i = 0;
    if (live) { s = (int)(m / g.mk); i = (int)(m - (long)s * g.mk); }
    const bool is00 = live && (i == 0 && s == 0);
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

    // ---------------- sweep A ----------------
    if (live) {
        cx<T> fm = zero, f0 = fld[m], fp;          // this system's field at j-1, j, j+1
        cx<T> thm = zero;                            // CTH at j-1 (U2 buoyancy)
        cx<T> cpm = zero;                            // CP at j-1 (U2 pressure)
        cx<T> gprev = zero;
        for (int j = 0; j <= NY + 1; ++j) {
            const long o = m + (long)j * ps;
            const double *t = tab + (long)j * kRow;
            const T mult = T(t[T_MULT]);
            const bool isbc = t[T_ISBC] != 0.0;
            const bool cfr = t[T_CFR] != 0.0;
            fp = (j <= NY) ? fld[o + ps] : zero;
            cx<T> nlv = zero, cfo = zero, cp0 = zero, aux1 = zero, aux2 = zero;
            if (cfr || !isbc) {
                nlv = make_cx<T>(T(nl[o].re), T(nl[o].im));
                if (A.rk_gt1) cfo = cf[o];
            }
            // explicit-term inputs
            if (q == SYS_U1 || q == SYS_U3) {
                aux1 = (q == SYS_U1) ? P.cu3[o] : P.cu1[o];   // Coriolis partner
                if (NTH) aux2 = P.cth[o];
                cp0 = P.cp[o];
            } else if (q == SYS_U2) {
                if (NTH) aux2 = P.cth[o];
                cp0 = P.cp[o];
            } else {
                aux1 = P.cu1[o];
                aux2 = P.cu3[o];
            }
            // new explicit term CF (for/channel.f:175-290 + nonlinear terms)
            cx<T> cfn = neg_scale(visc, f0);
            if (q == SYS_U1) {
                cfn = fma_c(T(A.i_ro), aux1, cfn);
                if (NTH) cfn = fma_c(T(A.rgx), aux2, cfn);
            } else if (q == SYS_U3) {
                cfn = fma_c(-T(A.i_ro), aux1, cfn);
                if (NTH) cfn = fma_c(T(A.rgz), aux2, cfn);
            } else if (q == SYS_U2) {
                if (NTH && !is00) cfn = cfn + make_cx<T>(T(t[T_B0]) * aux2.re + T(t[T_BM]) * thm.re,
                                                          T(t[T_B0]) * aux2.im + T(t[T_BM]) * thm.im);
            } else {
                cfn = cfn - make_cx<T>(aux1.re * T(A.drhodx) + aux2.re * T(A.drhodz),
                                       aux1.im * T(A.drhodx) + aux2.im * T(A.drhodz));
            }
            cfn = cfn + nlv;
            cx<T> vec;
            if (isbc) {
                vec = is00 ? make_cx<T>(T(t[T_BC00]), 0) : zero;
            } else {
                cx<T> cr = f0;
                if (A.rk_gt1) cr = fma_c(temp3, cfo, cr);
                if (q == SYS_U1) {
                    cr = cr - make_cx<T>(-temp4 * (kx * cp0.im), temp4 * (kx * cp0.re));
                    if (is00 && (A.f_type == 1 || A.f_type == 2)) cr.re = cr.re - temp4 * T(A.force00);
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
            const cx<T> gj = (j == 0) ? vec : fma_c(mult, gprev, vec);
            if (cfr) cf[o] = cfn;
            gs[o] = gj;
            gprev = gj;
            fm = f0;
            f0 = fp;
            thm = aux2;
            cpm = cp0;
        }
    }
    __syncthreads();
    // ---------------- sweep B ----------------
    if (live) {
        cx<T> x = zero;
        for (int j = NY + 1; j >= 0; --j) {
            const long o = m + (long)j * ps;
            const double *t = tab + (long)j * kRow;
            const cx<T> gj = gs[o];
            const T ip = T(t[T_INVPIV]);
            if (j == NY + 1) {
                x = make_cx<T>(gj.re * ip, gj.im * ip);
            } else {
                const T u = T(t[T_UP]);
                x = make_cx<T>((gj.re - u * x.re) * ip, (gj.im - u * x.im) * ip);
            }
            if (t[T_CPY] != 0.0) fld[o] = x;
        }
    }
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
template <typename T, typename NLT>
__global__ void __launch_bounds__(64) k_remdiv(RdArgs A, CnPtrs<T, NLT> P)
{
    const Geo &g = A.g;
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.ps) return;
    const int s = (int)(m / g.mk), i = (int)(m - (long)s * g.mk);
    const bool is00 = (i == 0 && s == 0);
    const int NY = g.ny;
    const long ps = g.ps;
    const T kx = T(A.kx[i]), kz = T(A.kz[s]);
    const cx<T> zero = make_cx<T>(0, 0);
    cx<T> *scr = P.gs[0];  // forward-eliminated right-hand side
    // sweep C: rows 2..NY (row 1 is phi(1) = 0 / phi(1) - phi(2) = 0, RHS 0)
    {
        cx<T> gprev = zero;
        cx<T> c2 = P.cu2[m + 2 * ps];
        scr[m + ps] = zero;
        for (int j = 2; j <= NY; ++j) {
            const long o = m + (long)j * ps;
            const cx<T> c2p = P.cu2[o + ps];
            cx<T> vec = zero;
            if (j < NY) {
                const cx<T> u1 = P.cu1[o], u3 = P.cu3[o];
                const T idf = T(A.rd[(long)j * 4 + 1]);
                // (ikx CU1 + (CU2(j+1)-CU2(j))/DYF(j)) + ikz CU3
                vec = make_cx<T>((-kx * u1.im + (c2p.re - c2.re) * idf) + (-kz * u3.im),
                                 (kx * u1.re + (c2p.im - c2.im) * idf) + (kz * u3.re));
            }
            const T mu = T(A.pmult[o]);
            const cx<T> gj = fma_c(mu, gprev, vec);
            scr[o] = gj;
            gprev = gj;
            c2 = c2p;
        }
    }
    // sweep D: back substitution (j = NY..1), projection, CP update
    {
        const T u1row = is00 ? T(0) : T(-1);
        const T itemp4 = T(A.inv_temp4);
        cx<T> phip = zero;
        for (int j = NY; j >= 1; --j) {
            const long o = m + (long)j * ps;
            const cx<T> gj = scr[o];
            const T ip = T(A.pinv[o]);
            cx<T> phi;
            if (j == NY) {
                phi = make_cx<T>(gj.re * ip, gj.im * ip);
            } else {
                const T U = (j == 1) ? u1row : T(A.rd[(long)j * 4 + 0]);
                phi = make_cx<T>((gj.re - U * phip.re) * ip, (gj.im - U * phip.im) * ip);
            }
            const cx<T> a = P.cu1[o], b = P.cu3[o];
            P.cu1[o] = make_cx<T>(a.re + kx * phi.im, a.im - kx * phi.re);
            P.cu3[o] = make_cx<T>(b.re + kz * phi.im, b.im - kz * phi.re);
            if (j + 1 <= NY) {
                const T idy = T(A.rd[(long)(j + 1) * 4 + 2]);
                const cx<T> c = P.cu2[o + ps];
                P.cu2[o + ps] = make_cx<T>(c.re - (phip.re - phi.re) * idy, c.im - (phip.im - phi.im) * idy);
            }
            if (A.update_cp && j >= A.jstart && j <= A.jend) {
                const cx<T> p = P.cp[o];
                P.cp[o] = make_cx<T>(p.re + phi.re * itemp4, p.im + phi.im * itemp4);
            }
            phip = phi;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, typename NLT>
void launch_cn(hipStream_t s, const CnArgs &a, const CnPtrs<T, NLT> &p)
{
    const long blocks = (a.g.ps + 63) / 64;
    if (a.g.nth == 0)
        hipLaunchKernelGGL((k_cn<T, 0, NLT>), dim3(blocks), dim3(256), 0, s, a, p);
    else
        hipLaunchKernelGGL((k_cn<T, 1, NLT>), dim3(blocks), dim3(256), 0, s, a, p);
}

template <typename T, typename NLT>
void launch_remdiv(hipStream_t s, const RdArgs &a, const CnPtrs<T, NLT> &p)
{
    const long blocks = (a.g.ps + 63) / 64;
    hipLaunchKernelGGL((k_remdiv<T, NLT>), dim3(blocks), dim3(64), 0, s, a, p);
}

void launch_pois_factor(hipStream_t s, const Geo &g, const double *dy, const double *dyf, const double *kx2,
                        const double *kz2, double *pmult, double *pinv)
{
    const long blocks = (g.ps + 255) / 256;
    hipLaunchKernelGGL(k_pois_factor, dim3(blocks), dim3(256), 0, s, g, dy, dyf, kx2, kz2, pmult, pinv);
}

template void launch_cn<double, double>(hipStream_t, const CnArgs &, const CnPtrs<double, double> &);
template void launch_cn<double, float>(hipStream_t, const CnArgs &, const CnPtrs<double, float> &);
template void launch_remdiv<double, double>(hipStream_t, const RdArgs &, const CnPtrs<double, double> &);
template void launch_remdiv<double, float>(hipStream_t, const RdArgs &, const CnPtrs<double, float> &);

}  // namespace dgpu
