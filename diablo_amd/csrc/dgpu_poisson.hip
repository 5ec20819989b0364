// dgpu_poisson.hip -- POISSON_P_CHAN (for/channel.f:1217-1453) on the device.
//
// CP from CU1-3 (and CTH): the right-hand side sum_ij (du_i/dx_j)(du_j/dx_i)
// is assembled with the reference's four groups of terms (nine transforms to
// physical space and six back), then one tridiagonal solve per (kx,kz) mode.
// The solve rows that are read (1..NY) carry the same matrix as REM_DIV_CHAN's
// (for/channel.f:1126-1143 vs 1399-1440, boundary rows for/mpi.f:1293-1369),
// so the context's precomputed Poisson factorisation is reused.
//
// Work arrays are Fortran-layout planes (the layout the standalone transform
// kernels take); this routine runs once per restart, not per time step.
#include "dgpu_kernels.hpp"

namespace dgpu {

// Spectral inputs of one group of terms into Fortran-layout complex planes.
// group 0: (i kx CU1, (CU2(j+1)-CU2(j))/DYF(j), i kz CU3)
// group 1: ((CU1(j+1)-CU1(j-1))/(2 DYF(j)), i kx 0.5 (CU2(j)+CU2(j+1)))
// group 2: ((CU3(j+1)-CU3(j-1))/(2 DYF(j)), i kz 0.5 (CU2(j)+CU2(j+1)))
// group 3: (i kz CU1, i kx CU3)
// Planes 2..NY get the values on the kept slots; every other slot is zero.
__global__ void k_pp_spec(int group, Geo g, const cx<double> *__restrict__ cu1, const cx<double> *__restrict__ cu2,
                          const cx<double> *__restrict__ cu3, const double *__restrict__ dyf,
                          const double *__restrict__ kxv, const double *__restrict__ kzv, double *fa, double *fb,
                          double *fc)
{
    const int j = blockIdx.y;
    const int nxp = g.nx / 2;
    const long plane = (long)(nxp + 1) * (g.nz + 2);
    const long ps = g.ps;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < plane; e += (long)gridDim.x * blockDim.x) {
        const int K = (int)(e / (nxp + 1)), i = (int)(e - (long)K * (nxp + 1));
        cx<double> a = make_cx<double>(0, 0), b = a, c = a;
        if (i < g.mk && K < g.mz && j >= 2 && j <= g.ny) {
            const long o = (long)j * ps + (long)K * g.mk + i;
            const double kx = kxv[i], kz = kzv[K];
            if (group == 0) {
                const cx<double> u1 = cu1[o], u3 = cu3[o], d = cu2[o + ps] - cu2[o];
                a = make_cx<double>(-kx * u1.im, kx * u1.re);
                b = make_cx<double>(d.re / dyf[j], d.im / dyf[j]);
                c = make_cx<double>(-kz * u3.im, kz * u3.re);
            } else if (group == 1 || group == 2) {
                const cx<double> *u = (group == 1) ? cu1 : cu3;
                const double k = (group == 1) ? kx : kz;
                const cx<double> d = u[o + ps] - u[o - ps];
                a = make_cx<double>(d.re / (2. * dyf[j]), d.im / (2. * dyf[j]));
                const cx<double> s = cu2[o] + cu2[o + ps];
                b = make_cx<double>(-(k * 0.5) * s.im, (k * 0.5) * s.re);
            } else {
                const cx<double> u1 = cu1[o], u3 = cu3[o];
                a = make_cx<double>(-kz * u1.im, kz * u1.re);
                b = make_cx<double>(-kx * u3.im, kx * u3.re);
            }
        }
        const long idx = ((long)j * plane + e) * 2;
        fa[idx] = a.re; fa[idx + 1] = a.im;
        fb[idx] = b.re; fb[idx + 1] = b.im;
        if (fc) { fc[idx] = c.re; fc[idx + 1] = c.im; }
    }
}

// Physical-space products on planes 2..NY: op 0 squares a, b, c in place
// (for/channel.f:1246-1254); op 1 sets a = 2*a*b (for/channel.f:1283-1289).
__global__ void k_pp_phys(int op, Geo g, double *fa, const double *fb, double *fbw, double *fc)
{
    const int j = 2 + blockIdx.y;
    const long rowlen = g.nx + 2;
    const long plane = rowlen * (g.nz + 2);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < (long)g.nz * g.nx;
         e += (long)gridDim.x * blockDim.x) {
        const int z = (int)(e / g.nx), x = (int)(e - (long)z * g.nx);
        const long idx = (long)j * plane + (long)z * rowlen + x;
        if (op == 0) {
            fa[idx] = fa[idx] * fa[idx];
            fbw[idx] = fbw[idx] * fbw[idx];
            fc[idx] = fc[idx] * fc[idx];
        } else {
            fa[idx] = 2. * fa[idx] * fb[idx];
        }
    }
}

// cs (kept layout, planes 2..NY) = [cs +] a (+ b + c) from Fortran-layout
// Fourier planes (for/channel.f:1261-1367).
__global__ void k_pp_accum(Geo g, cx<double> *cs, const double *fa, const double *fb, const double *fc, int first)
{
    const int j = 2 + blockIdx.y;
    const int nxp = g.nx / 2;
    const long plane = (long)(nxp + 1) * (g.nz + 2);
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < g.ps; m += (long)gridDim.x * blockDim.x) {
        const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
        const long idx = ((long)j * plane + (long)K * (nxp + 1) + i) * 2;
        cx<double> v = make_cx<double>(fa[idx], fa[idx + 1]);
        if (fb) v = v + make_cx<double>(fb[idx], fb[idx + 1]);
        if (fc) v = v + make_cx<double>(fc[idx], fc[idx + 1]);
        cx<double> &o = cs[(long)j * g.ps + m];
        o = first ? v : o + v;
    }
}

// Buoyancy term, the tridiagonal solve and CP (for/channel.f:1373-1449).
__global__ void k_pp_solve(Geo g, const cx<double> *__restrict__ cs, const cx<double> *__restrict__ cth,
                           double ri, const double *__restrict__ gyf, const double *__restrict__ pmult,
                           const double *__restrict__ pinv, const double *__restrict__ rd, cx<double> *scr,
                           cx<double> *cp)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.ps) return;
    const bool is00 = (m == 0);
    const int NY = g.ny;
    const long ps = g.ps;
    cx<double> gprev = make_cx<double>(0, 0);
    scr[m + ps] = gprev;
    for (int j = 2; j <= NY; ++j) {
        const long o = m + (long)j * ps;
        cx<double> vec = make_cx<double>(0, 0);
        if (j < NY) {
            cx<double> s = cs[o];
            if (cth && !is00) {
                const cx<double> d = cth[o + ps] - cth[o - ps];
                const double den = gyf[j + 1] - gyf[j - 1];
                s = s + make_cx<double>((ri * d.re) / den, (ri * d.im) / den);
            }
            vec = make_cx<double>(-1. * s.re, -1. * s.im);
        }
        const double mu = pmult[o];
        const cx<double> gj = make_cx<double>(vec.re + mu * gprev.re, vec.im + mu * gprev.im);
        scr[o] = gj;
        gprev = gj;
    }
    cx<double> phip = make_cx<double>(0, 0);
    const double u1row = is00 ? 0.0 : -1.0;
    for (int j = NY; j >= 1; --j) {
        const long o = m + (long)j * ps;
        const cx<double> gj = scr[o];
        const double ip = pinv[o];
        cx<double> phi;
        if (j == NY) {
            phi = make_cx<double>(gj.re * ip, gj.im * ip);
        } else {
            const double U = (j == 1) ? u1row : rd[(long)j * 4 + 0];
            phi = make_cx<double>((gj.re - U * phip.re) * ip, (gj.im - U * phip.im) * ip);
        }
        cp[o] = phi;
        phip = phi;
    }
}

void launch_pp_spec(hipStream_t s, int group, const Geo &g, const cx<double> *cu1, const cx<double> *cu2,
                    const cx<double> *cu3, const double *dyf, const double *kx, const double *kz, double *fa,
                    double *fb, double *fc)
{
    const long plane = (long)(g.nx / 2 + 1) * (g.nz + 2);
    hipLaunchKernelGGL(k_pp_spec, dim3((unsigned)((plane + 255) / 256), g.ny + 2), dim3(256), 0, s, group, g, cu1,
                       cu2, cu3, dyf, kx, kz, fa, fb, fc);
}

void launch_pp_phys(hipStream_t s, int op, const Geo &g, double *fa, const double *fb, double *fbw, double *fc)
{
    const long n = (long)g.nz * g.nx;
    hipLaunchKernelGGL(k_pp_phys, dim3((unsigned)((n + 255) / 256), g.ny - 1), dim3(256), 0, s, op, g, fa, fb, fbw,
                       fc);
}

void launch_pp_accum(hipStream_t s, const Geo &g, cx<double> *cs, const double *fa, const double *fb,
                     const double *fc, int first)
{
    hipLaunchKernelGGL(k_pp_accum, dim3((unsigned)((g.ps + 255) / 256), g.ny - 1), dim3(256), 0, s, g, cs, fa, fb,
                       fc, first);
}

void launch_pp_solve(hipStream_t s, const Geo &g, const cx<double> *cs, const cx<double> *cth, double ri,
                     const double *gyf, const double *pmult, const double *pinv, const double *rd, cx<double> *scr,
                     cx<double> *cp)
{
    hipLaunchKernelGGL(k_pp_solve, dim3((unsigned)((g.ps + 63) / 64)), dim3(64), 0, s, g, cs, cth, ri, gyf, pmult,
                       pinv, rd, scr, cp);
}

}  // namespace dgpu
