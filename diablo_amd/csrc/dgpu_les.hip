// dgpu_les.hip -- les_chan's constant Smagorinsky model (for/les.f:1-331) on
// the device, in FP64, on Fortran-layout (0:NX+1, 0:NZ+1, 0:NY+1) planes:
//
//   k_les_bc      APPLY_BC_LES (for/les.f:805-864) on the kept modes
//   k_les_strain  S_ij of compute_strain_chan (for/les.f:526-571) in Fourier
//                 space, and CU1, CU3, written as Fortran-layout Fourier
//                 planes for the standalone transforms
//   k_les_nut     |S| on GYF and GY, |S| S_ij with the implicit parts taken
//                 out, NU_T on GY with GHOST_LES's zero rows
//                 (for/les.f:107-230, for/mpi.f:488-503)
//   k_les_cf      tau_ij = DELTA |S| S_ij and its divergence into CF1-3
//                 (for/les.f:236-301)
//   k_les_cross   NU_T dU/dy for the explicit cross terms, and their i k
//                 into CF2 (for/les.f:302-331)
#include "dgpu_kernels.hpp"

namespace dgpu {

namespace {

__device__ __forceinline__ long fidx3(int i, int k, int j, int nx, int nz)
{
    return (long)i + (long)(nx + 2) * ((long)k + (long)(nz + 2) * (long)j);
}

// Fourier-view slot (i, K) of plane j of a Fortran-layout field
__device__ __forceinline__ long cidx3(int i, int K, int j, int nxp, int nz)
{
    return (long)i + (long)(nxp + 1) * ((long)K + (long)(nz + 2) * (long)j);
}

__device__ __forceinline__ cx<double> i_times(double k, cx<double> a) { return make_cx<double>(-k * a.im, k * a.re); }

__device__ __forceinline__ double sq(double x) { return x * x; }

}  // namespace

__global__ void k_les_bc(Geo g, cx<double> *__restrict__ cu1, cx<double> *__restrict__ cu3, int u_min, int u_max,
                         int w_min, int w_max)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.ps) return;
    const int NY = g.ny;
    const long ps = g.ps;
    auto top = [&](cx<double> *a) {
        a[m + (long)NY * ps] = a[m + (long)(NY - 1) * ps];
        a[m + (long)(NY + 1) * ps] = a[m + (long)NY * ps];
    };
    auto bottom = [&](cx<double> *a) {
        a[m + ps] = a[m + 2 * ps];
        a[m] = a[m + ps];
    };
    if (u_max == 1) top(cu1);
    if (w_max == 1) top(cu3);
    if (u_min == 1) bottom(cu1);
    if (w_min == 1) bottom(cu3);
}

// out[q]: Fortran-layout Fourier planes; q = 0..5 S_11, S_22, S_33, S_12,
// S_13, S_23 (the reference's Sij1..Sij6), 6 CU1, 7 CU3.  Slots outside the
// kept modes and planes outside each S_ij's rows are zero.
__global__ void k_les_strain(Geo g, const cx<double> *__restrict__ cu1, const cx<double> *__restrict__ cu2,
                             const cx<double> *__restrict__ cu3, const double *__restrict__ kxv,
                             const double *__restrict__ kzv, const double *__restrict__ dy,
                             const double *__restrict__ dyf, LesPtrs out)
{
    const int nxp = g.nx / 2, nz = g.nz, NY = g.ny;
    const long plane = (long)(nxp + 1) * (nz + 2);
    const int j = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= plane) return;
    const int K = (int)(e / (nxp + 1)), i = (int)(e - (long)K * (nxp + 1));
    const bool kept = i < g.mk && K < g.mz;
    const long ps = g.ps;
    const cx<double> zero = make_cx<double>(0.0, 0.0);
    cx<double> s[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) s[q] = zero;
    if (kept) {
        const long m = (long)K * g.mk + i;
        const double kx = kxv[i], kz = kzv[K];
        auto at = [&](const cx<double> *f, int jj) { return f[m + (long)jj * ps]; };
        const cx<double> u1 = at(cu1, j), u3 = at(cu3, j), u2 = at(cu2, j);
        if (j >= 1 && j <= NY) {
            const cx<double> u2p = at(cu2, j + 1);
            s[0] = i_times(kx, u1);
            s[1] = make_cx<double>((u2p.re - u2.re) / dyf[j], (u2p.im - u2.im) / dyf[j]);
            s[2] = i_times(kz, u3);
            const cx<double> a = i_times(kz, u1), b = i_times(kx, u3);
            s[4] = make_cx<double>(0.5 * (a.re + b.re), 0.5 * (a.im + b.im));
        }
        if (j >= 1) {
            const cx<double> u1m = at(cu1, j - 1), u3m = at(cu3, j - 1);
            const cx<double> a = i_times(kx, u2), b = i_times(kz, u2);
            s[3] = make_cx<double>(0.5 * ((u1.re - u1m.re) / dy[j] + a.re), 0.5 * ((u1.im - u1m.im) / dy[j] + a.im));
            s[5] = make_cx<double>(0.5 * (b.re + (u3.re - u3m.re) / dy[j]), 0.5 * (b.im + (u3.im - u3m.im) / dy[j]));
        }
        s[6] = u1;
        s[7] = u3;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        double *p = out.p[q] + 2 * ((long)j * plane + e);
        p[0] = s[q].re;
        p[1] = s[q].im;
    }
}

// Physical planes: in[0..5] S_ij, in[6] U1, in[7] U3; out[0..4] the stress
// parts to transform (|S| S_11, |S| S_33, TEMP (S_12 - dU1/dy / 2),
// |S| S_13, TEMP (S_23 - dU3/dy / 2)), out[5] NU_T.  ubar, wbar: the (0,0)
// modes of CU1, CU3.  delta_y: -2 C^2 (DX DY(j) 2 DZ)^(2/3).
__global__ void k_les_nut(Geo g, LesPtrs in, LesPtrs out,
                          const cx<double> *__restrict__ cu1, const cx<double> *__restrict__ cu3,
                          const double *__restrict__ dy, const double *__restrict__ dyf,
                          const double *__restrict__ delta_y, int jstart, int jend)
{
    const int nx = g.nx, nz = g.nz, NY = g.ny;
    const long n = (long)nx * nz;
    const int j = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e % nx), k = (int)(e / nx);
    const long ps = g.ps;
    const long x = fidx3(i, k, j, nx, nz);
    const double *S11 = in.p[0], *S22 = in.p[1], *S33 = in.p[2], *S12 = in.p[3], *S13 = in.p[4], *S23 = in.p[5];
    const double *U1 = in.p[6], *U3 = in.p[7];
    auto ub = [&](int jj) { return cu1[(long)jj * ps].re; };
    auto wb = [&](int jj) { return cu3[(long)jj * ps].re; };
    double o11 = S11[x], o33 = S33[x], o13 = S13[x], o12 = S12[x], o23 = S23[x], nut = 0.0;
    if (j >= jstart && j <= jend) {  // |S| on GYF (for/les.f:107-130)
        const long xp = fidx3(i, k, j + 1, nx, nz);
        const double s1 = sqrt(2.0 * sq(S11[x])
                               + 4.0 * sq(0.5 * (S12[xp] + S12[x]) - 0.5 * (ub(j) - ub(j - 1)) / dy[j]
                                          - 0.5 * (ub(j + 1) - ub(j)) / dy[j + 1])
                               + 4.0 * sq(S13[x]) + 2.0 * sq(S22[x])
                               + 4.0 * sq(0.5 * (S23[xp] + S23[x]) - 0.5 * (wb(j) - wb(j - 1)) / dy[j]
                                          - 0.5 * (wb(j + 1) - wb(j)) / dy[j + 1])
                               + 2.0 * sq(S33[x]));
        o11 = s1 * S11[x];
        o13 = s1 * S13[x];
        o33 = s1 * S33[x];
    }
    if (j >= 2 && j <= NY) {  // |S| on GY (for/les.f:133-156), then NU_T
        const long xm = fidx3(i, k, j - 1, nx, nz);
        const double w0 = dyf[j - 1], w1 = dyf[j], d2 = 2.0 * dy[j];
        const double temp = sqrt(2.0 * sq((S11[x] * w0 + S11[xm] * w1) / d2)
                                 + 4.0 * sq(S12[x] - 0.5 * (ub(j) - ub(j - 1)) / dy[j])
                                 + 4.0 * sq((S13[x] * w0 + S13[xm] * w1) / d2)
                                 + 2.0 * sq((S22[x] * w0 + S22[xm] * w1) / d2)
                                 + 4.0 * sq(S23[x] - 0.5 * (wb(j) - wb(j - 1)) / dy[j])
                                 + 2.0 * sq((S33[x] * w0 + S33[xm] * w1) / d2));
        o12 = temp * (S12[x] - 0.5 * (U1[x] - U1[xm]) / dy[j]);
        o23 = temp * (S23[x] - 0.5 * (U3[x] - U3[xm]) / dy[j]);
        if (j >= 3 && j <= NY - 1) nut = -0.5 * delta_y[j] * temp;  // rows 1, 2, NY, NY+1 stay 0
    }
    out.p[0][x] = o11;
    out.p[1][x] = o33;
    out.p[2][x] = o12;
    out.p[3][x] = o13;
    out.p[4][x] = o23;
    out.p[5][x] = nut;
}

// CF -= div tau (for/les.f:236-301): t[0..4] the transformed parts of
// k_les_nut (Fortran-layout Fourier), scaled here by delta_yf / delta_y.
// fresh: cf1..cf3 start from zero and every row 0..NY+1 is written (the
// rows outside the force's ranges with zero), so the buffers need no clear.
__global__ void k_les_cf(Geo g, LesPtrs t, cx<double> *__restrict__ cf1,
                         cx<double> *__restrict__ cf2, cx<double> *__restrict__ cf3, const double *__restrict__ kxv,
                         const double *__restrict__ kzv, const double *__restrict__ dyf,
                         const double *__restrict__ delta_y, const double *__restrict__ delta_yf, int j1, int j2,
                         int fresh)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    if (m >= g.ps) return;
    const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
    const int nxp = g.nx / 2, nz = g.nz, NY = g.ny;
    const long ps = g.ps;
    const double kx = kxv[i], kz = kzv[K];
    auto T = [&](int q, int jj, double scale) {
        const cx<double> *a = reinterpret_cast<const cx<double> *>(t.p[q]);
        const cx<double> v = a[cidx3(i, K, jj, nxp, nz)];
        return make_cx<double>(scale * v.re, scale * v.im);
    };
    const long o = m + (long)j * ps;
    const cx<double> zero = make_cx<double>(0.0, 0.0);
    if (j >= j1 && j <= j2) {
        const cx<double> c11 = T(0, j, delta_yf[j]), c13 = T(3, j, delta_yf[j]), c33 = T(1, j, delta_yf[j]);
        const double dyj = j >= 2 ? delta_y[j] : 1.0;  // rows from 2 on are scaled (for/les.f:262-270)
        const cx<double> c12 = T(2, j, dyj), c12p = T(2, j + 1, delta_y[j + 1]);
        const cx<double> c23 = T(4, j, dyj), c23p = T(4, j + 1, delta_y[j + 1]);
        cx<double> f = fresh ? zero : cf1[o];
        const cx<double> a = i_times(kx, c11), b = i_times(kz, c13);
        f = make_cx<double>(f.re - a.re - (c12p.re - c12.re) / dyf[j] - b.re,
                            f.im - a.im - (c12p.im - c12.im) / dyf[j] - b.im);
        cf1[o] = f;
        cx<double> h = fresh ? zero : cf3[o];
        const cx<double> c = i_times(kx, c13), d = i_times(kz, c33);
        h = make_cx<double>(h.re - c.re - (c23p.re - c23.re) / dyf[j] - d.re,
                            h.im - c.im - (c23p.im - c23.im) / dyf[j] - d.im);
        cf3[o] = h;
    } else if (fresh) {
        cf1[o] = zero;
        cf3[o] = zero;
    }
    if (j >= 2 && j <= NY) {
        const cx<double> c12 = T(2, j, delta_y[j]), c23 = T(4, j, delta_y[j]);
        const cx<double> a = i_times(kx, c12), b = i_times(kz, c23);
        cx<double> f = fresh ? zero : cf2[o];
        f = make_cx<double>(f.re - a.re - b.re, f.im - a.im - b.im);
        cf2[o] = f;
    } else if (fresh) {
        cf2[o] = zero;
    }
}

// NU_T (U(j) - U(j-1)) / DY(j) on rows 2..NY (physical), for the cross terms
__global__ void k_les_cross_src(Geo g, const double *__restrict__ nut, const double *__restrict__ u,
                                const double *__restrict__ dy, double *__restrict__ out)
{
    const int nx = g.nx, nz = g.nz;
    const long n = (long)nx * nz;
    const int j = 2 + blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e % nx), k = (int)(e / nx);
    const long x = fidx3(i, k, j, nx, nz), xm = fidx3(i, k, j - 1, nx, nz);
    out[x] = nut[x] * (u[x] - u[xm]) / dy[j];
}

// CF2 += i k CT on rows 2..NY (k = kx for pass 0, kz for pass 1)
__global__ void k_les_cross_add(Geo g, const double *__restrict__ ct, cx<double> *__restrict__ cf2,
                                const double *__restrict__ kxv, const double *__restrict__ kzv, int use_kz)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = 2 + blockIdx.y;
    if (m >= g.ps) return;
    const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
    const int nxp = g.nx / 2, nz = g.nz;
    const double k = use_kz ? kzv[K] : kxv[i];
    const cx<double> v = reinterpret_cast<const cx<double> *>(ct)[cidx3(i, K, j, nxp, nz)];
    const cx<double> a = i_times(k, v);
    const long o = m + (long)j * g.ps;
    cx<double> f = cf2[o];
    f = make_cx<double>(f.re + a.re, f.im + a.im);
    cf2[o] = f;
}

// ---------------------------------------------------------------------------
// RK_CHAN_1 with LES: the Crank-Nicolson solves in physical space
// ---------------------------------------------------------------------------

// APPLY_BC_VEL_PHYS_LOWER / _UPPER (for/channel.f:2495-2640) stated on the
// kept modes: a value C in physical space is C on the (0,0) mode and 0 on
// the others, and the Neumann rows are linear in the field.
__global__ void k_les_bc_phys(Geo g, cx<double> *__restrict__ cu1, cx<double> *__restrict__ cu2,
                              cx<double> *__restrict__ cu3, LesBc b, const double *__restrict__ dy,
                              const double *__restrict__ dyf)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.ps) return;
    const int NY = g.ny;
    const long ps = g.ps;
    const double on = (m == 0) ? 1.0 : 0.0;
    auto R = [&](cx<double> *a, int j) -> cx<double> & { return a[m + (long)j * ps]; };
    auto cst = [&](double c) { return make_cx<double>(on * c, 0.0); };
    auto plus = [&](cx<double> a, double c) { return make_cx<double>(a.re + on * c, a.im); };
    for (int q = 0; q < 2; ++q) {
        cx<double> *u = q == 0 ? cu1 : cu3;
        const int lo = q == 0 ? b.u_lo : b.w_lo, hi = q == 0 ? b.u_hi : b.w_hi;
        const double clo = q == 0 ? b.u_lo_c : b.w_lo_c, chi = q == 0 ? b.u_hi_c : b.w_hi_c;
        if (lo == 0) {
            R(u, 0) = cst(clo);
            R(u, 1) = cst(clo);
        } else {
            R(u, 1) = plus(R(u, 2), -dy[2] * clo);
            R(u, 0) = plus(R(u, 1), -dy[1] * clo);
        }
        if (hi == 0) {
            R(u, NY) = cst(chi);
            R(u, NY + 1) = cst(chi);
        } else {
            R(u, NY) = plus(R(u, NY - 1), dy[NY] * chi);
            R(u, NY + 1) = plus(R(u, NY), dy[NY] * chi);
        }
    }
    if (b.v_lo == 0) {
        const cx<double> u2 = R(cu2, 2);
        R(cu2, 1) = make_cx<double>(2.0 * on * b.v_lo_c - u2.re, -u2.im);
        R(cu2, 0) = R(cu2, 1);
    } else {
        R(cu2, 1) = plus(R(cu2, 2), -dyf[1] * b.v_lo_c);
        R(cu2, 0) = plus(R(cu2, 1), -dyf[1] * b.v_lo_c);
    }
    if (b.v_hi == 0) {
        const cx<double> u2 = R(cu2, NY);
        R(cu2, NY + 1) = make_cx<double>(2.0 * on * b.v_hi_c - u2.re, -u2.im);
    } else {
        R(cu2, NY + 1) = plus(R(cu2, NY), dy[NY] * b.v_hi_c);
    }
}

// APPLY_BC_VEL_PHYS_LOWER / _UPPER's U1 and U3 rows (for/channel.f:2495-2640)
// on physical Fortran-layout planes.
__global__ void k_les_bc_phys_planes(Geo g, double *__restrict__ u1, double *__restrict__ u3, LesBc b,
                                     const double *__restrict__ dy)
{
    const long n = (long)g.nx * g.nz;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e % g.nx), k = (int)(e / g.nx), NY = g.ny;
    auto X = [&](int j) { return fidx3(i, k, j, g.nx, g.nz); };
    for (int q = 0; q < 2; ++q) {
        double *u = q == 0 ? u1 : u3;
        const int lo = q == 0 ? b.u_lo : b.w_lo, hi = q == 0 ? b.u_hi : b.w_hi;
        const double clo = q == 0 ? b.u_lo_c : b.w_lo_c, chi = q == 0 ? b.u_hi_c : b.w_hi_c;
        if (lo == 0) {
            u[X(0)] = clo;
            u[X(1)] = clo;
        } else {
            u[X(1)] = u[X(2)] - dy[2] * clo;
            u[X(0)] = u[X(1)] - dy[1] * clo;
        }
        if (hi == 0) {
            u[X(NY)] = chi;
            u[X(NY + 1)] = chi;
        } else {
            u[X(NY)] = u[X(NY - 1)] + dy[NY] * chi;
            u[X(NY + 1)] = u[X(NY)] + dy[NY] * chi;
        }
    }
}

void launch_les_bc_phys_planes(hipStream_t s, const Geo &g, double *u1, double *u3, const LesBc &b, const double *dy)
{
    const long n = (long)g.nx * g.nz;
    hipLaunchKernelGGL(k_les_bc_phys_planes, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, u1, u3, b, dy);
}

// Kept modes (plane stride ps) -> Fortran-layout Fourier planes, times i kx
// (kmul 1) or i kz (kmul 2); planes outside j0..j1 and slots outside the
// kept modes are zero.
__global__ void k_les_pack(Geo g, const cx<double> *__restrict__ src, double *__restrict__ fort,
                           const double *__restrict__ kxv, const double *__restrict__ kzv, int kmul, int j0, int j1)
{
    const int nxp = g.nx / 2, nz = g.nz;
    const long plane = (long)(nxp + 1) * (nz + 2);
    const int j = blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= plane) return;
    const int K = (int)(e / (nxp + 1)), i = (int)(e - (long)K * (nxp + 1));
    cx<double> v = make_cx<double>(0.0, 0.0);
    if (i < g.mk && K < g.mz && j >= j0 && j <= j1) {
        v = src[(long)K * g.mk + i + (long)j * g.ps];
        if (kmul == 1) v = i_times(kxv[i], v);
        else if (kmul == 2) v = i_times(kzv[K], v);
    }
    double *p = fort + 2 * ((long)j * plane + e);
    p[0] = v.re;
    p[1] = v.im;
}

// Physical planes j0..j1: a *= nut
__global__ void k_les_mul(Geo g, double *__restrict__ a, const double *__restrict__ nut, int j0)
{
    const long n = (long)g.nx * g.nz;
    const int j = j0 + blockIdx.y;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i = (int)(e % g.nx), k = (int)(e / g.nx);
    const long x = fidx3(i, k, j, g.nx, g.nz);
    a[x] = nut[x] * a[x];
}

// dst (kept modes) += (i kx | i kz | 1) * Fortran-layout Fourier planes, j0..j1
__global__ void k_les_acc(Geo g, cx<double> *__restrict__ dst, const double *__restrict__ fort,
                          const double *__restrict__ kxv, const double *__restrict__ kzv, int kmul, int j0)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int j = j0 + blockIdx.y;
    if (m >= g.ps) return;
    const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
    cx<double> v = reinterpret_cast<const cx<double> *>(fort)[cidx3(i, K, j, g.nx / 2, g.nz)];
    if (kmul == 1) v = i_times(kxv[i], v);
    else if (kmul == 2) v = i_times(kzv[K], v);
    const long o = m + (long)j * g.ps;
    cx<double> f = dst[o];
    f = make_cx<double>(f.re + v.re, f.im + v.im);
    dst[o] = f;
}

// One implicit Crank-Nicolson solve in physical space per (x, z) column
// (for/channel.f:762-1029 with the LES terms): rows jlo..jhi from the
// molecular and eddy parts, the boundary rows of APPLY_BC_*_LOWER/UPPER
// (for/channel.f:1660-2282), THOMAS_REAL at NPROCY = 1 (for/mpi.f:630-809),
// and rows clo..chi written to out.  bm, gw: per-column scratch planes.
// With EXPL the right-hand side of rows jlo..jhi also takes the explicit
// eddy-viscosity (diffusivity) half, temp2 d/dy(nut du/dy)
// (for/channel.f:628-659, 748-760), from the field u at rows j-1..j+1: on
// GYF rows with NU_T on the faces (U1, U3, TH), or on GY rows with NU_T
// averaged onto the GYF faces (U2).
template <bool EXPL>
__device__ __forceinline__ void les_row(const LesCn &a, const double *__restrict__ nut, const double *u, long x,
                                        long xp, long xm, const double *__restrict__ dy,
                                        const double *__restrict__ dyf, const double *__restrict__ rhs, int j,
                                        double &L, double &D, double &U, double &V)
{
    const int NY = a.ny;
    L = 0.0; D = 1.0; U = 0.0; V = 0.0;
    if (j >= a.jlo && j <= a.jhi) {
        const double c = a.cmol;
        if (a.sys == LES_SYS_TH) {
            L = -c / (dy[j] * dyf[j]);
            D = 1. + c / (dy[j + 1] * dyf[j]) + c / (dy[j] * dyf[j]);
            U = -c / (dy[j + 1] * dyf[j]);
        } else if (a.sys == LES_SYS_U2) {
            L = -c / (dyf[j - 1] * dy[j]);
            D = 1. + c / (dyf[j] * dy[j]) + c / (dyf[j - 1] * dy[j]);
            U = -c / (dyf[j] * dy[j]);
        } else {
            L = -c / (dy[j] * dyf[j]);
            D = 1. - c * (-1. / (dy[j + 1] * dyf[j]) - 1. / (dy[j] * dyf[j]));
            U = -c / (dy[j + 1] * dyf[j]);
        }
        V = rhs[x];
        const double t2 = a.temp2;
        if constexpr (EXPL) {
            double t;
            if (a.sys == LES_SYS_U2)
                t = (0.5 * (nut[x] + nut[xp]) * (u[xp] - u[x]) / dyf[j] -
                     0.5 * (nut[x] + nut[xm]) * (u[x] - u[xm]) / dyf[j - 1]) /
                    dy[j];
            else
                t = (nut[xp] * (u[xp] - u[x]) / dy[j + 1] - nut[x] * (u[x] - u[xm]) / dy[j]) / dyf[j];
            V = V + t2 * t;
        }
        if (a.sys == LES_SYS_U2) {
            const double nm = nut[x] + nut[xm], np = nut[x] + nut[xp];
            L = L - t2 * 0.5 * nm / (dyf[j - 1] * dy[j]);
            D = D + t2 * 0.5 * np / (dyf[j] * dy[j]) + t2 * 0.5 * nm / (dyf[j - 1] * dy[j]);
            U = U - t2 * 0.5 * np / (dyf[j] * dy[j]);
        } else {
            const double n0 = nut[x], n1 = nut[xp];
            L = L - t2 * n0 / (dy[j] * dyf[j]);
            D = D + t2 * n1 / (dy[j + 1] * dyf[j]) + t2 * n0 / (dy[j] * dyf[j]);
            U = U - t2 * n1 / (dy[j + 1] * dyf[j]);
        }
    }
    auto set = [&](double l, double d, double u, double v) { L = l; D = d; U = u; V = v; };
    if (a.sys == LES_SYS_U2) {
        if (j == 0) set(0., 1., 0., 0.);
        if (a.lo == 0) {
            if (j == 1 || j == 2) set(0., 1., 0., a.clo_c);
        } else if (j == 1) {
            set(0., -1., 1., dyf[1] * a.clo_c);
        }
        if (a.hi == 0) {
            if (j == NY || j == NY + 1) set(0., 1., 0., a.chi_c);
        } else if (j == NY + 1) {
            set(-1., 1., 0., dyf[NY] * a.chi_c);
        }
    } else {
        if (j == 0) set(0., 1., 0., 0.);
        if (j == 1) {
            if (a.lo == 0) set(0., 1., 0., a.clo_c);
            else set(0., -1., 1., dy[2] * a.clo_c);
        }
        if (j == NY + 1) set(0., 1., 0., 0.);
        if (j == NY) {
            if (a.hi == 0) set(0., 1., 0., a.chi_c);
            else set(-1., 1., 0., dy[NY] * a.chi_c);
        }
    }
}

// grid (column blocks, systems): every system of the substep in one launch
// (they are independent), so the sweeps of all four fill the machine.  The
// system's field is read (the explicit half, forward sweep) and overwritten
// (back substitution) through out: a column is read and written by its own
// thread only, every read before the first write.
__global__ void k_les_cn(Geo g, LesCnSet set, const double *__restrict__ nut, const double *__restrict__ dy,
                         const double *__restrict__ dyf)
{
    const long n = (long)g.nx * g.nz;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int sy = blockIdx.y;
    const LesCn &a = set.a[sy];
    const double *__restrict__ rhs = set.rhs[sy];
    double *__restrict__ bm = set.bm[sy];
    double *__restrict__ gw = set.gw[sy];
    double *__restrict__ out = set.out[sy];
    const int i = (int)(e % g.nx), k = (int)(e / g.nx);
    const int NY = g.ny;
    auto X = [&](int j) { return fidx3(i, k, j, g.nx, g.nz); };
    // forward elimination, rows 1..NY+1 (THOMAS_FORWARD_REAL_MPI)
    double L, D, U, V;
    les_row<true>(a, nut, out, X(0), X(1), X(0), dy, dyf, rhs, 0, L, D, U, V);
    double bprev = D, cprev = U, gprev = V;
    bm[X(0)] = D;
    gw[X(0)] = V;
    for (int j = 1; j <= NY + 1; ++j) {
        const long x = X(j);
        les_row<true>(a, nut, out, x, X(j <= NY ? j + 1 : j), X(j - 1), dy, dyf, rhs, j, L, D, U, V);
        const double m = -L / bprev;
        D = D + m * cprev;
        V = V + m * gprev;
        bm[x] = D;
        gw[x] = V;
        bprev = D;
        cprev = U;
        gprev = V;
    }
    // back substitution with the zero-pivot guard (THOMAS_BACKWARD_REAL_MPI)
    double xn = gprev / bprev;
    if (NY + 1 >= a.clo && NY + 1 <= a.chi) out[X(NY + 1)] = xn;
    for (int j = NY; j >= 0; --j) {
        const long x = X(j);
        les_row<false>(a, nut, out, x, X(j + 1), X(j > 0 ? j - 1 : 0), dy, dyf, rhs, j, L, D, U, V);
        const double b = bm[x];
        xn = (b != 0.0) ? (gw[x] - U * xn) / b : 0.0;
        if (j >= a.clo && j <= a.chi) out[x] = xn;
    }
}

// Per-plane maxima of |U1|, |U2|, |U3| on physical planes 1..NY into
// out[q * NY + j - 1] (COURANT's input, for/courant.f:25-45)
__global__ void __launch_bounds__(256) k_les_vmax(Geo g, const double *__restrict__ u1, const double *__restrict__ u2,
                                                  const double *__restrict__ u3, double *__restrict__ out)
{
    __shared__ double red[3][256];
    const int j = 1 + blockIdx.x, tid = threadIdx.x;
    const long n = (long)g.nx * g.nz;
    double m[3] = {0.0, 0.0, 0.0};
    for (long e = tid; e < n; e += 256) {
        const int i = (int)(e % g.nx), k = (int)(e / g.nx);
        const long x = fidx3(i, k, j, g.nx, g.nz);
        m[0] = fmax(m[0], fabs(u1[x]));
        m[1] = fmax(m[1], fabs(u2[x]));
        m[2] = fmax(m[2], fabs(u3[x]));
    }
    for (int q = 0; q < 3; ++q) red[q][tid] = m[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h)
            for (int q = 0; q < 3; ++q) red[q][tid] = fmax(red[q][tid], red[q][tid + h]);
        __syncthreads();
    }
    if (tid < 3) out[(long)tid * g.ny + (j - 1)] = red[tid][0];
}

// The reference forms CF from the velocity before les_chan resets its
// boundary rows (for/channel.f:175-290 precede for/channel.f:303); the
// column kernel saw the reset rows.  On rows 1 and NY (pre: the saved rows
// of CU1 and CU3, plane 0 row 1, plane 1 row NY, then the same for CU3) the
// linear terms that read them are corrected: horizontal viscosity and
// Coriolis in CF1, CF3, the background-gradient terms in CFTH.
__global__ void k_les_cf_rows(Geo g, const cx<double> *__restrict__ pre, const cx<double> *__restrict__ cu1,
                              const cx<double> *__restrict__ cu3, cx<double> *__restrict__ cf1,
                              cx<double> *__restrict__ cf3, cx<double> *__restrict__ cfth,
                              const double *__restrict__ kx2, const double *__restrict__ kz2, LesCfFix f)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g.ps) return;
    const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
    const long ps = g.ps;
    const double k2 = kx2[i] + kz2[K];
    const double visc = f.nu * ((f.beta == 1.0) ? k2 : pow(k2, f.beta));
    for (int r = 0; r < 2; ++r) {
        const int j = r == 0 ? 1 : g.ny;
        const long o = m + (long)j * ps;
        const cx<double> a = pre[m + (long)r * ps], b = pre[m + (long)(2 + r) * ps];
        const cx<double> d1 = make_cx<double>(a.re - cu1[o].re, a.im - cu1[o].im);
        const cx<double> d3 = make_cx<double>(b.re - cu3[o].re, b.im - cu3[o].im);
        if (j >= f.jstart && j <= f.jend) {
            cx<double> x = cf1[o], z = cf3[o];
            x = make_cx<double>(x.re - visc * d1.re + f.i_ro * d3.re, x.im - visc * d1.im + f.i_ro * d3.im);
            z = make_cx<double>(z.re - visc * d3.re - f.i_ro * d1.re, z.im - visc * d3.im - f.i_ro * d1.im);
            cf1[o] = x;
            cf3[o] = z;
        }
        if (cfth && j >= f.jstart_th && j <= f.jend_th) {
            cx<double> t = cfth[o];
            t = make_cx<double>(t.re - (d1.re * f.drhodx + d3.re * f.drhodz), t.im - (d1.im * f.drhodx + d3.im * f.drhodz));
            cfth[o] = t;
        }
    }
}

void launch_les_cf_rows(hipStream_t s, const Geo &g, const cx<double> *pre, const cx<double> *cu1,
                        const cx<double> *cu3, cx<double> *cf1, cx<double> *cf3, cx<double> *cfth, const double *kx2,
                        const double *kz2, const LesCfFix &f)
{
    hipLaunchKernelGGL(k_les_cf_rows, dim3((unsigned)((g.ps + 255) / 256)), dim3(256), 0, s, g, pre, cu1, cu3, cf1,
                       cf3, cfth, kx2, kz2, f);
}

// tkebudget_chan_les (for/save_stats.f:887-960): per plane j = 1..NY the sum
// over the physical points of u_i times its implicit eddy-viscosity term;
// out[j - 1].  Fixed-order reduction.
__global__ void __launch_bounds__(256) k_les_eps(Geo g, const double *__restrict__ u1, const double *__restrict__ u2,
                                                 const double *__restrict__ u3, const double *__restrict__ nut,
                                                 const double *__restrict__ dy, const double *__restrict__ dyf,
                                                 double *__restrict__ out)
{
    __shared__ double red[256];
    const int j = 1 + blockIdx.x, tid = threadIdx.x;
    const long n = (long)g.nx * g.nz;
    double acc = 0.0;
    for (long e = tid; e < n; e += 256) {
        const int i = (int)(e % g.nx), k = (int)(e / g.nx);
        const long x = fidx3(i, k, j, g.nx, g.nz), xp = fidx3(i, k, j + 1, g.nx, g.nz),
                   xm = fidx3(i, k, j - 1, g.nx, g.nz);
        const double t = u1[x] * ((nut[xp] * (u1[xp] - u1[x]) / dy[j + 1] - nut[x] * (u1[x] - u1[xm]) / dy[j]) / dyf[j]) +
                         u3[x] * ((nut[xp] * (u3[xp] - u3[x]) / dy[j + 1] - nut[x] * (u3[x] - u3[xm]) / dy[j]) / dyf[j]) +
                         u2[x] * ((0.5 * (nut[x] + nut[xp]) * (u2[xp] - u2[x]) / dyf[j] -
                                   0.5 * (nut[x] + nut[xm]) * (u2[x] - u2[xm]) / dyf[j - 1]) /
                                  dy[j]);
        acc += t;
    }
    red[tid] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (tid < h) red[tid] += red[tid + h];
        __syncthreads();
    }
    if (tid == 0) out[j - 1] = red[0];
}

void launch_les_eps(hipStream_t s, const Geo &g, const double *u1, const double *u2, const double *u3,
                    const double *nut, const double *dy, const double *dyf, double *out)
{
    hipLaunchKernelGGL(k_les_eps, dim3(g.ny), dim3(256), 0, s, g, u1, u2, u3, nut, dy, dyf, out);
}

__global__ void k_les_add(cx<double> *__restrict__ dst, const cx<double> *__restrict__ src, long n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const cx<double> a = dst[e], b = src[e];
    dst[e] = make_cx<double>(a.re + b.re, a.im + b.im);
}

void launch_les_add(hipStream_t s, cx<double> *dst, const cx<double> *src, long n)
{
    hipLaunchKernelGGL(k_les_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
}

void launch_les_bc_phys(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu2, cx<double> *cu3,
                        const LesBc &b, const double *dy, const double *dyf)
{
    hipLaunchKernelGGL(k_les_bc_phys, dim3((unsigned)((g.ps + 255) / 256)), dim3(256), 0, s, g, cu1, cu2, cu3, b, dy,
                       dyf);
}

void launch_les_pack(hipStream_t s, const Geo &g, const cx<double> *src, double *fort, const double *kx,
                     const double *kz, int kmul, int j0, int j1)
{
    const long plane = (long)(g.nx / 2 + 1) * (g.nz + 2);
    hipLaunchKernelGGL(k_les_pack, dim3((unsigned)((plane + 255) / 256), g.ny + 2), dim3(256), 0, s, g, src, fort,
                       kx, kz, kmul, j0, j1);
}

void launch_les_mul(hipStream_t s, const Geo &g, double *a, const double *nut, int j0, int j1)
{
    const long n = (long)g.nx * g.nz;
    hipLaunchKernelGGL(k_les_mul, dim3((unsigned)((n + 255) / 256), j1 - j0 + 1), dim3(256), 0, s, g, a, nut, j0);
}

void launch_les_acc(hipStream_t s, const Geo &g, cx<double> *dst, const double *fort, const double *kx,
                    const double *kz, int kmul, int j0, int j1)
{
    hipLaunchKernelGGL(k_les_acc, dim3((unsigned)((g.ps + 255) / 256), j1 - j0 + 1), dim3(256), 0, s, g, dst, fort,
                       kx, kz, kmul, j0);
}

void launch_les_cn(hipStream_t s, const Geo &g, const LesCnSet &set, const double *nut, const double *dy,
                   const double *dyf)
{
    const long n = (long)g.nx * g.nz;
    hipLaunchKernelGGL(k_les_cn, dim3((unsigned)((n + 255) / 256), set.n), dim3(256), 0, s, g, set, nut, dy, dyf);
}

void launch_les_vmax(hipStream_t s, const Geo &g, const double *u1, const double *u2, const double *u3, double *out)
{
    hipLaunchKernelGGL(k_les_vmax, dim3(g.ny), dim3(256), 0, s, g, u1, u2, u3, out);
}

void launch_les_bc(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu3, int u_min, int u_max, int w_min,
                   int w_max)
{
    hipLaunchKernelGGL(k_les_bc, dim3((unsigned)((g.ps + 255) / 256)), dim3(256), 0, s, g, cu1, cu3, u_min, u_max,
                       w_min, w_max);
}

void launch_les_strain(hipStream_t s, const Geo &g, const cx<double> *cu1, const cx<double> *cu2,
                       const cx<double> *cu3, const double *kx, const double *kz, const double *dy, const double *dyf,
                       LesPtrs out)
{
    const long plane = (long)(g.nx / 2 + 1) * (g.nz + 2);
    hipLaunchKernelGGL(k_les_strain, dim3((unsigned)((plane + 255) / 256), g.ny + 2), dim3(256), 0, s, g, cu1, cu2,
                       cu3, kx, kz, dy, dyf, out);
}

void launch_les_nut(hipStream_t s, const Geo &g, LesPtrs in, LesPtrs out,
                    const cx<double> *cu1, const cx<double> *cu3, const double *dy, const double *dyf,
                    const double *delta_y, int jstart, int jend)
{
    const long n = (long)g.nx * g.nz;
    hipLaunchKernelGGL(k_les_nut, dim3((unsigned)((n + 255) / 256), g.ny + 2), dim3(256), 0, s, g, in, out, cu1, cu3,
                       dy, dyf, delta_y, jstart, jend);
}

void launch_les_cf(hipStream_t s, const Geo &g, LesPtrs t, cx<double> *cf1, cx<double> *cf2,
                   cx<double> *cf3, const double *kx, const double *kz, const double *dyf, const double *delta_y,
                   const double *delta_yf, int j1, int j2, bool fresh)
{
    hipLaunchKernelGGL(k_les_cf, dim3((unsigned)((g.ps + 255) / 256), g.ny + (fresh ? 2 : 1)), dim3(256), 0, s, g,
                       t, cf1, cf2, cf3, kx, kz, dyf, delta_y, delta_yf, j1, j2, fresh ? 1 : 0);
}

void launch_les_cross_src(hipStream_t s, const Geo &g, const double *nut, const double *u, const double *dy,
                          double *out)
{
    const long n = (long)g.nx * g.nz;
    hipLaunchKernelGGL(k_les_cross_src, dim3((unsigned)((n + 255) / 256), g.ny - 1), dim3(256), 0, s, g, nut, u, dy,
                       out);
}

void launch_les_cross_add(hipStream_t s, const Geo &g, const double *ct, cx<double> *cf2, const double *kx,
                          const double *kz, int use_kz)
{
    hipLaunchKernelGGL(k_les_cross_add, dim3((unsigned)((g.ps + 255) / 256), g.ny - 1), dim3(256), 0, s, g, ct, cf2,
                       kx, kz, use_kz);
}

}  // namespace dgpu
