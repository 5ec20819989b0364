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
__global__ void k_les_cf(Geo g, LesPtrs t, cx<double> *__restrict__ cf1,
                         cx<double> *__restrict__ cf2, cx<double> *__restrict__ cf3, const double *__restrict__ kxv,
                         const double *__restrict__ kzv, const double *__restrict__ dyf,
                         const double *__restrict__ delta_y, const double *__restrict__ delta_yf, int j1, int j2)
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
    if (j >= j1 && j <= j2) {
        const cx<double> c11 = T(0, j, delta_yf[j]), c13 = T(3, j, delta_yf[j]), c33 = T(1, j, delta_yf[j]);
        const double dyj = j >= 2 ? delta_y[j] : 1.0;  // rows from 2 on are scaled (for/les.f:262-270)
        const cx<double> c12 = T(2, j, dyj), c12p = T(2, j + 1, delta_y[j + 1]);
        const cx<double> c23 = T(4, j, dyj), c23p = T(4, j + 1, delta_y[j + 1]);
        cx<double> f = cf1[o];
        const cx<double> a = i_times(kx, c11), b = i_times(kz, c13);
        f = make_cx<double>(f.re - a.re - (c12p.re - c12.re) / dyf[j] - b.re,
                            f.im - a.im - (c12p.im - c12.im) / dyf[j] - b.im);
        cf1[o] = f;
        cx<double> h = cf3[o];
        const cx<double> c = i_times(kx, c13), d = i_times(kz, c33);
        h = make_cx<double>(h.re - c.re - (c23p.re - c23.re) / dyf[j] - d.re,
                            h.im - c.im - (c23p.im - c23.im) / dyf[j] - d.im);
        cf3[o] = h;
    }
    if (j >= 2 && j <= NY) {
        const cx<double> c12 = T(2, j, delta_y[j]), c23 = T(4, j, delta_y[j]);
        const cx<double> a = i_times(kx, c12), b = i_times(kz, c23);
        cx<double> f = cf2[o];
        f = make_cx<double>(f.re - a.re - b.re, f.im - a.im - b.im);
        cf2[o] = f;
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
                   const double *delta_yf, int j1, int j2)
{
    hipLaunchKernelGGL(k_les_cf, dim3((unsigned)((g.ps + 255) / 256), g.ny + 1), dim3(256), 0, s, g, t, cf1, cf2,
                       cf3, kx, kz, dyf, delta_y, delta_yf, j1, j2);
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
