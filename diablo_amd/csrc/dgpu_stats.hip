// dgpu_stats.hip -- SAVE_STATS_CHAN's statistics (for/save_stats.f:106-347,
// scalar part for/save_stats.f:420-522) on the device, in FP64.
//
// The reference transforms the fields to physical space and sums over the
// (x,z) points of each plane.  Every quantity is a plane mean of a product of
// two dealiased fields, so by Parseval it is a sum over the kept Fourier modes
// of the plane: (1/(NX NZ)) sum_x a b = sum_k w_k Re(a_k conj b_k), with
// w = 1 for I = 0 (both kz signs are stored) and 2 for I >= 1 (the -kx half
// is the conjugate).  Fluctuations leave out the (0,0) mode; full fields take
// its real part only, as the C2R of FFT_XZ_TO_PHYSICAL does.  One block per
// plane, fixed-order tree reduction: the result does not depend on timing.
#include "../../include/diablo_gpu.h"
#include "dgpu_kernels.hpp"

#include <mutex>

namespace dgpu {

namespace {

constexpr int kSums = 13;
enum { S_URMS, S_VRMS, S_WRMS, S_UV, S_UW, S_WV, S_SHEAR, S_OX, S_OY, S_OZ, S_THRMS, S_THV, S_PE };

__device__ __forceinline__ double abs2(cx<double> a) { return a.re * a.re + a.im * a.im; }
// Re(a conj b)
__device__ __forceinline__ double dotr(cx<double> a, cx<double> b) { return a.re * b.re + a.im * b.im; }
__device__ __forceinline__ cx<double> ax(double s, cx<double> a) { return make_cx<double>(s * a.re, s * a.im); }
__device__ __forceinline__ cx<double> ik(double k, cx<double> a) { return make_cx<double>(-k * a.im, k * a.re); }

}  // namespace

__global__ void __launch_bounds__(256) k_stats(Geo g, const cx<double> *__restrict__ cu1,
                                               const cx<double> *__restrict__ cu2, const cx<double> *__restrict__ cu3,
                                               const cx<double> *__restrict__ cp, const cx<double> *__restrict__ cth,
                                               const double *__restrict__ dy, const double *__restrict__ dyf,
                                               const double *__restrict__ gyf, const double *__restrict__ kxv,
                                               const double *__restrict__ kzv, double *__restrict__ out)
{
    __shared__ double red[kSums][256];
    const int j = 1 + blockIdx.x;
    const int NY = g.ny;
    const long ps = g.ps;
    double acc[kSums];
#pragma unroll
    for (int q = 0; q < kSums; ++q) acc[q] = 0.0;
    const double ayj = dyf[j - 1] / (2.0 * dy[j]), bmj = dyf[j] / (2.0 * dy[j]);
    const double i2f = 1.0 / (2.0 * dyf[j]);
    const double igt = 1.0 / (gyf[j + 1] - gyf[j - 1]);
    for (long m = threadIdx.x; m < ps; m += blockDim.x) {
        const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
        const bool mean = (m == 0);
        const double w = (i == 0) ? 1.0 : 2.0;
        const double kx = kxv[i], kz = kzv[K];
        auto ld = [&](const cx<double> *f, int jj) {
            cx<double> v = f[m + (long)jj * ps];
            if (mean) v.im = 0.0;
            return v;
        };
        const cx<double> u1m = ld(cu1, j - 1), u1 = ld(cu1, j), u1p = ld(cu1, j + 1);
        const cx<double> u2 = ld(cu2, j), u2p = ld(cu2, j + 1);
        const cx<double> u3m = ld(cu3, j - 1), u3 = ld(cu3, j), u3p = ld(cu3, j + 1);
        const cx<double> du1 = ax(i2f, u1p - u1m), du3 = ax(i2f, u3p - u3m);
        const cx<double> v2 = ax(0.5, u2p + u2);
        if (!mean) {
            acc[S_URMS] += w * abs2(u1);
            acc[S_VRMS] += w * (0.5 * (abs2(u2) + abs2(u2p)));
            acc[S_WRMS] += w * abs2(u3);
            acc[S_UV] += w * dotr(ax(ayj, u1) + ax(bmj, u1m), u2);  // U1 at GY times U2
            acc[S_WV] += w * dotr(ax(ayj, u3) + ax(bmj, u3m), u2);
            acc[S_UW] += w * dotr(u1, u3);
        }
        acc[S_SHEAR] += w * (abs2(du1) + abs2(du3));
        acc[S_OX] += w * abs2(du3 - ik(kz, v2));                     // dW/dy - d/dz V
        acc[S_OY] += w * abs2(ik(kz, u1) - ik(kx, u3));              // dU/dz - dW/dx
        acc[S_OZ] += w * abs2(ik(kx, v2) - du1);                     // dV/dx - dU/dy
        if (cth) {
            const cx<double> tm = ld(cth, j - 1), t0 = ld(cth, j), tp = ld(cth, j + 1);
            if (!mean) {
                acc[S_THRMS] += w * abs2(t0);
                acc[S_THV] += w * dotr(t0, v2);
            }
            acc[S_PE] += w * (abs2(ik(kx, t0)) + abs2(ax(igt, tp - tm)) + abs2(ik(kz, t0)));
        }
    }
#pragma unroll
    for (int q = 0; q < kSums; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int q = 0; q < kSums; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        auto put = [&](int q, double v) { out[(long)q * NY + (j - 1)] = v; };
        const double ume = cu1[(long)j * ps].re, umem = cu1[(long)(j - 1) * ps].re;
        const double wme = cu3[(long)j * ps].re, wmem = cu3[(long)(j - 1) * ps].re;
        put(DGPU_ST_UME, ume);
        put(DGPU_ST_VME, cu2[(long)j * ps].re);
        put(DGPU_ST_WME, wme);
        put(DGPU_ST_URMS, sqrt(red[S_URMS][0]));
        put(DGPU_ST_VRMS, sqrt(red[S_VRMS][0]));
        put(DGPU_ST_WRMS, sqrt(red[S_WRMS][0]));
        put(DGPU_ST_UV, red[S_UV][0]);
        put(DGPU_ST_UW, red[S_UW][0]);
        put(DGPU_ST_WV, red[S_WV][0]);
        put(DGPU_ST_DUDY, (ume - umem) / (gyf[j] - gyf[j - 1]));
        put(DGPU_ST_DWDY, (wme - wmem) / (gyf[j] - gyf[j - 1]));
        put(DGPU_ST_CP, cp[(long)j * ps].re);
        put(DGPU_ST_SHEAR, red[S_SHEAR][0]);
        put(DGPU_ST_OMEGA_X, sqrt(red[S_OX][0]));
        put(DGPU_ST_OMEGA_Y, sqrt(red[S_OY][0]));
        put(DGPU_ST_OMEGA_Z, sqrt(red[S_OZ][0]));
        if (cth) {
            const double tme = cth[(long)j * ps].re, tmem = cth[(long)(j - 1) * ps].re;
            put(DGPU_ST_THME, tme);
            put(DGPU_ST_THRMS, sqrt(red[S_THRMS][0]));
            put(DGPU_ST_THV, red[S_THV][0]);
            put(DGPU_ST_DTHDY, (tme - tmem) / (gyf[j] - gyf[j - 1]));
            put(DGPU_ST_PE_DISS, red[S_PE][0]);
        } else {
            for (int q = DGPU_ST_THME; q <= DGPU_ST_PE_DISS; ++q) put(q, 0.0);
        }
    }
}

// tkebudget_chan's epsilon (for/save_stats.f:965-1159), per plane j = 1..NY,
// as Parseval sums the host combines: out[q * NY + j - 1] for
//   q = 0..5  sum w kx^2 |CU1|^2, kx^2 |CU2|^2, kx^2 |CU3|^2, kz^2 |CU1|^2,
//             kz^2 |CU2|^2, kz^2 |CU3|^2 of plane j (the x and z derivative
//             blocks; the host forms the DYF-weighted averages onto GY)
//   q = 6, 7  sum over fluctuating modes of w |CU1(j) - CU1(j-1)|^2 and of
//             CU3 (du/dy, dw/dy before the 1/DY(j)^2)
//   q = 8     sum over fluctuating modes of w |CU2(j+1) - CU2(j-1)|^2,
//             j = 2..NY-1 (dv/dy), 0 elsewhere
//   q = 9     Re CU2(0,0,j) (vme)
__global__ void __launch_bounds__(256) k_tke(Geo g, const cx<double> *__restrict__ cu1,
                                             const cx<double> *__restrict__ cu2, const cx<double> *__restrict__ cu3,
                                             const double *__restrict__ kxv, const double *__restrict__ kzv,
                                             double *__restrict__ out)
{
    constexpr int kT = 9;
    __shared__ double red[kT][256];
    const int j = 1 + blockIdx.x;
    const int NY = g.ny;
    const long ps = g.ps;
    double acc[kT];
#pragma unroll
    for (int q = 0; q < kT; ++q) acc[q] = 0.0;
    const bool dv = (j >= 2 && j <= NY - 1);
    for (long m = threadIdx.x; m < ps; m += blockDim.x) {
        const int K = (int)(m / g.mk), i = (int)(m - (long)K * g.mk);
        const bool mean = (m == 0);
        const double w = (i == 0) ? 1.0 : 2.0;
        const double kx2 = kxv[i] * kxv[i], kz2 = kzv[K] * kzv[K];
        const cx<double> a = cu1[m + (long)j * ps], am = cu1[m + (long)(j - 1) * ps];
        const cx<double> v = cu2[m + (long)j * ps];
        const cx<double> c = cu3[m + (long)j * ps], cm = cu3[m + (long)(j - 1) * ps];
        const double a2 = abs2(a), v2 = abs2(v), c2 = abs2(c);
        acc[0] += w * kx2 * a2;
        acc[1] += w * kx2 * v2;
        acc[2] += w * kx2 * c2;
        acc[3] += w * kz2 * a2;
        acc[4] += w * kz2 * v2;
        acc[5] += w * kz2 * c2;
        if (!mean) {
            acc[6] += w * abs2(a - am);
            acc[7] += w * abs2(c - cm);
            if (dv) acc[8] += w * abs2(cu2[m + (long)(j + 1) * ps] - cu2[m + (long)(j - 1) * ps]);
        }
    }
#pragma unroll
    for (int q = 0; q < kT; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
#pragma unroll
            for (int q = 0; q < kT; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < kT; ++q) out[(long)q * NY + (j - 1)] = red[q][0];
        out[(long)kT * NY + (j - 1)] = cu2[(long)j * ps].re;
    }
}

void launch_tke(hipStream_t s, const Geo &g, const cx<double> *cu1, const cx<double> *cu2, const cx<double> *cu3,
                const double *kx, const double *kz, double *out)
{
    hipLaunchKernelGGL(k_tke, dim3(g.ny), dim3(256), 0, s, g, cu1, cu2, cu3, kx, kz, out);
}

void launch_stats(hipStream_t s, const Geo &g, const cx<double> *cu1, const cx<double> *cu2, const cx<double> *cu3,
                  const cx<double> *cp, const cx<double> *cth, const double *dy, const double *dyf,
                  const double *gyf, const double *kx, const double *kz, double *out)
{
    hipLaunchKernelGGL(k_stats, dim3(g.ny), dim3(256), 0, s, g, cu1, cu2, cu3, cp, cth, dy, dyf, gyf, kx, kz, out);
}

// The per-plane statistics of dgpu_stats_chan_noalloc: DGPU_NSTAT rows of
// up to DGPU_NOALLOC_MAX_NY planes.  k_stats writes every entry of the NY
// planes it covers once per call, so nothing has to be cleared between
// calls.  Host code must never pass the array's name to a kernel or a copy:
// on the host it is only a stand-in; the device address of each GPU's copy
// comes from hipGetSymbolAddress below.
static __device__ __align__(256) double scratch[(long)DGPU_NSTAT * DGPU_NOALLOC_MAX_NY];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mtx;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_stats_scratch(double **out)
{
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mtx);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

}  // namespace dgpu
