// dgpu_kernels.hip -- Stockham-transform kernels of the standalone entry points
// FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL (for/mpi.f:1530-1674) and of
// POISSON_P_CHAN, plus APPLY_BC_VEL on kept-mode fields.  The physical-space
// side of these transforms is in natural order (Fortran-layout planes).  The
// substep's own transforms are in dgpu_hot_z.hip and dgpu_hot_x.hip.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include "dgpu_kernels.hpp"

namespace dgpu {

// Stockham plans the standalone transforms are also built for (FID > 0).
__device__ constexpr FftLen kDevConstFft[] = {fft_len_of(kConstFftN[0]), fft_len_of(kConstFftN[1]),
                                              fft_len_of(kConstFftN[2])};
static_assert(sizeof(kDevConstFft) / sizeof(FftLen) == (size_t)kNumConstFft, "one entry per compile-time plan");

template <int FID>
__device__ __forceinline__ const FftLen &fft_ref(const FftLen &rt)
{
    if constexpr (FID == 0)
        return rt;
    else
        return kDevConstFft[FID];
}

// -------------------------------------------------------------------------
// helpers
// -------------------------------------------------------------------------

__device__ __forceinline__ int kzslot_to_z(int s, int nkz, int tnkz, int nz)
{
    // packed slot K <= NKZ holds kz = K, slot K > NKZ holds kz = K - (TNKZ+1)
    // (for/mpi.f:1602-1606); its index in the length-NZ transform:
    return s <= nkz ? s : nz - (tnkz + 1 - s);
}

// The inverse: the packed slot of transform index z, or -1 in the gap.
__device__ __forceinline__ int z_to_kzslot(int z, int nkz, int tnkz, int nz)
{
    if (z <= nkz) return z;
    return z >= nz - tnkz + nkz ? z - nz + tnkz + 1 : -1;
}

template <typename TO, typename FROM>
__device__ __forceinline__ cx<TO> cvt(cx<FROM> a) { return make_cx<TO>(TO(a.re), TO(a.im)); }

// -------------------------------------------------------------------------
// z-inverse: kept slots -> length-NZ columns (FFT_XZ_TO_PHYSICAL z part,
// for/mpi.f:1627-1641): unpack, zero the gap, unnormalised backward C2C.
// grid (8 * col-chunks * plane groups of 8, 1, fields), in plane_chunk's
// order; block W*T threads; LDS W*NZ complex.
// -------------------------------------------------------------------------
template <typename T, int E, typename TIN, int FID>
__global__ void k_zinv(FftLen fz_rt, const cx<T> *__restrict__ tw, const cx<TIN> *__restrict__ in, long in_fs,
                       long in_ps, int in_mk, int ncols, cx<T> *__restrict__ out, long out_fs, long out_ps,
                       int out_mk, int j0, int nkz, int tnkz, int lw, int nchunks, int nplanes)
{
    const FftLen &fz = fft_ref<FID>(fz_rt);
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int W = 1 << lw;
    const int nz = fz.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    int chunk, jp;
    if (!plane_chunk(blockIdx.x, nchunks, nplanes, chunk, jp)) return;
    const int i0 = chunk * W;
    const int j = j0 + jp;
    const int f = blockIdx.z;
    const cx<TIN> *src = in + f * in_fs + (long)j * in_ps;
    cx<T> *dst = out + f * out_fs + (long)j * out_ps;
    const cx<T> zero = make_cx<T>(0, 0);
    // every transform index once: the kept slots, zero in the gap
    for (int idx = tid; idx < W * nz; idx += nthr) {
        const int z = idx >> lw, b = idx & (W - 1);
        const int i = i0 + b, s = z_to_kzslot(z, nkz, tnkz, nz);
        lds[idx] = (i < ncols && s >= 0) ? cvt<T>(src[(long)s * in_mk + i]) : zero;
    }
    __syncthreads();
    // batch index fastest across lanes: LDS accesses of a wavefront are contiguous
    const int b = tid & (W - 1), th = tid >> lw;
    fft_lds<T, E, +1>(lds + b, W, fz, th, th < fz.t, tw);
    for (int idx = tid; idx < W * nz; idx += nthr) {
        const int z = idx >> lw, bb = idx & (W - 1);
        const int i = i0 + bb;
        if (i < ncols) dst[(long)z * out_mk + i] = lds[idx];
    }
}

// -------------------------------------------------------------------------
// Standalone transforms on Fortran-layout planes (host entry points)
// -------------------------------------------------------------------------

// R2C along x of the physical rows of Fortran-layout planes; output
// xt[j][z][0..NKX] scaled 1/NX (for/mpi.f:1547-1561). Pairs of rows per FFT.
template <typename T, int E, int FID>
__global__ void k_x_r2c(FftLen fx_rt, const cx<T> *__restrict__ tw, const double *__restrict__ phys, int nzp2,
                        cx<T> *__restrict__ xt, int mk, int nz, int j0)
{
    const FftLen &fx = fft_ref<FID>(fx_rt);
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int nx = fx.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int j = j0 + blockIdx.y;
    const int npair = blockDim.x / fx.t;
    const int z0 = blockIdx.x * 2 * npair;
    const long rowlen = nx + 2;
    const double *pl = phys + (long)j * nzp2 * rowlen;
    for (int idx = tid; idx < npair * nx; idx += nthr) {
        const int p = idx / nx, x = idx - p * nx;
        const int za = z0 + 2 * p, zb = za + 1;
        const T a = za < nz ? T(pl[(long)za * rowlen + x]) : T(0);
        const T b = zb < nz ? T(pl[(long)zb * rowlen + x]) : T(0);
        lds[idx] = make_cx<T>(a, b);
    }
    __syncthreads();
    const int p = tid / fx.t, th = tid - p * fx.t;
    // rows of T consecutive threads: inside one wavefront when T divides 64
    if (64 % fx.t == 0) fft_lds<T, E, -1, true>(lds + p * nx, 1, fx, th, p < npair, tw);
    else fft_lds<T, E, -1>(lds + p * nx, 1, fx, th, p < npair, tw);
    __syncthreads();  // the outputs below read every row
    const T h = T(0.5) / T(nx);
    for (int idx = tid; idx < npair * mk; idx += nthr) {
        const int pp = idx / mk, k = idx - pp * mk;
        const cx<T> Z = lds[pp * nx + k];
        const cx<T> Zn = lds[pp * nx + (k == 0 ? 0 : nx - k)];
        const int za = z0 + 2 * pp, zb = za + 1;
        cx<T> *o = xt + ((long)j * nz) * mk;
        if (za < nz) o[(long)za * mk + k] = make_cx<T>((Z.re + Zn.re) * h, (Z.im - Zn.im) * h);
        if (zb < nz) o[(long)zb * mk + k] = make_cx<T>((Z.im + Zn.im) * h, -(Z.re - Zn.re) * h);
    }
}

// z-forward of xt columns, scale 1/NZ, pack kept slots into Fortran layout;
// every other slot of the plane is set to zero.
template <typename T, int E, int FID>
__global__ void k_zfwd_pack(FftLen fz_rt, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ xt, int mk,
                            double *__restrict__ fort, int nxp, int nkz, int tnkz, int j0, int lw, int nchunks,
                            int nplanes)
{
    const FftLen &fz = fft_ref<FID>(fz_rt);
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int W = 1 << lw;
    const int nz = fz.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    int chunk, jp;
    if (!plane_chunk(blockIdx.x, nchunks, nplanes, chunk, jp)) return;
    const int i0 = chunk * W;
    const int j = j0 + jp;
    const cx<T> *src = xt + (long)j * nz * mk;
    for (int idx = tid; idx < W * nz; idx += nthr) {
        const int z = idx >> lw, b = idx & (W - 1);
        const int i = i0 + b;
        lds[idx] = (i < mk) ? src[(long)z * mk + i] : make_cx<T>(0, 0);
    }
    __syncthreads();
    const int b = tid & (W - 1), th = tid >> lw;
    fft_lds<T, E, -1>(lds + b, W, fz, th, th < fz.t, tw);
    // Fortran complex view (0:NXP, 0:NZ+1) per plane
    double *pl = fort + (long)j * (long)(nxp + 1) * (nz + 2) * 2;
    const T inv = T(1) / T(nz);
    for (int idx = tid; idx < W * (nz + 2); idx += nthr) {
        const int K = idx >> lw, bb = idx & (W - 1);
        const int i = i0 + bb;
        if (i > nxp) continue;
        cx<T> v = make_cx<T>(0, 0);
        if (i < mk && K <= tnkz) v = scale(lds[kzslot_to_z(K, nkz, tnkz, nz) * W + bb], inv);
        const long e = ((long)K * (nxp + 1) + i) * 2;
        pl[e] = double(v.re);
        pl[e + 1] = double(v.im);
    }
}

// C2R along x: zt[j][z][0..mkin-1] (Nyquist zeroed, Im of DC ignored) into the
// physical rows of Fortran-layout planes (for/mpi.f:1643-1672). The x pads of
// rows 0..NZ-1 are set to zero; rows NZ, NZ+1 are not touched.
template <typename T, int E, int FID>
__global__ void k_x_c2r(FftLen fx_rt, const cx<T> *__restrict__ tw, const cx<T> *__restrict__ zt, int mkin,
                        double *__restrict__ fort, int nz, int j0)
{
    const FftLen &fx = fft_ref<FID>(fx_rt);
    extern __shared__ unsigned char smem_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(smem_raw);
    const int nx = fx.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int j = j0 + blockIdx.y;
    const int npair = blockDim.x / fx.t;
    const int z0 = blockIdx.x * 2 * npair;
    const cx<T> *src = zt + (long)j * nz * mkin;
    for (int idx = tid; idx < npair * nx; idx += nthr) {
        const int p = idx / nx, k = idx - p * nx;
        const int za = z0 + 2 * p, zb = za + 1;
        cx<T> A = make_cx<T>(0, 0), B = make_cx<T>(0, 0);
        int kk = -1;
        bool cj = false;
        if (k < mkin && k < nx / 2) kk = k;
        else if (k > nx / 2 && nx - k < mkin) { kk = nx - k; cj = true; }
        if (kk >= 0) {
            if (za < nz) A = src[(long)za * mkin + kk];
            if (zb < nz) B = src[(long)zb * mkin + kk];
            if (kk == 0) { A.im = 0; B.im = 0; }
            if (cj) { A.im = -A.im; B.im = -B.im; }
        }
        lds[idx] = make_cx<T>(A.re - B.im, A.im + B.re);
    }
    __syncthreads();
    const int p = tid / fx.t, th = tid - p * fx.t;
    if (64 % fx.t == 0) fft_lds<T, E, +1, true>(lds + p * nx, 1, fx, th, p < npair, tw);
    else fft_lds<T, E, +1>(lds + p * nx, 1, fx, th, p < npair, tw);
    __syncthreads();  // the outputs below read every row
    const long rowlen = nx + 2;
    double *pl = fort + (long)j * (nz + 2) * rowlen;
    for (int idx = tid; idx < npair * (nx + 2); idx += nthr) {
        const int pp = idx / (nx + 2), x = idx - pp * (nx + 2);
        const int za = z0 + 2 * pp, zb = za + 1;
        const cx<T> v = x < nx ? lds[pp * nx + x] : make_cx<T>(0, 0);
        if (za < nz) pl[(long)za * rowlen + x] = double(v.re);
        if (zb < nz) pl[(long)zb * rowlen + x] = double(v.im);
    }
}

// Fortran complex layout -> kept-mode layout.
__global__ void k_unpack_fortran(const double *__restrict__ fort, cx<double> *__restrict__ kept, int nxp, int nz,
                                 int mk, int mz, int j0)
{
    const int j = j0 + blockIdx.y;
    const long plane = (long)mk * mz;
    for (long m = (long)blockIdx.x * blockDim.x + threadIdx.x; m < plane; m += (long)gridDim.x * blockDim.x) {
        const int s = (int)(m / mk), i = (int)(m - (long)s * mk);
        const double *p = fort + (((long)j * (nz + 2) + s) * (nxp + 1) + i) * 2;
        kept[(long)j * plane + m] = make_cx<double>(p[0], p[1]);
    }
}

// APPLY_BC_VEL_LOWER / UPPER (for/channel.f:2285-2492) on kept-mode fields.
// One thread per mode (i,K) of the rows 0, 1, NY, NY+1; the (0,0) mode and
// the V code-0 upper rule are handled by the thread of mode 0.
__global__ void k_apply_bc_vel(cx<double> *cu1, cx<double> *cu2, cx<double> *cu3, long ps, int ny, const int *codes,
                               const double *c1, const double *dy, const double *dyf)
{
    const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= ps) return;
    const bool is00 = (m == 0);
    auto at = [&](cx<double> *f, int j) -> cx<double> & { return f[m + (long)j * ps]; };
    const cx<double> z0 = make_cx<double>(0, 0);
    // codes: u_min, v_min, w_min, u_max, v_max, w_max
    cx<double> *hf[2] = {cu1, cu3};
    const int hmin[2] = {codes[0], codes[2]}, hmax[2] = {codes[3], codes[5]};
    const double cmin[2] = {c1[0], c1[2]}, cmax[2] = {c1[3], c1[5]};
    for (int c = 0; c < 2; ++c) {
        cx<double> *f = hf[c];
        if (hmin[c] == 0) {
            at(f, 0) = z0; at(f, 1) = z0;
            if (is00) { at(f, 1) = make_cx<double>(cmin[c], 0); at(f, 0) = make_cx<double>(cmin[c], 0); }
        } else {
            at(f, 1) = at(f, 2); at(f, 0) = at(f, 1);
            if (is00) {
                at(f, 1) = at(f, 2) - make_cx<double>(dy[2] * cmin[c], 0);
                at(f, 0) = at(f, 1) - make_cx<double>(dy[1] * cmin[c], 0);
            }
        }
    }
    if (codes[1] == 0) {
        at(cu2, 1) = make_cx<double>(2.0 * c1[1], 0) - at(cu2, 2);
        at(cu2, 0) = at(cu2, 1);
    } else if (codes[1] == 1) {
        at(cu2, 1) = at(cu2, 2); at(cu2, 0) = at(cu2, 1);
        if (is00) {
            at(cu2, 1) = at(cu2, 2) - make_cx<double>(dyf[1] * c1[1], 0);
            at(cu2, 0) = at(cu2, 1) - make_cx<double>(dyf[1] * c1[1], 0);
        }
    } else if (codes[1] == 2) {
        if (is00) at(cu2, 1) = make_cx<double>(-at(cu2, 2).re, -at(cu2, 2).im);
    }
    for (int c = 0; c < 2; ++c) {
        cx<double> *f = hf[c];
        if (hmax[c] == 0) {
            at(f, ny) = z0; at(f, ny + 1) = z0;
            if (is00) { at(f, ny) = make_cx<double>(cmax[c], 0); at(f, ny + 1) = make_cx<double>(cmax[c], 0); }
            if (c == 1 && is00) at(f, ny + 1) = z0;  // for/channel.f:2445
        } else {
            at(f, ny) = at(f, ny - 1); at(f, ny + 1) = at(f, ny);
            if (is00) {
                at(f, ny) = at(f, ny - 1) + make_cx<double>(dy[ny] * cmax[c], 0);
                at(f, ny + 1) = at(f, ny) + make_cx<double>(dy[ny] * cmax[c], 0);
            }
        }
    }
    if (codes[4] == 0) {
        if (is00) at(cu2, ny + 1) = make_cx<double>(2.0 * c1[4], 0) - at(cu2, ny);
    } else if (codes[4] == 1) {
        at(cu2, ny) = at(cu2, ny - 1); at(cu2, ny + 1) = at(cu2, ny);
        if (is00) at(cu2, ny + 1) = at(cu2, ny) + make_cx<double>(dy[ny] * c1[4], 0);
    } else if (codes[4] == 2) {
        if (is00) at(cu2, ny + 1) = make_cx<double>(-at(cu2, ny).re, -at(cu2, ny).im);
    }
}

// -------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------

// Columns per block for the z transforms: the largest power of two with
// W * T <= threads (T threads per transform); returns log2(W).
static inline int chunk_lw(int t, int threads)
{
    int lw = 0;
    while ((2 << lw) * t <= threads) ++lw;
    return lw;
}

// Id of the compile-time Stockham plan equal to f, or 0
// (DGPU_NO_CONST_PLANS=1: always 0).
static int const_fft_id(const FftLen &f)
{
    static const bool off = std::getenv("DGPU_NO_CONST_PLANS") != nullptr;
    if (off) return 0;
    for (int id = 1; id < kNumConstFft; ++id) {
        const FftLen c = fft_len_of(kConstFftN[id]);
        if (std::memcmp(&c, &f, sizeof(FftLen)) == 0) return id;
    }
    return 0;
}

// Calls fn(std::integral_constant<int, FID>) for a compile-time plan id;
// false for 0.
template <class F>
static bool with_fid(int fid, F &&fn)
{
    static_assert(kNumConstFft == 3, "one case per compile-time plan");
    switch (fid) {
    case 1: fn(std::integral_constant<int, 1>{}); return true;
    case 2: fn(std::integral_constant<int, 2>{}); return true;
    default: return false;
    }
}

template <typename T, int E, typename TIN, int FID>
static void zinv_e(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<TIN> *in, long in_fs, long in_ps,
                   int in_mk, int ncols, cx<T> *out, long out_fs, long out_ps, int out_mk, int j0, int nplanes,
                   int nfields)
{
    const int lw = chunk_lw(fz.len.t, 256), W = 1 << lw;
    const int nchunks = (ncols + W - 1) / W;
    dim3 grid((unsigned)(8 * nchunks * ((nplanes + 7) / 8)), 1, nfields);
    const size_t lds = sizeof(cx<T>) * W * fz.len.n;
    hipLaunchKernelGGL((k_zinv<T, E, TIN, FID>), grid, dim3(W * fz.len.t), lds, s, fz.len, fz.tw, in, in_fs, in_ps, in_mk,
                       ncols, out, out_fs, out_ps, out_mk, j0, g.nkz, g.tnkz, lw, nchunks, nplanes);
}

#define DGPU_DISPATCH_E(E_, CALL)                                  \
    switch (E_) {                                                  \
    case 4: { constexpr int EE = 4; CALL; } break;                 \
    case 6: { constexpr int EE = 6; CALL; } break;                 \
    case 8: { constexpr int EE = 8; CALL; } break;                 \
    case 12: { constexpr int EE = 12; CALL; } break;               \
    case 16: { constexpr int EE = 16; CALL; } break;               \
    default: break;                                                \
    }

template <typename T, typename TIN>
void launch_zinv(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<TIN> *in, long in_fs, long in_ps, int in_mk,
                 int ncols, cx<T> *out, long out_fs, long out_ps, int out_mk, int j0, int nplanes, int nfields)
{
    if (with_fid(const_fft_id(fz.len), [&](auto fc) {
            constexpr int FI = decltype(fc)::value;
            zinv_e<T, fft_len_of(kConstFftN[FI]).e, TIN, FI>(s, g, fz, in, in_fs, in_ps, in_mk, ncols, out, out_fs,
                                                              out_ps, out_mk, j0, nplanes, nfields);
        }))
        return;
    DGPU_DISPATCH_E(fz.len.e, (zinv_e<T, EE, TIN, 0>(s, g, fz, in, in_fs, in_ps, in_mk, ncols, out, out_fs, out_ps,
                                                out_mk, j0, nplanes, nfields)));
}

template <typename T, int E, int FID>
static void x_r2c_e(hipStream_t s, const Geo &g, const Fft<T> &fx, const double *phys, cx<T> *xt, int j0,
                    int nplanes)
{
    int npair = 256 / fx.len.t;
    if (npair < 1) npair = 1;
    const int zper = 2 * npair;
    dim3 grid((g.nz + zper - 1) / zper, nplanes);
    hipLaunchKernelGGL((k_x_r2c<T, E, FID>), grid, dim3(npair * fx.len.t), sizeof(cx<T>) * npair * fx.len.n, s, fx.len,
                       fx.tw, phys, g.nz + 2, xt, g.mk, g.nz, j0);
}

template <typename T>
void launch_x_r2c(hipStream_t s, const Geo &g, const Fft<T> &fx, const double *phys, cx<T> *xt, int j0, int nplanes)
{
    if (with_fid(const_fft_id(fx.len), [&](auto fc) {
            constexpr int FI = decltype(fc)::value;
            x_r2c_e<T, fft_len_of(kConstFftN[FI]).e, FI>(s, g, fx, phys, xt, j0, nplanes);
        }))
        return;
    DGPU_DISPATCH_E(fx.len.e, (x_r2c_e<T, EE, 0>(s, g, fx, phys, xt, j0, nplanes)));
}

template <typename T, int E, int FID>
static void zfwd_pack_e(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<T> *xt, double *fort, int j0,
                        int nplanes)
{
    const int lw = chunk_lw(fz.len.t, 256), W = 1 << lw;
    const int ncols = g.nx / 2 + 1;  // 0..NXP, zero beyond NKX
    const int nchunks = (ncols + W - 1) / W;
    dim3 grid((unsigned)(8 * nchunks * ((nplanes + 7) / 8)), 1);
    hipLaunchKernelGGL((k_zfwd_pack<T, E, FID>), grid, dim3(W * fz.len.t), sizeof(cx<T>) * W * fz.len.n, s, fz.len, fz.tw,
                       xt, g.mk, fort, g.nx / 2, g.nkz, g.tnkz, j0, lw, nchunks, nplanes);
}

template <typename T>
void launch_zfwd_pack(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<T> *xt, double *fort, int j0,
                      int nplanes)
{
    if (with_fid(const_fft_id(fz.len), [&](auto fc) {
            constexpr int FI = decltype(fc)::value;
            zfwd_pack_e<T, fft_len_of(kConstFftN[FI]).e, FI>(s, g, fz, xt, fort, j0, nplanes);
        }))
        return;
    DGPU_DISPATCH_E(fz.len.e, (zfwd_pack_e<T, EE, 0>(s, g, fz, xt, fort, j0, nplanes)));
}

template <typename T, int E, int FID>
static void x_c2r_e(hipStream_t s, const Geo &g, const Fft<T> &fx, const cx<T> *zt, int mkin, double *fort, int j0,
                    int nplanes)
{
    int npair = 256 / fx.len.t;
    if (npair < 1) npair = 1;
    const int zper = 2 * npair;
    dim3 grid((g.nz + zper - 1) / zper, nplanes);
    hipLaunchKernelGGL((k_x_c2r<T, E, FID>), grid, dim3(npair * fx.len.t), sizeof(cx<T>) * npair * fx.len.n, s, fx.len,
                       fx.tw, zt, mkin, fort, g.nz, j0);
}

template <typename T>
void launch_x_c2r(hipStream_t s, const Geo &g, const Fft<T> &fx, const cx<T> *zt, int mkin, double *fort, int j0,
                  int nplanes)
{
    if (with_fid(const_fft_id(fx.len), [&](auto fc) {
            constexpr int FI = decltype(fc)::value;
            x_c2r_e<T, fft_len_of(kConstFftN[FI]).e, FI>(s, g, fx, zt, mkin, fort, j0, nplanes);
        }))
        return;
    DGPU_DISPATCH_E(fx.len.e, (x_c2r_e<T, EE, 0>(s, g, fx, zt, mkin, fort, j0, nplanes)));
}

void launch_unpack_fortran(hipStream_t s, const Geo &g, const double *fort, cx<double> *kept, int j0, int nplanes)
{
    const long plane = g.ps;
    const int threads = 256;
    int bx = (int)((plane + threads - 1) / threads);
    hipLaunchKernelGGL(k_unpack_fortran, dim3(bx, nplanes), dim3(threads), 0, s, fort, kept, g.nx / 2, g.nz, g.mk,
                       g.mz, j0);
}

void launch_apply_bc_vel(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu2, cx<double> *cu3,
                         const int *codes, const double *c1, const double *dy, const double *dyf)
{
    const int threads = 256;
    const long blocks = (g.ps + threads - 1) / threads;
    hipLaunchKernelGGL(k_apply_bc_vel, dim3(blocks), dim3(threads), 0, s, cu1, cu2, cu3, g.ps, g.ny, codes, c1, dy,
                       dyf);
}

// ---- restart-file layout (WriteHDF5 / ReadHDF5, for/hdf5.f:172-268, 470-534) ----
namespace {

__device__ __forceinline__ long fidx(int x, int z, int y, int nx, int nz)
{
    return (long)x + (long)(nx + 2) * ((long)z + (long)(nz + 2) * (long)y);
}

// SWAPZY (for/hdf5.f:1243-1258) of the physical field, planes 1..NY, into
// out(1:NX, 1:NY, 1:NZ); with g2gf, G2GF first (for/hdf5.f:1176-1192): plane
// y is 0.5 (v'(y) + v'(y+1)) with v'(1) = v(2), v'(NY+1) = v(NY).
__global__ void k_to_file(const double *__restrict__ fort, double *__restrict__ out, int nx, int ny, int nz,
                          int g2gf)
{
    const long n = (long)nx * ny * nz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % nx);
        const long r = e / nx;
        const int y = 1 + (int)(r % ny), z = (int)(r / ny);
        double v;
        if (g2gf) {
            const double a = fort[fidx(x, z, y == 1 ? 2 : y, nx, nz)];
            const double b = fort[fidx(x, z, y == ny ? ny : y + 1, nx, nz)];
            v = 0.5 * (a + b);
        } else {
            v = fort[fidx(x, z, y, nx, nz)];
        }
        out[e] = v;
    }
}

// SWAPYZ (for/hdf5.f:1262-1278): every element of the Fortran-layout field,
// zero outside x < NX, z < NZ, 1 <= y <= NY.
__global__ void k_from_file(const double *__restrict__ in, double *__restrict__ fort, int nx, int ny, int nz)
{
    const long n = (long)(nx + 2) * (nz + 2) * (ny + 2);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int x = (int)(e % (nx + 2));
        const long r = e / (nx + 2);
        const int z = (int)(r % (nz + 2)), y = (int)(r / (nz + 2));
        double v = 0.0;
        if (x < nx && z < nz && y >= 1 && y <= ny) v = in[(long)x + (long)nx * ((long)(y - 1) + (long)ny * z)];
        fort[e] = v;
    }
}

// GF2G (for/hdf5.f:1196-1238), one (x, z) column per thread: v(NY+1) = v(NY),
// then v(y) = 2 v(y) - v(y+1) for y = NY..1.
__global__ void k_gf2g_cols(double *__restrict__ fort, int nx, int ny, int nz)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nx * nz) return;
    const int x = e % nx, z = e / nx;
    double up = fort[fidx(x, z, ny, nx, nz)];
    fort[fidx(x, z, ny + 1, nx, nz)] = up;
    for (int y = ny; y >= 1; --y) {
        const long q = fidx(x, z, y, nx, nz);
        const double v = 2.0 * fort[q] - up;
        fort[q] = v;
        up = v;
    }
}

// ... then the zero-mass-flux ghost planes over whole planes, pads included:
// v(:,:,1) = -v(:,:,2), v(:,:,NY+1) = -v(:,:,NY).
__global__ void k_gf2g_walls(double *__restrict__ fort, int nx, int ny, int nz)
{
    const long plane = (long)(nx + 2) * (nz + 2);
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < plane; e += (long)gridDim.x * blockDim.x) {
        fort[plane + e] = -fort[2 * plane + e];
        fort[(long)(ny + 1) * plane + e] = -fort[(long)ny * plane + e];
    }
}

}  // namespace

void launch_to_file(hipStream_t s, const Geo &g, const double *fort, double *out, bool g2gf)
{
    const long n = (long)g.nx * g.ny * g.nz;
    const int blocks = (int)std::min<long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_file, dim3(blocks), dim3(256), 0, s, fort, out, g.nx, g.ny, g.nz, g2gf ? 1 : 0);
}

void launch_from_file(hipStream_t s, const Geo &g, const double *in, double *fort, bool gf2g)
{
    const long n = (long)(g.nx + 2) * (g.nz + 2) * (g.ny + 2);
    const int blocks = (int)std::min<long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_from_file, dim3(blocks), dim3(256), 0, s, in, fort, g.nx, g.ny, g.nz);
    if (gf2g) {
        hipLaunchKernelGGL(k_gf2g_cols, dim3((g.nx * g.nz + 255) / 256), dim3(256), 0, s, fort, g.nx, g.ny, g.nz);
        const long plane = (long)(g.nx + 2) * (g.nz + 2);
        hipLaunchKernelGGL(k_gf2g_walls, dim3((unsigned)((plane + 255) / 256)), dim3(256), 0, s, fort, g.nx, g.ny,
                           g.nz);
    }
}

// explicit instantiations: FP64 path and the mixed path (FP32 transforms)
template void launch_zinv<double, double>(hipStream_t, const Geo &, const Fft<double> &, const cx<double> *, long,
                                          long, int, int, cx<double> *, long, long, int, int, int, int);
template void launch_x_r2c<double>(hipStream_t, const Geo &, const Fft<double> &, const double *, cx<double> *, int,
                                   int);
template void launch_zfwd_pack<double>(hipStream_t, const Geo &, const Fft<double> &, const cx<double> *, double *,
                                       int, int);
template void launch_x_c2r<double>(hipStream_t, const Geo &, const Fft<double> &, const cx<double> *, int, double *,
                                   int, int);

}  // namespace dgpu
