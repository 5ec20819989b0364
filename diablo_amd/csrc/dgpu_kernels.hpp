// dgpu_kernels.hpp -- host-visible launch interface of the channel kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "dgpu_fft.hpp"

namespace dgpu {

constexpr int kMaxTh = 4;

// Sizes shared by all kernels.  Fourier fields keep only the dealiased modes
// I = 0..NKX, K = 0..TNKZ of each plane, stored [j][K][I] (mode index
// m = K*mk + I), so a y-column of mode m is m + j*ps.
struct Geo {
    int nx, ny, nz, nth;
    int nkx, nkz, tnkz;
    int mk, mz;   // NKX+1, TNKZ+1
    long ps;      // mk*mz
};

// Per-length FFT description plus its twiddle table (device pointer).
template <typename T>
struct Fft {
    FftLen len;
    const cx<T> *tw;
};

// Solve tables for one tridiagonal CN system (rows 0..NY+1): the reference's
// matrix after THOMAS_FORWARD_REAL_MPI (multiplier A', pivot B', upper C),
// which rows are boundary rows and their right-hand side for the (0,0) mode.
struct CnSys {
    const double *mult, *piv, *up, *bc00, *isbc;
    int lo, hi;  // rows copied back into the field
};

enum { SYS_U1 = 0, SYS_U3 = 1, SYS_U2 = 2, SYS_TH0 = 3 };

struct ColArgs {
    Geo g;
    int mode;  // 0: full RK substep, 1: REM_DIV_CHAN only
    int rk_gt1;
    int jstart, jend;
    int jstart_th[kMaxTh], jend_th[kMaxTh];
    double temp1, temp3, temp4, temp5;
    double nu, beta, i_ro, grav_x, grav_y, grav_z, nuv;
    double force00;  // PX0 (F_TYPE 1) or PX0 + AMP*cos(OMEGA0*TIME) (F_TYPE 2)
    int f_type;
    double ri[kMaxTh], pr[kMaxTh], drhodx[kMaxTh], drhodz[kMaxTh];
    const double *dy, *dyf, *kx, *kz, *kx2, *kz2;
    CnSys sys[3 + kMaxTh];
};

template <typename T>
struct ColPtrs {
    cx<T> *cu1, *cu2, *cu3, *cp;
    cx<T> *cth[kMaxTh];
    cx<T> *cf1, *cf2, *cf3;
    cx<T> *cfth[kMaxTh];
    cx<T> *nl[3 + kMaxTh];  // nonlinear accumulators, reused as solve scratch
    cx<T> *pois;            // Poisson elimination scratch (complex)
    T *piv;                 // Poisson pivots
};

// ---- launchers (dgpu_kernels.hip) ----
template <typename T>
void launch_zinv(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<T> *in, long in_fs, long in_ps,
                 int in_mk, int ncols, cx<T> *out, long out_fs, long out_ps, int out_mk, int j0, int nplanes,
                 int nfields);

template <typename T>
void launch_xfused(hipStream_t s, const Geo &g, const Fft<T> &fx, const cx<T> *zt, long zt_fs, cx<T> *wt,
                   long wt_fs, const double *dy, const double *dyf, const double *kx, int jlo, int jhi);

template <typename T>
void launch_zfwd_nl(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<T> *wt, long wt_fs, cx<T> *const *nl,
                    const double *kz, int jlo, int jhi, int nout);

template <typename T>
void launch_columns(hipStream_t s, const ColArgs &a, const ColPtrs<T> &p);

// Standalone FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL on Fortran-layout planes.
template <typename T>
void launch_x_r2c(hipStream_t s, const Geo &g, const Fft<T> &fx, const double *phys, cx<T> *xt, int j0,
                  int nplanes);
template <typename T>
void launch_zfwd_pack(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<T> *xt, double *fort, int j0,
                      int nplanes);
template <typename T>
void launch_x_c2r(hipStream_t s, const Geo &g, const Fft<T> &fx, const cx<T> *zt, int mkin, double *fort, int j0,
                  int nplanes);
void launch_unpack_fortran(hipStream_t s, const Geo &g, const double *fort, cx<double> *kept, int j0, int nplanes);
void launch_apply_bc_vel(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu2, cx<double> *cu3,
                         const int *codes, const double *c1, const double *dy, const double *dyf);

}  // namespace dgpu
