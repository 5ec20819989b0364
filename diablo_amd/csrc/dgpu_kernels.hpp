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

enum { SYS_U1 = 0, SYS_U3 = 1, SYS_U2 = 2, SYS_TH0 = 3 };

// Per-row constants of one CN system for one RK stage (doubles, kRow per row):
// the matrix after THOMAS_FORWARD_REAL_MPI's elimination (multiplier,
// reciprocal pivot, upper diagonal), whether the row is a boundary row and its
// right-hand side for the (0,0) mode, the explicit y-diffusion stencil weights
// (TEMP1*NU_V_SCALE folded in), the pressure-gradient weight TEMP4/DY (U2),
// the buoyancy interpolation weights (U2), and whether the row is in the
// explicit-term range (CF written) and in the copy-back range.
enum { T_MULT = 0, T_INVPIV, T_UP, T_BC00, T_ISBC, T_WP, T_W0, T_WM, T_CPW, T_B0, T_BM, T_CFR, T_CPY, kRow = 16 };

struct CnArgs {
    Geo g;
    int rk_gt1, f_type;
    double temp3, temp4, temp5;
    double nu, nupr, beta, i_ro, force00;
    double rgx, rgz, drhodx, drhodz;  // RI*GRAV_X, RI*GRAV_Z, background gradients
    const double *tab;                // [sys][NY+2][kRow] for this stage
    const double *kx, *kz, *kx2, *kz2;
    // Lane m solves mode map[m] (null: mode m) of state arrays whose plane
    // stride is g.ps; the nonlinear terms are indexed by that mode with plane
    // stride nl_ps.  The mixed path runs its FP64 low-|k| columns this way.
    const int *map;
    long nl_ps;
    // 1: sweep A only, with no elimination: the solve scratch receives each
    // row's right-hand side and the field is left as it was (the LES path
    // solves in physical space)
    int rhs_only;
};

template <typename T, typename NLT = T>
struct CnPtrs {
    cx<T> *cu1, *cu2, *cu3, *cp, *cth;
    cx<T> *cf1, *cf2, *cf3, *cfth;
    const cx<NLT> *nl[4];  // nonlinear term per system (index SYS_*)
    cx<T> *gs[4];          // solve scratch per system (may alias nl when NLT == T)
};

// REM_DIV_CHAN: per-mode Poisson factors and per-row constants
// rd[j*4 + {0,1,2}] = {1/(DY(j+1)*DYF(j)), 1/DYF(j), 1/DY(j)}.
struct RdArgs {
    Geo g;
    int update_cp, jstart, jend;
    double inv_temp4;
    const double *pmult, *pinv, *rd, *kx, *kz;
    const float *pmultf, *pinvf;  // FP32 copies of the factors (FP32 state)
    const int *map;               // as CnArgs::map; factors have plane stride fac_ps
    long fac_ps;
};

// Column kernels of one ensemble member (batched launches: g.ps and the
// block shape are shared by the members of a launch; A.g.ps is the lane count).
template <typename T, typename NLT>
struct ColDesc {
    CnArgs A;
    RdArgs R;
    CnPtrs<T, NLT> P;
};
// Low-mode columns -> full state, per member (batched scatter).
template <typename TS, typename TF>
struct ScatDesc {
    const cx<TS> *side;
    cx<TF> *full;
    long side_fs, full_fs;
    const int *map;
    int planes;
};
template <typename TS, typename TF>
void launch_side_scatter_batch(hipStream_t s, const ScatDesc<TS, TF> *d, int nmem, int nside, long ps,
                               int max_planes, int nf);
template <typename T, typename NLT>
void launch_cn_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem);
// Segmented column kernels (dgpu_colseg.hip): the CN solutions go to P.gs[q]
// (complete new fields), k_remdiv_seg projects them into the state.  nw:
// waves per block (4, 8 or 16).
template <typename T>
bool colseg_supported(int ny, int nw);
template <typename T, typename NLT>
void launch_cn_seg(hipStream_t s, int nw, const CnArgs &a, const CnPtrs<T, NLT> &p);
template <typename T, typename NLT>
void launch_remdiv_seg(hipStream_t s, int nw, const RdArgs &a, const CnPtrs<T, NLT> &p);
template <typename T, typename NLT>
void launch_cn_seg_batch(hipStream_t s, int nw, const Geo &g, int max_ny, const ColDesc<T, NLT> *d, int nmem);
template <typename T, typename NLT>
void launch_remdiv_seg_batch(hipStream_t s, int nw, const Geo &g, int max_ny, const ColDesc<T, NLT> *d, int nmem);
// The fused column stage (FP32 or FP64 state): the CN solves, the
// divergence, the Poisson solve, the projection and the CP update in one
// kernel, no scratch.
template <typename T>
bool colfused_supported(int ny, int nw);
template <typename T, typename NLT>
void launch_colfused(hipStream_t s, int nw, const CnArgs &a, const RdArgs &r, const CnPtrs<T, NLT> &p);
template <typename T, typename NLT>
void launch_colfused_batch(hipStream_t s, int nw, const Geo &g, int max_ny, const ColDesc<T, NLT> *d, int nmem);
template <typename T, typename NLT>
void launch_remdiv_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem);
template <typename T, typename NLT>
void launch_cn_side_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem);
template <typename T, typename NLT>
void launch_remdiv_side_batch(hipStream_t s, const Geo &g, const ColDesc<T, NLT> *d, int nmem);

// ---- launchers (dgpu_kernels.hip) ----
template <typename T, typename TIN>
void launch_zinv(hipStream_t s, const Geo &g, const Fft<T> &fz, const cx<TIN> *in, long in_fs, long in_ps,
                 int in_mk, int ncols, cx<T> *out, long out_fs, long out_ps, int out_mk, int j0, int nplanes,
                 int nfields);


// ---- per-member descriptors of the batched (ensemble) launches ----
// Members of one launch share the transform lengths (NX, NZ); NY, the
// parameters and every array differ.
template <typename T, typename TIN>
struct ZinvDesc {
    const cx<TIN> *in;   // state (fields at in_fs)
    cx<T> *out;          // Zt (fields at out_fs)
    long in_fs, out_fs;
    int nplanes;         // NY+2
};
template <typename T>
struct XfDesc {
    const cx<T> *zt;
    cx<T> *wt;
    long zt_fs, wt_fs;
    const double *dy, *dyf, *kx;
    int jlo, jhi;
};
template <typename T, typename TOUT>
struct ZfDesc {
    const cx<T> *wt;
    long wt_fs;
    cx<TOUT> *const *nl;  // device array of the member's nonlinear-term slots
    const double *kz;
    int jlo, nplanes;
};

// ---- substep transforms on the in-place plans (dgpu_hot_z.hip, dgpu_hot_x.hip) ----
// Zt/Wt rows are ld complex values long; tw points at the plan's twiddles.
template <typename T, typename TIN>
void launch_zinv_ip(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const cx<TIN> *in, long in_fs,
                    long in_ps, cx<T> *out, long out_fs, long out_ps, int ld, int j0, int nplanes, int nfields);
template <typename T>
void launch_xfused_ip(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, int ld, const cx<T> *zt,
                      long zt_fs, cx<T> *wt, long wt_fs, const double *dy, const double *dyf, const double *kx,
                      int jlo, int jhi, int chunk);
template <typename T, typename TOUT>
void launch_zfwd_nl_ip(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const cx<T> *wt, long wt_fs,
                       int ld, cx<TOUT> *const *nl, const double *kz, int jlo, int jhi, int nout);
template <typename T, typename TIN>
void launch_zinv_ip_batch(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const ZinvDesc<T, TIN> *d,
                          int nmem, int max_planes, int ld, int nfields);
template <typename T>
void launch_xfused_ip_batch(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, const XfDesc<T> *d,
                            int nmem, int max_span, int ld, int chunk);
template <typename T, typename TOUT>
void launch_zfwd_nl_ip_batch(hipStream_t s, const Geo &g, const IpLen &fz, const cx<T> *tw, const ZfDesc<T, TOUT> *d,
                             int nmem, int max_planes, int ld, int nout);
// COURANT's velocity maxima from the z-backward rows of CU1..CU3 (zt, planes
// 1..NY): out[q * NY + j - 1] = max |u_{q+1}| over plane j.  part holds
// vmax_part_count doubles.
template <typename T>
void launch_vmax_ip(hipStream_t s, const Geo &g, const IpLen &fx, const cx<T> *tw, int ld, const cx<T> *zt,
                    long zt_fs, double *part, double *out);
int vmax_rows_per_block(const IpLen &fx);
size_t vmax_part_count(const Geo &g, const IpLen &fx);
size_t zip_lds_bytes(const IpLen &fz, size_t elem);

template <typename T, typename NLT>
void launch_cn(hipStream_t s, const CnArgs &a, const CnPtrs<T, NLT> &p);
// The mixed path's low-mode columns (few lanes): deeper prefetch variants.
template <typename T, typename NLT>
void launch_cn_side(hipStream_t s, const CnArgs &a, const CnPtrs<T, NLT> &p);
template <typename T, typename NLT>
void launch_remdiv_side(hipStream_t s, const RdArgs &a, const CnPtrs<T, NLT> &p);
template <typename T, typename NLT>
void launch_remdiv(hipStream_t s, const RdArgs &a, const CnPtrs<T, NLT> &p);
void launch_pois_factor(hipStream_t s, const Geo &g, const double *dy, const double *dyf, const double *kx2,
                        const double *kz2, double *pmult, double *pinv);

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
// POISSON_P_CHAN pieces (dgpu_poisson.hip)
void launch_pp_spec(hipStream_t s, int group, const Geo &g, const cx<double> *cu1, const cx<double> *cu2,
                    const cx<double> *cu3, const double *dyf, const double *kx, const double *kz, double *fa,
                    double *fb, double *fc);
void launch_pp_phys(hipStream_t s, int op, const Geo &g, double *fa, const double *fb, double *fbw, double *fc);
void launch_pp_accum(hipStream_t s, const Geo &g, cx<double> *cs, const double *fa, const double *fb,
                     const double *fc, int first);
void launch_pp_solve(hipStream_t s, const Geo &g, const cx<double> *cs, const cx<double> *cth, double ri,
                     const double *gyf, const double *pmult, const double *pinv, const double *rd, cx<double> *scr,
                     cx<double> *cp);
// SAVE_STATS_CHAN statistics (dgpu_stats.hip): out[q * NY + (j - 1)]
// tkebudget_chan's epsilon as per-plane Parseval sums (10 * NY doubles)
void launch_tke(hipStream_t s, const Geo &g, const cx<double> *cu1, const cx<double> *cu2, const cx<double> *cu3,
                const double *kx, const double *kz, double *out);
void launch_stats(hipStream_t s, const Geo &g, const cx<double> *cu1, const cx<double> *cu2, const cx<double> *cu3,
                  const cx<double> *cp, const cx<double> *cth, const double *dy, const double *dyf,
                  const double *gyf, const double *kx, const double *kz, double *out);
// The fixed statistics array of dgpu_stats_chan_noalloc (dgpu_stats.hip):
// device address of this GPU's copy, resolved once per device with the
// device current (hipGetSymbolAddress).  Returns hipSuccess or the error.
hipError_t noalloc_stats_scratch(double **out);
void launch_unpack_fortran(hipStream_t s, const Geo &g, const double *fort, cx<double> *kept, int j0, int nplanes);
// Restart-file layout: out(1:NX, 1:NY, 1:NZ) from the physical Fortran-layout
// field (G2GF first for V), and back (GF2G for V).
void launch_to_file(hipStream_t s, const Geo &g, const double *fort, double *out, bool g2gf);
void launch_from_file(hipStream_t s, const Geo &g, const double *in, double *fort, bool gf2g);
// Mixed path state copies (dgpu_columns.hip): dst[k] = src[k], k < n.
template <typename TO, typename FROM>
void launch_convert(hipStream_t s, const cx<FROM> *src, cx<TO> *dst, long n);
void launch_convert_real(hipStream_t s, const double *src, float *dst, long n);
// side[f][j][l] <-> full[f*full_fs + j*full_ps + map[l]], f < nf, j < planes
template <typename TS, typename TF>
void launch_side_gather(hipStream_t s, const cx<TF> *full, long full_fs, long full_ps, cx<TS> *side, long side_fs,
                        const int *map, int nside, int planes, int nf);
template <typename TS, typename TF>
void launch_side_scatter(hipStream_t s, const cx<TS> *side, long side_fs, cx<TF> *full, long full_fs, long full_ps,
                         const int *map, int nside, int planes, int nf);
void launch_apply_bc_vel(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu2, cx<double> *cu3,
                         const int *codes, const double *c1, const double *dy, const double *dyf);

// les_chan's constant Smagorinsky model (dgpu_les.hip).  Planes are
// Fortran-layout (0:NX+1, 0:NZ+1, 0:NY+1) doubles, passed by value.
struct LesPtrs {
    double *p[8];
};
void launch_les_bc(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu3, int u_min, int u_max, int w_min,
                   int w_max);
void launch_les_strain(hipStream_t s, const Geo &g, const cx<double> *cu1, const cx<double> *cu2,
                       const cx<double> *cu3, const double *kx, const double *kz, const double *dy, const double *dyf,
                       LesPtrs out);
void launch_les_nut(hipStream_t s, const Geo &g, LesPtrs in, LesPtrs out, const cx<double> *cu1,
                    const cx<double> *cu3, const double *dy, const double *dyf, const double *delta_y, int jstart,
                    int jend);
void launch_les_cf(hipStream_t s, const Geo &g, LesPtrs t, cx<double> *cf1, cx<double> *cf2, cx<double> *cf3,
                   const double *kx, const double *kz, const double *dyf, const double *delta_y,
                   const double *delta_yf, int j1, int j2, bool fresh);
void launch_les_cross_src(hipStream_t s, const Geo &g, const double *nut, const double *u, const double *dy,
                          double *out);
void launch_les_cross_add(hipStream_t s, const Geo &g, const double *ct, cx<double> *cf2, const double *kx,
                          const double *kz, int use_kz);

// z transforms as split-BF16 DFT-as-GEMM on the matrix cores (dgpu_zdft.hip):
// weights fragment-major, hi and lo halves, nmt M tiles of 32 rows, nks
// k-steps of 16.
struct ZdftTab {
    const uint4 *hi, *lo;
    int nmt, nks;
};
template <typename TIN>
void launch_zdft_inv(hipStream_t s, const Geo &g, const ZdftTab &t, const cx<TIN> *in, long in_fs, long in_ps,
                     cx<float> *out, long out_fs, long out_ps, int ld, int j0, int nplanes, int nf);
void launch_zdft_fwd(hipStream_t s, const Geo &g, const ZdftTab &t, const cx<float> *wt, long wt_fs, int ld,
                     cx<float> *const *nl, int jlo, int jhi, int nout);

// RK_CHAN_1 with LES (dgpu_les.hip): the velocity and scalar Crank-Nicolson
// solves in physical space with the eddy viscosity.
struct LesBc {
    int u_lo, u_hi, w_lo, w_hi, v_lo, v_hi;
    double u_lo_c, u_hi_c, w_lo_c, w_hi_c, v_lo_c, v_hi_c;
};
enum { LES_SYS_U13 = 0, LES_SYS_U2 = 1, LES_SYS_TH = 2 };
struct LesCn {
    int sys, ny, jlo, jhi, clo, chi, lo, hi;  // rows, copied rows, BC codes
    double cmol, temp2, clo_c, chi_c;         // molecular TEMP1*NU_V_SCALE (/PR), H_BAR/2, BC values
};
void launch_les_bc_phys(hipStream_t s, const Geo &g, cx<double> *cu1, cx<double> *cu2, cx<double> *cu3,
                        const LesBc &b, const double *dy, const double *dyf);
void launch_les_bc_phys_planes(hipStream_t s, const Geo &g, double *u1, double *u3, const LesBc &b, const double *dy);
void launch_les_pack(hipStream_t s, const Geo &g, const cx<double> *src, double *fort, const double *kx,
                     const double *kz, int kmul, int j0, int j1);
void launch_les_mul(hipStream_t s, const Geo &g, double *a, const double *nut, int j0, int j1);
void launch_les_acc(hipStream_t s, const Geo &g, cx<double> *dst, const double *fort, const double *kx,
                    const double *kz, int kmul, int j0, int j1);
// Up to four independent solves (TH, U2, U1, U3), each with its right-hand
// side, output and two scratch planes.
struct LesCnSet {
    LesCn a[4];
    const double *rhs[4];
    double *out[4], *bm[4], *gw[4];
    int n;
};
void launch_les_cn(hipStream_t s, const Geo &g, const LesCnSet &set, const double *nut, const double *dy,
                   const double *dyf);
void launch_les_vmax(hipStream_t s, const Geo &g, const double *u1, const double *u2, const double *u3, double *out);
void launch_les_add(hipStream_t s, cx<double> *dst, const cx<double> *src, long n);
void launch_les_eps(hipStream_t s, const Geo &g, const double *u1, const double *u2, const double *u3,
                    const double *nut, const double *dy, const double *dyf, double *out);
struct LesCfFix {
    double nu, beta, i_ro, drhodx, drhodz;
    int jstart, jend, jstart_th, jend_th;
};
void launch_les_cf_rows(hipStream_t s, const Geo &g, const cx<double> *pre, const cx<double> *cu1,
                        const cx<double> *cu3, cx<double> *cf1, cx<double> *cf3, cx<double> *cfth, const double *kx2,
                        const double *kz2, const LesCfFix &f);

// Block L of a plane-major grid -> (column chunk, plane): the hardware deals
// consecutive blocks to the 8 XCDs in turn, so L & 7 picks the plane within
// a group of 8 and every chunk of a plane runs on one XCD.  A chunk's W
// columns are W * 16 bytes of a row that need not start on a cache line;
// the line two chunks share is then fetched (and written back) once, from
// that XCD's L2.  false: a block past the last plane.
__device__ __forceinline__ bool plane_chunk(int L, int nchunks, int nplanes, int &chunk, int &plane)
{
    const int r = L >> 3;
    chunk = r % nchunks;
    plane = 8 * (r / nchunks) + (L & 7);
    return plane < nplanes;
}


}  // namespace dgpu
