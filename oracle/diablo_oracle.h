/*
 * diablo_oracle.h -- CPU FP64 restatement of DIABLO 2.0's channel time-stepper.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker that the HIP product
 * path is compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (diablo_amd/, libdiablo_gpu.so)
 * never links, imports or falls back to it.
 *
 * Layout follows the reference exactly (for/header:214-263): every field is one
 * REAL*8 array (0:NX+1, 0:NZ+1, 0:NY+1), column-major (i fastest, then k, then
 * j), and its Fourier view is COMPLEX*16 (0:NX/2, 0:NZ+1, 0:NY+1) over the same
 * bytes (the EQUIVALENCE at for/header:247-257).  The configuration is the MPI
 * code path with NPROCS = NPROCY = NPROCZ = 1 (for/mpi.f:868-1135).
 */
#ifndef DIABLO_ORACLE_H
#define DIABLO_ORACLE_H

#include <stdint.h>

#define DOR_MAX_TH 4

/* Runtime parameters: the subset of /INPUT/, /SCALAR_VARS/, /GEOPHYS_VARS/,
 * /OSCILL_VARS/, /VARIABLE_VISC/ and /DRHODX_VARS/ (for/header:83-301) that the
 * channel path reads. */
typedef struct dor_params {
    int32_t nx, ny, nz, n_th;
    int32_t f_type;
    int32_t u_bc_ymin, v_bc_ymin, w_bc_ymin;
    int32_t u_bc_ymax, v_bc_ymax, w_bc_ymax;
    int32_t th_bc_ymin[DOR_MAX_TH], th_bc_ymax[DOR_MAX_TH];
    double nu, beta, lx, ly, lz, nu_v_scale, delta_t;
    double i_ro, grav_x, grav_y, grav_z;
    double px0, omega0, amp_omega0;
    double u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1;
    double u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1;
    double th_bc_ymin_c1[DOR_MAX_TH], th_bc_ymax_c1[DOR_MAX_TH];
    double ri[DOR_MAX_TH], pr[DOR_MAX_TH];
    double drhodx[DOR_MAX_TH], drhodz[DOR_MAX_TH];
} dor_params;

/* Field identifiers for dor_field(). */
enum {
    DOR_U1 = 0, DOR_U2, DOR_U3, DOR_P, DOR_R1, DOR_R2, DOR_R3,
    DOR_F1, DOR_F2, DOR_F3, DOR_S1,
    DOR_NUT = 11,      /* NU_T (physical, LES) */
    DOR_TH0 = 16,      /* TH(:,:,:,n)   = DOR_TH0   + n (0-based n) */
    DOR_KAPPA0 = 20,   /* KAPPA_T(:,:,:,n) (physical, LES) */
    DOR_FTH0 = 24,     /* FTH, unused by the channel path but kept   */
    DOR_RTH0 = 32,     /* RTH(:,:,:,n)  */
    DOR_CTH0 = 40,     /* CTH  (Fourier, not aliased with TH)        */
    DOR_CFTH0 = 48,    /* CFTH */
    DOR_CRTH0 = 56     /* CRTH */
};

/* Grid / wavenumber vector identifiers for dor_vec(). */
enum {
    DOR_GY = 0, DOR_GYF, DOR_DY, DOR_DYF, DOR_KX, DOR_KZ, DOR_KX2, DOR_KZ2,
    DOR_GX, DOR_GZ
};

typedef struct dor_state dor_state;

/* Creates the state: INIT_CHAN_MPI index ranges, grid metrics from GY(1..NY+1)
 * (ReadGridHDF5 + CREATE_GRID_CHAN), INIT_FFT wavenumbers and the RKW3 table.
 * gy_1 holds NY+1 values GY(1..NY+1). Returns NULL on bad sizes. */
dor_state *dor_create(const dor_params *p, const double *gy_1);
void dor_destroy(dor_state *s);

double *dor_field(dor_state *s, int id);
double *dor_vec(dor_state *s, int id);
int dor_info(const dor_state *s, int what);  /* 0 jstart 1 jend 2 nkx 3 nkz 4 tnkz 5 nxp 10+n jstart_th 20+n jend_th */
void dor_set_time(dor_state *s, double time);
void dor_set_dt(dor_state *s, double dt);
void dor_set_threads(int n);
double dor_get_dt(const dor_state *s);

/* COURANT (for/courant.f:1-65) on the current physical U1..U3 arrays: sets
 * and returns DELTA_T and sets H_BAR.  dor_courant_request makes the next
 * dor_rk_chan_1(s, 3) call COURANT where RK_CHAN_1 does under VARIABLE_DT
 * (for/channel.f:1037-1040), on the velocity before the projection. */
double dor_courant(dor_state *s, double cfl);
void dor_courant_request(dor_state *s, double cfl);

/* LES switches of RK_CHAN_1: les = input.dat's LES (the eddy viscosity and
 * diffusivity enter the Crank-Nicolson terms, for/channel.f:628-659, 748-760,
 * 793-806, 862-874, 927-939, 991-1003); model = the branch condition of
 * for/channel.f:303 (les_chan is called this substep, then KAPPA_T = NU_T and
 * the scalars' horizontal eddy-diffusion terms, for/channel.f:303-375).
 * NU_T and KAPPA_T start at zero. */
void dor_set_les(dor_state *s, int les, int model);

/* The reference call surface (for/mpi.f:1530-1674, for/channel.f:62-1453). */
void dor_fft_xz_to_fourier(dor_state *s, double *v, int jmin, int jmax);
void dor_fft_xz_to_physical(dor_state *s, double *vv, int jmin, int jmax);
void dor_rk_chan_1(dor_state *s, int rk_step);
void dor_rem_div_chan(dor_state *s);
void dor_poisson_p_chan(dor_state *s);
void dor_apply_bc_vel(dor_state *s);
void dor_apply_bc_scalar(dor_state *s);

/* Building blocks exposed for unit tests. */
/* Statistics of SAVE_STATS_CHAN (for/save_stats.f:106-347 and the scalar
 * part for/save_stats.f:420-522, non-FINAL branch, NPROCY = 1), without the
 * running means, movies and file output: out[q * NY + (j - 1)], j = 1..NY,
 * q = DOR_ST_*, for scalar 1; *ubulk = INTEGRATE_Y_VAR of ume
 * (for/mpi.f:1676-1698).  The state is not modified. */
enum {
    DOR_ST_UME = 0, DOR_ST_VME, DOR_ST_WME, DOR_ST_URMS, DOR_ST_VRMS, DOR_ST_WRMS,
    DOR_ST_UV, DOR_ST_UW, DOR_ST_WV, DOR_ST_DUDY, DOR_ST_DWDY, DOR_ST_CP, DOR_ST_SHEAR,
    DOR_ST_OMEGA_X, DOR_ST_OMEGA_Y, DOR_ST_OMEGA_Z,
    DOR_ST_THME, DOR_ST_THRMS, DOR_ST_THV, DOR_ST_DTHDY, DOR_ST_PE_DISS, DOR_NSTAT
};
void dor_stats_chan(dor_state *s, double *out, double *ubulk);
/* tkebudget_chan's epsilon (for/save_stats.f:965-1159) as SAVE_STATS_CHAN
 * calls it: eps[j - 1], j = 1..NY.  Uses the state's S1 array as the
 * reference uses CS1/S1 (its plane 0 is read unset, so eps[0] depends on
 * what S1 held); the velocity state is not modified. */
void dor_tke_chan(dor_state *s, double *eps);
void dor_tke_les(dor_state *s, double *eps);

/* WriteHDF5 / ReadHDF5's field layout (for/hdf5.f:172-268, 470-534): id is
 * DOR_U1, DOR_U2, DOR_U3, DOR_P or DOR_CTH0 + n; the file array is
 * tmp(1:NX, 1:NY, 1:NZ), x fastest; V is on the fractional grid (G2GF on
 * writing, GF2G on reading).  Writing leaves the state unchanged. */
void dor_flow_to_file(dor_state *s, int id, double *out);
void dor_flow_from_file(dor_state *s, int id, const double *in);

/* les_chan, constant Smagorinsky (for/les.f:1-331): BCs, strain, NU_T on GY,
 * the SGS stress divergence into CF1-3 and the explicit cross terms into
 * CF2.  nu_t (field layout, may be NULL) receives NU_T. */
void dor_les_chan(dor_state *s, double *nu_t);

void dor_thomas_real(double *a, double *b, double *c, double *g, int ny, int nx);
void dor_thomas_complex(double *a, double *b, double *c, double *g_c, int ny, int nxp);
void dor_fft_c2c(double *data, int n, int sign);   /* unnormalised, sign -1 forward */
void dor_r2c(const double *x, double *X, int n);    /* X: n/2+1 complex           */
void dor_c2r(const double *X, double *x, int n);    /* FFTW-style C2R, unnormalised */

#endif
