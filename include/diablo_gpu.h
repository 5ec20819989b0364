/*
 * diablo_gpu.h -- C-ABI of the MI355X channel time-stepper (libdiablo_gpu.so).
 *
 * The library replaces the arithmetic of DIABLO 2.0's channel path
 * (diablo_2.0/for/channel.f, for/mpi.f, for/fft.f) behind the reference's own
 * call surface.  A Fortran bridge (diablo_amd/fortran/diablo_gpu_bridge.f)
 * keeps the reference subroutine names RK_CHAN_1, REM_DIV_CHAN,
 * POISSON_P_CHAN, FFT_XZ_TO_FOURIER and FFT_XZ_TO_PHYSICAL and forwards to the
 * entry points below through ISO_C_BINDING.
 *
 * Conventions
 *  - Plain C types only; every array argument is a host pointer into a
 *    Fortran-layout array: a field is REAL*8 (0:NX+1, 0:NZ+1, 0:NY+1),
 *    column-major, aliased with COMPLEX*16 (0:NX/2, 0:NZ+1, 0:NY+1)
 *    (for/header:214-263).
 *  - The Fourier state (CU1..3, CP, CTH, CF1..3, CFTH) lives in device
 *    memory between calls.  dgpu_set_field / dgpu_get_field move it to and
 *    from the Fortran arrays; only the dealiased slots I = 0..NKX,
 *    K = 0..TNKZ are carried (every other slot reads back as zero).
 *  - Every function returns 0 on success and a negative DGPU_E_* code on
 *    failure; dgpu_last_error() gives the message.  No C++ exception crosses
 *    the boundary.  Calls on one context are not thread-safe.
 *  - Work is enqueued on the context's own HIP stream.  Functions that take or
 *    return host arrays synchronise before returning; dgpu_rk_chan_1,
 *    dgpu_rem_div_chan and dgpu_apply_bc_vel only enqueue (call
 *    dgpu_synchronize before reading results on the host).
 */
#ifndef DIABLO_GPU_H
#define DIABLO_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGPU_ABI_VERSION 2
/* Active scalars the device path carries: N_TH = 0 or 1.  The reference
 * loops over N_TH scalars (DO N = 1, N_TH, e.g. for/channel.f:203, 314);
 * the fused kernels here carry one (products, transforms and column systems
 * are laid out for U1, U2, U3, TH), and dgpu_create refuses n_th > 1. */
#define DGPU_MAX_TH 1

enum dgpu_status {
    DGPU_OK = 0,
    DGPU_E_ARG = -1,      /* bad argument or unsupported configuration */
    DGPU_E_HIP = -2,      /* HIP runtime error (message has the HIP error) */
    DGPU_E_NOMEM = -3,    /* device or host allocation failed */
    DGPU_E_STATE = -4     /* call not valid in the current state */
};

/* DGPU_FP64: every kernel in FP64.
 * DGPU_MIXED: the x-z transforms, the nonlinear products and terms, and the
 * per-mode Crank-Nicolson, projection and CP solves run in FP32 on an FP32
 * working copy of the state; the low modes I <= 2, |kz index| <= 2 -- the
 * (0,0) mean-flow mode and pressure gauge among them, and the low-wavenumber
 * Poisson modes -- are solved in FP64 on FP64 columns.  The FP64 state that
 * every other entry point reads and writes is refreshed from the working
 * copy on demand.
 * DGPU_MIXED_F64STATE: FP32 transforms, products and nonlinear terms; the
 * state and every solve stay FP64. */
enum dgpu_precision { DGPU_FP64 = 0, DGPU_MIXED = 1, DGPU_MIXED_F64STATE = 2 };

/* Profiles of dgpu_stats_chan, in this order (the quantities SAVE_STATS_CHAN
 * writes to mean.h5 / mean.txt, for/save_stats.f:350-430, and its scalar
 * part, for/save_stats.f:420-522, for scalar 1). */
enum dgpu_stat {
    DGPU_ST_UME = 0, DGPU_ST_VME, DGPU_ST_WME, DGPU_ST_URMS, DGPU_ST_VRMS, DGPU_ST_WRMS,
    DGPU_ST_UV, DGPU_ST_UW, DGPU_ST_WV, DGPU_ST_DUDY, DGPU_ST_DWDY, DGPU_ST_CP, DGPU_ST_SHEAR,
    DGPU_ST_OMEGA_X, DGPU_ST_OMEGA_Y, DGPU_ST_OMEGA_Z,
    DGPU_ST_THME, DGPU_ST_THRMS, DGPU_ST_THV, DGPU_ST_DTHDY, DGPU_ST_PE_DISS,
    DGPU_NSTAT
};

/* Field identifiers for dgpu_set_field / dgpu_get_field. */
enum dgpu_field {
    DGPU_CU1 = 0, DGPU_CU2 = 1, DGPU_CU3 = 2, DGPU_CP = 3,
    DGPU_CF1 = 4, DGPU_CF2 = 5, DGPU_CF3 = 6,
    DGPU_CTH0 = 16,   /* CTH(:,:,:,n+1)  = DGPU_CTH0  + n */
    DGPU_CFTH0 = 24   /* CFTH(:,:,:,n+1) = DGPU_CFTH0 + n */
};

/*
 * Run parameters: the values of /INPUT/, /SCALAR_VARS/, /GEOPHYS_VARS/,
 * /OSCILL_VARS/, /VARIABLE_VISC/ and /DRHODX_VARS/ (for/header:83-301) that the
 * channel path reads, plus grid_def's sizes.
 *
 * gy points at GY(1..NY+1) (the /grids/y dataset, for/hdf5.f:656).  The
 * library derives GYF, DY, DYF as ReadGridHDF5 + CREATE_GRID_CHAN do
 * (for/hdf5.f:676-703, for/channel.f:1646-1652) and KX, KZ, KX2, KZ2 as
 * INIT_FFT does (for/fft.f:72-105).  A caller that already holds those
 * arrays (the Fortran bridge) may pass them instead: any non-NULL pointer in
 * gyf/dy/dyf (0:NY+1), kx/kx2 (0:NX/2) or kz/kz2 (0:TNKZ) is used verbatim.
 */
typedef struct dgpu_params {
    int32_t nx, ny, nz, n_th;
    int32_t precision;
    int32_t f_type;
    int32_t u_bc_ymin, v_bc_ymin, w_bc_ymin;
    int32_t u_bc_ymax, v_bc_ymax, w_bc_ymax;
    int32_t th_bc_ymin[DGPU_MAX_TH], th_bc_ymax[DGPU_MAX_TH];
    double nu, beta, lx, ly, lz, nu_v_scale;
    double i_ro, grav_x, grav_y, grav_z;
    double px0, omega0, amp_omega0;
    double u_bc_ymin_c1, v_bc_ymin_c1, w_bc_ymin_c1;
    double u_bc_ymax_c1, v_bc_ymax_c1, w_bc_ymax_c1;
    double th_bc_ymin_c1[DGPU_MAX_TH], th_bc_ymax_c1[DGPU_MAX_TH];
    double ri[DGPU_MAX_TH], pr[DGPU_MAX_TH];
    double drhodx[DGPU_MAX_TH], drhodz[DGPU_MAX_TH];
    const double *gy;
    const double *gyf, *dy, *dyf;
    const double *kx, *kz, *kx2, *kz2;
} dgpu_params;

typedef struct dgpu_ctx dgpu_ctx;

/* Library version (DGPU_ABI_VERSION) and a one-line build description. */
int32_t dgpu_abi_version(void);
const char *dgpu_build_info(void);

/* Creates a context on the current HIP device: INIT_CHAN_MPI index ranges
 * (for/mpi.f:1080-1100), grid metrics, wavenumbers, FFT plans and device
 * buffers.  Replaces the set-up half of INITIALIZE for this path
 * (for/diablo_io.f:69-124).  The state starts at zero. */
int32_t dgpu_create(const dgpu_params *p, dgpu_ctx **out);
int32_t dgpu_destroy(dgpu_ctx *ctx);
const char *dgpu_last_error(const dgpu_ctx *ctx);

/* Copies the kept Fourier slots of one Fortran-layout field to the device
 * (set) or back (get).  get writes every slot of planes 0..NY+1: kept slots
 * get the state, the rest zero. */
int32_t dgpu_set_field(dgpu_ctx *ctx, int32_t field, const double *fortran_array);
int32_t dgpu_get_field(dgpu_ctx *ctx, int32_t field, double *fortran_array);

/* RK_CHAN_1 (for/channel.f:62-1073): one RK3 / Crank-Nicolson substep of the
 * device-resident state, then REM_DIV_CHAN and the CP update.  rk_step is
 * RK_STEP (1..3); time is TIME (read only when F_TYPE = 2); h_bar, beta_bar
 * and zeta_bar are the three-entry /RKW3_PARAMS/ arrays (for/header:175). */
int32_t dgpu_rk_chan_1(dgpu_ctx *ctx, int32_t rk_step, double time, const double *h_bar,
                       const double *beta_bar, const double *zeta_bar);

/* Substep graphs: by default dgpu_rk_chan_1 records its kernels once per RK
 * stage into a HIP graph and replays it (the stage constants are part of the
 * key; F_TYPE = 2 and kernel timing run without graphs).  enable = 0 switches
 * graphs off for this context. */
int32_t dgpu_set_graphs(dgpu_ctx *ctx, int32_t enable);

/* REM_DIV_CHAN (for/channel.f:1103-1214) on the device-resident CU1..3. */
int32_t dgpu_rem_div_chan(dgpu_ctx *ctx);

/* POISSON_P_CHAN (for/channel.f:1217-1453): CP from the device-resident
 * CU1..3 and CTH.  Rows 1..NY of CP are written (as in the reference); the
 * reference's use of CF1..3 as scratch is not reproduced (device scratch is
 * used instead, CF is left unchanged).  Synchronises before returning. */
int32_t dgpu_poisson_p_chan(dgpu_ctx *ctx);

/* APPLY_BC_VEL_MPI (for/mpi.f:1371-1385 -> for/channel.f:2285-2492) on the
 * device-resident CU1..3. */
int32_t dgpu_apply_bc_vel(dgpu_ctx *ctx);

/* FFT_XZ_TO_FOURIER(V, VV, JMIN, JMAX) (for/mpi.f:1530-1609) on a host
 * Fortran-layout array, in place (V and VV are the same memory, as in every
 * reference call).  Planes outside JMIN..JMAX are not touched.  Kept slots
 * are as in the reference; I > NKX, the padding column and rows NZ..NZ+1 are
 * zero; the slots K = TNKZ+1..NZ-1, which the reference leaves holding
 * unscaled intermediate values that no caller reads, are zero here. */
int32_t dgpu_fft_xz_to_fourier(dgpu_ctx *ctx, double *v, int32_t jmin, int32_t jmax);

/* FFT_XZ_TO_PHYSICAL(VV, V, JMIN, JMAX) (for/mpi.f:1613-1674) on a host
 * Fortran-layout array, in place.  Uses slots I = 0..NX/2-1, K = 0..TNKZ of
 * each plane (x modes above NKX are transformed, not zeroed, as in the
 * reference); writes x = 0..NX-1 of rows 0..NZ-1, zeroes the two x pads of
 * those rows and leaves rows NZ..NZ+1 untouched. */
int32_t dgpu_fft_xz_to_physical(dgpu_ctx *ctx, double *vv, int32_t jmin, int32_t jmax);

/* Restart files: the per-field conversion of WriteHDF5 and ReadHDF5
 * (for/hdf5.f:172-268, 470-534; the HDF5 calls, the Resolution and Time
 * attributes and the file itself stay with the caller).  field is DGPU_CU1
 * (dataset U), DGPU_CU2 (V), DGPU_CU3 (W), DGPU_CP (P) or DGPU_CTH0 + n
 * (TH<n+1>); the file array is tmp(1:NX, 1:NY, 1:NZ), x fastest, NX*NY*NZ
 * doubles.  get: FFT_XZ_TO_PHYSICAL of the device field, G2GF for V, SWAPZY;
 * the device state is not changed (the reference's round trip through
 * physical space changes it by round-off only).  set: SWAPYZ, GF2G for V
 * (ghost planes V(1) = -V(2), V(NY+1) = -V(NY)), FFT_XZ_TO_FOURIER into the
 * device field's kept slots, planes 0..NY+1. */
int32_t dgpu_get_flow_physical(dgpu_ctx *ctx, int32_t field, double *out);
int32_t dgpu_set_flow_physical(dgpu_ctx *ctx, int32_t field, const double *in);

/* COURANT (for/courant.f:1-65), for VARIABLE_DT runs.  The reference calls it
 * inside RK_CHAN_1 at RK_STEP = 3 when MOD(TIME_STEP, UPDATE_DT) = 0
 * (for/channel.f:1037-1040), on the physical velocity after the implicit
 * solves and before the projection; that step still uses the old DELTA_T.
 * dgpu_courant_request makes the next dgpu_rk_chan_1 call with rk_step = 3
 * record the per-plane maxima of |U1|, |U2|, |U3| of that velocity (in the
 * precision of the run's transforms; the call runs without a graph).
 * dgpu_courant then applies COURANT's limits -- viscous/diffusive, rotation,
 * buoyancy, cfl * grid spacing / |velocity|, the 0.0001 and 999 caps -- and
 * returns DELTA_T and H_BAR(1..3) (h_bar: 3 doubles) for the caller's next
 * steps; it waits for the step-3 call to finish.  DGPU_E_STATE when no
 * maxima were recorded since the request. */
int32_t dgpu_courant_request(dgpu_ctx *ctx);
int32_t dgpu_courant(dgpu_ctx *ctx, double cfl, double *delta_t, double *h_bar);

/* Waits for all work enqueued on the context. */
int32_t dgpu_synchronize(dgpu_ctx *ctx);

/* The statistics of SAVE_STATS_CHAN (for/save_stats.f:106-347, scalar part
 * for/save_stats.f:420-522; the running means, movies and file output stay
 * with the Fortran driver) from the device state, in FP64.  profiles:
 * DGPU_NSTAT * NY doubles, profiles[q * NY + (J - 1)] for J = 1..NY,
 * q = DGPU_ST_*; *ubulk: INTEGRATE_Y_VAR of ume (for/mpi.f:1676-1698).  The
 * scalar profiles are for scalar 1 (zero when N_TH = 0).  omega_z is the
 * z vorticity (the reference's mean.h5 branch writes omega_y under that
 * name, for/save_stats.f:404-406). */
int32_t dgpu_stats_chan(dgpu_ctx *ctx, double *profiles, double *ubulk);

/* dgpu_stats_chan without any allocation (no counterpart in the reference).
 * For callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc: the same
 * kernel and the same results as dgpu_stats_chan, but the per-plane sums go
 * through a fixed device array of the library (DGPU_NSTAT *
 * DGPU_NOALLOC_MAX_NY doubles, 688 KiB, one copy per GPU, held until the
 * process ends) instead of a buffer allocated on the first call.  NY above
 * DGPU_NOALLOC_MAX_NY is refused with DGPU_E_ARG.  Every context on one GPU
 * shares that array: calls that overlap on one GPU (contexts on different
 * streams, or host threads) race and must be serialised by the caller.
 * Synchronises the context's stream before returning, like dgpu_stats_chan. */
#define DGPU_NOALLOC_MAX_NY 4096
int32_t dgpu_stats_chan_noalloc(dgpu_ctx *ctx, double *profiles, double *ubulk);

/* tkebudget_chan's pseudo-dissipation epsilon = NU <du_i/dx_j du_i/dx_j>
 * (for/save_stats.f:965-1159), as SAVE_STATS_CHAN calls it, from the device
 * state in FP64: epsilon[J - 1], J = 1..NY (tke.h5's 'epsilon').  Planes
 * 2..NY follow the reference, including its dv/dy mean removal and the
 * du/dz values its scratch still holds on plane NY; on plane 1 the reference
 * also adds terms from a scratch plane it never sets, which are left out. */
int32_t dgpu_tke_chan(dgpu_ctx *ctx, double *epsilon);

/* les_chan with LES_MODEL_TYPE = 1, the constant Smagorinsky model
 * (for/les.f:1-331, called from RK_CHAN_1 at for/channel.f:303-307), in FP64
 * on the device state at NPROCY = 1: APPLY_BC_VEL and APPLY_BC_LES
 * (for/les.f:805-864) on CU1-CU3, then CF1, CF3 -= d tau_1j/dx_j, d tau_3j/dx_j
 * on planes JSTART..NY-1 and CF2 -= d tau_2j/dx_j on planes 2..NY with
 * tau_ij = -2 C^2 DELTA^2 |S| S_ij (C = 0.13, DELTA^3 = DX DY 2 DZ), the
 * implicitly treated d<U>/dy parts left out of tau_12, tau_23, and the
 * explicit cross terms d(NU_T dU_i/dy)/dx_i added to CF2 (for/les.f:302-331).
 * nu_t (may be NULL): NU_T in physical space, field layout (NY+2) x (NZ+2) x
 * (NX+2), zero on planes 0, 1, 2, NY, NY+1 (GHOST_LES_MPI).  The reference's
 * implicit NU_T terms in the Crank-Nicolson solves are not part of this call. */
int32_t dgpu_les_chan(dgpu_ctx *ctx, double *nu_t);

/* RK_CHAN_1's LES switches (FP64 contexts only).  les: input.dat's LES =
 * .TRUE.: dgpu_rk_chan_1 then solves the Crank-Nicolson systems in physical
 * space with the eddy viscosity NU_T (and KAPPA_T = NU_T for the scalar) in
 * both halves (for/channel.f:628-659, 748-760, 793-806, 862-874, 927-939,
 * 991-1003).  model: the branch condition of for/channel.f:303
 * (LES .AND. (.NOT.CREATE_NEW_FLOW .OR. TIME_STEP > 100)), set by the caller
 * before each step: les_chan runs in the substep, then the physical-space
 * velocity BCs (for/channel.f:2495-2640), KAPPA_T = NU_T and the scalar's
 * horizontal eddy diffusion (for/channel.f:313-375).  NU_T starts at zero and
 * keeps its last value while the model is off.  LES substeps are not
 * captured in graphs.  The first call with les = 1 allocates 19
 * Fortran-layout work planes, (NX+2)(NZ+2)(NY+2) doubles each (5.9 GB at
 * 384 x 257 x 384), held until dgpu_destroy. */
int32_t dgpu_set_les(dgpu_ctx *ctx, int32_t les, int32_t model);

/* tkebudget_chan_les (for/save_stats.f:887-960), as SAVE_STATS_CHAN calls it
 * under LES: eps_sgs[J - 1], J = 1..NY, the sum over the physical points of
 * plane J of u_i times its implicitly treated eddy-viscosity term, from the
 * device state and NU_T, in FP64 (tke.h5's 'epsilon_sgs').  eps_sgs[0] is
 * NaN, as in the reference: its U2 term divides 0 by DYF(0), which is never
 * set.  LES must be on (dgpu_set_les). */
int32_t dgpu_tke_les(dgpu_ctx *ctx, double *eps_sgs);

/* The optional DFT-as-GEMM z transforms of a DGPU_MIXED / DGPU_MIXED_F64STATE
 * context (no counterpart in the reference, which transforms with FFTW,
 * for/mpi.f:1530-1674): the substep's z backward and z forward (with the
 * i kz derivative folded in) as real GEMMs on the matrix cores with split-BF16
 * operands (three products) and FP32 accumulation.  NX <= 256, NZ <= 256 and a
 * multiple of 16; the x transforms stay FFTs.  on = 0 returns to the FFTs. */
int32_t dgpu_set_zdft(dgpu_ctx *ctx, int32_t on);

/* Column-stage launch shape (no counterpart in the reference): the
 * Crank-Nicolson and REM_DIV column solves of RK_CHAN_1 run as segmented
 * solves (each y column cut into 32 segments in FP64, 16 in FP32, whose
 * carries pass through LDS) in blocks of `waves` wavefronts: 4, 8 or 16
 * (modes per block 2 * waves in FP64, 4 * waves in FP32); 0 selects the
 * whole-column sweeps.  The default is 8, what dgpu_autotune measures best
 * at 384 x 257 x 384 (environment: DGPU_COLSEG_NW, DGPU_COLSEG=0).  A shape
 * the column does not fit (more than 17 rows per segment, NY > 542 in
 * FP64) is refused with DGPU_E_ARG; a context whose NY is outside the
 * segmented kernels' range runs the sweeps whatever the setting.  The
 * results are the same bits for 4, 8 and 16 (the same operations in the
 * same order for each mode); the sweeps round differently (a different
 * elimination order), within round-off.
 * fused (contexts with the FP32 state, DGPU_MIXED, and FP64 contexts): 1
 * runs the whole column stage -- the four CN solves, the divergence, the
 * Poisson solve, the projection and the CP update -- as one kernel that
 * reads and writes the state once (30 segments, NY <= 268; 4 or 8 waves in
 * FP64); 0 keeps the two kernels; -1 (the default; environment
 * DGPU_COLFUSE) fuses from NY = 160 up (in FP64 only at 8 waves) and, with
 * the FP32 state, up to NY = 142 (16 segments there, 22 up to NY = 196,
 * else 30).  With the
 * FP32 state the fused kernel rounds differently from the two kernels (30
 * segments, not 16), within FP32 round-off; in FP64 (30 against 32
 * segments) they agree to round-off.  An RK stage 3 with a COURANT request runs the
 * two kernels (the maxima are taken between them).
 * dgpu_get_column_shape reports the waves in use (0 for the sweeps) and
 * whether the fused kernel runs (fused may be NULL). */
int32_t dgpu_set_column_shape(dgpu_ctx *ctx, int32_t waves, int32_t fused);
int32_t dgpu_get_column_shape(const dgpu_ctx *ctx, int32_t *waves, int32_t *fused);

/* Launch-shape autotuner (no counterpart in the reference): times `steps` RK3
 * steps of size delta_t (after one untimed step) of the context's own state
 * for each candidate -- planes per block of the fused x kernel (4..64), and
 * for mixed contexts with NX, NZ <= 256 the FFT or the MFMA z transforms;
 * then, with those fixed, the column shape (dgpu_set_column_shape: 4, 8 or
 * 16 waves per block, unless the sweeps are selected; with the FP32 state
 * and in FP64 also the fused column kernel on and off) -- keeps the fastest and puts
 * the state back as it was (the run does not advance).
 * Outputs (may be NULL): the chosen planes per block, whether the MFMA z
 * transforms are on, and the best ms per step.  Not with a COURANT request
 * pending.  Allocates a copy of the state while it runs. */
int32_t dgpu_autotune(dgpu_ctx *ctx, int32_t steps, double delta_t, int32_t *xchunk, int32_t *zdft,
                      double *ms_per_step);

/* Kernel timing: when enabled, every kernel launch is bracketed by HIP events
 * on the context's stream.  dgpu_kernel_times returns, per kernel kind
 * (0 zinv, 1 xfused, 2 zfwd_nl, 3 cn, 4 remdiv), the summed milliseconds and the
 * number of launches since the last reset.  Timing syncs the stream. */
int32_t dgpu_set_kernel_timing(dgpu_ctx *ctx, int32_t enable);
int32_t dgpu_kernel_times(dgpu_ctx *ctx, double *ms, int64_t *count, int32_t n);

/*
 * Ensembles (beyond the reference surface: its sizes are compile-time
 * PARAMETERs, for/header:8-13).  n independent runs, each with its own grid
 * sizes, parameters and state, advanced together: dgpu_ensemble_rk_chan_1
 * runs RK_CHAN_1 on every member, as one HIP graph per RK stage whose
 * branches are the members.  h_bar, beta_bar and zeta_bar hold 3 values per
 * member (member-major).  Members are ordinary contexts for every other call
 * (dgpu_ensemble_member): set/get fields, initial transforms, REM_DIV.
 * Members of one (NX, NZ, N_TH) group share batched launches (one per
 * kernel kind); with several groups each group's launches run on a stream
 * of its own, forked from and joined into the ensemble's stream
 * (environment DGPU_ENS_STREAMS=0 at dgpu_ensemble_create: one stream; the
 * same bits either way).
 * Ordering: the ensemble's launches run on streams of its own.  Call
 * dgpu_ensemble_synchronize before any call on a member after
 * dgpu_ensemble_rk_chan_1, and finish work enqueued on a member
 * (dgpu_synchronize on it) before the next dgpu_ensemble_rk_chan_1.  A new
 * h_bar for a stage waits for the ensemble's earlier launches before the
 * stage's Crank-Nicolson tables are rebuilt.  A member's COURANT request
 * (dgpu_courant_request) is served as for a single context: the rk_step = 3
 * call then runs every member through its own launches (no graphs, no
 * batched launches) and records the maxima for dgpu_courant.
 */
typedef struct dgpu_ensemble dgpu_ensemble;
int32_t dgpu_ensemble_create(const dgpu_params *members, int32_t n, dgpu_ensemble **out);
int32_t dgpu_ensemble_destroy(dgpu_ensemble *ens);
int32_t dgpu_ensemble_size(const dgpu_ensemble *ens);
dgpu_ctx *dgpu_ensemble_member(dgpu_ensemble *ens, int32_t i);
const char *dgpu_ensemble_last_error(const dgpu_ensemble *ens);
int32_t dgpu_ensemble_set_graphs(dgpu_ensemble *ens, int32_t enable);
int32_t dgpu_ensemble_rk_chan_1(dgpu_ensemble *ens, int32_t rk_step, const double *h_bar,
                                const double *beta_bar, const double *zeta_bar);
int32_t dgpu_ensemble_synchronize(dgpu_ensemble *ens);

/* Ensemble launch-shape autotuner (no counterpart in the reference): times
 * `steps` RK3 steps (after one untimed step, wall clock between ensemble
 * synchronisations) of the members' own states for each candidate -- planes
 * per block of the batched x kernel (8..64, every member) and the column
 * shape (dgpu_set_column_shape's waves: 4, 8, 16, every member; the fused
 * setting of each member stays) -- keeps the fastest and puts every
 * member's state back (the run does not advance).  h_bar, beta_bar and
 * zeta_bar as for dgpu_ensemble_rk_chan_1.  Outputs (may be NULL): the
 * planes per block, the waves, the best ms per step.  Not with a member's
 * COURANT request pending.  Allocates a copy of every member's state while
 * it runs. */
int32_t dgpu_ensemble_autotune(dgpu_ensemble *ens, int32_t steps, const double *h_bar, const double *beta_bar,
                               const double *zeta_bar, int32_t *xchunk, int32_t *waves, double *ms_per_step);

#ifdef __cplusplus
}
#endif

#endif /* DIABLO_GPU_H */
