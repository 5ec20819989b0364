/*
 * diablo_oracle.c -- CPU FP64 restatement of the DIABLO 2.0 channel path.
 *
 * TEST INFRASTRUCTURE ONLY (see diablo_oracle.h).  It is the checker for the
 * HIP path and the "port" CPU baseline, never part of the product.
 *
 * Every routine below restates one reference routine loop for loop, with the
 * same index ranges, the same statement order and the same order of
 * floating-point operations inside each statement.  The reference files are
 * under diablo_2.0/for/ (cited as for/<file>:<line>).  Configuration: the MPI
 * code path at NPROCS = NPROCY = NPROCZ = 1, LES off, USE_MPI true.
 *
 * Deliberate differences, none of which changes a kept value:
 *   - FFTW 2.1.5 is replaced by our own mixed-radix FFT (radix 4/2/3/5/generic,
 *     recursive decimation in time).  Round-off differs from FFTW's; the
 *     arithmetic definition (unnormalised, forward sign -1, C2R ignoring the
 *     imaginary part of the DC term) is FFTW's.
 *   - The x-direction real transforms are done two rows at a time through one
 *     complex FFT.
 *   - Complex / integer scalings (for/mpi.f:1552, 1589, 1604) are done
 *     component-wise, i.e. (a+ib)/N = a/N + i b/N.
 *   - The padding column I = NXP of the complex Thomas solves (uninitialised
 *     in the reference, for/header:228-230) is skipped; its result is never
 *     copied out (for/channel.f:1187-1190).
 *   - z transforms of all-zero columns are skipped (their result is zero).
 */
#include "diablo_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* Planted mistakes.  tests/test_oracle_rk_terms.py builds this file with
 * -DDOR_PLANT=n (n >= 1) into a throw-away library and checks that the pins
 * of RK_CHAN_1's explicit terms reject it; each PLANT(n, wrong, right) site
 * below marks the term a pin must guard.  DOR_PLANT = 0, the default build,
 * is the restatement itself: every site takes its right-hand operand. */
#ifndef DOR_PLANT
#define DOR_PLANT 0
#endif
#define PLANT(n, wrong, right) (DOR_PLANT == (n) ? (wrong) : (right))

/* ------------------------------------------------------------------------- */
/* FFT: mixed-radix recursive decimation in time                            */
/* ------------------------------------------------------------------------- */

typedef struct fplan {
    int n, nf;
    int fac[64];
    cplx *tw; /* tw[m] = exp(-2 pi i m / n), m = 0..n-1 */
} fplan;

static void fplan_init(fplan *pl, int n)
{
    int m = n, nf = 0;
    pl->n = n;
    while (m % 4 == 0) { pl->fac[nf++] = 4; m /= 4; }
    while (m % 2 == 0) { pl->fac[nf++] = 2; m /= 2; }
    while (m % 3 == 0) { pl->fac[nf++] = 3; m /= 3; }
    while (m % 5 == 0) { pl->fac[nf++] = 5; m /= 5; }
    for (int p = 7; m > 1; p += 2)
        while (m % p == 0) { pl->fac[nf++] = p; m /= p; }
    pl->nf = nf;
    pl->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    /* Twiddles in extended precision, rounded once to double. */
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int j = 0; j < n; j++) {
        long double a = -twopi * (long double)j / (long double)n;
        pl->tw[j] = (double)cosl(a) + I * (double)sinl(a);
    }
}

static void fplan_free(fplan *pl) { free(pl->tw); pl->tw = NULL; }

static inline cplx tw_get(const fplan *pl, long e, int sign)
{
    cplx w = pl->tw[e % pl->n];
    return sign < 0 ? w : conj(w);
}

static void fft_rec(const fplan *pl, const cplx *in, long istride, cplx *out,
                    int n, int fi, int sign)
{
    if (n == 1) { out[0] = in[0]; return; }
    const int r = pl->fac[fi], m = n / r;
    for (int q = 0; q < r; q++)
        fft_rec(pl, in + q * istride, istride * r, out + (long)q * m, m, fi + 1, sign);
    const long tstep = pl->n / n;
    const long rstep = pl->n / r;
    cplx t[64];
    for (int k = 0; k < m; k++) {
        t[0] = out[k];
        for (int q = 1; q < r; q++)
            t[q] = out[(long)q * m + k] * tw_get(pl, (long)q * k * tstep, sign);
        if (r == 2) {
            out[k] = t[0] + t[1];
            out[m + k] = t[0] - t[1];
        } else if (r == 4) {
            cplx a0 = t[0] + t[2], a1 = t[0] - t[2];
            cplx b0 = t[1] + t[3], b1 = t[1] - t[3];
            /* multiply b1 by -i (forward) or +i (inverse) */
            cplx b1r = sign < 0 ? (cimag(b1) - I * creal(b1)) : (-cimag(b1) + I * creal(b1));
            out[k] = a0 + b0;
            out[m + k] = a1 + b1r;
            out[2 * m + k] = a0 - b0;
            out[3 * m + k] = a1 - b1r;
        } else {
            for (int p = 0; p < r; p++) {
                cplx s = t[0];
                for (int q = 1; q < r; q++)
                    s += t[q] * tw_get(pl, (long)((q * p) % r) * rstep, sign);
                out[(long)p * m + k] = s;
            }
        }
    }
}

/* In-place unnormalised transform of a contiguous complex vector. */
static void fft_run(const fplan *pl, cplx *data, cplx *work, int sign)
{
    memcpy(work, data, sizeof(cplx) * (size_t)(unsigned)pl->n);
    fft_rec(pl, work, 1, data, pl->n, 0, sign);
}

/* Two real rows a, b -> their half spectra A, B (k = 0..n/2). */
static void r2c_pair(const fplan *pl, const double *a, const double *b,
                     cplx *A, cplx *B, cplx *z, cplx *w)
{
    const int n = pl->n;
    for (int j = 0; j < n; j++) z[j] = a[j] + I * (b ? b[j] : 0.0);
    fft_run(pl, z, w, -1);
    for (int k = 0; k <= n / 2; k++) {
        cplx zk = z[k], zn = conj(z[(n - k) % n]);
        A[k] = 0.5 * (zk + zn);
        if (B) {
            cplx d = 0.5 * (zk - zn);
            B[k] = cimag(d) - I * creal(d);
        }
    }
}

/* Two half spectra -> two real rows, FFTW C2R convention: the imaginary parts
 * of the k = 0 and k = n/2 terms are ignored, no normalisation. */
static void c2r_pair(const fplan *pl, const cplx *A, const cplx *B,
                     double *a, double *b, cplx *z, cplx *w)
{
    const int n = pl->n, h = n / 2;
    for (int k = 0; k <= h; k++) {
        cplx ak = A[k], bk = B ? B[k] : 0.0;
        if (k == 0 || k == h) { ak = creal(ak); bk = creal(bk); }
        z[k] = ak + I * bk;
        if (k > 0 && k < h) z[n - k] = conj(ak) + I * conj(bk);
    }
    fft_run(pl, z, w, +1);
    for (int j = 0; j < n; j++) {
        a[j] = creal(z[j]);
        if (b) b[j] = cimag(z[j]);
    }
}

/* ------------------------------------------------------------------------- */
/* State                                                                      */
/* ------------------------------------------------------------------------- */

struct dor_state {
    dor_params p;
    int NX, NY, NZ, NTH, NXP, NZP, NKX, NKZ, TNKZ, NXM;
    int JSTART, JEND, JSTART_TH[DOR_MAX_TH], JEND_TH[DOR_MAX_TH];
    double TIME;
    double H_BAR[4], BETA_BAR[4], ZETA_BAR[4]; /* 1-based */
    int les, les_model;                        /* LES, and les_chan called this substep */
    int courant_pending;                       /* COURANT at the next RK_STEP 3 */
    double courant_cfl;
    double *GX, *GZ, *GY, *GYF, *DY, *DYF;     /* 0:N+1 */
    double *KX, *KZ, *KX2, *KZ2;               /* KX 0:NXP, KZ 0:TNKZ */
    double *fld[64];
    size_t fsize; /* doubles per field */
    fplan px, pz;
};

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

#define NX_ (s->NX)
#define NY_ (s->NY)
#define NZ_ (s->NZ)
#define NXP_ (s->NXP)
#define RIDX(i, k, j) ((size_t)(i) + (size_t)(NX_ + 2) * ((size_t)(k) + (size_t)(NZ_ + 2) * (size_t)(j)))
#define CIDX(i, k, j) ((size_t)(i) + (size_t)(NXP_ + 1) * ((size_t)(k) + (size_t)(NZ_ + 2) * (size_t)(j)))

static int good_fft_size(int n)
{
    if (n < 2 || (n & 1)) return 0;
    return 1;
}

dor_state *dor_create(const dor_params *p, const double *gy_1)
{
    if (!p || p->nx < 4 || p->ny < 4 || p->nz < 4 || p->n_th < 0 || p->n_th > DOR_MAX_TH) return NULL;
    if (!good_fft_size(p->nx) || !good_fft_size(p->nz)) return NULL;
    dor_state *s = (dor_state *)calloc(1, sizeof(dor_state));
    s->p = *p;
    s->NX = p->nx; s->NY = p->ny; s->NZ = p->nz; s->NTH = p->n_th;
    s->NXP = p->nx / 2;               /* for/header_mpi:13, NPROCZ = 1 */
    s->NZP = p->nz;                   /* for/header_mpi:14 */
    s->NKX = p->nx / 3;               /* for/header:22 */
    s->NKZ = p->nz / 3;
    s->TNKZ = 2 * s->NKZ;             /* for/header:28 */
    s->NXM = p->nx - 1;
    const int NY = s->NY;

    /* INIT_CHAN_MPI at RANKY = 0 = NPROCY-1 (for/mpi.f:1080-1100). */
    s->JEND = NY;
    s->JSTART = 2;
    for (int n = 0; n < s->NTH; n++) {
        s->JEND_TH[n] = NY;
        s->JSTART_TH[n] = (p->th_bc_ymin[n] == 1) ? 1 : 2;
    }

    /* Grid: ReadGridHDF5 (for/hdf5.f:656-703) + CREATE_GRID_CHAN
     * (for/channel.f:1572-1652).  GY(0), DY(0), DYF(0) stay 0. */
    s->GX = zalloc(s->NX + 2); s->GZ = zalloc(s->NZ + 2);
    s->GY = zalloc(NY + 2); s->GYF = zalloc(NY + 2);
    s->DY = zalloc(NY + 2); s->DYF = zalloc(NY + 2);
    for (int i = 0; i <= s->NX; i++) s->GX[i] = ((double)i * p->lx) / (double)s->NX;
    for (int k = 0; k <= s->NZ; k++) s->GZ[k] = ((double)k * p->lz) / (double)s->NZ;
    for (int j = 1; j <= NY + 1; j++) s->GY[j] = gy_1[j - 1];
    for (int j = 1; j <= NY; j++) s->GYF[j] = 0.5 * (s->GY[j] + s->GY[j + 1]);
    s->GYF[0] = 2.0 * s->GYF[1] - s->GYF[2];
    s->GYF[NY + 1] = 2.0 * s->GYF[NY] - s->GYF[NY - 1];
    for (int j = 1; j <= NY + 1; j++) s->DY[j] = s->GYF[j] - s->GYF[j - 1];
    for (int j = 1; j <= NY; j++) s->DYF[j] = s->GY[j + 1] - s->GY[j];
    s->DYF[NY + 1] = s->DYF[NY];

    /* INIT_FFT wavenumbers (for/fft.f:72-105): PI = 4*atan(1d0),
     * KX(I) = I*(2*PI)/LX for I = 0..NXPP-1, KZ packed. */
    const double PI = 4.0 * atan(1.0);
    s->KX = zalloc(s->NXP + 1); s->KX2 = zalloc(s->NXP + 1);
    s->KZ = zalloc(s->TNKZ + 1); s->KZ2 = zalloc(s->TNKZ + 1);
    for (int i = 0; i <= s->NXP - 1; i++) {
        s->KX[i] = ((double)i * (2.0 * PI)) / p->lx;
        s->KX2[i] = s->KX[i] * s->KX[i];
    }
    for (int k = 0; k <= s->NKZ; k++) s->KZ[k] = ((double)k * (2.0 * PI)) / p->lz;
    for (int k = 1; k <= s->NKZ; k++) s->KZ[s->TNKZ + 1 - k] = ((double)(-k) * (2.0 * PI)) / p->lz;
    for (int k = 0; k <= s->TNKZ; k++) s->KZ2[k] = s->KZ[k] * s->KZ[k];

    /* RKW3 table (for/diablo_io.f:116-124). */
    dor_set_dt(s, p->delta_t);
    s->BETA_BAR[1] = 1.0; s->BETA_BAR[2] = 25.0 / 8.0; s->BETA_BAR[3] = 9.0 / 4.0;
    s->ZETA_BAR[1] = 0.0; s->ZETA_BAR[2] = -17.0 / 8.0; s->ZETA_BAR[3] = -5.0 / 4.0;

    s->fsize = (size_t)(s->NX + 2) * (size_t)(s->NZ + 2) * (size_t)(NY + 2);
    for (int f = DOR_U1; f <= DOR_S1; f++) s->fld[f] = zalloc(s->fsize);
    for (int n = 0; n < s->NTH; n++) {
        s->fld[DOR_TH0 + n] = zalloc(s->fsize);
        s->fld[DOR_FTH0 + n] = zalloc(s->fsize);
        s->fld[DOR_RTH0 + n] = zalloc(s->fsize);
        s->fld[DOR_CTH0 + n] = zalloc(s->fsize);
        s->fld[DOR_CFTH0 + n] = zalloc(s->fsize);
        s->fld[DOR_CRTH0 + n] = zalloc(s->fsize);
    }
    fplan_init(&s->px, s->NX);
    fplan_init(&s->pz, s->NZ);
    return s;
}

void dor_destroy(dor_state *s)
{
    if (!s) return;
    for (int f = 0; f < 64; f++) free(s->fld[f]);
    free(s->GX); free(s->GZ); free(s->GY); free(s->GYF); free(s->DY); free(s->DYF);
    free(s->KX); free(s->KZ); free(s->KX2); free(s->KZ2);
    fplan_free(&s->px); fplan_free(&s->pz);
    free(s);
}

double *dor_field(dor_state *s, int id) { return (id >= 0 && id < 64) ? s->fld[id] : NULL; }

double *dor_vec(dor_state *s, int id)
{
    switch (id) {
    case DOR_GY: return s->GY;
    case DOR_GYF: return s->GYF;
    case DOR_DY: return s->DY;
    case DOR_DYF: return s->DYF;
    case DOR_KX: return s->KX;
    case DOR_KZ: return s->KZ;
    case DOR_KX2: return s->KX2;
    case DOR_KZ2: return s->KZ2;
    case DOR_GX: return s->GX;
    case DOR_GZ: return s->GZ;
    }
    return NULL;
}

int dor_info(const dor_state *s, int what)
{
    if (what == 0) return s->JSTART;
    if (what == 1) return s->JEND;
    if (what == 2) return s->NKX;
    if (what == 3) return s->NKZ;
    if (what == 4) return s->TNKZ;
    if (what == 5) return s->NXP;
    if (what >= 10 && what < 10 + DOR_MAX_TH) return s->JSTART_TH[what - 10];
    if (what >= 20 && what < 20 + DOR_MAX_TH) return s->JEND_TH[what - 20];
    return -1;
}

void dor_set_time(dor_state *s, double time) { s->TIME = time; }

void dor_set_dt(dor_state *s, double dt)
{
    s->p.delta_t = dt;
    s->H_BAR[1] = dt * (8.0 / 15.0);
    s->H_BAR[2] = dt * (2.0 / 15.0);
    s->H_BAR[3] = dt * (5.0 / 15.0);
}

double dor_get_dt(const dor_state *s) { return s->p.delta_t; }

void dor_courant_request(dor_state *s, double cfl)
{
    s->courant_pending = 1;
    s->courant_cfl = cfl;
}

/* COURANT (for/courant.f:1-65) on the physical U1, U2, U3 arrays as they
 * stand: the smallest of the viscous/diffusive, rotation and buoyancy limits
 * and cfl * (grid spacing) / |velocity| over j = 1..NY, k = 0..NZP-1,
 * i = 0..NXM; sets DELTA_T (0.0001 if dt <= 0, 999 if dt >= 999) and H_BAR.
 * A zero velocity gives an infinite quotient, which the minimum ignores. */
double dor_courant(dor_state *s, double cfl)
{
    const int NY = s->NY, NZP = s->NZP, NXM = s->NXM, NTH = s->NTH;
    const double NU = s->p.nu;
    const double PI = 4.0 * atan(1.0);
    const double dx = s->p.lx / s->NX, dz = s->p.lz / s->NZ;   /* DX(I), DZ(K) (for/channel.f:1574-1583) */
    const double *DY = s->DY;
    const double *U1 = s->fld[DOR_U1], *U2 = s->fld[DOR_U2], *U3 = s->fld[DOR_U3];
    double dt = 999.0;
    const double h = dx < DY[1] ? dx : DY[1];
    dt = fmin(dt, 0.5 * pow(h, 2.0) / NU);
    for (int n = 0; n < NTH; n++) dt = fmin(dt, dt * NU / (NU / s->p.pr[n]));
    if (s->p.i_ro != 0.0) dt = fmin(dt, 2.0 * PI / fabs(s->p.i_ro) / 20.0);
    for (int n = 0; n < NTH; n++) {
        if (s->p.ri[n] != 0.0) {
            const double Nmax = sqrt(fabs(s->p.ri[n] * s->p.th_bc_ymin_c1[n]));
            /* the literal 0.1 is default (single precision) REAL */
            dt = fmin(dt, (double)0.1f * 2.0 * PI / Nmax);
        }
    }
    for (int j = 1; j <= NY; j++)
        for (int k = 0; k <= NZP - 1; k++)
            for (int i = 0; i <= NXM; i++) {
                const size_t x = RIDX(i, k, j);
                const double dt_x = cfl * dx / fabs(U1[x]);
                const double dt_y = cfl * DY[j] / fabs(U2[x]);
                const double dt_z = cfl * dz / fabs(U3[x]);
                dt = fmin(dt, fmin(dt_x, fmin(dt_y, dt_z)));
            }
    double DELTA_T;
    if (dt <= 0.0) DELTA_T = 0.0001;
    else if (dt >= 999.0) DELTA_T = 999.0;
    else DELTA_T = dt;
    dor_set_dt(s, DELTA_T);
    return DELTA_T;
}

void dor_set_les(dor_state *s, int les, int model)
{
    s->les = les != 0;
    s->les_model = s->les && model != 0;
    if (s->les && !s->fld[DOR_NUT]) {
        s->fld[DOR_NUT] = zalloc(s->fsize);
        for (int n = 0; n < s->NTH; n++) s->fld[DOR_KAPPA0 + n] = zalloc(s->fsize);
    }
}

void dor_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------------- */
/* FFT_XZ_TO_FOURIER / FFT_XZ_TO_PHYSICAL (for/mpi.f:1530-1674)               */
/* ------------------------------------------------------------------------- */

void dor_fft_xz_to_fourier(dor_state *s, double *v, int jmin, int jmax)
{
    const int NX = s->NX, NZ = s->NZ, NXP = s->NXP, NKX = s->NKX, NKZ = s->NKZ;
    const int H = NX / 2;
    cplx *vv = (cplx *)v;
#pragma omp parallel
    {
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(H + 1) * NZ);
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
        cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
        cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)NZ);
#pragma omp for schedule(static)
        for (int J = jmin; J <= jmax; J++) {
            /* R2C along x over the NZP rows of plane J (for/mpi.f:1548-1549). */
            for (int K = 0; K < NZ; K += 2) {
                const double *a = v + RIDX(0, K, J);
                const double *b = (K + 1 < NZ) ? v + RIDX(0, K + 1, J) : NULL;
                r2c_pair(&s->px, a, b, tmp + (size_t)K * (H + 1),
                         b ? tmp + (size_t)(K + 1) * (H + 1) : NULL, z, w);
            }
            /* Scale kept modes by 1/NX, zero the rest (for/mpi.f:1550-1561). */
            for (int K = 0; K < NZ; K++) {
                cplx *t = tmp + (size_t)K * (H + 1);
                for (int Ii = 0; Ii <= NKX; Ii++)
                    t[Ii] = creal(t[Ii]) / (double)NX + I * (cimag(t[Ii]) / (double)NX);
                for (int Ii = NKX + 1; Ii <= H; Ii++) t[Ii] = 0.0;
            }
            /* Transpose (a copy at NPROCZ = 1) and pad zeroing
             * (for/mpi.f:1567-1580). */
            for (int K = 0; K < NZ; K++) {
                for (int Ii = 0; Ii <= NXP - 1; Ii++) vv[CIDX(Ii, K, J)] = tmp[(size_t)K * (H + 1) + Ii];
                vv[CIDX(NXP, K, J)] = 0.0;
            }
            for (int K = NZ; K <= NZ + 1; K++)
                for (int Ii = 0; Ii <= NXP; Ii++) vv[CIDX(Ii, K, J)] = 0.0;
            /* C2C along z, in place, columns I = 0..NXP-1 (for/mpi.f:1585).
             * Columns I > NKX are zero, their transform is zero. */
            for (int Ii = 0; Ii <= NKX && Ii <= NXP - 1; Ii++) {
                for (int K = 0; K < NZ; K++) col[K] = vv[CIDX(Ii, K, J)];
                fft_run(&s->pz, col, w, -1);
                for (int K = 0; K < NZ; K++) vv[CIDX(Ii, K, J)] = col[K];
            }
            /* Scale and pack (for/mpi.f:1587-1606). */
            for (int K = 0; K <= NKZ; K++)
                for (int Ii = 0; Ii <= NXP; Ii++) {
                    cplx t = vv[CIDX(Ii, K, J)];
                    vv[CIDX(Ii, K, J)] = creal(t) / (double)NZ + I * (cimag(t) / (double)NZ);
                }
            for (int K = 1; K <= NKZ; K++)
                for (int Ii = 0; Ii <= NXP; Ii++) {
                    cplx t = vv[CIDX(Ii, NZ - 1 + K - NKZ, J)];
                    vv[CIDX(Ii, NKZ + K, J)] = creal(t) / (double)NZ + I * (cimag(t) / (double)NZ);
                }
        }
        free(tmp); free(z); free(w); free(col);
    }
}

void dor_fft_xz_to_physical(dor_state *s, double *vout, int jmin, int jmax)
{
    const int NX = s->NX, NZ = s->NZ, NXP = s->NXP, NKZ = s->NKZ;
    const int H = NX / 2;
    cplx *vv = (cplx *)vout;
    double *v = vout;
#pragma omp parallel
    {
        cplx *tmp = (cplx *)malloc(sizeof(cplx) * (size_t)(H + 1) * NZ);
        cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
        cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)(NX > NZ ? NX : NZ));
        cplx *col = (cplx *)malloc(sizeof(cplx) * (size_t)NZ);
        double *row = (double *)malloc(sizeof(double) * (size_t)NX * 2);
#pragma omp for schedule(static)
        for (int J = jmin; J <= jmax; J++) {
            /* Unpack, descending (for/mpi.f:1629-1633); zero the gap
             * (for/mpi.f:1634-1638). */
            for (int K = NKZ; K >= 1; K--)
                for (int Ii = 0; Ii <= NXP; Ii++)
                    vv[CIDX(Ii, NZ - 1 + K - NKZ, J)] = vv[CIDX(Ii, NKZ + K, J)];
            for (int K = NKZ + 1; K <= NZ - NKZ - 1; K++)
                for (int Ii = 0; Ii <= NXP; Ii++) vv[CIDX(Ii, K, J)] = 0.0;
            /* C2C backward along z, columns I = 0..NXP-1 (for/mpi.f:1639). */
            for (int Ii = 0; Ii <= NXP - 1; Ii++) {
                int nz_any = 0;
                for (int K = 0; K < NZ; K++) {
                    col[K] = vv[CIDX(Ii, K, J)];
                    nz_any |= (col[K] != 0.0);
                }
                if (!nz_any) continue;
                fft_run(&s->pz, col, w, +1);
                for (int K = 0; K < NZ; K++) vv[CIDX(Ii, K, J)] = col[K];
            }
            /* Transpose into TMP, zero the x-Nyquist slot (for/mpi.f:1643-1661). */
            for (int K = 0; K < NZ; K++) {
                for (int Ii = 0; Ii <= H - 1; Ii++) tmp[(size_t)K * (H + 1) + Ii] = vv[CIDX(Ii, K, J)];
                tmp[(size_t)K * (H + 1) + H] = 0.0;
            }
            /* C2R along x into V rows (for/mpi.f:1669-1671). */
            for (int K = 0; K < NZ; K += 2) {
                const int two = (K + 1 < NZ);
                c2r_pair(&s->px, tmp + (size_t)K * (H + 1), two ? tmp + (size_t)(K + 1) * (H + 1) : NULL,
                         row, two ? row + NX : NULL, z, w);
                memcpy(v + RIDX(0, K, J), row, sizeof(double) * NX);
                if (two) memcpy(v + RIDX(0, K + 1, J), row + NX, sizeof(double) * NX);
            }
        }
        free(tmp); free(z); free(w); free(col); free(row);
    }
}

/* ------------------------------------------------------------------------- */
/* Thomas solvers (for/mpi.f:630-865) at NPROCY = 1                           */
/* ------------------------------------------------------------------------- */

/* A,B,C,G are (0:NX-1, 0:NY+1) column-major: element (i,j) at i + nx*j. */
void dor_thomas_real(double *A, double *B, double *C, double *G, int INY, int INX)
{
#define M2(X, i, j) X[(size_t)(i) + (size_t)INX * (size_t)(j)]
    /* THOMAS_FORWARD_REAL_MPI, lowest and highest process (for/mpi.f:673-699). */
    for (int J = 1; J <= INY; J++)
        for (int i = 0; i <= INX - 1; i++) {
            M2(A, i, J) = -M2(A, i, J) / M2(B, i, J - 1);
            M2(B, i, J) = M2(B, i, J) + M2(A, i, J) * M2(C, i, J - 1);
            M2(G, i, J) = M2(G, i, J) + M2(A, i, J) * M2(G, i, J - 1);
        }
    {
        int J = INY + 1;
        for (int i = 0; i <= INX - 1; i++) {
            M2(A, i, J) = -M2(A, i, J) / M2(B, i, J - 1);
            M2(B, i, J) = M2(B, i, J) + M2(A, i, J) * M2(C, i, J - 1);
            M2(G, i, J) = M2(G, i, J) + M2(A, i, J) * M2(G, i, J - 1);
        }
    }
    /* THOMAS_BACKWARD_REAL_MPI (for/mpi.f:784-807). */
    for (int i = 0; i <= INX - 1; i++) {
        M2(G, i, INY + 1) = M2(G, i, INY + 1) / M2(B, i, INY + 1);
        for (int J = INY; J >= 0; J--) {
            if (M2(B, i, J) != 0.0)
                M2(G, i, J) = (M2(G, i, J) - M2(C, i, J) * M2(G, i, J + 1)) / M2(B, i, J);
            else
                M2(G, i, J) = 0.0;
        }
    }
#undef M2
}

/* Complex RHS, real matrix, (0:INX, 0:NY+1) with INX = NXP; column INX is
 * skipped (see file header). */
void dor_thomas_complex(double *A, double *B, double *C, double *Gd, int INY, int INX)
{
    cplx *G = (cplx *)Gd;
    const size_t ld = (size_t)INX + 1;
#define M2(X, i, j) X[(size_t)(i) + ld * (size_t)(j)]
    for (int i = 0; i <= INX - 1; i++) {
        /* THOMAS_FORWARD_COMPLEX_MPI (for/mpi.f:741-749). */
        for (int J = 2; J <= INY; J++) {
            if (M2(B, i, J - 1) != 0.0)
                M2(A, i, J) = -M2(A, i, J) / M2(B, i, J - 1);
            else
                M2(A, i, J) = 0.0;
            M2(B, i, J) = M2(B, i, J) + M2(A, i, J) * M2(C, i, J - 1);
            M2(G, i, J) = M2(G, i, J) + M2(A, i, J) * M2(G, i, J - 1);
        }
        /* THOMAS_BACKWARD_COMPLEX_MPI (for/mpi.f:842-855). */
        if (M2(B, i, INY) != 0.0)
            M2(G, i, INY) = M2(G, i, INY) / M2(B, i, INY);
        else
            M2(G, i, INY) = 0.0;
        for (int J = INY - 1; J >= 1; J--) {
            if (M2(B, i, J) != 0.0)
                M2(G, i, J) = (M2(G, i, J) - M2(C, i, J) * M2(G, i, J + 1)) / M2(B, i, J);
            else
                M2(G, i, J) = 0.0;
        }
    }
#undef M2
}

void dor_fft_c2c(double *data, int n, int sign)
{
    fplan pl;
    fplan_init(&pl, n);
    cplx *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    fft_run(&pl, (cplx *)data, w, sign < 0 ? -1 : +1);
    free(w);
    fplan_free(&pl);
}

void dor_r2c(const double *x, double *X, int n)
{
    fplan pl;
    fplan_init(&pl, n);
    cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)n), *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    r2c_pair(&pl, x, NULL, (cplx *)X, NULL, z, w);
    free(z); free(w);
    fplan_free(&pl);
}

void dor_c2r(const double *X, double *x, int n)
{
    fplan pl;
    fplan_init(&pl, n);
    cplx *z = (cplx *)malloc(sizeof(cplx) * (size_t)n), *w = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    c2r_pair(&pl, (const cplx *)X, NULL, x, NULL, z, w);
    free(z); free(w);
    fplan_free(&pl);
}

/* ------------------------------------------------------------------------- */
/* Boundary-condition row builders (for/channel.f:1660-2282) via the           */
/* APPLY_BC_*_MPI wrappers (for/mpi.f:1137-1290).  Arrays (0:NX-1, 0:NY+1).    */
/* ------------------------------------------------------------------------- */

#define MT(X, i, j) X[(size_t)(i) + (size_t)nx * (size_t)(j)]

static void bc_u_lower(const dor_state *s, int code, double c1, double *L, double *D, double *U, double *V)
{
    /* APPLY_BC_1_LOWER / APPLY_BC_3_LOWER */
    const int nx = s->NX;
    for (int i = 0; i < nx; i++) {
        MT(L, i, 0) = 0.; MT(D, i, 0) = 1.; MT(U, i, 0) = 0.; MT(V, i, 0) = 0.;
        if (code == 0) {
            MT(L, i, 1) = 0.; MT(D, i, 1) = 1.; MT(U, i, 1) = 0.; MT(V, i, 1) = c1;
        } else {
            MT(L, i, 1) = 0.; MT(D, i, 1) = -1.; MT(U, i, 1) = 1.; MT(V, i, 1) = s->DY[2] * c1;
        }
    }
}

static void bc_u_upper(const dor_state *s, int code, double c1, double *L, double *D, double *U, double *V)
{
    /* APPLY_BC_1_UPPER / APPLY_BC_3_UPPER */
    const int nx = s->NX, NY = s->NY;
    for (int i = 0; i < nx; i++) {
        if (code == 0) {
            MT(L, i, NY + 1) = 0.; MT(D, i, NY + 1) = 1.; MT(U, i, NY + 1) = 0.; MT(V, i, NY + 1) = 0.;
            MT(L, i, NY) = 0.; MT(D, i, NY) = 1.; MT(U, i, NY) = 0.; MT(V, i, NY) = c1;
        } else {
            MT(L, i, NY) = -1.; MT(D, i, NY) = 1.; MT(U, i, NY) = 0.; MT(V, i, NY) = s->DY[NY] * c1;
            MT(L, i, NY + 1) = 0.; MT(D, i, NY + 1) = 1.; MT(U, i, NY + 1) = 0.; MT(V, i, NY + 1) = 0.;
        }
    }
}

static void bc_v_lower(const dor_state *s, double *L, double *D, double *U, double *V)
{
    /* APPLY_BC_2_LOWER */
    const int nx = s->NX;
    const int code = s->p.v_bc_ymin;
    const double c1 = s->p.v_bc_ymin_c1;
    for (int i = 0; i < nx; i++) {
        if (code == 0) {
            MT(L, i, 1) = 0.; MT(D, i, 1) = 1.; MT(U, i, 1) = 0.; MT(V, i, 1) = c1;
            MT(L, i, 2) = 0.; MT(D, i, 2) = 1.; MT(U, i, 2) = 0.; MT(V, i, 2) = c1;
        } else if (code == 1) {
            MT(D, i, 1) = -1.; MT(U, i, 1) = 1.; MT(L, i, 1) = 0.; MT(V, i, 1) = s->DYF[1] * c1;
        }
        MT(L, i, 0) = 0.; MT(D, i, 0) = 1.; MT(U, i, 0) = 0.; MT(V, i, 0) = 0.;
    }
}

static void bc_v_upper(const dor_state *s, double *L, double *D, double *U, double *V)
{
    /* APPLY_BC_2_UPPER */
    const int nx = s->NX, NY = s->NY;
    const int code = s->p.v_bc_ymax;
    const double c1 = s->p.v_bc_ymax_c1;
    for (int i = 0; i < nx; i++) {
        if (code == 0) {
            MT(L, i, NY + 1) = 0.; MT(D, i, NY + 1) = 1.; MT(U, i, NY + 1) = 0.; MT(V, i, NY + 1) = c1;
            MT(L, i, NY) = 0.; MT(D, i, NY) = 1.; MT(U, i, NY) = 0.; MT(V, i, NY) = c1;
        } else if (code == 1) {
            MT(L, i, NY + 1) = -1.; MT(D, i, NY + 1) = 1.; MT(U, i, NY + 1) = 0.; MT(V, i, NY + 1) = s->DYF[NY] * c1;
        }
    }
}

static void bc_th_lower(const dor_state *s, int n, double *L, double *D, double *U, double *V)
{
    /* APPLY_BC_TH_LOWER */
    const int nx = s->NX;
    const double c1 = s->p.th_bc_ymin_c1[n];
    for (int i = 0; i < nx; i++) {
        if (s->p.th_bc_ymin[n] == 0) {
            MT(L, i, 0) = 0.; MT(D, i, 0) = 1.; MT(U, i, 0) = 0.; MT(V, i, 0) = 0.;
            MT(L, i, 1) = 0.; MT(D, i, 1) = 1.; MT(U, i, 1) = 0.; MT(V, i, 1) = c1;
        } else {
            MT(L, i, 1) = 0.; MT(D, i, 1) = -1.; MT(U, i, 1) = 1.; MT(V, i, 1) = s->DY[2] * c1;
            MT(L, i, 0) = 0.; MT(D, i, 0) = 1.; MT(U, i, 0) = 0.; MT(V, i, 0) = 0.;
        }
    }
}

static void bc_th_upper(const dor_state *s, int n, double *L, double *D, double *U, double *V)
{
    /* APPLY_BC_TH_UPPER */
    const int nx = s->NX, NY = s->NY;
    const double c1 = s->p.th_bc_ymax_c1[n];
    for (int i = 0; i < nx; i++) {
        if (s->p.th_bc_ymax[n] == 0) {
            MT(L, i, NY + 1) = 0.; MT(D, i, NY + 1) = 1.; MT(U, i, NY + 1) = 0.; MT(V, i, NY + 1) = 0.;
            MT(L, i, NY) = 0.; MT(D, i, NY) = 1.; MT(U, i, NY) = 0.; MT(V, i, NY) = c1;
        } else {
            MT(L, i, NY) = -1.; MT(D, i, NY) = 1.; MT(U, i, NY) = 0.; MT(V, i, NY) = s->DY[NY] * c1;
            MT(L, i, NY + 1) = 0.; MT(D, i, NY + 1) = 1.; MT(U, i, NY + 1) = 0.; MT(V, i, NY + 1) = 0.;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* RK_CHAN_1 (for/channel.f:62-1073)                                          */
/* ------------------------------------------------------------------------- */

#define C(f) ((cplx *)s->fld[f])
#define R(f) (s->fld[f])

enum { SYS_TH, SYS_U2, SYS_U1, SYS_U3 };

/* APPLY_BC_VEL_PHYS_LOWER / _UPPER (for/channel.f:2495-2640) on the physical
 * U1..U3 arrays. */
static void apply_bc_vel_phys(dor_state *s)
{
    const int NY = s->NY, NZP = s->NZP, NXM = s->NXM;
    const dor_params *p = &s->p;
    const double *DY = s->DY, *DYF = s->DYF;
    double *U1 = R(DOR_U1), *U2 = R(DOR_U2), *U3 = R(DOR_U3);
    for (int K = 0; K <= NZP - 1; K++)
        for (int i = 0; i <= NXM; i++) {
#define AT2(A, J) A[RIDX(i, K, (J))]
            for (int c = 0; c < 2; c++) {
                double *UU = c == 0 ? U1 : U3;
                const int lo = c == 0 ? p->u_bc_ymin : p->w_bc_ymin, hi = c == 0 ? p->u_bc_ymax : p->w_bc_ymax;
                const double clo = c == 0 ? p->u_bc_ymin_c1 : p->w_bc_ymin_c1;
                const double chi = c == 0 ? p->u_bc_ymax_c1 : p->w_bc_ymax_c1;
                if (lo == 0) {
                    AT2(UU, 0) = clo;
                    AT2(UU, 1) = clo;
                } else {
                    AT2(UU, 1) = AT2(UU, 2) - DY[2] * clo;
                    AT2(UU, 0) = AT2(UU, 1) - DY[1] * clo;
                }
                if (hi == 0) {
                    AT2(UU, NY) = chi;
                    AT2(UU, NY + 1) = chi;
                } else {
                    AT2(UU, NY) = AT2(UU, NY - 1) + DY[NY] * chi;
                    AT2(UU, NY + 1) = AT2(UU, NY) + DY[NY] * chi;
                }
            }
            if (p->v_bc_ymin == 0) {
                AT2(U2, 1) = 2.0 * p->v_bc_ymin_c1 - AT2(U2, 2);
                AT2(U2, 0) = AT2(U2, 1);
            } else {
                AT2(U2, 1) = AT2(U2, 2) - DYF[1] * p->v_bc_ymin_c1;
                AT2(U2, 0) = AT2(U2, 1) - DYF[1] * p->v_bc_ymin_c1;
            }
            if (p->v_bc_ymax == 0)
                AT2(U2, NY + 1) = 2.0 * p->v_bc_ymax_c1 - AT2(U2, NY);
            else
                AT2(U2, NY + 1) = AT2(U2, NY) + DY[NY] * p->v_bc_ymax_c1;
#undef AT2
        }
}

/* One implicit CN solve of a physical field, looping over K planes
 * (for/channel.f:771-833, 842-901, 907-966, 969-1029). */
static void cn_solve_phys(dor_state *s, int sys, int n, double TEMP1, double TEMP2)
{
    const int NX = s->NX, NY = s->NY, NZP = s->NZP, nx = NX;
    const double NUV = s->p.nu_v_scale;
    const double *DY = s->DY, *DYF = s->DYF;
    double *rhs, *out;
    int jlo, jhi, clo, chi;
    if (sys == SYS_TH) {
        rhs = R(DOR_RTH0 + n); out = R(DOR_TH0 + n);
        jlo = s->JSTART_TH[n]; jhi = s->JEND_TH[n]; clo = jlo - 1; chi = jhi + 1;
    } else if (sys == SYS_U2) {
        rhs = R(DOR_R2); out = R(DOR_U2);
        jlo = 2; jhi = NY; clo = 1; chi = NY + 1;
    } else if (sys == SYS_U1) {
        rhs = R(DOR_R1); out = R(DOR_U1);
        jlo = s->JSTART; jhi = s->JEND; clo = jlo - 1; chi = jhi + 1;
    } else {
        rhs = R(DOR_R3); out = R(DOR_U3);
        jlo = s->JSTART; jhi = s->JEND; clo = jlo - 1; chi = jhi + 1;
    }
    const size_t msz = (size_t)NX * (size_t)(NY + 2);
#pragma omp parallel
    {
        double *L = zalloc(msz), *D = zalloc(msz), *U = zalloc(msz), *V = zalloc(msz);
        for (int J = 0; J <= NY + 1; J++)
            for (int i = 0; i < NX; i++) { MT(L, i, J) = 0.; MT(D, i, J) = 1.; MT(U, i, J) = 0.; MT(V, i, J) = 0.; }
#pragma omp for schedule(static)
        for (int K = 0; K <= NZP - 1; K++) {
            for (int J = jlo; J <= jhi; J++)
                for (int i = 0; i < NX; i++) {
                    if (sys == SYS_TH) {
                        const double PRn = s->p.pr[n];
                        MT(L, i, J) = -(TEMP1 / PRn * NUV) / (DY[J] * DYF[J]);
                        MT(D, i, J) = 1. + (TEMP1 / PRn * NUV) / (DY[J + 1] * DYF[J]) + (TEMP1 / PRn * NUV) / (DY[J] * DYF[J]);
                        MT(U, i, J) = -(TEMP1 / PRn * NUV) / (DY[J + 1] * DYF[J]);
                    } else if (sys == SYS_U2) {
                        MT(L, i, J) = -TEMP1 * NUV / (DYF[J - 1] * DY[J]);
                        MT(D, i, J) = 1. + TEMP1 * NUV / (DYF[J] * DY[J]) + TEMP1 * NUV / (DYF[J - 1] * DY[J]);
                        MT(U, i, J) = -TEMP1 * NUV / (DYF[J] * DY[J]);
                    } else {
                        MT(L, i, J) = -TEMP1 * NUV / (DY[J] * DYF[J]);
                        MT(D, i, J) = 1. - TEMP1 * NUV * (-1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]));
                        MT(U, i, J) = -TEMP1 * NUV / (DY[J + 1] * DYF[J]);
                    }
                    MT(V, i, J) = rhs[RIDX(i, K, J)];
                    if (s->les) {
                        /* the eddy viscosity / diffusivity, implicitly (for/channel.f:793-806,
                         * 862-874, 927-939, 991-1003); NU_T and KAPPA_T live on GY */
                        if (sys == SYS_U2) {
                            const double *NT = R(DOR_NUT);
                            const double nm = NT[RIDX(i, K, J)] + NT[RIDX(i, K, J - 1)];
                            const double np = NT[RIDX(i, K, J)] + NT[RIDX(i, K, J + 1)];
                            MT(L, i, J) = MT(L, i, J) - TEMP2 * 0.5 * nm / (DYF[J - 1] * DY[J]);
                            MT(D, i, J) = MT(D, i, J) + TEMP2 * 0.5 * np / (DYF[J] * DY[J])
                                          + TEMP2 * 0.5 * nm / (DYF[J - 1] * DY[J]);
                            MT(U, i, J) = MT(U, i, J) - TEMP2 * 0.5 * np / (DYF[J] * DY[J]);
                        } else {
                            const double *NT = (sys == SYS_TH) ? R(DOR_KAPPA0 + n) : R(DOR_NUT);
                            const double n0 = NT[RIDX(i, K, J)], n1 = NT[RIDX(i, K, J + 1)];
                            MT(L, i, J) = MT(L, i, J) - TEMP2 * n0 / (DY[J] * DYF[J]);
                            MT(D, i, J) = MT(D, i, J) + TEMP2 * n1 / (DY[J + 1] * DYF[J]) + TEMP2 * n0 / (DY[J] * DYF[J]);
                            MT(U, i, J) = MT(U, i, J) - TEMP2 * n1 / (DY[J + 1] * DYF[J]);
                        }
                    }
                }
            if (sys == SYS_TH) {
                bc_th_lower(s, n, L, D, U, V);
                bc_th_upper(s, n, L, D, U, V);
            } else if (sys == SYS_U2) {
                bc_v_lower(s, L, D, U, V);
                bc_v_upper(s, L, D, U, V);
            } else if (sys == SYS_U1) {
                bc_u_lower(s, s->p.u_bc_ymin, s->p.u_bc_ymin_c1, L, D, U, V);
                bc_u_upper(s, s->p.u_bc_ymax, s->p.u_bc_ymax_c1, L, D, U, V);
            } else {
                bc_u_lower(s, s->p.w_bc_ymin, s->p.w_bc_ymin_c1, L, D, U, V);
                bc_u_upper(s, s->p.w_bc_ymax, s->p.w_bc_ymax_c1, L, D, U, V);
            }
            dor_thomas_real(L, D, U, V, NY, NX);
            for (int J = clo; J <= chi; J++)
                for (int i = 0; i < NX; i++) out[RIDX(i, K, J)] = MT(V, i, J);
        }
        free(L); free(D); free(U); free(V);
    }
}

void dor_rk_chan_1(dor_state *s, int RK_STEP)
{
    const int NY = s->NY, NZP = s->NZP, NXP = s->NXP, TNKZ = s->TNKZ, NXM = s->NXM;
    const int JSTART = s->JSTART, JEND = s->JEND;
    const double NU = s->p.nu, BETA = s->p.beta;
    const double *DY = s->DY, *DYF = s->DYF, *KX2 = s->KX2, *KZ2 = s->KZ2, *KX = s->KX, *KZ = s->KZ;
    const double TEMP1 = NU * s->H_BAR[RK_STEP] / 2.0;
    const double TEMP2 = s->H_BAR[RK_STEP] / 2.0;
    const double TEMP3 = s->ZETA_BAR[RK_STEP] * s->H_BAR[RK_STEP];
    const double TEMP4 = s->H_BAR[RK_STEP];
    const double TEMP5 = s->BETA_BAR[RK_STEP] * s->H_BAR[RK_STEP];
    cplx *CU1 = C(DOR_U1), *CU2 = C(DOR_U2), *CU3 = C(DOR_U3), *CP = C(DOR_P);
    cplx *CR1 = C(DOR_R1), *CR2 = C(DOR_R2), *CR3 = C(DOR_R3);
    cplx *CF1 = C(DOR_F1), *CF2 = C(DOR_F2), *CF3 = C(DOR_F3), *CS1 = C(DOR_S1);
    double *U1 = R(DOR_U1), *U2 = R(DOR_U2), *U3 = R(DOR_U3);
    double *R1 = R(DOR_R1), *R2 = R(DOR_R2), *R3 = R(DOR_R3), *S1 = R(DOR_S1);
#define CIKX(i) (I * KX[i])
#define CIKZ(k) (I * KZ[k])

    /* for/channel.f:88-102 */
#pragma omp parallel for schedule(static)
    for (int J = JSTART; J <= JEND; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                CR1[CIDX(i, K, J)] = CU1[CIDX(i, K, J)];
                CR3[CIDX(i, K, J)] = CU3[CIDX(i, K, J)];
            }
#pragma omp parallel for schedule(static)
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) CR2[CIDX(i, K, J)] = CU2[CIDX(i, K, J)];

    /* for/channel.f:105-121 */
    if (RK_STEP > 1) {
#pragma omp parallel for schedule(static)
        for (int J = JSTART; J <= JEND; J++)
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) {
                    CR1[CIDX(i, K, J)] = CR1[CIDX(i, K, J)] + TEMP3 * CF1[CIDX(i, K, J)];
                    CR3[CIDX(i, K, J)] = CR3[CIDX(i, K, J)] + TEMP3 * CF3[CIDX(i, K, J)];
                }
#pragma omp parallel for schedule(static)
        for (int J = 2; J <= NY; J++)
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++)
                    CR2[CIDX(i, K, J)] = CR2[CIDX(i, K, J)] + TEMP3 * CF2[CIDX(i, K, J)];
    }

    /* for/channel.f:124-148 */
#pragma omp parallel for schedule(static)
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                cplx d = CP[CIDX(i, K, J)] - CP[CIDX(i, K, J - 1)];
                const int JD = PLANT(2, J + 1, J);
                CS1[CIDX(i, K, J)] = creal(d) / DY[JD] + I * (cimag(d) / DY[JD]);
            }
#pragma omp parallel for schedule(static)
    for (int J = JSTART; J <= JEND; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                CR1[CIDX(i, K, J)] = PLANT(1, CR1[CIDX(i, K, J)] + TEMP4 * (CIKX(i) * CP[CIDX(i, K, J)]),
                                           CR1[CIDX(i, K, J)] - TEMP4 * (CIKX(i) * CP[CIDX(i, K, J)]));
                CR3[CIDX(i, K, J)] = CR3[CIDX(i, K, J)] - TEMP4 * (PLANT(3, CIKX(i), CIKZ(K)) * CP[CIDX(i, K, J)]);
            }
#pragma omp parallel for schedule(static)
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++)
                CR2[CIDX(i, K, J)] = CR2[CIDX(i, K, J)] - TEMP4 * CS1[CIDX(i, K, J)];

    /* Mean forcing (for/channel.f:155-167), RANKZ = 0. */
    if (s->p.f_type == 1) {
        for (int J = JSTART; J <= JEND; J++) CR1[CIDX(0, 0, J)] = CR1[CIDX(0, 0, J)] - TEMP4 * PLANT(29, -s->p.px0, s->p.px0);
    } else if (s->p.f_type == 2) {
        for (int J = JSTART; J <= JEND; J++)
            CR1[CIDX(0, 0, J)] = CR1[CIDX(0, 0, J)] - TEMP4 * (s->p.px0 + s->p.amp_omega0 * PLANT(4, sin(s->p.omega0 * s->TIME), cos(s->p.omega0 * s->TIME)));
    }

    /* Horizontal viscous terms (for/channel.f:175-189). */
#pragma omp parallel for schedule(static)
    for (int J = JSTART; J <= JEND; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const double f = NU * PLANT(5, (KX2[i] + KZ2[K]) * BETA, pow(KX2[i] + KZ2[K], BETA));
                CF1[CIDX(i, K, J)] = -(f * CU1[CIDX(i, K, J)]);
                CF3[CIDX(i, K, J)] = -(f * CU3[CIDX(i, K, J)]);
            }
#pragma omp parallel for schedule(static)
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const double f = NU * pow(KX2[i] + KZ2[K], BETA);
                CF2[CIDX(i, K, J)] = -(f * CU2[CIDX(i, K, J)]);
            }

    /* Coriolis (for/channel.f:193-200). */
#pragma omp parallel for schedule(static)
    for (int K = 0; K <= TNKZ; K++)
        for (int i = 0; i <= NXP - 1; i++)
            for (int J = JSTART; J <= JEND; J++) {
                CF1[CIDX(i, K, J)] = CF1[CIDX(i, K, J)] + PLANT(6, -s->p.i_ro, s->p.i_ro) * CU3[CIDX(i, K, J)];
                CF3[CIDX(i, K, J)] = CF3[CIDX(i, K, J)] - PLANT(7, -s->p.i_ro, s->p.i_ro) * CU1[CIDX(i, K, J)];
            }

    for (int n = 0; n < s->NTH; n++) {
        cplx *CTH = C(DOR_CTH0 + n), *CRTH = C(DOR_CRTH0 + n), *CFTH = C(DOR_CFTH0 + n);
        const double RI = s->p.ri[n], PR = s->p.pr[n];
        const int JS = s->JSTART_TH[n], JE = s->JEND_TH[n];
        /* Buoyancy (for/channel.f:209-243). */
#pragma omp parallel for schedule(static)
        for (int J = 2; J <= NY; J++) {
            for (int K = 1; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) {
                    cplx t = CTH[CIDX(i, K, J)] * DYF[PLANT(8, J, J - 1)] + CTH[CIDX(i, K, J - 1)] * DYF[PLANT(8, J - 1, J)];
                    cplx b = (RI * s->p.grav_y) * t;
                    CF2[CIDX(i, K, J)] = CF2[CIDX(i, K, J)] + (creal(b) / (2.0 * DY[J]) + I * (cimag(b) / (2.0 * DY[J])));
                }
            for (int i = PLANT(9, 0, 1); i <= NXP - 1; i++) {
                const int K = 0;
                cplx t = CTH[CIDX(i, K, J)] * DYF[PLANT(8, J, J - 1)] + CTH[CIDX(i, K, J - 1)] * DYF[PLANT(8, J - 1, J)];
                cplx b = (RI * s->p.grav_y) * t;
                CF2[CIDX(i, K, J)] = CF2[CIDX(i, K, J)] + (creal(b) / (2.0 * DY[J]) + I * (cimag(b) / (2.0 * DY[J])));
            }
        }
#pragma omp parallel for schedule(static)
        for (int J = JSTART; J <= JEND; J++)
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) {
                    CF1[CIDX(i, K, J)] = CF1[CIDX(i, K, J)] + (RI * PLANT(10, s->p.grav_z, s->p.grav_x)) * CTH[CIDX(i, K, J)];
                    CF3[CIDX(i, K, J)] = CF3[CIDX(i, K, J)] + (RI * s->p.grav_z) * CTH[CIDX(i, K, J)];
                }
        /* Scalar RHS (for/channel.f:254-290). */
#pragma omp parallel for schedule(static)
        for (int J = JS; J <= JE; J++)
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) {
                    CRTH[CIDX(i, K, J)] = CTH[CIDX(i, K, J)];
                    if (RK_STEP > 1) CRTH[CIDX(i, K, J)] = CRTH[CIDX(i, K, J)] + TEMP3 * CFTH[CIDX(i, K, J)];
                    const double f = PLANT(11, NU * PR, NU / PR) * pow(KX2[i] + KZ2[K], BETA);
                    CFTH[CIDX(i, K, J)] = -(f * CTH[CIDX(i, K, J)]);
                    CFTH[CIDX(i, K, J)] = CFTH[CIDX(i, K, J)] - CU1[CIDX(i, K, J)] * PLANT(12, -s->p.drhodx[n], s->p.drhodx[n])
                                          - PLANT(13, CU1[CIDX(i, K, J)], CU3[CIDX(i, K, J)]) * s->p.drhodz[n];
                }
    }
    /* USER_RHS_CHAN_FOURIER is an empty hook (for/user_rhs.f). */

    /* To physical space (for/channel.f:303-399).  With the subgrid model on,
     * les_chan sets the velocity BCs and adds the SGS force to CF1-3 first,
     * then the physical-space BCs are applied, KAPPA_T = NU_T, and the
     * scalars get their horizontal eddy-diffusion terms explicitly. */
    if (s->les_model) dor_les_chan(s, R(DOR_NUT));
    dor_fft_xz_to_physical(s, U1, 0, NY + 1);
    dor_fft_xz_to_physical(s, U2, 0, NY + 1);
    dor_fft_xz_to_physical(s, U3, 0, NY + 1);
    if (s->les_model) {
        apply_bc_vel_phys(s);
        for (int n = 0; n < s->NTH; n++) {
            double *KT = R(DOR_KAPPA0 + n);
            const double *NT = R(DOR_NUT);
            cplx *CTH = C(DOR_CTH0 + n), *CFTH = C(DOR_CFTH0 + n);
            for (int J = 1; J <= NY + 1; J++)
                for (int K = 0; K <= NZP - 1; K++)
                    for (int i = 0; i <= NXM; i++) KT[RIDX(i, K, J)] = 1.0 * NT[RIDX(i, K, J)];
            for (int dir = 0; dir < 2; dir++) {
                for (int J = 1; J <= NY + 1; J++)
                    for (int K = 0; K <= TNKZ; K++)
                        for (int i = 0; i <= NXP - 1; i++)
                            CS1[CIDX(i, K, J)] = (dir == 0 ? CIKX(i) : CIKZ(K)) * CTH[CIDX(i, K, J)];
                dor_fft_xz_to_physical(s, S1, 0, NY + 1);
                for (int J = 1; J <= NY + 1; J++)
                    for (int K = 0; K <= NZP - 1; K++)
                        for (int i = 0; i <= NXM; i++) S1[RIDX(i, K, J)] = KT[RIDX(i, K, J)] * S1[RIDX(i, K, J)];
                dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
                for (int J = s->JSTART_TH[n]; J <= s->JEND_TH[n]; J++)
                    for (int K = 0; K <= TNKZ; K++)
                        for (int i = 0; i <= NXP - 1; i++)
                            CFTH[CIDX(i, K, J)] = CFTH[CIDX(i, K, J)] + (dir == 0 ? CIKX(i) : CIKZ(K)) * CS1[CIDX(i, K, J)];
            }
        }
    }
    for (int n = 0; n < s->NTH; n++) {
        memcpy(S1, R(DOR_CTH0 + n), sizeof(double) * s->fsize);
        dor_fft_xz_to_physical(s, S1, 0, NY + 1);
        memcpy(R(DOR_TH0 + n), S1, sizeof(double) * s->fsize);
    }

    /* Nonlinear terms (for/channel.f:410-571). */
#define PROD_LOOP(JA, JB, EXPR)                                                 \
    _Pragma("omp parallel for schedule(static)")                                \
    for (int J = (JA); J <= (JB); J++)                                          \
        for (int K = 0; K <= NZP - 1; K++)                                      \
            for (int i = 0; i <= NXM; i++) S1[RIDX(i, K, J)] = (EXPR);
#define ACC_LOOP(JA, JB, BODY)                                                  \
    _Pragma("omp parallel for schedule(static)")                                \
    for (int J = (JA); J <= (JB); J++)                                          \
        for (int K = 0; K <= TNKZ; K++)                                         \
            for (int i = 0; i <= NXP - 1; i++) { BODY; }
#define AT(A, dj) A[RIDX(i, K, J + (dj))]
#define CA(A) A[CIDX(i, K, J)]

    /* U1*U3 */
    PROD_LOOP(JSTART, JEND, AT(U3, 0) * AT(U1, 0));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(JSTART, JEND, CA(CF1) = CA(CF1) - PLANT(14, CIKX(i), CIKZ(K)) * CA(CS1); CA(CF3) = CA(CF3) - CIKX(i) * CA(CS1));
    /* U1*U1 */
    PROD_LOOP(JSTART, JEND, AT(U1, 0) * AT(U1, 0));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(JSTART, JEND, CA(CF1) = CA(CF1) - PLANT(30, CIKZ(K), CIKX(i)) * CA(CS1));
    /* U3*U3 */
    PROD_LOOP(JSTART, JEND, AT(U3, 0) * AT(U3, 0));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(JSTART, JEND, CA(CF3) = PLANT(15, CA(CF3) + CIKZ(K) * CA(CS1), CA(CF3) - CIKZ(K) * CA(CS1)));
    /* U1*U2 at GY */
    PROD_LOOP(2, NY, ((DYF[PLANT(16, J - 1, J)] * AT(U1, 0) + DYF[PLANT(16, J, J - 1)] * AT(U1, -1)) / (2. * DY[J]))
                         * AT(U2, 0));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CF2) = CA(CF2) - CIKX(i) * CA(CS1));
    /* U3*U2 at GY */
    PROD_LOOP(2, NY, ((DYF[PLANT(17, J - 1, J)] * AT(U3, 0) + DYF[PLANT(17, J, J - 1)] * AT(U3, -1)) / (2. * DY[J]))
                         * AT(U2, 0));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CF2) = CA(CF2) - CIKZ(K) * CA(CS1));
    /* d/dy(U1*U2) */
    PROD_LOOP(JSTART, JEND,
              (AT(U1, 1) * AT(U2, 1) + PLANT(18, AT(U1, 1) * AT(U2, 0), AT(U1, 0) * AT(U2, 1)) - AT(U1, 0) * AT(U2, 0)
               - AT(U1, -1) * AT(U2, 0))
                  / (2.0 * DYF[J]));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(JSTART, JEND, CA(CF1) = CA(CF1) - CA(CS1));
    /* d/dy(U3*U2) */
    PROD_LOOP(JSTART, JEND,
              (AT(U3, 1) * AT(U2, 1) + AT(U3, 0) * AT(U2, 1) - AT(U3, 0) * AT(U2, 0) - AT(U3, -1) * AT(U2, 0))
                  / (2.0 * PLANT(19, DY[J], DYF[J])));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(JSTART, JEND, CA(CF3) = CA(CF3) - CA(CS1));
    /* d/dy(U2*U2) */
    PROD_LOOP(2, NY,
              (0.25 * ((AT(U2, 0) + AT(U2, 1)) * (AT(U2, 0) + AT(U2, 1)))
               - 0.25 * ((AT(U2, 0) + AT(U2, -1)) * (AT(U2, 0) + AT(U2, -1)))) / PLANT(20, DYF[J], DY[J]));
    dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CF2) = CA(CF2) - CA(CS1));

    /* USER_RHS_CHAN_PHYSICAL is an empty hook. */
    /* for/channel.f:581-595 */
    ACC_LOOP(JSTART, JEND, CA(CR1) = CA(CR1) + TEMP5 * CA(CF1); CA(CR3) = CA(CR3) + TEMP5 * CA(CF3));
    ACC_LOOP(2, NY, CA(CR2) = CA(CR2) + TEMP5 * CA(CF2));

    /* for/channel.f:598-600 */
    dor_fft_xz_to_physical(s, R1, 0, NY + 1);
    dor_fft_xz_to_physical(s, R2, 2, NY);
    dor_fft_xz_to_physical(s, R3, 0, NY + 1);

    /* Explicit half of CN (for/channel.f:606-626). */
    const double NUV = s->p.nu_v_scale;
#pragma omp parallel for schedule(static)
    for (int J = JSTART; J <= JEND; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++) {
                AT(R1, 0) = AT(R1, 0) + TEMP1 * PLANT(27, 1.0, NUV)
                    * (((AT(U1, 1) - AT(U1, 0)) / DY[PLANT(24, J, J + 1)] - (AT(U1, 0) - AT(U1, -1)) / DY[J]) / DYF[J]);
                AT(R3, 0) = AT(R3, 0) + TEMP1 * NUV
                    * (((AT(U3, 1) - AT(U3, 0)) / DY[J + 1] - (AT(U3, 0) - AT(U3, -1)) / DY[J]) / DYF[J]);
            }
#pragma omp parallel for schedule(static)
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++)
                AT(R2, 0) = AT(R2, 0) + TEMP1 * NUV
                    * (((AT(U2, 1) - AT(U2, 0)) / DYF[J] - (AT(U2, 0) - AT(U2, -1)) / DYF[PLANT(25, J, J - 1)]) / DY[J]);
    if (s->les) {
        /* the eddy viscosity, explicit half (for/channel.f:628-659) */
        const double *NT = R(DOR_NUT);
#pragma omp parallel for schedule(static)
        for (int J = JSTART; J <= JEND; J++)
            for (int K = 0; K <= NZP - 1; K++)
                for (int i = 0; i <= NXM; i++) {
                    AT(R1, 0) = AT(R1, 0) + TEMP2 * ((AT(NT, 1) * (AT(U1, 1) - AT(U1, 0)) / DY[J + 1]
                                                      - AT(NT, 0) * (AT(U1, 0) - AT(U1, -1)) / DY[J]) / DYF[J]);
                    AT(R3, 0) = AT(R3, 0) + TEMP2 * ((AT(NT, 1) * (AT(U3, 1) - AT(U3, 0)) / DY[J + 1]
                                                      - AT(NT, 0) * (AT(U3, 0) - AT(U3, -1)) / DY[J]) / DYF[J]);
                }
#pragma omp parallel for schedule(static)
        for (int J = 2; J <= NY; J++)
            for (int K = 0; K <= NZP - 1; K++)
                for (int i = 0; i <= NXM; i++)
                    AT(R2, 0) = AT(R2, 0) + TEMP2 * ((0.5 * (AT(NT, 0) + AT(NT, 1)) * (AT(U2, 1) - AT(U2, 0)) / DYF[J]
                                                      - 0.5 * (AT(NT, 0) + AT(NT, -1)) * (AT(U2, 0) - AT(U2, -1)) / DYF[J - 1])
                                                     / DY[J]);
    }

    /* Scalars (for/channel.f:665-836). */
    for (int n = 0; n < s->NTH; n++) {
        double *TH = R(DOR_TH0 + n), *RTH = R(DOR_RTH0 + n);
        cplx *CFTH = C(DOR_CFTH0 + n), *CRTH = C(DOR_CRTH0 + n);
        const int JS = s->JSTART_TH[n], JE = s->JEND_TH[n];
        PROD_LOOP(JS, JE, AT(TH, 0) * AT(U1, 0));
        dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
        ACC_LOOP(JS, JE, CA(CFTH) = CA(CFTH) - PLANT(21, CIKZ(K), CIKX(i)) * CA(CS1));
        PROD_LOOP(JS, JE, AT(TH, 0) * AT(U3, 0));
        dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
        ACC_LOOP(JS, JE, CA(CFTH) = PLANT(22, CA(CFTH) + CIKZ(K) * CA(CS1), CA(CFTH) - CIKZ(K) * CA(CS1)));
        PROD_LOOP(JS, JE,
                  (AT(TH, 1) * AT(U2, 1) + AT(TH, 0) * AT(U2, 1) - AT(TH, 0) * AT(U2, 0)
                   - AT(TH, -1) * AT(U2, PLANT(23, -1, 0)))
                      / (2.0 * DYF[J]));
        dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
        ACC_LOOP(JS, JE, CA(CFTH) = CA(CFTH) - CA(CS1));
        ACC_LOOP(JS, JE, CA(CRTH) = CA(CRTH) + TEMP5 * CA(CFTH));
        memcpy(S1, R(DOR_CRTH0 + n), sizeof(double) * s->fsize);
        dor_fft_xz_to_physical(s, S1, 0, NY + 1);
        memcpy(RTH, S1, sizeof(double) * s->fsize);
        const double PR = s->p.pr[n];
#pragma omp parallel for schedule(static)
        for (int J = JS; J <= JE; J++)
            for (int K = 0; K <= NZP - 1; K++)
                for (int i = 0; i <= NXM; i++)
                    AT(RTH, 0) = AT(RTH, 0) + PLANT(26, TEMP1 * PR, TEMP1 / PR) * NUV
                        * (((AT(TH, 1) - AT(TH, 0)) / DY[J + 1] - (AT(TH, 0) - AT(TH, -1)) / DY[J]) / DYF[J]);
        if (s->les) {
            /* the eddy diffusivity, explicit half (for/channel.f:748-760) */
            const double *KT = R(DOR_KAPPA0 + n);
#pragma omp parallel for schedule(static)
            for (int J = JS; J <= JE; J++)
                for (int K = 0; K <= NZP - 1; K++)
                    for (int i = 0; i <= NXM; i++)
                        AT(RTH, 0) = AT(RTH, 0) + TEMP2 * ((AT(KT, 1) * (AT(TH, 1) - AT(TH, 0)) / DY[J + 1]
                                                            - AT(KT, 0) * (AT(TH, 0) - AT(TH, -1)) / DY[J]) / DYF[J]);
        }
        /* Solve the scalar before velocity (for/channel.f:762-833). */
        cn_solve_phys(s, SYS_TH, n, TEMP1, TEMP2);
    }

    cn_solve_phys(s, SYS_U2, 0, TEMP1, TEMP2);
    cn_solve_phys(s, SYS_U1, 0, TEMP1, TEMP2);
    cn_solve_phys(s, SYS_U3, 0, TEMP1, TEMP2);
    /* VARIABLE_DT: COURANT on the physical velocity after the implicit solves,
     * before the projection (for/channel.f:1037-1040); the caller asks for it
     * when MOD(TIME_STEP, UPDATE_DT) = 0.  TEMP1..TEMP5 keep this step's dt. */
    if (RK_STEP == 3 && s->courant_pending) {
        dor_courant(s, s->courant_cfl);
        s->courant_pending = 0;
    }

    /* Back to Fourier space (for/channel.f:1043-1050). */
    dor_fft_xz_to_fourier(s, U1, 0, NY + 1);
    dor_fft_xz_to_fourier(s, U2, 0, NY + 1);
    dor_fft_xz_to_fourier(s, U3, 0, NY + 1);
    for (int n = 0; n < s->NTH; n++) {
        memcpy(S1, R(DOR_TH0 + n), sizeof(double) * s->fsize);
        dor_fft_xz_to_fourier(s, S1, 0, NY + 1);
        memcpy(R(DOR_CTH0 + n), S1, sizeof(double) * s->fsize);
    }

    dor_rem_div_chan(s);

    /* CP update (for/channel.f:1059-1065); phi is in CR1. */
    ACC_LOOP(JSTART, JEND, {
        cplx t = CA(CR1);
        const double T4 = PLANT(28, TEMP5, TEMP4);
        CA(CP) = CA(CP) + (creal(t) / T4 + I * (cimag(t) / T4));
    });
    /* GHOST_CHAN_MPI is a no-op at NPROCY = 1. */
#undef AT
#undef CA
}

/* ------------------------------------------------------------------------- */
/* REM_DIV_CHAN (for/channel.f:1103-1214)                                     */
/* ------------------------------------------------------------------------- */

void dor_rem_div_chan(dor_state *s)
{
    const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ;
    const double *DY = s->DY, *DYF = s->DYF, *KX = s->KX, *KZ = s->KZ, *KX2 = s->KX2, *KZ2 = s->KZ2;
    cplx *CU1 = C(DOR_U1), *CU2 = C(DOR_U2), *CU3 = C(DOR_U3), *CR1 = C(DOR_R1);
    const size_t ld = (size_t)NXP + 1;
    const size_t msz = ld * (size_t)(NY + 2);
#define MC(X, i, j) X[(size_t)(i) + ld * (size_t)(j)]
#pragma omp parallel
    {
        double *L = zalloc(msz), *D = zalloc(msz), *U = zalloc(msz);
        cplx *V = (cplx *)calloc(msz, sizeof(cplx));
        for (int J = 0; J <= NY + 1; J++)
            for (int i = 0; i <= NXP - 1; i++) { MC(L, i, J) = 0.; MC(D, i, J) = 1.; MC(U, i, J) = 0.; MC(V, i, J) = 0.; }
#pragma omp for schedule(static)
        for (int K = 0; K <= TNKZ; K++) {
            for (int J = 1; J <= NY; J++)
                for (int i = 0; i <= NXP - 1; i++) {
                    MC(L, i, J) = 1. / (DY[J] * DYF[J]);
                    MC(D, i, J) = -KX2[i] - KZ2[K] - 1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]);
                    MC(U, i, J) = 1. / (DY[J + 1] * DYF[J]);
                }
            for (int J = 1; J <= NY; J++)
                for (int i = 0; i <= NXP - 1; i++) {
                    cplx d = CU2[CIDX(i, K, J + 1)] - CU2[CIDX(i, K, J)];
                    MC(V, i, J) = (CIKX(i) * CU1[CIDX(i, K, J)] + (creal(d) / DYF[J] + I * (cimag(d) / DYF[J])))
                                  + CIKZ(K) * CU3[CIDX(i, K, J)];
                }
            /* APPLY_BC_REM_DIV_MPI (for/mpi.f:1293-1332). */
            for (int i = 0; i <= NXP - 1; i++) {
                if (K == 0 && i == 0) {
                    MC(L, i, 1) = 0.; MC(D, i, 1) = 1.; MC(U, i, 1) = 0.; MC(V, i, 1) = 0.;
                } else {
                    MC(L, i, 1) = 0.; MC(D, i, 1) = 1.; MC(U, i, 1) = -1.; MC(V, i, 1) = 0.;
                }
                MC(L, i, NY) = 1.; MC(D, i, NY) = -1.; MC(U, i, NY) = 0.; MC(V, i, NY) = 0.;
            }
            dor_thomas_complex(L, D, U, (double *)V, NY, NXP);
            for (int J = 1; J <= NY; J++)
                for (int i = 0; i <= NXP - 1; i++) CR1[CIDX(i, K, J)] = MC(V, i, J);
        }
        free(L); free(D); free(U); free(V);
    }
    /* Projection (for/channel.f:1196-1211). */
#pragma omp parallel for schedule(static)
    for (int J = 1; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                CU1[CIDX(i, K, J)] = CU1[CIDX(i, K, J)] - CIKX(i) * CR1[CIDX(i, K, J)];
                CU3[CIDX(i, K, J)] = CU3[CIDX(i, K, J)] - CIKZ(K) * CR1[CIDX(i, K, J)];
            }
#pragma omp parallel for schedule(static)
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                cplx d = CR1[CIDX(i, K, J)] - CR1[CIDX(i, K, J - 1)];
                CU2[CIDX(i, K, J)] = CU2[CIDX(i, K, J)] - (creal(d) / DY[J] + I * (cimag(d) / DY[J]));
            }
#undef MC
}

/* ------------------------------------------------------------------------- */
/* POISSON_P_CHAN (for/channel.f:1217-1453)                                   */
/* ------------------------------------------------------------------------- */

void dor_poisson_p_chan(dor_state *s)
{
    const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ, NZP = s->NZP, NXM = s->NXM;
    const double *DY = s->DY, *DYF = s->DYF, *KX = s->KX, *KZ = s->KZ, *KX2 = s->KX2, *KZ2 = s->KZ2, *GYF = s->GYF;
    cplx *CU1 = C(DOR_U1), *CU2 = C(DOR_U2), *CU3 = C(DOR_U3), *CP = C(DOR_P);
    cplx *CF1 = C(DOR_F1), *CF2 = C(DOR_F2), *CF3 = C(DOR_F3), *CS1 = C(DOR_S1);
    double *F1 = R(DOR_F1), *F2 = R(DOR_F2), *F3 = R(DOR_F3);
#define CA(A) A[CIDX(i, K, J)]
#define CAJ(A, dj) A[CIDX(i, K, J + (dj))]
#define RA(A) A[RIDX(i, K, J)]
#define DIVR(z, r) (creal(z) / (r) + I * (cimag(z) / (r)))
#define PHYS_LOOP(BODY)                                                          \
    _Pragma("omp parallel for schedule(static)")                                 \
    for (int J = 2; J <= NY; J++)                                                \
        for (int K = 0; K <= NZP - 1; K++)                                       \
            for (int i = 0; i <= NXM; i++) { BODY; }

    ACC_LOOP(2, NY, CA(CF1) = CIKX(i) * CA(CU1); CA(CF2) = DIVR(CAJ(CU2, 1) - CA(CU2), DYF[J]);
             CA(CF3) = CIKZ(K) * CA(CU3));
    dor_fft_xz_to_physical(s, F1, 0, NY + 1);
    dor_fft_xz_to_physical(s, F2, 0, NY + 1);
    dor_fft_xz_to_physical(s, F3, 0, NY + 1);
    PHYS_LOOP(RA(F1) = RA(F1) * RA(F1); RA(F2) = RA(F2) * RA(F2); RA(F3) = RA(F3) * RA(F3));
    dor_fft_xz_to_fourier(s, F1, 0, NY + 1);
    dor_fft_xz_to_fourier(s, F2, 0, NY + 1);
    dor_fft_xz_to_fourier(s, F3, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CS1) = CA(CF1) + CA(CF2) + CA(CF3));

    ACC_LOOP(2, NY, CA(CF1) = DIVR(CAJ(CU1, 1) - CAJ(CU1, -1), 2. * DYF[J]);
             CA(CF2) = CIKX(i) * (0.5 * (CA(CU2) + CAJ(CU2, 1))));
    dor_fft_xz_to_physical(s, F1, 0, NY + 1);
    dor_fft_xz_to_physical(s, F2, 0, NY + 1);
    PHYS_LOOP(RA(F1) = 2. * RA(F1) * RA(F2));
    dor_fft_xz_to_fourier(s, F1, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CS1) = CA(CS1) + CA(CF1));

    ACC_LOOP(2, NY, CA(CF1) = DIVR(CAJ(CU3, 1) - CAJ(CU3, -1), 2. * DYF[J]);
             CA(CF2) = CIKZ(K) * (0.5 * (CA(CU2) + CAJ(CU2, 1))));
    dor_fft_xz_to_physical(s, F1, 0, NY + 1);
    dor_fft_xz_to_physical(s, F2, 0, NY + 1);
    PHYS_LOOP(RA(F1) = 2. * RA(F1) * RA(F2));
    dor_fft_xz_to_fourier(s, F1, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CS1) = CA(CS1) + CA(CF1));

    ACC_LOOP(2, NY, CA(CF1) = CIKZ(K) * CA(CU1); CA(CF2) = CIKX(i) * CA(CU3));
    dor_fft_xz_to_physical(s, F1, 0, NY + 1);
    dor_fft_xz_to_physical(s, F2, 0, NY + 1);
    PHYS_LOOP(RA(F1) = 2. * RA(F1) * RA(F2));
    dor_fft_xz_to_fourier(s, F1, 0, NY + 1);
    ACC_LOOP(2, NY, CA(CS1) = CA(CS1) + CA(CF1));

    /* Buoyancy contribution, (0,0) skipped (for/channel.f:1373-1384). */
    for (int n = 0; n < s->NTH; n++) {
        cplx *CTH = C(DOR_CTH0 + n);
        const double RI = s->p.ri[n];
        ACC_LOOP(2, NY, if (i != 0 || K != 0) CA(CS1) = CA(CS1) + DIVR(RI * (CAJ(CTH, 1) - CAJ(CTH, -1)), GYF[J + 1] - GYF[J - 1]));
    }

    const size_t ld = (size_t)NXP + 1;
    const size_t msz = ld * (size_t)(NY + 2);
#define MC(X, i, j) X[(size_t)(i) + ld * (size_t)(j)]
#pragma omp parallel
    {
        double *L = zalloc(msz), *D = zalloc(msz), *U = zalloc(msz);
        cplx *V = (cplx *)calloc(msz, sizeof(cplx));
        for (int J = 0; J <= NY + 1; J++)
            for (int i = 0; i <= NXP - 1; i++) { MC(L, i, J) = 0.; MC(D, i, J) = 1.; MC(U, i, J) = 0.; MC(V, i, J) = 0.; }
#pragma omp for schedule(static)
        for (int K = 0; K <= TNKZ; K++) {
            for (int J = 2; J <= NY; J++)
                for (int i = 0; i <= NXP - 1; i++) {
                    MC(L, i, J) = 1. / (DY[J] * DYF[J]);
                    MC(D, i, J) = -KX2[i] - KZ2[K] - 1. / (DY[J + 1] * DYF[J]) - 1. / (DY[J] * DYF[J]);
                    MC(U, i, J) = 1. / (DY[J + 1] * DYF[J]);
                    MC(V, i, J) = -1. * CS1[CIDX(i, K, J)];
                }
            /* APPLY_BC_POISSON_MPI (for/mpi.f:1334-1369). */
            for (int i = 0; i <= NXP - 1; i++) {
                if (i == 0 && K == 0) {
                    MC(D, i, 1) = 1.; MC(U, i, 1) = 0.; MC(V, i, 1) = 0.;
                } else {
                    MC(D, i, 1) = 1.; MC(U, i, 1) = -1.; MC(V, i, 1) = 0.;
                }
                MC(D, i, NY) = -1.; MC(L, i, NY) = 1.; MC(V, i, NY) = 0.;
            }
            dor_thomas_complex(L, D, U, (double *)V, NY, NXP);
            for (int J = 1; J <= NY; J++)
                for (int i = 0; i <= NXP - 1; i++) CP[CIDX(i, K, J)] = MC(V, i, J);
        }
        free(L); free(D); free(U); free(V);
    }
#undef MC
#undef CA
#undef CAJ
#undef RA
}

/* ------------------------------------------------------------------------- */
/* APPLY_BC_VEL_LOWER/UPPER (for/channel.f:2285-2492), SCALAR (2723-2807)     */
/* ------------------------------------------------------------------------- */

void dor_apply_bc_vel(dor_state *s)
{
    const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ;
    const double *DY = s->DY, *DYF = s->DYF;
    cplx *CU1 = C(DOR_U1), *CU2 = C(DOR_U2), *CU3 = C(DOR_U3);
    const dor_params *p = &s->p;
#define X(A, i, k, j) A[CIDX(i, k, j)]
    /* Lower wall. */
    cplx *CUH[2] = {CU1, CU3};
    const int hcode[2] = {p->u_bc_ymin, p->w_bc_ymin};
    const double hc1[2] = {p->u_bc_ymin_c1, p->w_bc_ymin_c1};
    for (int c = 0; c < 2; c++) {
        cplx *A = CUH[c];
        if (hcode[c] == 0) {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, 0) = 0.; X(A, i, K, 1) = 0.; }
            X(A, 0, 0, 1) = hc1[c];
            X(A, 0, 0, 0) = hc1[c];
        } else {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, 1) = X(A, i, K, 2); X(A, i, K, 0) = X(A, i, K, 1); }
            X(A, 0, 0, 1) = X(A, 0, 0, 2) - DY[2] * hc1[c];
            X(A, 0, 0, 0) = X(A, 0, 0, 1) - DY[1] * hc1[c];
        }
    }
    if (p->v_bc_ymin == 0) {
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                X(CU2, i, K, 1) = 2.0 * p->v_bc_ymin_c1 - X(CU2, i, K, 2);
                X(CU2, i, K, 0) = X(CU2, i, K, 1);
            }
    } else if (p->v_bc_ymin == 1) {
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) { X(CU2, i, K, 1) = X(CU2, i, K, 2); X(CU2, i, K, 0) = X(CU2, i, K, 1); }
        X(CU2, 0, 0, 1) = X(CU2, 0, 0, 2) - DYF[1] * p->v_bc_ymin_c1;
        X(CU2, 0, 0, 0) = X(CU2, 0, 0, 1) - DYF[1] * p->v_bc_ymin_c1;
    } else if (p->v_bc_ymin == 2) {
        X(CU2, 0, 0, 1) = -X(CU2, 0, 0, 2);
    }
    /* Upper wall. */
    const int ucode[2] = {p->u_bc_ymax, p->w_bc_ymax};
    const double uc1[2] = {p->u_bc_ymax_c1, p->w_bc_ymax_c1};
    for (int c = 0; c < 2; c++) {
        cplx *A = CUH[c];
        if (ucode[c] == 0) {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, NY) = 0.; X(A, i, K, NY + 1) = 0.; }
            X(A, 0, 0, NY) = uc1[c];
            X(A, 0, 0, NY + 1) = uc1[c];
            if (c == 1) X(A, 0, 0, NY + 1) = 0.0; /* for/channel.f:2445 */
        } else {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, NY) = X(A, i, K, NY - 1); X(A, i, K, NY + 1) = X(A, i, K, NY); }
            X(A, 0, 0, NY) = X(A, 0, 0, NY - 1) + DY[NY] * uc1[c];
            X(A, 0, 0, NY + 1) = X(A, 0, 0, NY) + DY[NY] * uc1[c];
        }
    }
    if (p->v_bc_ymax == 0) {
        X(CU2, 0, 0, NY + 1) = 2.0 * p->v_bc_ymax_c1 - X(CU2, 0, 0, NY);
    } else if (p->v_bc_ymax == 1) {
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) { X(CU2, i, K, NY) = X(CU2, i, K, NY - 1); X(CU2, i, K, NY + 1) = X(CU2, i, K, NY); }
        X(CU2, 0, 0, NY + 1) = X(CU2, 0, 0, NY) + DY[NY] * p->v_bc_ymax_c1;
    } else if (p->v_bc_ymax == 2) {
        X(CU2, 0, 0, NY + 1) = -X(CU2, 0, 0, NY);
    }
}

void dor_apply_bc_scalar(dor_state *s)
{
    const int NY = s->NY, NXP = s->NXP, TNKZ = s->TNKZ;
    const double *DY = s->DY;
    for (int n = 0; n < s->NTH; n++) {
        cplx *A = C(DOR_CTH0 + n);
        const double c0 = s->p.th_bc_ymin_c1[n], c1 = s->p.th_bc_ymax_c1[n];
        if (s->p.th_bc_ymin[n] == 0) {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, 0) = 0.; X(A, i, K, 1) = 0.; }
            X(A, 0, 0, 1) = c0; X(A, 0, 0, 0) = c0;
        } else {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, 1) = X(A, i, K, 2); X(A, i, K, 0) = X(A, i, K, 1); }
            X(A, 0, 0, 1) = X(A, 0, 0, 2) - DY[2] * c0;
            X(A, 0, 0, 0) = X(A, 0, 0, 1) - DY[1] * c0;
        }
        if (s->p.th_bc_ymax[n] == 0) {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, NY) = 0.; X(A, i, K, NY + 1) = 0.; }
            X(A, 0, 0, NY) = c1; X(A, 0, 0, NY + 1) = c1;
        } else {
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) { X(A, i, K, NY) = X(A, i, K, NY - 1); X(A, i, K, NY + 1) = X(A, i, K, NY); }
            X(A, 0, 0, NY) = X(A, 0, 0, NY - 1) + DY[NY] * c1;
            X(A, 0, 0, NY + 1) = X(A, 0, 0, NY) + DY[NY] * c1;
        }
    }
#undef X
}

/* ------------------------------------------------------------------------ */
/* SAVE_STATS_CHAN statistics (for/save_stats.f:106-347, scalar part        */
/* for/save_stats.f:420-522), non-FINAL branch at NPROCY = 1.  Physical     */
/* fields come from FFT_XZ_TO_PHYSICAL of copies, as the reference converts */
/* CU1-3, CS1, CTH and the CR gradients; sums run over k = 0..NZP-1,        */
/* i = 0..NXM in the reference's order.                                     */
/* ------------------------------------------------------------------------ */
void dor_stats_chan(dor_state *s, double *out, double *ubulk)
{
    const int NX = s->NX, NY = s->NY, NZ = s->NZ, NXP = s->NXP, TNKZ = s->TNKZ, NXM = s->NXM, NZP = s->NZP;
    const double *DY = s->DY, *DYF = s->DYF, *GYF = s->GYF, *KX = s->KX, *KZ = s->KZ;
    const size_t fs = s->fsize;
    double *u1 = zalloc(fs), *u2 = zalloc(fs), *u3 = zalloc(fs), *th = zalloc(fs), *w = zalloc(fs);
    double *w2 = zalloc(fs), *w3 = zalloc(fs);
    const cplx *CU1 = (const cplx *)s->fld[DOR_U1], *CU2 = (const cplx *)s->fld[DOR_U2];
    const cplx *CU3 = (const cplx *)s->fld[DOR_U3], *CP = (const cplx *)s->fld[DOR_P];
    const cplx *CTH = s->NTH > 0 ? (const cplx *)s->fld[DOR_CTH0] : NULL;
    double *ume = zalloc(NY + 2), *vme = zalloc(NY + 2), *wme = zalloc(NY + 2), *thme = zalloc(NY + 2);
    double *acc = zalloc((size_t)DOR_NSTAT * (NY + 2));
#define ST(q, j) acc[(size_t)(q) * (NY + 2) + (j)]
#define CI(i) (I * KX[i])
#define CK(k) (I * KZ[k])
    for (int j = 0; j <= NY + 1; j++) {   /* for/save_stats.f:146-151 */
        ume[j] = creal(CU1[CIDX(0, 0, j)]);
        vme[j] = creal(CU2[CIDX(0, 0, j)]);
        wme[j] = creal(CU3[CIDX(0, 0, j)]);
        if (CTH) thme[j] = creal(CTH[CIDX(0, 0, j)]);
    }
    /* INTEGRATE_Y_VAR (for/mpi.f:1676-1698) */
    double res = 0.;
    for (int j = 2; j <= NY; j++) res = res + 0.5 * (ume[j] + ume[j - 1]) * DY[j];
    *ubulk = res / s->p.ly;
    /* for/save_stats.f:163-165 */
    memcpy(u1, s->fld[DOR_U1], fs * sizeof(double));
    memcpy(u2, s->fld[DOR_U2], fs * sizeof(double));
    memcpy(u3, s->fld[DOR_U3], fs * sizeof(double));
    dor_fft_xz_to_physical(s, u1, 0, NY + 1);
    dor_fft_xz_to_physical(s, u2, 0, NY + 1);
    dor_fft_xz_to_physical(s, u3, 0, NY + 1);
    const double nn = (double)NZ * (double)NX;
    /* for/save_stats.f:172-193 */
    for (int j = 1; j <= NY; j++) {
        double ur = 0., vr = 0., wr = 0., uv = 0., wv = 0., uw = 0., sh = 0.;
        for (int k = 0; k <= NZP - 1; k++)
            for (int i = 0; i <= NXM; i++) {
                const double du = u1[RIDX(i, k, j)] - ume[j], dw = u3[RIDX(i, k, j)] - wme[j];
                ur = ur + pow(du, 2.);
                vr = vr + 0.5 * (pow(u2[RIDX(i, k, j)] - vme[j], 2.) + pow(u2[RIDX(i, k, j + 1)] - vme[j + 1], 2.));
                wr = wr + pow(dw, 2.);
                /* for/save_stats.f:205-218 */
                uv = uv + (DYF[j - 1] * du + DYF[j] * (u1[RIDX(i, k, j - 1)] - ume[j - 1])) / (2.0 * DY[j]) *
                              (u2[RIDX(i, k, j)] - vme[j]);
                wv = wv + (DYF[j - 1] * dw + DYF[j] * (u3[RIDX(i, k, j - 1)] - wme[j - 1])) / (2.0 * DY[j]) *
                              (u2[RIDX(i, k, j)] - vme[j]);
                uw = uw + du * dw;
                /* for/save_stats.f:242-250 */
                sh = sh + pow((u1[RIDX(i, k, j + 1)] - u1[RIDX(i, k, j - 1)]) / (2.0 * DYF[j]), 2.0) +
                     pow((u3[RIDX(i, k, j + 1)] - u3[RIDX(i, k, j - 1)]) / (2.0 * DYF[j]), 2.0);
            }
        ST(DOR_ST_URMS, j) = sqrt(ur / nn);
        ST(DOR_ST_VRMS, j) = sqrt(vr / nn);
        ST(DOR_ST_WRMS, j) = sqrt(wr / nn);
        ST(DOR_ST_UV, j) = uv / nn;
        ST(DOR_ST_WV, j) = wv / nn;
        ST(DOR_ST_UW, j) = uw / nn;
        ST(DOR_ST_SHEAR, j) = sh / (double)(NX * NZ);
        /* for/save_stats.f:235-238 */
        ST(DOR_ST_DUDY, j) = (ume[j] - ume[j - 1]) / (GYF[j] - GYF[j - 1]);
        ST(DOR_ST_DWDY, j) = (wme[j] - wme[j - 1]) / (GYF[j] - GYF[j - 1]);
        ST(DOR_ST_UME, j) = ume[j];
        ST(DOR_ST_VME, j) = vme[j];
        ST(DOR_ST_WME, j) = wme[j];
        ST(DOR_ST_CP, j) = creal(CP[CIDX(0, 0, j)]);
    }
    /* vorticity (for/save_stats.f:268-347): CS1 from the saved CR = CU */
    for (int comp = 0; comp < 3; comp++) {
        memset(w, 0, fs * sizeof(double));
        cplx *CS1 = (cplx *)w;
        for (int j = 1; j <= NY; j++)
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) {
                    cplx v;
                    if (comp == 0)
                        v = (CU3[CIDX(i, K, j + 1)] - CU3[CIDX(i, K, j - 1)]) / (2.0 * DYF[j]) -
                            CK(K) * 0.5 * (CU2[CIDX(i, K, j + 1)] + CU2[CIDX(i, K, j)]);
                    else if (comp == 1)
                        v = CK(K) * CU1[CIDX(i, K, j)] - CI(i) * CU3[CIDX(i, K, j)];
                    else
                        v = CI(i) * 0.5 * (CU2[CIDX(i, K, j + 1)] + CU2[CIDX(i, K, j)]) -
                            (CU1[CIDX(i, K, j + 1)] - CU1[CIDX(i, K, j - 1)]) / (2.0 * DYF[j]);
                    CS1[CIDX(i, K, j)] = v;
                }
        dor_fft_xz_to_physical(s, w, 0, NY + 1);
        for (int j = 1; j <= NY; j++) {
            double o = 0.;
            for (int k = 0; k <= NZP - 1; k++)
                for (int i = 0; i <= NXM; i++) o = o + pow(w[RIDX(i, k, j)], 2.0);
            ST(DOR_ST_OMEGA_X + comp, j) = sqrt(o / ((double)NX * (double)NZ));
        }
    }
    if (CTH) {
        /* scalar 1: TH, its rms, flux with U2, mean gradient, grad TH squared */
        memcpy(th, s->fld[DOR_CTH0], fs * sizeof(double));
        dor_fft_xz_to_physical(s, th, 0, NY + 1);
        memset(w, 0, fs * sizeof(double));
        memset(w2, 0, fs * sizeof(double));
        memset(w3, 0, fs * sizeof(double));
        cplx *R1c = (cplx *)w, *R2c = (cplx *)w2, *R3c = (cplx *)w3;
        for (int j = 1; j <= NY; j++)    /* for/save_stats.f:438-448 */
            for (int K = 0; K <= TNKZ; K++)
                for (int i = 0; i <= NXP - 1; i++) {
                    R1c[CIDX(i, K, j)] = CI(i) * CTH[CIDX(i, K, j)];
                    R2c[CIDX(i, K, j)] = (CTH[CIDX(i, K, j + 1)] - CTH[CIDX(i, K, j - 1)]) / (GYF[j + 1] - GYF[j - 1]);
                    R3c[CIDX(i, K, j)] = CK(K) * CTH[CIDX(i, K, j)];
                }
        dor_fft_xz_to_physical(s, w, 0, NY + 1);
        dor_fft_xz_to_physical(s, w2, 0, NY + 1);
        dor_fft_xz_to_physical(s, w3, 0, NY + 1);
        for (int j = 1; j <= NY; j++) {
            double tr = 0., tv = 0., pe = 0.;
            for (int k = 0; k <= NZP - 1; k++)
                for (int i = 0; i <= NXM; i++) {
                    tr = tr + pow(fabs(th[RIDX(i, k, j)] - thme[j]), 2.);
                    tv = tv + (th[RIDX(i, k, j)] - thme[j]) *
                                  (0.5 * (u2[RIDX(i, k, j)] + u2[RIDX(i, k, j + 1)]) - 0.5 * (vme[j] + vme[j + 1]));
                    pe = pe + pow(w[RIDX(i, k, j)], 2.0) + pow(w2[RIDX(i, k, j)], 2.0) + pow(w3[RIDX(i, k, j)], 2.0);
                }
            ST(DOR_ST_THRMS, j) = sqrt(tr / nn);
            ST(DOR_ST_THV, j) = tv / nn;
            ST(DOR_ST_PE_DISS, j) = pe / (double)(NX * NZ);
            ST(DOR_ST_DTHDY, j) = (thme[j] - thme[j - 1]) / (GYF[j] - GYF[j - 1]);
            ST(DOR_ST_THME, j) = thme[j];
        }
    }
    for (int q = 0; q < DOR_NSTAT; q++)
        for (int j = 1; j <= NY; j++) out[(size_t)q * NY + (j - 1)] = ST(q, j);
#undef ST
#undef CI
#undef CK
    free(u1); free(u2); free(u3); free(th); free(w); free(w2); free(w3);
    free(ume); free(vme); free(wme); free(thme); free(acc);
}

/* ------------------------------------------------------------------------ */
/* Restart-file layout of WriteHDF5 / ReadHDF5 (for/hdf5.f:172-268,          */
/* 470-534) at NPROCY = NPROCZ = 1: the physical field, SWAPZY'd to          */
/* tmp(1:NX, 1:NY, 1:NZ) (x fastest, then y, then z), V interpolated to the  */
/* fractional grid by G2GF first (for/hdf5.f:1176-1192); reading SWAPYZ's    */
/* the file array back (zeroing the pads and planes 0 and NY+1), applies     */
/* GF2G to V (for/hdf5.f:1196-1238) and transforms to Fourier space.         */
/* The field id selects the Fourier array: U1/U2/U3/P or CTH0 + n.           */
/* ------------------------------------------------------------------------ */
static void g2gf(dor_state *s, double *v)
{
    const int NX = s->NX, NY = s->NY, NZP = s->NZP;
    for (int x = 0; x <= NX - 1; x++)
        for (int z = 0; z <= NZP - 1; z++) {
            v[RIDX(x, z, 1)] = v[RIDX(x, z, 2)];
            v[RIDX(x, z, NY + 1)] = v[RIDX(x, z, NY)];
            for (int y = 1; y <= NY; y++) v[RIDX(x, z, y)] = 0.5 * (v[RIDX(x, z, y)] + v[RIDX(x, z, y + 1)]);
        }
}

static void gf2g(dor_state *s, double *v)
{
    const int NX = s->NX, NY = s->NY, NZP = s->NZP, NZ = s->NZ;
    for (int x = 0; x <= NX - 1; x++)
        for (int z = 0; z <= NZP - 1; z++) v[RIDX(x, z, NY + 1)] = v[RIDX(x, z, NY)];
    for (int x = 0; x <= NX - 1; x++)
        for (int z = 0; z <= NZP - 1; z++)
            for (int y = NY; y >= 1; y--) v[RIDX(x, z, y)] = 2.0 * v[RIDX(x, z, y)] - v[RIDX(x, z, y + 1)];
    /* var(:,:,1) = -var(:,:,2) over the whole plane, pads included */
    for (int z = 0; z <= NZ + 1; z++)
        for (int x = 0; x <= NX + 1; x++) {
            v[RIDX(x, z, 1)] = -v[RIDX(x, z, 2)];
            v[RIDX(x, z, NY + 1)] = -v[RIDX(x, z, NY)];
        }
}

void dor_flow_to_file(dor_state *s, int id, double *out)
{
    const int NX = s->NX, NY = s->NY, NZP = s->NZP;
    double *w = zalloc(s->fsize);
    memcpy(w, s->fld[id], sizeof(double) * s->fsize);
    dor_fft_xz_to_physical(s, w, 0, NY + 1);
    if (id == DOR_U2) g2gf(s, w);
    for (int x = 0; x <= NX - 1; x++)        /* SWAPZY */
        for (int y = 1; y <= NY; y++)
            for (int z = 0; z <= NZP - 1; z++)
                out[(size_t)x + (size_t)NX * ((size_t)(y - 1) + (size_t)NY * (size_t)z)] = w[RIDX(x, z, y)];
    free(w);
}

void dor_flow_from_file(dor_state *s, int id, const double *in)
{
    const int NX = s->NX, NY = s->NY, NZP = s->NZP;
    double *w = zalloc(s->fsize);           /* SWAPYZ: out = 0 first */
    for (int x = 0; x <= NX - 1; x++)
        for (int y = 1; y <= NY; y++)
            for (int z = 0; z <= NZP - 1; z++)
                w[RIDX(x, z, y)] = in[(size_t)x + (size_t)NX * ((size_t)(y - 1) + (size_t)NY * (size_t)z)];
    if (id == DOR_U2) gf2g(s, w);
    dor_fft_xz_to_fourier(s, w, 0, NY + 1);
    memcpy(s->fld[id], w, sizeof(double) * s->fsize);
    free(w);
}

/* ------------------------------------------------------------------------ */
/* tkebudget_chan's dissipation rate (for/save_stats.f:965-1159), as         */
/* SAVE_STATS_CHAN calls it (for/save_stats.f:134-169): CR1-3 hold copies    */
/* of CU1-3, U1-3 their physical fields, ume/vme/wme the (0,0) modes.  The  */
/* scratch CS1/S1 is the state's own S1 array with whatever it holds (the   */
/* reference's plane 0 of it is never set here, so epsilon(1) depends on     */
/* it), and each gradient block overwrites planes 1..NY only (dv/dy: 2..NY-1, */
/* the others keep du/dz's values).  eps[j-1], j = 1..NY.                    */
/* ------------------------------------------------------------------------ */
void dor_tke_chan(dor_state *s, double *eps)
{
    const int NX = s->NX, NY = s->NY, NZ = s->NZ, NXP = s->NXP, TNKZ = s->TNKZ, NXM = s->NXM, NZP = s->NZP;
    const double *DY = s->DY, *DYF = s->DYF, *GY = s->GY, *KX = s->KX, *KZ = s->KZ;
    const size_t fs = s->fsize;
    const cplx *CR1 = (const cplx *)s->fld[DOR_U1], *CR2 = (const cplx *)s->fld[DOR_U2];
    const cplx *CR3 = (const cplx *)s->fld[DOR_U3];
    double *U1 = zalloc(fs), *U2 = zalloc(fs), *U3 = zalloc(fs);
    memcpy(U1, s->fld[DOR_U1], sizeof(double) * fs);
    memcpy(U2, s->fld[DOR_U2], sizeof(double) * fs);
    memcpy(U3, s->fld[DOR_U3], sizeof(double) * fs);
    double *ume = zalloc(NY + 2), *vme = zalloc(NY + 2), *wme = zalloc(NY + 2), *e = zalloc(NY + 2);
    for (int j = 0; j <= NY + 1; j++) {
        ume[j] = creal(CR1[CIDX(0, 0, j)]);
        vme[j] = creal(CR2[CIDX(0, 0, j)]);
        wme[j] = creal(CR3[CIDX(0, 0, j)]);
    }
    dor_fft_xz_to_physical(s, U1, 0, NY + 1);
    dor_fft_xz_to_physical(s, U2, 0, NY + 1);
    dor_fft_xz_to_physical(s, U3, 0, NY + 1);
    double *S1 = s->fld[DOR_S1];
    cplx *CS1 = (cplx *)S1;
    /* CS1 = (i kx or i kz) * CR for j = 1..NY, then to physical space */
    #define SPEC(CR, KV, KIDX)                                                        \
        do {                                                                            \
            for (int j = 1; j <= NY; j++)                                               \
                for (int k = 0; k <= TNKZ; k++)                                         \
                    for (int i = 0; i <= NXP - 1; i++)                                  \
                        CS1[CIDX(i, k, j)] = (I * KV[KIDX]) * CR[CIDX(i, k, j)];        \
            dor_fft_xz_to_physical(s, S1, 0, NY + 1);                                   \
        } while (0)
    /* interpolated to GY: (DYF(j-1) S1(j)^2 + DYF(j) S1(j-1)^2) / (2 DY(j)) */
    #define ACC_GY()                                                                    \
        for (int j = 1; j <= NY; j++)                                                   \
            for (int k = 0; k <= NZP - 1; k++)                                          \
                for (int i = 0; i <= NXM; i++)                                          \
                    e[j] = e[j] + (DYF[j - 1] * pow(S1[RIDX(i, k, j)], 2.0) +          \
                                   DYF[j] * pow(S1[RIDX(i, k, j - 1)], 2.0)) / (2.0 * DY[j])
    #define ACC_SQ()                                                                    \
        for (int j = 1; j <= NY; j++)                                                   \
            for (int k = 0; k <= NZP - 1; k++)                                          \
                for (int i = 0; i <= NXM; i++) e[j] = e[j] + pow(S1[RIDX(i, k, j)], 2.0)
    SPEC(CR1, KX, i); ACC_GY();                                   /* du/dx */
    SPEC(CR2, KX, i); ACC_SQ();                                   /* dv/dx */
    for (int j = 1; j <= NY; j++)                                 /* du/dy */
        for (int k = 0; k <= NZP - 1; k++)
            for (int i = 0; i <= NXM; i++)
                S1[RIDX(i, k, j)] = ((U1[RIDX(i, k, j)] - ume[j]) - (U1[RIDX(i, k, j - 1)] - ume[j - 1])) / DY[j];
    ACC_SQ();
    SPEC(CR3, KX, i); ACC_GY();                                   /* dw/dx */
    SPEC(CR1, KZ, k); ACC_GY();                                   /* du/dz */
    for (int j = 2; j <= NY - 1; j++)                             /* dv/dy */
        for (int k = 0; k <= NZP - 1; k++)
            for (int i = 0; i <= NXM; i++)
                S1[RIDX(i, k, j)] = ((U2[RIDX(i, k, j + 1)] - vme[j]) - (U2[RIDX(i, k, j - 1)] - vme[j - 1]))
                                    / (GY[j + 1] - GY[j - 1]);
    ACC_SQ();
    for (int j = 1; j <= NY; j++)                                 /* dw/dy */
        for (int k = 0; k <= NZP - 1; k++)
            for (int i = 0; i <= NXM; i++)
                S1[RIDX(i, k, j)] = ((U3[RIDX(i, k, j)] - wme[j]) - (U3[RIDX(i, k, j - 1)] - wme[j - 1])) / DY[j];
    ACC_SQ();
    SPEC(CR2, KZ, k); ACC_SQ();                                   /* dv/dz */
    SPEC(CR3, KZ, k); ACC_GY();                                   /* dw/dz */
    #undef SPEC
    #undef ACC_GY
    #undef ACC_SQ
    for (int j = 1; j <= NY; j++) eps[j - 1] = s->p.nu * e[j] / (double)(NX * NZ);
    free(U1); free(U2); free(U3); free(ume); free(vme); free(wme); free(e);
}

/* ------------------------------------------------------------------------ */
/* les_chan (for/les.f:1-267, 274-331), LES_MODEL_TYPE = 1 (constant         */
/* Smagorinsky; the dynamic models stop under MPI, for/les.f:269-272), at    */
/* NPROCY = 1: APPLY_BC_VEL, APPLY_BC_LES (for/les.f:805-864),               */
/* compute_strain_chan (for/les.f:526-571), |S| on GYF and GY, NU_T on GY    */
/* with GHOST_LES_MPI's zero rows 1, 2, NY, NY+1 (for/mpi.f:488-503), the    */
/* stress tau_ij = DELTA * |S| S_ij in Fourier space and its divergence      */
/* subtracted from CF1-3, plus the explicit cross terms added to CF2.  The   */
/* mean-statistics block (SAVE_STATS_INT) writes files only and is left out. */
/* nu_t: NU_T (field layout), may be NULL.  The velocity state is changed    */
/* only by the two BC routines, as in the reference; it stays in Fourier     */
/* space here (the reference leaves U1-3 physical on return).                */
/* ------------------------------------------------------------------------ */
void dor_les_chan(dor_state *s, double *nu_t)
{
    const int NX = s->NX, NY = s->NY, NZ = s->NZ, NXP = s->NXP, TNKZ = s->TNKZ, NXM = s->NXM, NZP = s->NZP;
    const double *DY = s->DY, *DYF = s->DYF, *KX = s->KX, *KZ = s->KZ;
    const size_t fs = s->fsize;
    const dor_params *p = &s->p;
    const double C_SMAG = 0.13, alpha_sgs = 2.449, beta_sgs = 1.0;
    (void)alpha_sgs;
    const int J1 = s->JSTART, J2 = NY - 1;     /* RANKY = NPROCY-1 branch, for/les.f:46-50 */
    dor_apply_bc_vel(s);
    cplx *CU1 = (cplx *)s->fld[DOR_U1], *CU2 = (cplx *)s->fld[DOR_U2], *CU3 = (cplx *)s->fld[DOR_U3];
    /* APPLY_BC_LES */
    if (p->u_bc_ymax == 1)
        for (int k = 0; k <= TNKZ; k++) for (int i = 0; i <= NXP - 1; i++) {
            CU1[CIDX(i, k, NY)] = CU1[CIDX(i, k, NY - 1)];
            CU1[CIDX(i, k, NY + 1)] = CU1[CIDX(i, k, NY)];
        }
    if (p->w_bc_ymax == 1)
        for (int k = 0; k <= TNKZ; k++) for (int i = 0; i <= NXP - 1; i++) {
            CU3[CIDX(i, k, NY)] = CU3[CIDX(i, k, NY - 1)];
            CU3[CIDX(i, k, NY + 1)] = CU3[CIDX(i, k, NY)];
        }
    if (p->u_bc_ymin == 1)
        for (int k = 0; k <= TNKZ; k++) for (int i = 0; i <= NXP - 1; i++) {
            CU1[CIDX(i, k, 1)] = CU1[CIDX(i, k, 2)];
            CU1[CIDX(i, k, 0)] = CU1[CIDX(i, k, 1)];
        }
    if (p->w_bc_ymin == 1)
        for (int k = 0; k <= TNKZ; k++) for (int i = 0; i <= NXP - 1; i++) {
            CU3[CIDX(i, k, 1)] = CU3[CIDX(i, k, 2)];
            CU3[CIDX(i, k, 0)] = CU3[CIDX(i, k, 1)];
        }
    /* compute_strain_chan: S_ij in Fourier space, then physical */
    double *Sij[7];
    for (int q = 1; q <= 6; q++) Sij[q] = zalloc(fs);
    cplx *CS[7];
    for (int q = 1; q <= 6; q++) CS[q] = (cplx *)Sij[q];
    for (int J = 1; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const size_t x = CIDX(i, K, J);
                CS[1][x] = (I * KX[i]) * CU1[x];
                CS[2][x] = (CU2[CIDX(i, K, J + 1)] - CU2[x]) / DYF[J];
                CS[3][x] = (I * KZ[K]) * CU3[x];
                CS[5][x] = 0.5 * ((I * KZ[K]) * CU1[x] + (I * KX[i]) * CU3[x]);
            }
    for (int J = 1; J <= NY + 1; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const size_t x = CIDX(i, K, J);
                CS[4][x] = 0.5 * ((CU1[x] - CU1[CIDX(i, K, J - 1)]) / DY[J] + (I * KX[i]) * CU2[x]);
                CS[6][x] = 0.5 * ((I * KZ[K]) * CU2[x] + (CU3[x] - CU3[CIDX(i, K, J - 1)]) / DY[J]);
            }
    for (int q = 1; q <= 6; q++) dor_fft_xz_to_physical(s, Sij[q], 0, NY + 1);
    double *U1b = zalloc(NY + 2), *U3b = zalloc(NY + 2);
    for (int J = 0; J <= NY + 1; J++) {
        U1b[J] = creal(CU1[CIDX(0, 0, J)]);
        U3b[J] = creal(CU3[CIDX(0, 0, J)]);
    }
    double *U1 = zalloc(fs), *U3 = zalloc(fs);
    memcpy(U1, s->fld[DOR_U1], sizeof(double) * fs);
    memcpy(U3, s->fld[DOR_U3], sizeof(double) * fs);
    dor_fft_xz_to_physical(s, U1, 0, NY + 1);
    dor_fft_xz_to_physical(s, U3, 0, NY + 1);
    double *S1 = zalloc(fs), *TEMP = zalloc(fs), *NUT = zalloc(fs);
#define P2(v) pow((v), 2.0)
    for (int J = s->JSTART; J <= s->JEND; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++) {
                const size_t x = RIDX(i, K, J), xp = RIDX(i, K, J + 1);
                S1[x] = sqrt(2.0 * P2(Sij[1][x])
                             + 4.0 * P2(0.5 * (Sij[4][xp] + Sij[4][x]) - 0.5 * (U1b[J] - U1b[J - 1]) / DY[J]
                                        - 0.5 * (U1b[J + 1] - U1b[J]) / DY[J + 1])
                             + 4.0 * P2(Sij[5][x]) + 2.0 * P2(Sij[2][x])
                             + 4.0 * P2(0.5 * (Sij[6][xp] + Sij[6][x]) - 0.5 * (U3b[J] - U3b[J - 1]) / DY[J]
                                        - 0.5 * (U3b[J + 1] - U3b[J]) / DY[J + 1])
                             + 2.0 * P2(Sij[3][x]));
            }
    for (int J = 2; J <= NY + 1; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++) {
                const size_t x = RIDX(i, K, J), xm = RIDX(i, K, J - 1);
                TEMP[x] = sqrt(2.0 * P2((Sij[1][x] * DYF[J - 1] + Sij[1][xm] * DYF[J]) / (2.0 * DY[J]))
                               + 4.0 * P2(Sij[4][x] - 0.5 * (U1b[J] - U1b[J - 1]) / DY[J])
                               + 4.0 * P2((Sij[5][x] * DYF[J - 1] + Sij[5][xm] * DYF[J]) / (2.0 * DY[J]))
                               + 2.0 * P2((Sij[2][x] * DYF[J - 1] + Sij[2][xm] * DYF[J]) / (2.0 * DY[J]))
                               + 4.0 * P2(Sij[6][x] - 0.5 * (U3b[J] - U3b[J - 1]) / DY[J])
                               + 2.0 * P2((Sij[3][x] * DYF[J - 1] + Sij[3][xm] * DYF[J]) / (2.0 * DY[J])));
            }
#undef P2
    for (int J = s->JSTART; J <= s->JEND; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++) {
                const size_t x = RIDX(i, K, J);
                Sij[1][x] = S1[x] * Sij[1][x];
                Sij[5][x] = S1[x] * Sij[5][x];
                Sij[2][x] = 0.0;
                Sij[3][x] = S1[x] * Sij[3][x];
            }
    for (int J = 2; J <= NY + 1; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++) {
                const size_t x = RIDX(i, K, J), xm = RIDX(i, K, J - 1);
                Sij[4][x] = TEMP[x] * (Sij[4][x] - 0.5 * (U1[x] - U1[xm]) / DY[J]);
                Sij[6][x] = TEMP[x] * (Sij[6][x] - 0.5 * (U3[x] - U3[xm]) / DY[J]);
            }
    /* filter widths, -2 C_SMAG^2 absorbed (for/les.f:200-223) */
    const double dx1 = p->lx / NX, dz1 = p->lz / NZ;
    double *DELTA_YF = zalloc(NY + 2), *DELTA_Y = zalloc(NY + 2);
    for (int J = 1; J <= NY + 1; J++) {
        DELTA_YF[J] = -2.0 * pow(C_SMAG, 2.0) * pow(dx1 * beta_sgs * DYF[J] * 2.0 * dz1 * beta_sgs, 2.0 / 3.0);
        DELTA_Y[J] = -2.0 * pow(C_SMAG, 2.0) * pow(dx1 * beta_sgs * DY[J] * 2.0 * dz1 * beta_sgs, 2.0 / 3.0);
    }
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++) NUT[RIDX(i, K, J)] = -0.5 * DELTA_Y[J] * TEMP[RIDX(i, K, J)];
    for (int K = 0; K <= NZP - 1; K++)         /* GHOST_LES_MPI at NPROCY = 1 */
        for (int i = 0; i <= NXM; i++) {
            NUT[RIDX(i, K, 1)] = 0.0;
            NUT[RIDX(i, K, 2)] = 0.0;
            NUT[RIDX(i, K, NY)] = 0.0;
            NUT[RIDX(i, K, NY + 1)] = 0.0;
        }
    const int tq[5] = {1, 3, 4, 5, 6};
    for (int t = 0; t < 5; t++) dor_fft_xz_to_fourier(s, Sij[tq[t]], 0, NY + 1);
    for (int J = 1; J <= NY + 1; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const size_t x = CIDX(i, K, J);
                CS[1][x] = DELTA_YF[J] * CS[1][x];
                CS[5][x] = DELTA_YF[J] * CS[5][x];
                CS[3][x] = DELTA_YF[J] * CS[3][x];
            }
    for (int J = 2; J <= NY + 1; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const size_t x = CIDX(i, K, J);
                CS[4][x] = DELTA_Y[J] * CS[4][x];
                CS[6][x] = DELTA_Y[J] * CS[6][x];
            }
    /* the SGS forcing into CF1-3 (for/les.f:274-301) */
    cplx *CF1 = (cplx *)s->fld[DOR_F1], *CF2 = (cplx *)s->fld[DOR_F2], *CF3 = (cplx *)s->fld[DOR_F3];
    for (int J = J1; J <= J2; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const size_t x = CIDX(i, K, J), xp = CIDX(i, K, J + 1);
                CF1[x] = CF1[x] - (I * KX[i]) * CS[1][x] - (CS[4][xp] - CS[4][x]) / DYF[J] - (I * KZ[K]) * CS[5][x];
                CF3[x] = CF3[x] - (I * KX[i]) * CS[5][x] - (CS[6][xp] - CS[6][x]) / DYF[J] - (I * KZ[K]) * CS[3][x];
            }
    for (int J = 2; J <= NY; J++)
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++) {
                const size_t x = CIDX(i, K, J);
                CF2[x] = CF2[x] - (I * KX[i]) * CS[4][x] - (I * KZ[K]) * CS[6][x];
            }
    /* cross terms, explicit (for/les.f:302-331) */
    double *T2 = zalloc(fs);
    cplx *CT = (cplx *)T2;
    for (int pass = 0; pass < 2; pass++) {
        const double *UU = pass == 0 ? U1 : U3;
        for (int K = 0; K <= NZP - 1; K++)
            for (int i = 0; i <= NXM; i++)
                for (int J = 2; J <= NY; J++)
                    T2[RIDX(i, K, J)] = NUT[RIDX(i, K, J)] * (UU[RIDX(i, K, J)] - UU[RIDX(i, K, J - 1)]) / DY[J];
        dor_fft_xz_to_fourier(s, T2, 2, NY);
        for (int K = 0; K <= TNKZ; K++)
            for (int i = 0; i <= NXP - 1; i++)
                for (int J = 2; J <= NY; J++) {
                    const size_t x = CIDX(i, K, J);
                    CF2[x] = CF2[x] + (pass == 0 ? (I * KX[i]) : (I * KZ[K])) * CT[x];
                }
    }
    if (nu_t) memcpy(nu_t, NUT, sizeof(double) * fs);
    for (int q = 1; q <= 6; q++) free(Sij[q]);
    free(U1b); free(U3b); free(U1); free(U3); free(S1); free(TEMP); free(NUT);
    free(DELTA_YF); free(DELTA_Y); free(T2);
}

/* tkebudget_chan_les (for/save_stats.f:887-960): epsilon_sgs(J), J = 1..NY,
 * the sum over the physical points of plane J of u_i times the implicitly
 * treated eddy-viscosity term of u_i, from the velocity as SAVE_STATS_CHAN
 * takes it to physical space (for/save_stats.f:163-170) and the current
 * NU_T.  Needs LES on (NU_T allocated).  Plane 1 is NaN, as in the
 * reference: its U2 term divides 0 by DYF(0), which is never set. */
void dor_tke_les(dor_state *s, double *eps)
{
    const int NY = s->NY, NZP = s->NZP, NXM = s->NXM;
    const double *DY = s->DY, *DYF = s->DYF;
    const size_t fs = s->fsize;
    const double *NT = s->fld[DOR_NUT];
    double *U1 = zalloc(fs), *U2 = zalloc(fs), *U3 = zalloc(fs);
    memcpy(U1, s->fld[DOR_U1], sizeof(double) * fs);
    memcpy(U2, s->fld[DOR_U2], sizeof(double) * fs);
    memcpy(U3, s->fld[DOR_U3], sizeof(double) * fs);
    dor_fft_xz_to_physical(s, U1, 0, NY + 1);
    dor_fft_xz_to_physical(s, U2, 0, NY + 1);
    dor_fft_xz_to_physical(s, U3, 0, NY + 1);
#define AT(A, dj) A[RIDX(i, K, J + (dj))]
    for (int J = 1; J <= NY; J++) {
        double e = 0.0;
        for (int i = 0; i <= NXM; i++)
            for (int K = 0; K <= NZP - 1; K++) {
                const double t =
                    AT(U1, 0) * ((AT(NT, 1) * (AT(U1, 1) - AT(U1, 0)) / DY[J + 1]
                                  - AT(NT, 0) * (AT(U1, 0) - AT(U1, -1)) / DY[J]) / DYF[J])
                    + AT(U3, 0) * ((AT(NT, 1) * (AT(U3, 1) - AT(U3, 0)) / DY[J + 1]
                                    - AT(NT, 0) * (AT(U3, 0) - AT(U3, -1)) / DY[J]) / DYF[J])
                    + AT(U2, 0) * ((0.5 * (AT(NT, 0) + AT(NT, 1)) * (AT(U2, 1) - AT(U2, 0)) / DYF[J]
                                    - 0.5 * (AT(NT, 0) + AT(NT, -1)) * (AT(U2, 0) - AT(U2, -1)) / DYF[J - 1])
                                   / DY[J]);
                e = e + t;
            }
        eps[J - 1] = e;
    }
#undef AT
    free(U1); free(U2); free(U3);
}
