!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! diablo_gpu_bridge.f -- drop-in replacements of DIABLO 2.0's channel-path
! subroutines that forward to libdiablo_gpu.so (include/diablo_gpu.h) through
! ISO_C_BINDING.
!
! Subroutines (same names and arguments as the reference):
!   RK_CHAN_1                         (for/channel.f:62-1073)
!   REM_DIV_CHAN                      (for/channel.f:1103-1214)
!   POISSON_P_CHAN                    (for/channel.f:1217-1453)
!   FFT_XZ_TO_FOURIER(V,VV,JMIN,JMAX) (for/mpi.f:1530-1609)
!   FFT_XZ_TO_PHYSICAL(VV,V,JMIN,JMAX)(for/mpi.f:1613-1674)
! plus DGPU_SYNC_TO_HOST / DGPU_SYNC_TO_DEVICE for drivers that keep the
! state on the GPU across time steps.
!
! Compile with the reference's own include files on the include path
! ('header' and what it includes) and link against libdiablo_gpu.so; the five
! reference definitions are then left out of the link: channel.f's
! RK_CHAN_1, REM_DIV_CHAN and POISSON_P_CHAN and mpi.f's FFT_XZ_* go behind
! '#ifndef DIABLO_GPU' guards in the reference build.
!
! State coherence.  The Fourier state lives on the GPU during a time step.
! By default RK_CHAN_1 uploads CU1-3, CP, CTH and CF at RK_STEP = 1 and writes
! CU1-3, CP, CTH, CF back at RK_STEP = 3, so every Fortran routine that runs
! between time steps (SAVE_STATS, SAVE_FLOW, END_RUN, user hooks) sees the
! same arrays as with the CPU path.  Setting the environment variable
! DIABLO_GPU_RESIDENT=1 keeps the state on the GPU across steps; the driver
! must then call DGPU_SYNC_TO_HOST before any routine that reads the state and
! DGPU_SYNC_TO_DEVICE after any routine that changes it.
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|

      MODULE DGPU_BRIDGE
      USE ISO_C_BINDING
      IMPLICIT NONE

      INTEGER, PARAMETER :: DGPU_MAX_TH = 1
      INTEGER, PARAMETER :: DGPU_CU1 = 0, DGPU_CU2 = 1, DGPU_CU3 = 2
      INTEGER, PARAMETER :: DGPU_CP = 3, DGPU_CF1 = 4, DGPU_CF2 = 5
      INTEGER, PARAMETER :: DGPU_CF3 = 6, DGPU_CTH0 = 16
      INTEGER, PARAMETER :: DGPU_CFTH0 = 24

! Mirror of struct dgpu_params (include/diablo_gpu.h)
      TYPE, BIND(C) :: DGPU_PARAMS
        INTEGER(C_INT32_T) :: NX, NY, NZ, N_TH
        INTEGER(C_INT32_T) :: PRECISION, F_TYPE
        INTEGER(C_INT32_T) :: U_BC_YMIN, V_BC_YMIN, W_BC_YMIN
        INTEGER(C_INT32_T) :: U_BC_YMAX, V_BC_YMAX, W_BC_YMAX
        INTEGER(C_INT32_T) :: TH_BC_YMIN(DGPU_MAX_TH)
        INTEGER(C_INT32_T) :: TH_BC_YMAX(DGPU_MAX_TH)
        REAL(C_DOUBLE) :: NU, BETA, LX, LY, LZ, NU_V_SCALE
        REAL(C_DOUBLE) :: I_RO, GRAV_X, GRAV_Y, GRAV_Z
        REAL(C_DOUBLE) :: PX0, OMEGA0, AMP_OMEGA0
        REAL(C_DOUBLE) :: U_BC_YMIN_C1, V_BC_YMIN_C1, W_BC_YMIN_C1
        REAL(C_DOUBLE) :: U_BC_YMAX_C1, V_BC_YMAX_C1, W_BC_YMAX_C1
        REAL(C_DOUBLE) :: TH_BC_YMIN_C1(DGPU_MAX_TH)
        REAL(C_DOUBLE) :: TH_BC_YMAX_C1(DGPU_MAX_TH)
        REAL(C_DOUBLE) :: RI(DGPU_MAX_TH), PR(DGPU_MAX_TH)
        REAL(C_DOUBLE) :: DRHODX(DGPU_MAX_TH), DRHODZ(DGPU_MAX_TH)
        TYPE(C_PTR) :: GY, GYF, DY, DYF, KX, KZ, KX2, KZ2
      END TYPE DGPU_PARAMS

      INTERFACE
        INTEGER(C_INT32_T) FUNCTION dgpu_create(P, CTX)
     &      BIND(C, NAME='dgpu_create')
          IMPORT :: DGPU_PARAMS, C_PTR, C_INT32_T
          TYPE(DGPU_PARAMS), INTENT(IN) :: P
          TYPE(C_PTR), INTENT(OUT) :: CTX
        END FUNCTION
! Field arrays are passed as assumed-type (void *) so that the REAL*8 fields
! and the COMPLEX*16 scalar arrays CTH, CFTH (not aliased with a real array,
! for/header:258-260) go through the same entry points.
        INTEGER(C_INT32_T) FUNCTION dgpu_set_field(CTX, F, A)
     &      BIND(C, NAME='dgpu_set_field')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
          INTEGER(C_INT32_T), VALUE :: F
          TYPE(*), DIMENSION(*), INTENT(IN) :: A
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_get_field(CTX, F, A)
     &      BIND(C, NAME='dgpu_get_field')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
          INTEGER(C_INT32_T), VALUE :: F
          TYPE(*), DIMENSION(*), INTENT(INOUT) :: A
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_rk_chan_1(CTX, K, T, H, B, Z)
     &      BIND(C, NAME='dgpu_rk_chan_1')
          IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
          TYPE(C_PTR), VALUE :: CTX
          INTEGER(C_INT32_T), VALUE :: K
          REAL(C_DOUBLE), VALUE :: T
          REAL(C_DOUBLE), INTENT(IN) :: H(3), B(3), Z(3)
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_rem_div_chan(CTX)
     &      BIND(C, NAME='dgpu_rem_div_chan')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_poisson_p_chan(CTX)
     &      BIND(C, NAME='dgpu_poisson_p_chan')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_fft_xz_to_fourier(CTX, V, J0,
     &      J1) BIND(C, NAME='dgpu_fft_xz_to_fourier')
          IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
          TYPE(C_PTR), VALUE :: CTX
          REAL(C_DOUBLE), INTENT(INOUT) :: V(*)
          INTEGER(C_INT32_T), VALUE :: J0, J1
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_fft_xz_to_physical(CTX, V, J0,
     &      J1) BIND(C, NAME='dgpu_fft_xz_to_physical')
          IMPORT :: C_PTR, C_INT32_T, C_DOUBLE_COMPLEX
          TYPE(C_PTR), VALUE :: CTX
          COMPLEX(C_DOUBLE_COMPLEX), INTENT(INOUT) :: V(*)
          INTEGER(C_INT32_T), VALUE :: J0, J1
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_synchronize(CTX)
     &      BIND(C, NAME='dgpu_synchronize')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_stats_chan(CTX, P, U)
     &      BIND(C, NAME='dgpu_stats_chan')
          IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
          TYPE(C_PTR), VALUE :: CTX
          REAL(C_DOUBLE), INTENT(OUT) :: P(*)
          REAL(C_DOUBLE), INTENT(OUT) :: U
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_tke_chan(CTX, E)
     &      BIND(C, NAME='dgpu_tke_chan')
          IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
          TYPE(C_PTR), VALUE :: CTX
          REAL(C_DOUBLE), INTENT(OUT) :: E(*)
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_set_les(CTX, L, M)
     &      BIND(C, NAME='dgpu_set_les')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
          INTEGER(C_INT32_T), VALUE :: L, M
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_courant_request(CTX)
     &      BIND(C, NAME='dgpu_courant_request')
          IMPORT :: C_PTR, C_INT32_T
          TYPE(C_PTR), VALUE :: CTX
        END FUNCTION
        INTEGER(C_INT32_T) FUNCTION dgpu_courant(CTX, CFL, DT, H)
     &      BIND(C, NAME='dgpu_courant')
          IMPORT :: C_PTR, C_INT32_T, C_DOUBLE
          TYPE(C_PTR), VALUE :: CTX
          REAL(C_DOUBLE), VALUE :: CFL
          REAL(C_DOUBLE), INTENT(OUT) :: DT
          REAL(C_DOUBLE), INTENT(OUT) :: H(3)
        END FUNCTION
      END INTERFACE

      TYPE(C_PTR), SAVE :: DGPU_CTX = C_NULL_PTR
      LOGICAL, SAVE :: DGPU_ON_DEVICE = .FALSE.
      LOGICAL, SAVE :: DGPU_RESIDENT = .FALSE.

      CONTAINS

      SUBROUTINE DGPU_CHECK(RC, WHAT)
      INTEGER(C_INT32_T), INTENT(IN) :: RC
      CHARACTER(*), INTENT(IN) :: WHAT
! The reference reports unrecoverable errors with STOP (e.g. for/channel.f:2322);
! ERROR STOP also gives the shell a non-zero exit status
      IF (RC .NE. 0) THEN
        WRITE(*,*) 'libdiablo_gpu: ', WHAT, ' failed, code ', RC
        ERROR STOP 'DIABLO GPU path error'
      END IF
      END SUBROUTINE DGPU_CHECK

      END MODULE DGPU_BRIDGE

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE DGPU_ENSURE_CONTEXT
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! Creates the GPU context from the COMMON blocks on first use: sizes from
! grid_def, run parameters from /INPUT/ etc., and the grid and wavenumber
! arrays exactly as INITIALIZE computed them (passed verbatim).
      USE DGPU_BRIDGE
      INCLUDE 'header'
      TYPE(DGPU_PARAMS) :: PRM
      REAL(C_DOUBLE), TARGET, ALLOCATABLE :: GYC(:), GYFC(:), DYC(:),
     &     DYFC(:), KXC(:), KZC(:), KX2C(:), KZ2C(:)
      INTEGER N
      CHARACTER(LEN=8) :: ENVV
      INTEGER ENVLEN, ENVSTAT

      IF (C_ASSOCIATED(DGPU_CTX)) RETURN
      IF (N_TH .GT. DGPU_MAX_TH) ERROR STOP
     &  'DIABLO GPU path: N_TH > 1 is not supported'
      PRM%NX = NX
      PRM%NY = NY
      PRM%NZ = NZ
      PRM%N_TH = N_TH
! DIABLO_GPU_PRECISION: 0 FP64 (default), 1 mixed (FP32 transforms and
! solves, FP64 low modes), 2 mixed with the state and solves in FP64
      PRM%PRECISION = 0
      CALL GET_ENVIRONMENT_VARIABLE('DIABLO_GPU_PRECISION', ENVV,
     &     ENVLEN, ENVSTAT)
      IF (ENVSTAT .EQ. 0 .AND. ENVLEN .GE. 1) THEN
        IF (ENVV(1:1) .EQ. '1') PRM%PRECISION = 1
        IF (ENVV(1:1) .EQ. '2') PRM%PRECISION = 2
      END IF
      PRM%F_TYPE = F_TYPE
      PRM%U_BC_YMIN = U_BC_YMIN
      PRM%V_BC_YMIN = V_BC_YMIN
      PRM%W_BC_YMIN = W_BC_YMIN
      PRM%U_BC_YMAX = U_BC_YMAX
      PRM%V_BC_YMAX = V_BC_YMAX
      PRM%W_BC_YMAX = W_BC_YMAX
      PRM%TH_BC_YMIN = 0
      PRM%TH_BC_YMAX = 0
      PRM%TH_BC_YMIN_C1 = 0.d0
      PRM%TH_BC_YMAX_C1 = 0.d0
      PRM%RI = 0.d0
      PRM%PR = 1.d0
      PRM%DRHODX = 0.d0
      PRM%DRHODZ = 0.d0
      DO N = 1, N_TH
        PRM%TH_BC_YMIN(N) = TH_BC_YMIN(N)
        PRM%TH_BC_YMAX(N) = TH_BC_YMAX(N)
        PRM%TH_BC_YMIN_C1(N) = TH_BC_YMIN_C1(N)
        PRM%TH_BC_YMAX_C1(N) = TH_BC_YMAX_C1(N)
        PRM%RI(N) = RI(N)
        PRM%PR(N) = PR(N)
        PRM%DRHODX(N) = DRHODX(N)
        PRM%DRHODZ(N) = DRHODZ(N)
      END DO
      PRM%NU = NU
      PRM%BETA = BETA
      PRM%LX = LX
      PRM%LY = LY
      PRM%LZ = LZ
      PRM%NU_V_SCALE = NU_V_SCALE
      PRM%I_RO = I_RO
      PRM%GRAV_X = GRAV_X
      PRM%GRAV_Y = GRAV_Y
      PRM%GRAV_Z = GRAV_Z
      PRM%PX0 = PX0
      PRM%OMEGA0 = OMEGA0
      PRM%AMP_OMEGA0 = AMP_OMEGA0
      PRM%U_BC_YMIN_C1 = U_BC_YMIN_C1
      PRM%V_BC_YMIN_C1 = V_BC_YMIN_C1
      PRM%W_BC_YMIN_C1 = W_BC_YMIN_C1
      PRM%U_BC_YMAX_C1 = U_BC_YMAX_C1
      PRM%V_BC_YMAX_C1 = V_BC_YMAX_C1
      PRM%W_BC_YMAX_C1 = W_BC_YMAX_C1
! grid and wavenumbers exactly as INITIALIZE left them (dgpu_create copies)
      ALLOCATE(GYC(NY+1), GYFC(0:NY+1), DYC(0:NY+1), DYFC(0:NY+1))
      ALLOCATE(KXC(0:NXP), KZC(0:TNKZ), KX2C(0:NXP), KZ2C(0:TNKZ))
      GYC = GY(1:NY+1)
      GYFC = GYF(0:NY+1)
      DYC = DY(0:NY+1)
      DYFC = DYF(0:NY+1)
      KXC = KX(0:NXP)
      KZC = KZ(0:TNKZ)
      KX2C = KX2(0:NXP)
      KZ2C = KZ2(0:TNKZ)
      PRM%GY = C_LOC(GYC)
      PRM%GYF = C_LOC(GYFC)
      PRM%DY = C_LOC(DYC)
      PRM%DYF = C_LOC(DYFC)
      PRM%KX = C_LOC(KXC)
      PRM%KZ = C_LOC(KZC)
      PRM%KX2 = C_LOC(KX2C)
      PRM%KZ2 = C_LOC(KZ2C)
      CALL DGPU_CHECK(dgpu_create(PRM, DGPU_CTX), 'dgpu_create')
      DEALLOCATE(GYC, GYFC, DYC, DYFC, KXC, KZC, KX2C, KZ2C)
      CALL GET_ENVIRONMENT_VARIABLE('DIABLO_GPU_RESIDENT', ENVV, ENVLEN,
     &     ENVSTAT)
      DGPU_RESIDENT = (ENVSTAT .EQ. 0 .AND. ENVLEN .GE. 1 .AND.
     &     ENVV(1:1) .EQ. '1')
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE DGPU_SYNC_TO_DEVICE
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! Host COMMON state -> GPU (CU1-3, CP, CTH, CF1-3, CFTH).
      USE DGPU_BRIDGE
      INCLUDE 'header'
      INTEGER N
      CALL DGPU_ENSURE_CONTEXT
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU1, U1),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU2, U2),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU3, U3),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CP, P), 'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CF1, F1),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CF2, F2),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CF3, F3),'set')
      DO N = 1, N_TH
        CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CTH0 + N - 1,
     &       CTH(0,0,0,N)), 'set')
        CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CFTH0 + N - 1,
     &       CFTH(0,0,0,N)), 'set')
      END DO
      DGPU_ON_DEVICE = .TRUE.
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE DGPU_SYNC_TO_HOST
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! GPU -> host COMMON state (CU1-3, CP, CTH, CF1-3, CFTH).
      USE DGPU_BRIDGE
      INCLUDE 'header'
      INTEGER N
      IF (.NOT. C_ASSOCIATED(DGPU_CTX)) RETURN
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CU1, U1),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CU2, U2),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CU3, U3),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CP, P), 'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CF1, F1),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CF2, F2),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CF3, F3),'get')
      DO N = 1, N_TH
        CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CTH0 + N - 1,
     &       CTH(0,0,0,N)), 'get')
        CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CFTH0 + N - 1,
     &       CFTH(0,0,0,N)), 'get')
      END DO
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE RK_CHAN_1
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! One RK3/CN substep on the GPU (for/channel.f:62-1073).  Under VARIABLE_DT
! the device takes COURANT's velocity maxima inside RK_STEP 3 and DELTA_T and
! H_BAR are updated after it, for the next step (for/channel.f:1037-1040).
! With LES the eddy viscosity NU_T stays on the device.
      USE DGPU_BRIDGE
      INCLUDE 'header'
      LOGICAL DO_CFL
      CALL DGPU_ENSURE_CONTEXT
      IF (.NOT. DGPU_ON_DEVICE) CALL DGPU_SYNC_TO_DEVICE
      DO_CFL = VARIABLE_DT .AND. (RK_STEP .EQ. 3) .AND.
     &     (MOD(TIME_STEP, UPDATE_DT) .EQ. 0)
      IF (DO_CFL) CALL DGPU_CHECK(dgpu_courant_request(DGPU_CTX),
     &     'dgpu_courant_request')
! LES: the eddy viscosity in the CN terms; les_chan once past the spin-up
! of a new flow (for/channel.f:303)
      IF (LES) CALL DGPU_CHECK(dgpu_set_les(DGPU_CTX, 1_C_INT32_T,
     &     MERGE(1_C_INT32_T, 0_C_INT32_T,
     &     (.NOT.CREATE_NEW_FLOW).OR.(TIME_STEP.GT.100))),
     &     'dgpu_set_les')
      CALL DGPU_CHECK(dgpu_rk_chan_1(DGPU_CTX, INT(RK_STEP, C_INT32_T),
     &     REAL(TIME, C_DOUBLE), H_BAR, BETA_BAR, ZETA_BAR),
     &     'dgpu_rk_chan_1')
      IF (DO_CFL) CALL DGPU_CHECK(dgpu_courant(DGPU_CTX,
     &     REAL(CFL, C_DOUBLE), DELTA_T, H_BAR), 'dgpu_courant')
      IF (RK_STEP .EQ. 3 .AND. .NOT. DGPU_RESIDENT) THEN
        CALL DGPU_SYNC_TO_HOST
        DGPU_ON_DEVICE = .FALSE.
      END IF
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE REM_DIV_CHAN
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! Projection of the host arrays CU1-3 (for/channel.f:1103-1214); called by
! CREATE_FLOW_CHAN (for/set_ics.f:168).  RK_CHAN_1 does its own projection on
! the GPU and does not call this.
      USE DGPU_BRIDGE
      INCLUDE 'header'
      CALL DGPU_ENSURE_CONTEXT
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU1, U1),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU2, U2),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU3, U3),'set')
      CALL DGPU_CHECK(dgpu_rem_div_chan(DGPU_CTX), 'dgpu_rem_div_chan')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CU1, U1),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CU2, U2),'get')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CU3, U3),'get')
      DGPU_ON_DEVICE = .FALSE.
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE POISSON_P_CHAN
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! CP from CU1-3 and CTH (for/channel.f:1217-1453); called by ReadHDF5 when a
! restart file has no pressure (for/hdf5.f:536).
      USE DGPU_BRIDGE
      INCLUDE 'header'
      INTEGER N
      CALL DGPU_ENSURE_CONTEXT
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU1, U1),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU2, U2),'set')
      CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CU3, U3),'set')
      DO N = 1, N_TH
        CALL DGPU_CHECK(dgpu_set_field(DGPU_CTX, DGPU_CTH0 + N - 1,
     &       CTH(0,0,0,N)), 'set')
      END DO
      CALL DGPU_CHECK(dgpu_poisson_p_chan(DGPU_CTX),
     &     'dgpu_poisson_p_chan')
      CALL DGPU_CHECK(dgpu_get_field(DGPU_CTX, DGPU_CP, P), 'get')
      DGPU_ON_DEVICE = .FALSE.
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE FFT_XZ_TO_FOURIER(V, VV, JMIN, JMAX)
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! In-place x-z transform to Fourier space of planes JMIN..JMAX
! (for/mpi.f:1530-1609).  V and VV are the same array in every reference call.
      USE DGPU_BRIDGE
      INCLUDE 'header'
      REAL*8 V(0:NX+1, 0:NZP+1, 0:NY+1)
      COMPLEX*16 VV(0:NXP, 0:NZ+1, 0:NY+1)
      INTEGER JMIN, JMAX
      CALL DGPU_ENSURE_CONTEXT
      CALL DGPU_CHECK(dgpu_fft_xz_to_fourier(DGPU_CTX, V,
     &     INT(JMIN, C_INT32_T), INT(JMAX, C_INT32_T)),
     &     'dgpu_fft_xz_to_fourier')
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE FFT_XZ_TO_PHYSICAL(VV, V, JMIN, JMAX)
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! In-place x-z transform to physical space of planes JMIN..JMAX
! (for/mpi.f:1613-1674).
      USE DGPU_BRIDGE
      INCLUDE 'header'
      COMPLEX*16 VV(0:NXP, 0:NZ+1, 0:NY+1)
      REAL*8 V(0:NX+1, 0:NZP+1, 0:NY+1)
      INTEGER JMIN, JMAX
      CALL DGPU_ENSURE_CONTEXT
      CALL DGPU_CHECK(dgpu_fft_xz_to_physical(DGPU_CTX, VV,
     &     INT(JMIN, C_INT32_T), INT(JMAX, C_INT32_T)),
     &     'dgpu_fft_xz_to_physical')
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE DGPU_SAVE_STATS(PROF, UB)
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! The statistics block of SAVE_STATS_CHAN (for/save_stats.f:126-347 and its
! scalar part, 420-522) computed on the GPU from its own state in FP64:
! PROF(J, Q) for J = 1..NY, Q = 1..21 in the order of DGPU_ST_* of
! include/diablo_gpu.h, and UB = UBULK.  The host state is uploaded first
! unless the GPU copy is current.
      USE DGPU_BRIDGE
      INCLUDE 'header'
      REAL*8 PROF(NY, 21), UB
      CALL DGPU_ENSURE_CONTEXT
      IF (.NOT. DGPU_ON_DEVICE) CALL DGPU_SYNC_TO_DEVICE
      CALL DGPU_CHECK(dgpu_stats_chan(DGPU_CTX, PROF, UB),
     &     'dgpu_stats_chan')
      RETURN
      END

!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE DGPU_TKE_EPSILON(TKEEPS)
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! tkebudget_chan's epsilon (for/save_stats.f:965-1159) on the GPU in FP64:
! TKEEPS(J), J = 1..NY, as tke.h5 / tke.txt hold it.  The host state is
! uploaded first unless the GPU copy is current.
      USE DGPU_BRIDGE
      INCLUDE 'header'
      REAL*8 TKEEPS(NY)
      CALL DGPU_ENSURE_CONTEXT
      IF (.NOT. DGPU_ON_DEVICE) CALL DGPU_SYNC_TO_DEVICE
      CALL DGPU_CHECK(dgpu_tke_chan(DGPU_CTX, TKEEPS),
     &     'dgpu_tke_chan')
      RETURN
      END
