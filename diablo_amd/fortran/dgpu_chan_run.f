!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
! dgpu_chan_run.f -- a minimal Fortran program over the drop-in subroutines
! of diablo_gpu_bridge.f, compiled against the reference's 'header' (its
! COMMON blocks and array shapes) with this run's grid_def.
!
! It does what the reference's main program does around the channel path,
! in the same order (the time loop of for/diablo.f:66-84, RK_STEP = 1..3
! per TIME_STEP, TIME advanced after each step), with the set-up of
! INITIALIZE replaced by a run file that the caller writes:
!
!   run file (stream, native byte order):
!     int32   NSTEPS
!     real*8  NU LX LY LZ DELTA_T BETA NU_V_SCALE I_RO GRAV_X GRAV_Y GRAV_Z
!             PX0 OMEGA0 AMP_OMEGA0 U_BC_YMIN_C1 V_BC_YMIN_C1 W_BC_YMIN_C1
!             U_BC_YMAX_C1 V_BC_YMAX_C1 W_BC_YMAX_C1 TH_BC_YMIN_C1(1)
!             TH_BC_YMAX_C1(1) RI(1) PR(1) DRHODX(1) DRHODZ(1)
!     int32   F_TYPE U_BC_YMIN V_BC_YMIN W_BC_YMIN U_BC_YMAX V_BC_YMAX
!             W_BC_YMAX TH_BC_YMIN(1) TH_BC_YMAX(1)
!     real*8  GY GYF DY DYF (0:NY+1 each), KX KX2 (0:NXP), KZ KZ2 (0:TNKZ)
!     real*8  U1 U2 U3 P CTH(:,:,:,1)   Fourier state, field layout
!     real*8  S1                        a physical field (transform check)
!     real*8  R1                        a Fourier field (transform check)
!
! Sequence: FFT_XZ_TO_FOURIER of S1 and FFT_XZ_TO_PHYSICAL of R1 (planes
! 0..NY+1), REM_DIV_CHAN of the state, then NSTEPS time steps of RK_CHAN_1.
! The output file holds S1, R1, then U1 U2 U3 P CTH F1 F2 F3 CFTH after the
! steps, then the wall time of the time loop in seconds (real*8).  By
! default the bridge copies the state to the host arrays after every step;
! with DIABLO_GPU_RESIDENT=1 it stays on the GPU during the loop and is
! fetched once after it (DGPU_SYNC_TO_HOST), inside the timed region.
!
!   usage: dgpu_chan_run RUNFILE OUTFILE
!----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      PROGRAM DGPU_CHAN_RUN
      USE DGPU_BRIDGE, ONLY: DGPU_RESIDENT
      INCLUDE 'header'
      INTEGER(KIND=4) NSTEPS, IV(9)
      REAL*8 RV(26)
      CHARACTER(LEN=512) RUNF, OUTF
      INTEGER(KIND=8) C0, C1, CRATE
      REAL*8 WALL

      CALL GET_COMMAND_ARGUMENT(1, RUNF)
      CALL GET_COMMAND_ARGUMENT(2, OUTF)
      IF (N_TH .NE. 1) STOP 'dgpu_chan_run: built for N_TH = 1'

      OPEN(10, FILE=TRIM(RUNF), ACCESS='STREAM', FORM='UNFORMATTED',
     &     STATUS='OLD')
      READ(10) NSTEPS
      READ(10) RV
      READ(10) IV
      NU = RV(1)
      LX = RV(2)
      LY = RV(3)
      LZ = RV(4)
      DELTA_T = RV(5)
      BETA = RV(6)
      NU_V_SCALE = RV(7)
      I_RO = RV(8)
      GRAV_X = RV(9)
      GRAV_Y = RV(10)
      GRAV_Z = RV(11)
      PX0 = RV(12)
      OMEGA0 = RV(13)
      AMP_OMEGA0 = RV(14)
      U_BC_YMIN_C1 = RV(15)
      V_BC_YMIN_C1 = RV(16)
      W_BC_YMIN_C1 = RV(17)
      U_BC_YMAX_C1 = RV(18)
      V_BC_YMAX_C1 = RV(19)
      W_BC_YMAX_C1 = RV(20)
      TH_BC_YMIN_C1(1) = RV(21)
      TH_BC_YMAX_C1(1) = RV(22)
      RI(1) = RV(23)
      PR(1) = RV(24)
      DRHODX(1) = RV(25)
      DRHODZ(1) = RV(26)
      F_TYPE = IV(1)
      U_BC_YMIN = IV(2)
      V_BC_YMIN = IV(3)
      W_BC_YMIN = IV(4)
      U_BC_YMAX = IV(5)
      V_BC_YMAX = IV(6)
      W_BC_YMAX = IV(7)
      TH_BC_YMIN(1) = IV(8)
      TH_BC_YMAX(1) = IV(9)
      READ(10) GY, GYF, DY, DYF, KX(0:NXP), KX2(0:NXP), KZ(0:TNKZ),
     &     KZ2(0:TNKZ)
      READ(10) U1, U2, U3, P
      READ(10) CTH(:,:,:,1)
      READ(10) S1
      READ(10) R1
      CLOSE(10)

! the RKW3 table of INITIALIZE (for/diablo_io.f:116-124)
      H_BAR(1) = DELTA_T*(8.d0/15.d0)
      H_BAR(2) = DELTA_T*(2.d0/15.d0)
      H_BAR(3) = DELTA_T*(5.d0/15.d0)
      BETA_BAR(1) = 1.d0
      BETA_BAR(2) = 25.d0/8.d0
      BETA_BAR(3) = 9.d0/4.d0
      ZETA_BAR(1) = 0.d0
      ZETA_BAR(2) = -17.d0/8.d0
      ZETA_BAR(3) = -5.d0/4.d0
      TIME = 0.d0
      TIME_STEP = 0
      VARIABLE_DT = .FALSE.
      UPDATE_DT = 1
      LES = .FALSE.
      CREATE_NEW_FLOW = .TRUE.
      F1 = 0.d0
      F2 = 0.d0
      F3 = 0.d0
      CFTH = (0.d0, 0.d0)

      CALL FFT_XZ_TO_FOURIER(S1, CS1, 0, NY+1)
      CALL FFT_XZ_TO_PHYSICAL(CR1, R1, 0, NY+1)
      CALL REM_DIV_CHAN

      CALL SYSTEM_CLOCK(C0, CRATE)
      DO TIME_STEP = 1, NSTEPS
        DO RK_STEP = 1, 3
          CALL RK_CHAN_1
        END DO
        TIME = TIME + DELTA_T
      END DO
      IF (DGPU_RESIDENT) CALL DGPU_SYNC_TO_HOST
      CALL SYSTEM_CLOCK(C1)
      WALL = DBLE(C1 - C0) / DBLE(CRATE)

      OPEN(11, FILE=TRIM(OUTF), ACCESS='STREAM', FORM='UNFORMATTED',
     &     STATUS='REPLACE')
      WRITE(11) S1
      WRITE(11) R1
      WRITE(11) U1, U2, U3, P
      WRITE(11) CTH(:,:,:,1)
      WRITE(11) F1, F2, F3
      WRITE(11) CFTH(:,:,:,1)
      WRITE(11) WALL
      CLOSE(11)
      END
