! ORACLE -- test infrastructure only.  Nothing in the product uses it.
!
! ref_nest_driver.f90: runs the reference's own nest-exchange arithmetic
! (myWAveModel/WAM src/mod/wam_interface_module.f90: TOTAL_ENERGY, FEMEAN,
! MEAN_DIRECTION, INTSPEC with ROTSPEC and STRSPEC), built from the reference
! sources by oracle/Makefile.ref, on spectra given in a binary file, so that
! golden vectors for the CPU restatement can be written (tests/golden/).
! Set-up: SET_FRE_DIR (the frequency and direction grid, the tail factors).
!
! Input (stream, native endianness):
!   int32  KL ML IREF NB NPAIR NROT NSTR
!   real4  FR1
!   real4  F(KL,ML,NB)                       the spectra
!   NPAIR x (int32 I1 I2, real4 DEL12 DEL1L) INTSPEC pairs (1-based)
!   NROT  x (int32 I, real4 RTHET)           ROTSPEC cases
!   NSTR  x (int32 I, real4 GAMMA)           STRSPEC cases
! Output (stream): real4 EMEAN(NB), THQ(NB), FMEAN(NB) of BOUNDARY_OUTPUT's
!   calls (wam_boundary_module.f90:359-362); per pair FL(KL,ML), FMEAN, EMEAN,
!   THETM; per ROTSPEC and STRSPEC case F_OUT(KL,ML).
PROGRAM REF_NEST_DRIVER

USE WAM_FRE_DIR_MODULE, ONLY: SET_FRE_DIR
USE WAM_INTERFACE_MODULE, ONLY: TOTAL_ENERGY, FEMEAN, MEAN_DIRECTION, INTSPEC, ROTSPEC, STRSPEC

IMPLICIT NONE

INTEGER :: KL, ML, IREF, NB, NPAIR, NROT, NSTR, I, I1, I2, M, K
REAL    :: FR1, DEL12, DEL1L, FMEAN, EMEAN, THETM, X
REAL, ALLOCATABLE :: F(:,:,:), FB(:,:,:), EM(:), TH(:), FM(:), FL(:,:)
CHARACTER(LEN=512) :: FIN, FOUT
INTEGER :: UIN, UOUT

CALL GET_COMMAND_ARGUMENT(1, FIN)
CALL GET_COMMAND_ARGUMENT(2, FOUT)
OPEN (NEWUNIT=UIN, FILE=TRIM(FIN), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
READ (UIN) KL, ML, IREF, NB, NPAIR, NROT, NSTR
READ (UIN) FR1
ALLOCATE (F(KL,ML,NB), FB(NB,KL,ML), EM(NB), TH(NB), FM(NB), FL(KL,ML))
READ (UIN) F
CALL SET_FRE_DIR (N_DIR=KL, N_FRE=ML, FR1=FR1, IREF=IREF)
OPEN (NEWUNIT=UOUT, FILE=TRIM(FOUT), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')

DO M = 1, ML
   DO K = 1, KL
      FB(:,K,M) = F(K,M,:)
   END DO
END DO
CALL TOTAL_ENERGY (FB, EM)
CALL FEMEAN (FB, EM, FM=FM)
CALL MEAN_DIRECTION (FB, THQ=TH)
WRITE (UOUT) EM, TH, FM

DO I = 1, NPAIR
   READ (UIN) I1, I2, DEL12, DEL1L
   CALL INTSPEC (DEL12, DEL1L, F(:,:,I1), FM(I1), EM(I1), TH(I1), F(:,:,I2), FM(I2), EM(I2), TH(I2), &
&                FL, FMEAN, EMEAN, THETM)
   WRITE (UOUT) FL, FMEAN, EMEAN, THETM
END DO
DO I = 1, NROT
   READ (UIN) I1, X
   CALL ROTSPEC (F(:,:,I1), FL, X)
   WRITE (UOUT) FL
END DO
DO I = 1, NSTR
   READ (UIN) I1, X
   CALL STRSPEC (F(:,:,I1), FL, X)
   WRITE (UOUT) FL
END DO
CLOSE (UIN)
CLOSE (UOUT)

END PROGRAM REF_NEST_DRIVER
