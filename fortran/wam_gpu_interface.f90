! ---------------------------------------------------------------------------- !
!  WAM_GPU_INTERFACE - Fortran binding of the MI355X hot path (wam_gpu.h).     !
!                                                                              !
!  ISO_C_BINDING interfaces to the C entry points, and drop-in shims with the  !
!  argument lists of the reference's IMPLSCH (wam_source_module.f90:374,       !
!  422-431) and PROPAGS (wam_propagation_module.f90:635,687).  A WAM host      !
!  replaces                                                                    !
!     CALL PROPAGS (FL3)                                 (wamodel.f90:332)     !
!     CALL IMPLSCH (FL3, U10, UDIR, TAUW, USTAR, Z0, ROAIRN, WSTAR,            !
!                   DEPTH(NIJS:NIJL), INDEP(NIJS:NIJL))   (wamodel.f90:354)    !
!  by PROPAGS_GPU / IMPLSCH_GPU after one WAM_GPU_SETUP and one                !
!  WAM_GPU_PREPARE_PROPAGATION (initmdl.f90:204, wamodel.f90:326).  A non-zero !
!  status from the library stops the program, as ABORT1 does in the model.    !
! ---------------------------------------------------------------------------- !
MODULE WAM_GPU_INTERFACE

USE, INTRINSIC :: ISO_C_BINDING
IMPLICIT NONE
PRIVATE

TYPE, BIND(C), PUBLIC :: WAM_CONFIG
   INTEGER(C_INT32_T) :: KL, ML
   REAL(C_DOUBLE)     :: FR1
   INTEGER(C_INT32_T) :: IREF, IPHYS, SHALLOW, SPHERICAL, REFR_D, REFR_C, IPER
   INTEGER(C_INT32_T) :: ISNONLIN, IDELPRO, IDELT, NDEPTH
   REAL(C_DOUBLE)     :: DEPTHA, DEPTHD, BETAMAX, ZALP, ALPHA, CDFAC
   INTEGER(C_INT32_T) :: NX, NY
   INTEGER(C_INT64_T) :: AMOSOP, AMONOP, AMOWEP, AMOEAP, XDELLA, XDELLO
END TYPE WAM_CONFIG

INTERFACE
   INTEGER(C_INT) FUNCTION c_wam_gpu_init(ctx, cfg, mask) BIND(C, NAME='wam_gpu_init')
      IMPORT :: C_INT, C_PTR, WAM_CONFIG
      TYPE(C_PTR), INTENT(OUT) :: ctx
      TYPE(WAM_CONFIG), INTENT(IN) :: cfg
      TYPE(C_PTR), VALUE :: mask
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_prepare_propagation(ctx, depth, u, v) &
&     BIND(C, NAME='wam_gpu_prepare_propagation')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, depth, u, v
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_prepare_propagation_dev(ctx, d_depth, d_u, d_v) &
&     BIND(C, NAME='wam_gpu_prepare_propagation_dev')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ctx, d_depth, d_u, d_v
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_propags(ctx, nijs, nijl, kl, ml, f3) BIND(C, NAME='wam_gpu_propags')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, f3
      INTEGER(C_INT32_T), VALUE :: nijs, nijl, kl, ml
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_implsch(ctx, nsea, kl, ml, fl3, u10, udir, tauw, ustar, z0, roairn, &
&     wstar, depth, indep, lcflx, phiaw, fluxes) BIND(C, NAME='wam_gpu_implsch')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, fl3, u10, udir, tauw, ustar, z0, roairn, wstar, depth, indep, phiaw, fluxes
      INTEGER(C_INT32_T), VALUE :: nsea, kl, ml, lcflx
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_output_parameters(ctx, nsea, kl, ml, fl3, u10, udir, ustar, tauw, z0, &
&     orient, out) BIND(C, NAME='wam_gpu_output_parameters')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, fl3, u10, udir, ustar, tauw, z0, out
      INTEGER(C_INT32_T), VALUE :: nsea, kl, ml, orient
   END FUNCTION
   FUNCTION c_wam_gpu_last_error(ctx) BIND(C, NAME='wam_gpu_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
      TYPE(C_PTR) :: c_wam_gpu_last_error
   END FUNCTION
   REAL(C_DOUBLE) FUNCTION c_wam_gpu_get_scalar(ctx, name) BIND(C, NAME='wam_gpu_get_scalar')
      IMPORT :: C_DOUBLE, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: name
   END FUNCTION
   SUBROUTINE c_wam_gpu_finalize(ctx) BIND(C, NAME='wam_gpu_finalize')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
   END SUBROUTINE
   FUNCTION c_strlen(s) BIND(C, NAME='strlen')
      IMPORT :: C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: s
      INTEGER(C_SIZE_T) :: c_strlen
   END FUNCTION
END INTERFACE

TYPE(C_PTR), SAVE :: CTX = C_NULL_PTR
INTEGER, SAVE, PUBLIC :: NSEA_GPU = 0

PUBLIC :: WAM_GPU_SETUP, WAM_GPU_PREPARE_PROPAGATION, WAM_GPU_PREPARE_PROPAGATION_DEV, PROPAGS_GPU, &
&         IMPLSCH_GPU, OUTPUT_PARAMETERS_GPU, WAM_GPU_END

CONTAINS

SUBROUTINE CHECK (RC, WHERE)
   INTEGER(C_INT), INTENT(IN) :: RC
   CHARACTER(LEN=*), INTENT(IN) :: WHERE
   CHARACTER(KIND=C_CHAR), POINTER :: MSG(:)
   TYPE(C_PTR) :: P
   INTEGER :: N
   IF (RC .EQ. 0) RETURN
   P = c_wam_gpu_last_error(CTX)
   N = INT(c_strlen(P))
   CALL C_F_POINTER(P, MSG, [N])
   WRITE (*,*) ' *** WAM GPU ERROR in ', WHERE, ': code ', RC, ' : ', MSG(1:N)
   ERROR STOP 1
END SUBROUTINE CHECK

SUBROUTINE WAM_GPU_SETUP (CFG, MASK)
   TYPE(WAM_CONFIG), INTENT(IN) :: CFG
   INTEGER(C_INT8_T), INTENT(IN), TARGET, CONTIGUOUS :: MASK(:,:)     !! (NX, NY), non-zero = sea
   CALL CHECK(c_wam_gpu_init(CTX, CFG, C_LOC(MASK)), 'WAM_GPU_SETUP')
   NSEA_GPU = NINT(c_wam_gpu_get_scalar(CTX, 'NSEA'//C_NULL_CHAR))
END SUBROUTINE WAM_GPU_SETUP

SUBROUTINE WAM_GPU_PREPARE_PROPAGATION (DEPTH, U, V)
   REAL(C_FLOAT), INTENT(IN), TARGET, CONTIGUOUS :: DEPTH(:)
   REAL(C_FLOAT), INTENT(IN), TARGET, CONTIGUOUS, OPTIONAL :: U(:), V(:)
   TYPE(C_PTR) :: PU, PV
   PU = C_NULL_PTR; PV = C_NULL_PTR
   IF (PRESENT(U)) PU = C_LOC(U)
   IF (PRESENT(V)) PV = C_LOC(V)
   CALL CHECK(c_wam_gpu_prepare_propagation(CTX, C_LOC(DEPTH), PU, PV), 'WAM_GPU_PREPARE_PROPAGATION')
END SUBROUTINE WAM_GPU_PREPARE_PROPAGATION

!  PREPARE_PROPAGATION from depth and currents already in device memory (NSEA
!  each; D_U, D_V may be C_NULL_PTR without current refraction), e.g. from a
!  coupled model that keeps them on the GPU.
SUBROUTINE WAM_GPU_PREPARE_PROPAGATION_DEV (D_DEPTH, D_U, D_V)
   TYPE(C_PTR), INTENT(IN) :: D_DEPTH, D_U, D_V
   CALL CHECK(c_wam_gpu_prepare_propagation_dev(CTX, D_DEPTH, D_U, D_V), 'WAM_GPU_PREPARE_PROPAGATION_DEV')
END SUBROUTINE WAM_GPU_PREPARE_PROPAGATION_DEV

!  Drop-in for PROPAGS (F3) -- F3(NIJS:NIJL,KL,ML).
SUBROUTINE PROPAGS_GPU (F3)
   REAL(C_FLOAT), INTENT(INOUT), TARGET, CONTIGUOUS :: F3(:,:,:)
   CALL CHECK(c_wam_gpu_propags(CTX, 1_C_INT32_T, INT(SIZE(F3,1), C_INT32_T), INT(SIZE(F3,2), C_INT32_T), &
&             INT(SIZE(F3,3), C_INT32_T), C_LOC(F3)), 'PROPAGS_GPU')
END SUBROUTINE PROPAGS_GPU

!  Drop-in for IMPLSCH (FL3, U10, UDIR, TAUW, USTAR, Z0, ROAIRN, WSTAR, DEPTH, INDEP).
SUBROUTINE IMPLSCH_GPU (FL3, U10, UDIR, TAUW, USTAR, Z0, ROAIRN, WSTAR, DEPTH, INDEP)
   REAL(C_FLOAT),      INTENT(INOUT), TARGET, CONTIGUOUS :: FL3(:,:,:)
   REAL(C_FLOAT),      INTENT(IN),    TARGET, CONTIGUOUS :: U10(:), UDIR(:), ROAIRN(:), WSTAR(:)
   REAL(C_FLOAT),      INTENT(INOUT), TARGET, CONTIGUOUS :: TAUW(:), Z0(:)
   REAL(C_FLOAT),      INTENT(OUT),   TARGET, CONTIGUOUS :: USTAR(:)
   REAL(C_FLOAT),      INTENT(IN),    TARGET, CONTIGUOUS :: DEPTH(:)
   INTEGER(C_INT32_T), INTENT(IN),    TARGET, CONTIGUOUS :: INDEP(:)
   CALL CHECK(c_wam_gpu_implsch(CTX, INT(SIZE(FL3,1), C_INT32_T), INT(SIZE(FL3,2), C_INT32_T), &
&             INT(SIZE(FL3,3), C_INT32_T), C_LOC(FL3), C_LOC(U10), C_LOC(UDIR), C_LOC(TAUW), C_LOC(USTAR), &
&             C_LOC(Z0), C_LOC(ROAIRN), C_LOC(WSTAR), C_LOC(DEPTH), C_LOC(INDEP), 0_C_INT32_T, C_NULL_PTR, &
&             C_NULL_PTR), 'IMPLSCH_GPU')
END SUBROUTINE IMPLSCH_GPU

!  Integrated parameters of COMPUTE_OUTPUT_PARAMETER (wam_output_module.f90:481-601):
!  BOUT(IJ,P) for the 17 parameters 1-16 and 32, in that order.  ORIENT = 1 gives
!  directions "coming from", 0 "going to" (orientation_of_directions).
SUBROUTINE OUTPUT_PARAMETERS_GPU (FL3, U10, UDIR, USTAR, TAUW, Z0, ORIENT, BOUT)
   REAL(C_FLOAT), INTENT(IN),  TARGET, CONTIGUOUS :: FL3(:,:,:)
   REAL(C_FLOAT), INTENT(IN),  TARGET, CONTIGUOUS :: U10(:), UDIR(:), USTAR(:), TAUW(:), Z0(:)
   INTEGER,       INTENT(IN)                      :: ORIENT
   REAL(C_FLOAT), INTENT(OUT), TARGET, CONTIGUOUS :: BOUT(:,:)
   IF (SIZE(BOUT,1) .NE. SIZE(FL3,1) .OR. SIZE(BOUT,2) .LT. 17) THEN
      WRITE (*,*) ' *** WAM GPU ERROR in OUTPUT_PARAMETERS_GPU: BOUT must be (NSEA, 17)'
      ERROR STOP 1
   END IF
   CALL CHECK(c_wam_gpu_output_parameters(CTX, INT(SIZE(FL3,1), C_INT32_T), INT(SIZE(FL3,2), C_INT32_T), &
&             INT(SIZE(FL3,3), C_INT32_T), C_LOC(FL3), C_LOC(U10), C_LOC(UDIR), C_LOC(USTAR), C_LOC(TAUW), &
&             C_LOC(Z0), INT(ORIENT, C_INT32_T), C_LOC(BOUT)), 'OUTPUT_PARAMETERS_GPU')
END SUBROUTINE OUTPUT_PARAMETERS_GPU

SUBROUTINE WAM_GPU_END
   IF (C_ASSOCIATED(CTX)) CALL c_wam_gpu_finalize(CTX)
   CTX = C_NULL_PTR
END SUBROUTINE WAM_GPU_END

END MODULE WAM_GPU_INTERFACE
