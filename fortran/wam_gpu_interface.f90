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
   INTEGER(C_INT) FUNCTION c_wam_gpu_boundary_output(ctx, nsea, kl, ml, fl3, nb, ijarc, fbc, mean) &
&     BIND(C, NAME='wam_gpu_boundary_output')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx, fl3, ijarc, fbc, mean
      INTEGER(C_INT32_T), VALUE :: nsea, kl, ml, nb
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_boundary_input(ctx, nsea, kl, ml, fl3, nbinp, f1, m1, f2, m2, which, &
&     del12, del1l, nbounf, ijarf, ibfl, ibfr, bfw, fi, mi) BIND(C, NAME='wam_gpu_boundary_input')
      IMPORT :: C_INT, C_INT32_T, C_PTR, C_FLOAT
      TYPE(C_PTR), VALUE :: ctx, fl3, f1, m1, f2, m2, ijarf, ibfl, ibfr, bfw, fi, mi
      INTEGER(C_INT32_T), VALUE :: nsea, kl, ml, nbinp, which, nbounf
      REAL(C_FLOAT), VALUE :: del12, del1l
   END FUNCTION
   FUNCTION c_wam_gpu_last_error(ctx) BIND(C, NAME='wam_gpu_last_error')
      IMPORT :: C_PTR
      TYPE(C_PTR), VALUE :: ctx
      TYPE(C_PTR) :: c_wam_gpu_last_error
   END FUNCTION
   INTEGER(C_INT64_T) FUNCTION c_wam_gpu_get_table(ctx, name, out, n) BIND(C, NAME='wam_gpu_get_table')
      IMPORT :: C_INT64_T, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx, out
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: name
      INTEGER(C_INT64_T), VALUE :: n
   END FUNCTION
   REAL(C_DOUBLE) FUNCTION c_wam_gpu_get_scalar(ctx, name) BIND(C, NAME='wam_gpu_get_scalar')
      IMPORT :: C_DOUBLE, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: ctx
      CHARACTER(KIND=C_CHAR), DIMENSION(*) :: name
   END FUNCTION
   INTEGER(C_INT) FUNCTION c_wam_gpu_set_source_options(ctx, brk, phil) BIND(C, NAME='wam_gpu_set_source_options')
      IMPORT :: C_INT, C_INT32_T, C_PTR
      TYPE(C_PTR), VALUE :: ctx
      INTEGER(C_INT32_T), VALUE :: brk, phil
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

PUBLIC :: WAM_GPU_SETUP, WAM_GPU_SET_SOURCE_OPTIONS, WAM_GPU_PREPARE_PROPAGATION, WAM_GPU_PREPARE_PROPAGATION_DEV, &
&         PROPAGS_GPU, &
&         IMPLSCH_GPU, OUTPUT_PARAMETERS_GPU, BOUNDARY_OUTPUT_GPU, BOUNDARY_INPUT_GPU, WAM_GPU_END

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

!  SET_MODEL_OPTION's WAVE_BREAKING and PHILLIPS switches (wam_timopt_module.f90:176-177):
!  call after WAM_GPU_SETUP with the host's WAVE_BREAKING_RUN and PHILLIPS_RUN.
SUBROUTINE WAM_GPU_SET_SOURCE_OPTIONS (WAVE_BREAKING, PHILLIPS)
   LOGICAL, INTENT(IN) :: WAVE_BREAKING, PHILLIPS
   INTEGER(C_INT32_T) :: B, P
   B = 0_C_INT32_T
   P = 0_C_INT32_T
   IF (WAVE_BREAKING) B = 1_C_INT32_T
   IF (PHILLIPS) P = 1_C_INT32_T
   CALL CHECK(c_wam_gpu_set_source_options(CTX, B, P), 'WAM_GPU_SET_SOURCE_OPTIONS')
END SUBROUTINE WAM_GPU_SET_SOURCE_OPTIONS

SUBROUTINE WAM_GPU_PREPARE_PROPAGATION (DEPTH, U, V, INDEP)
   REAL(C_FLOAT), INTENT(IN), TARGET, CONTIGUOUS :: DEPTH(:)
   REAL(C_FLOAT), INTENT(IN), TARGET, CONTIGUOUS, OPTIONAL :: U(:), V(:)
   !! the depth table index of every sea point, which the reference's
   !! PREPARE_PROPAGATION stores in the module array INDEP
   !! (wam_propagation_module.f90:334-336) and WAMODEL passes to IMPLSCH
   INTEGER(C_INT32_T), INTENT(OUT), TARGET, CONTIGUOUS, OPTIONAL :: INDEP(:)
   TYPE(C_PTR) :: PU, PV
   INTEGER(C_INT64_T) :: N
   PU = C_NULL_PTR; PV = C_NULL_PTR
   IF (PRESENT(U)) PU = C_LOC(U)
   IF (PRESENT(V)) PV = C_LOC(V)
   CALL CHECK(c_wam_gpu_prepare_propagation(CTX, C_LOC(DEPTH), PU, PV), 'WAM_GPU_PREPARE_PROPAGATION')
   IF (PRESENT(INDEP)) THEN
      N = c_wam_gpu_get_table(CTX, 'INDEP'//C_NULL_CHAR, C_LOC(INDEP), INT(SIZE(INDEP), C_INT64_T))
      IF (N .NE. SIZE(INDEP)) THEN
         WRITE (*,*) ' *** WAM GPU ERROR in WAM_GPU_PREPARE_PROPAGATION: INDEP must have NSEA elements'
         ERROR STOP 1
      END IF
   END IF
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
!  DEPTH and INDEP must be those of the last WAM_GPU_PREPARE_PROPAGATION, as in
!  the reference.  The reference's side outputs live in module variables: PHIAW
!  and, when LCFLX is set (wamodel.f90:342, output times), PHIOC, TAUOC_X,
!  TAUOC_Y, PHIBOT, TAUBOT_X, TAUBOT_Y of wam_flux_module.f90:15-21 (WNFLUXES).
!  Here they are optional arguments, so a host passes its own module arrays:
!     CALL IMPLSCH_GPU (..., DEPTH, INDEP, LCFLX=LCFLX, PHIAW=PHIAW, PHIOC=PHIOC, ...)
SUBROUTINE IMPLSCH_GPU (FL3, U10, UDIR, TAUW, USTAR, Z0, ROAIRN, WSTAR, DEPTH, INDEP, LCFLX, PHIAW, PHIOC, &
&                       TAUOC_X, TAUOC_Y, PHIBOT, TAUBOT_X, TAUBOT_Y)
   REAL(C_FLOAT),      INTENT(INOUT), TARGET, CONTIGUOUS :: FL3(:,:,:)
   REAL(C_FLOAT),      INTENT(IN),    TARGET, CONTIGUOUS :: U10(:), UDIR(:), ROAIRN(:), WSTAR(:)
   REAL(C_FLOAT),      INTENT(INOUT), TARGET, CONTIGUOUS :: TAUW(:), Z0(:)
   REAL(C_FLOAT),      INTENT(OUT),   TARGET, CONTIGUOUS :: USTAR(:)
   REAL(C_FLOAT),      INTENT(IN),    TARGET, CONTIGUOUS :: DEPTH(:)
   INTEGER(C_INT32_T), INTENT(IN),    TARGET, CONTIGUOUS :: INDEP(:)
   LOGICAL,            INTENT(IN),                      OPTIONAL :: LCFLX
   REAL(C_FLOAT),      INTENT(OUT),   TARGET, CONTIGUOUS, OPTIONAL :: PHIAW(:)
   REAL(C_FLOAT),      INTENT(OUT),                     OPTIONAL :: PHIOC(:), TAUOC_X(:), TAUOC_Y(:), PHIBOT(:), &
&                                                                   TAUBOT_X(:), TAUBOT_Y(:)
   REAL(C_FLOAT), ALLOCATABLE, TARGET :: FLUX(:,:)
   TYPE(C_PTR) :: PPH, PFX
   INTEGER(C_INT32_T) :: LC
   INTEGER :: N
   N = SIZE(FL3,1)
   LC = 0_C_INT32_T
   IF (PRESENT(LCFLX)) THEN
      IF (LCFLX) LC = 1_C_INT32_T
   END IF
   PPH = C_NULL_PTR; PFX = C_NULL_PTR
   IF (PRESENT(PHIAW)) PPH = C_LOC(PHIAW)
   IF (LC .NE. 0) THEN
      ALLOCATE (FLUX(N, 6))
      PFX = C_LOC(FLUX)
   END IF
   CALL CHECK(c_wam_gpu_implsch(CTX, INT(N, C_INT32_T), INT(SIZE(FL3,2), C_INT32_T), &
&             INT(SIZE(FL3,3), C_INT32_T), C_LOC(FL3), C_LOC(U10), C_LOC(UDIR), C_LOC(TAUW), C_LOC(USTAR), &
&             C_LOC(Z0), C_LOC(ROAIRN), C_LOC(WSTAR), C_LOC(DEPTH), C_LOC(INDEP), LC, PPH, PFX), 'IMPLSCH_GPU')
   IF (LC .NE. 0) THEN
      IF (PRESENT(PHIOC))    PHIOC    = FLUX(:, 1)
      IF (PRESENT(TAUOC_X))  TAUOC_X  = FLUX(:, 2)
      IF (PRESENT(TAUOC_Y))  TAUOC_Y  = FLUX(:, 3)
      IF (PRESENT(PHIBOT))   PHIBOT   = FLUX(:, 4)
      IF (PRESENT(TAUBOT_X)) TAUBOT_X = FLUX(:, 5)
      IF (PRESENT(TAUBOT_Y)) TAUBOT_Y = FLUX(:, 6)
   END IF
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

!  BOUNDARY_OUTPUT's gather and mean values (wam_boundary_module.f90:350-362) at
!  the sea points IJARC(1:NB): FBC(:,:,I) = FL3(IJARC(I),:,:), one spectrum per
!  point as the boundary file's record holds it (WRITE (IU) FBC(:,:,I)), and
!  EMEANC, THQC, FMEANC of TOTAL_ENERGY, MEAN_DIRECTION and FEMEAN.
SUBROUTINE BOUNDARY_OUTPUT_GPU (FL3, IJARC, FBC, EMEANC, THQC, FMEANC)
   REAL(C_FLOAT),      INTENT(IN),  TARGET, CONTIGUOUS :: FL3(:,:,:)
   INTEGER(C_INT32_T), INTENT(IN),  TARGET, CONTIGUOUS :: IJARC(:)
   REAL(C_FLOAT),      INTENT(OUT), TARGET, CONTIGUOUS :: FBC(:,:,:)
   REAL(C_FLOAT),      INTENT(OUT)                     :: EMEANC(:), THQC(:), FMEANC(:)
   REAL(C_FLOAT), ALLOCATABLE, TARGET :: MEAN(:,:)
   INTEGER :: NB
   NB = SIZE(IJARC)
   IF (SIZE(FBC,1) .NE. SIZE(FL3,2) .OR. SIZE(FBC,2) .NE. SIZE(FL3,3) .OR. SIZE(FBC,3) .LT. NB .OR.        &
&      SIZE(EMEANC) .LT. NB .OR. SIZE(THQC) .LT. NB .OR. SIZE(FMEANC) .LT. NB) THEN
      WRITE (*,*) ' *** WAM GPU ERROR in BOUNDARY_OUTPUT_GPU: FBC must be (KL, ML, NB), the means (NB)'
      ERROR STOP 1
   END IF
   ALLOCATE (MEAN(NB,3))
   CALL CHECK(c_wam_gpu_boundary_output(CTX, INT(SIZE(FL3,1), C_INT32_T), INT(SIZE(FL3,2), C_INT32_T),    &
&             INT(SIZE(FL3,3), C_INT32_T), C_LOC(FL3), INT(NB, C_INT32_T), C_LOC(IJARC), C_LOC(FBC),        &
&             C_LOC(MEAN)), 'BOUNDARY_OUTPUT_GPU')
   EMEANC(1:NB) = MEAN(:,1)
   THQC(1:NB) = MEAN(:,2)
   FMEANC(1:NB) = MEAN(:,3)
END SUBROUTINE BOUNDARY_OUTPUT_GPU

!  BOUNDARY_INPUT after its reads (wam_boundary_module.f90:226-285): the set at
!  the model time (WHICH = 1: F1, 2: F2, 0: INTSPEC at DEL1L of DEL12) into FI
!  and its means, then FL3(IJARF(I),:,:) = INTSPEC(1, BFW(I), FI(:,:,IBFL(I)),
!  FI(:,:,IBFR(I))).  F1, F2, FI are (KL, ML, NBINP) as in the reference.
SUBROUTINE BOUNDARY_INPUT_GPU (FL3, F1, EMEAN1, THQ1, FMEAN1, F2, EMEAN2, THQ2, FMEAN2, WHICH, DEL12, DEL1L, &
&                              IJARF, IBFL, IBFR, BFW, FI, EMEANI, THQI, FMEANI)
   REAL(C_FLOAT),      INTENT(INOUT), TARGET, CONTIGUOUS :: FL3(:,:,:)
   REAL(C_FLOAT),      INTENT(IN),    TARGET, CONTIGUOUS :: F1(:,:,:), F2(:,:,:)
   REAL(C_FLOAT),      INTENT(IN)                       :: EMEAN1(:), THQ1(:), FMEAN1(:)
   REAL(C_FLOAT),      INTENT(IN)                       :: EMEAN2(:), THQ2(:), FMEAN2(:)
   INTEGER,            INTENT(IN)                       :: WHICH
   REAL,               INTENT(IN)                       :: DEL12, DEL1L
   INTEGER(C_INT32_T), INTENT(IN),    TARGET, CONTIGUOUS :: IJARF(:), IBFL(:), IBFR(:)
   REAL(C_FLOAT),      INTENT(IN),    TARGET, CONTIGUOUS :: BFW(:)
   REAL(C_FLOAT),      INTENT(OUT),   TARGET, CONTIGUOUS :: FI(:,:,:)
   REAL(C_FLOAT),      INTENT(OUT)                      :: EMEANI(:), THQI(:), FMEANI(:)
   REAL(C_FLOAT), ALLOCATABLE, TARGET :: M1(:,:), M2(:,:), MI(:,:)
   INTEGER :: NBINP
   NBINP = SIZE(F1,3)
   IF (SIZE(F2,3) .NE. NBINP .OR. SIZE(FI,3) .NE. NBINP .OR. SIZE(F1,1) .NE. SIZE(FL3,2) .OR.              &
&      SIZE(F1,2) .NE. SIZE(FL3,3) .OR. SIZE(IBFL) .NE. SIZE(IJARF) .OR. SIZE(IBFR) .NE. SIZE(IJARF) .OR.   &
&      SIZE(BFW) .NE. SIZE(IJARF)) THEN
      WRITE (*,*) ' *** WAM GPU ERROR in BOUNDARY_INPUT_GPU: F1, F2, FI must be (KL, ML, NBINP)'
      ERROR STOP 1
   END IF
   ALLOCATE (M1(NBINP,3), M2(NBINP,3), MI(NBINP,3))
   M1(:,1) = EMEAN1(1:NBINP) ; M1(:,2) = THQ1(1:NBINP) ; M1(:,3) = FMEAN1(1:NBINP)
   M2(:,1) = EMEAN2(1:NBINP) ; M2(:,2) = THQ2(1:NBINP) ; M2(:,3) = FMEAN2(1:NBINP)
   CALL CHECK(c_wam_gpu_boundary_input(CTX, INT(SIZE(FL3,1), C_INT32_T), INT(SIZE(FL3,2), C_INT32_T),     &
&             INT(SIZE(FL3,3), C_INT32_T), C_LOC(FL3), INT(NBINP, C_INT32_T), C_LOC(F1), C_LOC(M1),         &
&             C_LOC(F2), C_LOC(M2), INT(WHICH, C_INT32_T), REAL(DEL12, C_FLOAT), REAL(DEL1L, C_FLOAT),      &
&             INT(SIZE(IJARF), C_INT32_T), C_LOC(IJARF), C_LOC(IBFL), C_LOC(IBFR), C_LOC(BFW), C_LOC(FI),   &
&             C_LOC(MI)), 'BOUNDARY_INPUT_GPU')
   EMEANI(1:NBINP) = MI(:,1)
   THQI(1:NBINP) = MI(:,2)
   FMEANI(1:NBINP) = MI(:,3)
END SUBROUTINE BOUNDARY_INPUT_GPU

SUBROUTINE WAM_GPU_END
   IF (C_ASSOCIATED(CTX)) CALL c_wam_gpu_finalize(CTX)
   CTX = C_NULL_PTR
END SUBROUTINE WAM_GPU_END

END MODULE WAM_GPU_INTERFACE
