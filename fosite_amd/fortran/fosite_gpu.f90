!##############################################################################
!# fosite_gpu_mod: Fortran 2008 host of the MI355X finite-volume step        #
!##############################################################################
!> Object API in the shape of fosite's driver (fosite.f90: Setup, FirstStep,
!! Step, Run, Finalize) with the Mesh, Physics, Fluxes, Boundary, Sources and
!! Timedisc parameters as plain components.  Setup serialises them once into
!! a device context through the C ABI (fosite_fv_c); all array work then runs
!! on the GPU and the host only exchanges scalars, except for InitData and
!! GetState.  Integer ids are the reference's (physics_base.f90:496-498,
!! geometry_base.f90:345-357, boundary_base.f90:110-130,
!! reconstruction_linear.f90:71-78, timedisc_base.f90:187-191).
!!
!! Errors follow the reference's convention (logging_base.f90:338-360): the
!! message is printed and the program stops; a numerical break is reported
!! through Timedisc%break and the return value of Step.
MODULE fosite_gpu_mod
  USE, INTRINSIC :: iso_c_binding
  USE fosite_fv_c
  IMPLICIT NONE
  PRIVATE

  INTEGER, PARAMETER :: DP = c_double
  INTEGER, PARAMETER, PUBLIC :: GNUM = 2
  INTEGER, PARAMETER, PUBLIC :: EULER_ISOTHERM = 16, EULER = 17
  INTEGER, PARAMETER, PUBLIC :: CARTESIAN = 1, LOGCYLINDRICAL = 31
  INTEGER, PARAMETER, PUBLIC :: NO_GRADIENTS = 1, PERIODIC = 2, ABSORBING = 13
  INTEGER, PARAMETER, PUBLIC :: MINMOD = 1, MONOCENT = 2, SWEBY = 3, SUPERBEE = 4, &
       OSPRE = 5, VANLEER = 7, NOLIMIT = 8
  INTEGER, PARAMETER, PUBLIC :: PRIMITIVE = 1, CONSERVATIVE = 2
  INTEGER, PARAMETER, PUBLIC :: RK_FEHLBERG = 2, CASH_KARP = 3, DORMAND_PRINCE = 4, SSPRK = 5

  TYPE, PUBLIC :: mesh_gpu
    INTEGER  :: geometry = CARTESIAN, inum = 32, jnum = 32
    REAL(DP) :: xmin = 0, xmax = 1, ymin = 0, ymax = 1, gparam = 1
    !> index ranges incl. ghost cells, set by Setup (mesh_base.f90:535-549)
    INTEGER  :: IGMIN = 0, IGMAX = -1, JGMIN = 0, JGMAX = -1
  END TYPE mesh_gpu

  TYPE, PUBLIC :: physics_gpu
    INTEGER  :: problem = EULER
    REAL(DP) :: gamma = 1.4_DP
    INTEGER  :: VNUM = 0       !< set by Setup
  END TYPE physics_gpu

  TYPE, PUBLIC :: fluxes_gpu
    INTEGER  :: limiter = MINMOD, variables = CONSERVATIVE
    REAL(DP) :: theta = 1
  END TYPE fluxes_gpu

  TYPE, PUBLIC :: boundary_gpu
    INTEGER :: western = NO_GRADIENTS, eastern = NO_GRADIENTS, &
               southern = NO_GRADIENTS, northern = NO_GRADIENTS
  END TYPE boundary_gpu

  !> point mass at the origin, Newtonian; accretion through boundary
  !! outbound (1..4, 0 off) with limit acclimit (< 0 off)
  !! (gravity_pointmass.f90:185-195, 285-299)
  TYPE, PUBLIC :: sources_gpu
    LOGICAL  :: pointmass = .FALSE.
    REAL(DP) :: mass = 1, GN = 1
    INTEGER  :: outbound = 0
    REAL(DP) :: acclimit = -1
    !> rotating frame with angular speed omega about the origin
    !! (sources_rotframe.f90:73-247): 0 off, 1 on, 2 disable_centaccel
    INTEGER  :: rotframe = 0
    REAL(DP) :: omega = 0
    !> viscosity (sources_viscosity.f90:147-165): 0 off, 1 MOLECULAR, 2 ALPHA
    !! (isothermal), its constants and the viscous Courant number
    INTEGER  :: vismodel = 0
    REAL(DP) :: dynconst = 0.1_DP, bulkconst = -2.0_DP/3.0_DP*0.1_DP, cvis = 0.5_DP
  END TYPE sources_gpu

  TYPE, PUBLIC :: timedisc_gpu
    INTEGER  :: method = SSPRK, order = 5, fargo = 0
    REAL(DP) :: cfl = 0.4_DP, tol_rel = 0.01_DP, tol_abs(4) = 0.001_DP
    REAL(DP) :: stoptime = 1, dtlimit = -1, beta = 0.08_DP
    !> CheckData limits (<= 0: TINY) and start time (timedisc_base.f90:295, 353-359)
    REAL(DP) :: pmin = -1, rhomin = -1, starttime = 0
    !> outputs between starttime and stoptime (fileio_base.f90:420-470)
    INTEGER  :: count = 1
    !> scalars after the last call, as Timedisc holds them
    REAL(DP) :: time = 0, dt = 0, dtold = 0, maxerrold = 0
    INTEGER  :: n_adj = 0
    LOGICAL  :: break = .FALSE.
  END TYPE timedisc_gpu

  TYPE, PUBLIC :: fosite_gpu
    TYPE(mesh_gpu)     :: Mesh
    TYPE(physics_gpu)  :: Physics
    TYPE(fluxes_gpu)   :: Fluxes
    TYPE(boundary_gpu) :: Boundary
    TYPE(sources_gpu)  :: Sources
    TYPE(timedisc_gpu) :: Timedisc
    INTEGER            :: iter = 0
    TYPE(c_ptr), PRIVATE :: ctx = c_null_ptr
  CONTAINS
    PROCEDURE :: Setup
    PROCEDURE :: SetSoundSpeeds
    PROCEDURE :: SetBackgroundVelocity
    PROCEDURE :: InitData
    PROCEDURE :: FirstStep
    PROCEDURE :: Step
    PROCEDURE :: Run
    PROCEDURE :: GetState
    PROCEDURE :: GetBackgroundVelocity
    PROCEDURE :: GetBoundaryFluxes
    PROCEDURE :: GetBoundaryFlux
    PROCEDURE :: GetPointMass
    PROCEDURE :: MeshArray
    PROCEDURE :: LastRunMs
    PROCEDURE :: Finalize
  END TYPE fosite_gpu

CONTAINS

  FUNCTION Config(this) RESULT(c)
    CLASS(fosite_gpu), INTENT(IN) :: this
    TYPE(fv_config) :: c
    c%physics = this%Physics%problem
    c%gamma = this%Physics%gamma
    c%geometry = this%Mesh%geometry
    c%inum = this%Mesh%inum
    c%jnum = this%Mesh%jnum
    c%xmin = this%Mesh%xmin
    c%xmax = this%Mesh%xmax
    c%ymin = this%Mesh%ymin
    c%ymax = this%Mesh%ymax
    c%gparam = this%Mesh%gparam
    c%bc = [this%Boundary%western, this%Boundary%eastern, this%Boundary%southern, &
            this%Boundary%northern]
    c%limiter = this%Fluxes%limiter
    c%theta = this%Fluxes%theta
    c%recvars = this%Fluxes%variables
    c%method = this%Timedisc%method
    c%order = this%Timedisc%order
    c%cfl = this%Timedisc%cfl
    c%tol_rel = this%Timedisc%tol_rel
    c%tol_abs = this%Timedisc%tol_abs
    c%stoptime = this%Timedisc%stoptime
    c%dtlimit = this%Timedisc%dtlimit
    c%beta = this%Timedisc%beta
    c%fargo = this%Timedisc%fargo
    c%gravity = MERGE(1, 0, this%Sources%pointmass)
    c%mass = this%Sources%mass
    c%GN = this%Sources%GN
    c%outbound = this%Sources%outbound
    c%acclimit = this%Sources%acclimit
    c%pmin = this%Timedisc%pmin
    c%rhomin = this%Timedisc%rhomin
    c%count = this%Timedisc%count
    c%starttime = this%Timedisc%starttime
    c%omega = this%Sources%omega
    c%rotframe = this%Sources%rotframe
    c%vismodel = this%Sources%vismodel
    c%dynconst = this%Sources%dynconst
    c%bulkconst = this%Sources%bulkconst
    c%cvis = this%Sources%cvis
    c%dtype = 0
  END FUNCTION Config

  SUBROUTINE Check(this, rc, routine)
    CLASS(fosite_gpu), INTENT(IN) :: this
    INTEGER(c_int), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: routine
    IF (rc.NE.0) THEN
      WRITE(*,'(A,A,A,I0,A,A)') "ERROR in fosite_gpu::", routine, " (code ", rc, "): ", &
           c_string(fv_last_error(this%ctx))
      ERROR STOP 99
    END IF
  END SUBROUTINE Check

  !> fosite%Setup (fosite.f90:152-255): validate, build the metric, allocate
  !! the device-resident state
  SUBROUTINE Setup(this)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    TYPE(fv_config) :: c
    INTEGER(c_int) :: rc, ni, nj, v
    c = Config(this)
    rc = fv_create(this%ctx, c)
    IF (rc.NE.0) THEN
      WRITE(*,'(A,I0,A)') "ERROR in fosite_gpu::Setup (code ", rc, "): see the message above"
      ERROR STOP 99
    END IF
    CALL Check(this, fv_dims(this%ctx, ni, nj, v), "Setup")
    this%Mesh%IGMIN = 1 - GNUM
    this%Mesh%IGMAX = this%Mesh%inum + GNUM
    this%Mesh%JGMIN = 1 - GNUM
    this%Mesh%JGMAX = this%Mesh%jnum + GNUM
    this%Physics%VNUM = v
    this%iter = 0
  END SUBROUTINE Setup

  !> physics_eulerisotherm%SetSoundSpeeds (physics_eulerisotherm.f90:288-309)
  SUBROUTINE SetSoundSpeeds(this, bccsound, fcsound)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP), INTENT(IN), CONTIGUOUS :: bccsound(:,:), fcsound(:,:,:)
    CALL Check(this, fv_set_sound_speeds(this%ctx, bccsound, fcsound), "SetSoundSpeeds")
  END SUBROUTINE SetSoundSpeeds

  !> Timedisc%w for FARGO type 2, one value per x-ring IGMIN..IGMAX
  SUBROUTINE SetBackgroundVelocity(this, w)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP), INTENT(IN), CONTIGUOUS :: w(:)
    CALL Check(this, fv_set_background_velocity(this%ctx, w), "SetBackgroundVelocity")
  END SUBROUTINE SetBackgroundVelocity

  !> InitData: Timedisc%pvar = pvar, then Convert2Conservative
  SUBROUTINE InitData(this, pvar)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP), INTENT(IN), CONTIGUOUS :: pvar(:,:,:)
    CALL Check(this, fv_set_pvar(this%ctx, pvar), "InitData")
  END SUBROUTINE InitData

  SUBROUTINE Pull(this)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    TYPE(fv_scalars) :: s
    CALL Check(this, fv_get_scalars(this%ctx, s), "Pull")
    this%iter = s%iter
    this%Timedisc%n_adj = s%n_adj
    this%Timedisc%break = s%brk.NE.0
    this%Timedisc%time = s%time
    this%Timedisc%dt = s%dt
    this%Timedisc%dtold = s%dtold
    this%Timedisc%maxerrold = s%maxerrold
  END SUBROUTINE Pull

  !> fosite%FirstStep (fosite.f90:258-348)
  SUBROUTINE FirstStep(this)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    INTEGER(c_int) :: brk
    CALL Check(this, fv_first_step(this%ctx, brk), "FirstStep")
    CALL Pull(this)
  END SUBROUTINE FirstStep

  !> fosite%Step (fosite.f90:367-464); .TRUE. when the run ended (stop time
  !! reached or numerical break)
  LOGICAL FUNCTION Step(this) RESULT(done)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    INTEGER(c_int) :: d
    CALL Check(this, fv_step(this%ctx, d), "Step")
    CALL Pull(this)
    done = d.NE.0
  END FUNCTION Step

  !> nsteps x Step in device-controlled batches; returns the steps taken
  INTEGER FUNCTION Run(this, nsteps) RESULT(taken)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    INTEGER, INTENT(IN) :: nsteps
    INTEGER(c_int) :: t
    CALL Check(this, fv_run(this%ctx, INT(nsteps, c_int), t), "Run")
    CALL Pull(this)
    taken = t
  END FUNCTION Run

  !> Timedisc%cvar and Convert2Primitive(cvar), all cells incl. ghosts
  SUBROUTINE GetState(this, cvar, pvar)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP), INTENT(OUT), CONTIGUOUS, TARGET, OPTIONAL :: cvar(:,:,:), pvar(:,:,:)
    TYPE(c_ptr) :: pc, pp
    pc = c_null_ptr
    pp = c_null_ptr
    IF (PRESENT(cvar)) pc = C_LOC(cvar)
    IF (PRESENT(pvar)) pp = C_LOC(pvar)
    CALL Check(this, fv_get_state(this%ctx, pc, pp), "GetState")
  END SUBROUTINE GetState

  !> Timedisc%w, one value per x-ring IGMIN..IGMAX (FARGO type 1: the value
  !! CalcBackgroundVelocity computed at the start of the last step)
  SUBROUTINE GetBackgroundVelocity(this, w)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP), INTENT(OUT), CONTIGUOUS :: w(:)
    CALL Check(this, fv_get_background_velocity(this%ctx, w), "GetBackgroundVelocity")
  END SUBROUTINE GetBackgroundVelocity

  !> Fluxes%bxflux(JGMIN:JGMAX,1,2,VNUM) and Fluxes%byflux(1,IGMIN:IGMAX,2,VNUM)
  !! (fluxes_base.f90:142-143)
  SUBROUTINE GetBoundaryFluxes(this, bxflux, byflux)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP), INTENT(OUT), CONTIGUOUS, TARGET, OPTIONAL :: bxflux(:,:,:,:), byflux(:,:,:,:)
    TYPE(c_ptr) :: px, py
    px = c_null_ptr
    py = c_null_ptr
    IF (PRESENT(bxflux)) px = C_LOC(bxflux)
    IF (PRESENT(byflux)) py = C_LOC(byflux)
    CALL Check(this, fv_get_boundary_flux(this%ctx, px, py), "GetBoundaryFluxes")
  END SUBROUTINE GetBoundaryFluxes

  !> Fluxes%GetBoundaryFlux (fluxes_base.f90:220-300): the VNUM totals through
  !! one boundary, direction WEST=1, EAST=2, SOUTH=3, NORTH=4
  FUNCTION GetBoundaryFlux(this, direction) RESULT(bflux)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    INTEGER, INTENT(IN) :: direction
    REAL(DP) :: bflux(4)
    bflux(:) = 0.0_DP
    CALL Check(this, fv_get_boundary_flux_total(this%ctx, INT(direction, c_int), bflux), "GetBoundaryFlux")
  END FUNCTION GetBoundaryFlux

  !> the point mass: mass, accrate, massloss, mdot (gravity_pointmass.f90:79-85)
  FUNCTION GetPointMass(this) RESULT(pm)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    REAL(DP) :: pm(4)
    CALL Check(this, fv_get_point_mass(this%ctx, pm), "GetPointMass")
  END FUNCTION GetPointMass

  !> one metric array of the configured mesh (ids as fv_mesh_array)
  SUBROUTINE MeshArray(this, id, a)
    CLASS(fosite_gpu), INTENT(IN) :: this
    INTEGER, INTENT(IN) :: id
    REAL(DP), INTENT(OUT), CONTIGUOUS :: a(:,:)
    CALL Check(this, fv_mesh_array(Config(this), INT(id, c_int), a), "MeshArray")
  END SUBROUTINE MeshArray

  !> device time of the last Run, ms
  REAL(DP) FUNCTION LastRunMs(this)
    CLASS(fosite_gpu), INTENT(IN) :: this
    LastRunMs = fv_last_run_ms(this%ctx)
  END FUNCTION LastRunMs

  !> fosite%Finalize (fosite.f90:680-691)
  SUBROUTINE Finalize(this)
    CLASS(fosite_gpu), INTENT(INOUT) :: this
    INTEGER(c_int) :: rc
    IF (C_ASSOCIATED(this%ctx)) rc = fv_destroy(this%ctx)
    this%ctx = c_null_ptr
  END SUBROUTINE Finalize

END MODULE fosite_gpu_mod
