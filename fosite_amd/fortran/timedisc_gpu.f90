!##############################################################################
!# timedisc_gpu: fosite's Timedisc on one MI355X, as a timedisc_base        #
!#               extension compiled against fosite's own modules            #
!##############################################################################
!> The binding a fosite build adds to run its time integration on the GPU:
!!
!!  * timedisc_gpu EXTENDS timedisc_base (timedisc/timedisc_base.f90:84-177)
!!    and overrides IntegrationStep, ComputeRHS, CalcTimestep, SolveODE and
!!    Finalize.  Mesh, Physics, Fluxes, Boundary and Sources stay fosite's
!!    own objects; their parameters are handed to the device context once, at
!!    the first ComputeRHS (FirstStep), when all of them exist.
!!  * new_timedisc_device is new_timedisc (timedisc/timedisc_generic.f90:
!!    45-88) with one more method id, GPU_TIMEDISC; every other id goes to
!!    fosite's own factory.
!!  * fosite_device EXTENDS fosite and overrides Setup (fosite.f90:152-255)
!!    only to call that factory, so a program written for fosite runs on the
!!    device by allocating a fosite_device instead of a fosite and setting
!!    timedisc/method = GPU_TIMEDISC.  A fosite build that patches
!!    new_timedisc does not need it.
!!
!! Keys of the timedisc dictionary besides fosite's own: device_method (the
!! Runge-Kutta scheme the device integrates with: RK_FEHLBERG, CASH_KARP,
!! DORMAND_PRINCE or SSPRK, default SSPRK), order (default 5), dtype (0 FP64,
!! 1 the FP32 precision form), device_sync (1: the host state vectors are
!! refreshed after every step, as fosite's output and any host code expect;
!! 0: only when the caller asks through SyncHost).
!!
!! FARGO frames: fosite's WriteDataset moves Timedisc%pvar/cvar to the real
!! frame on the host (io/fileio_base.f90:523-533).  The next IntegrationStep
!! sees that in the host state vectors' fargo_transformation_applied flags
!! and applies the same AddBackgroundVelocityY to the device state, so the
!! device takes the round trips fosite takes.
MODULE timedisc_gpu_mod
  USE, INTRINSIC :: iso_c_binding
  USE fosite_fv_c
  USE logging_base_mod
  USE common_dict
  USE mesh_base_mod
  USE mesh_generic_mod
  USE marray_compound_mod
  USE marray_base_mod
  USE physics_base_mod
  USE physics_generic_mod
  USE physics_eulerisotherm_mod, ONLY : physics_eulerisotherm, statevector_eulerisotherm
  USE physics_euler_mod, ONLY : physics_euler
  USE fluxes_base_mod
  USE fluxes_generic_mod
  USE reconstruction_base_mod
  USE reconstruction_generic_mod
  USE reconstruction_linear_mod, ONLY : reconstruction_linear, SWEBY, MONOCENT
  USE geometry_base_mod
  USE geometry_generic_mod
  USE boundary_base_mod
  USE boundary_generic_mod
  USE sources_base_mod
  USE sources_generic_mod
  USE sources_gravity_mod, ONLY : sources_gravity
  USE gravity_pointmass_mod, ONLY : gravity_pointmass
  USE sources_rotframe_mod, ONLY : sources_rotframe
  USE sources_viscosity_mod, ONLY : sources_viscosity
  USE timedisc_base_mod
  USE timedisc_generic_mod
  USE fileio_generic_mod
  USE fosite_mod
  IMPLICIT NONE
  PRIVATE

  !> method id of the device Timedisc (unused by fosite's own schemes)
  INTEGER, PARAMETER, PUBLIC :: GPU_TIMEDISC = 21

  TYPE, EXTENDS(timedisc_base), PUBLIC :: timedisc_gpu
    TYPE(c_ptr) :: ctx = c_null_ptr
    INTEGER     :: scheme = SSPRK, dtype = 0, sync = 1
    INTEGER     :: nadj_seen = 0
  CONTAINS
    PROCEDURE :: InitTimedisc_gpu
    PROCEDURE :: IntegrationStep => IntegrationStep_gpu
    PROCEDURE :: ComputeRHS => ComputeRHS_gpu
    PROCEDURE :: CalcTimestep => CalcTimestep_gpu
    PROCEDURE :: SolveODE => SolveODE_gpu
    PROCEDURE :: SyncHost
    PROCEDURE :: Finalize => Finalize_gpu
  END TYPE timedisc_gpu

  TYPE, EXTENDS(fosite), PUBLIC :: fosite_device
  CONTAINS
    PROCEDURE :: Setup => Setup_device
  END TYPE fosite_device

  PUBLIC :: new_timedisc_device

CONTAINS

  !> new_timedisc (timedisc/timedisc_generic.f90:45-88) with GPU_TIMEDISC
  SUBROUTINE new_timedisc_device(Timedisc,Mesh,Physics,config,IO)
    CLASS(timedisc_base), ALLOCATABLE   :: Timedisc
    CLASS(mesh_base),     INTENT(INOUT) :: Mesh
    CLASS(physics_base),  INTENT(IN)    :: Physics
    TYPE(Dict_TYP),       POINTER       :: config, IO
    INTEGER                             :: method
    CALL GetAttr(config, "method", method)
    IF (method.NE.GPU_TIMEDISC) THEN
      CALL new_timedisc(Timedisc,Mesh,Physics,config,IO)
      RETURN
    END IF
    ALLOCATE(timedisc_gpu::Timedisc)
    SELECT TYPE(obj => Timedisc)
    TYPE IS (timedisc_gpu)
      CALL obj%InitTimedisc_gpu(Mesh,Physics,config,IO)
    END SELECT
  END SUBROUTINE new_timedisc_device

  SUBROUTINE InitTimedisc_gpu(this,Mesh,Physics,config,IO)
    CLASS(timedisc_gpu), INTENT(INOUT) :: this
    CLASS(mesh_base),    INTENT(INOUT) :: Mesh
    CLASS(physics_base), INTENT(IN)    :: Physics
    TYPE(Dict_TYP),      POINTER       :: config, IO
    CHARACTER(LEN=32)                  :: tname
    tname = "MI355X device"
    CALL GetAttr(config, "device_method", this%scheme, SSPRK)
    CALL GetAttr(config, "order", this%order, 5)
    CALL GetAttr(config, "dtype", this%dtype, 0)
    CALL GetAttr(config, "device_sync", this%sync, 1)
    CALL this%InitTimedisc(Mesh,Physics,config,IO,GPU_TIMEDISC,tname)
  END SUBROUTINE InitTimedisc_gpu

  !> fv_config from fosite's objects (Mesh, Physics, Fluxes, Boundary,
  !! Sources) and this Timedisc's parameters
  SUBROUTINE DeviceConfig(this,Mesh,Physics,Sources,Fluxes,c)
    CLASS(timedisc_gpu), INTENT(INOUT) :: this
    CLASS(mesh_base),    INTENT(IN)    :: Mesh
    CLASS(physics_base), INTENT(INOUT) :: Physics
    CLASS(sources_list), ALLOCATABLE, INTENT(INOUT) :: Sources
    CLASS(fluxes_base),  INTENT(INOUT) :: Fluxes
    TYPE(fv_config),     INTENT(OUT)   :: c
    CLASS(sources_base), POINTER       :: sp
    LOGICAL                            :: vis_seen
    INTEGER                            :: d
    c%physics = Physics%GetType()
    SELECT TYPE(p => Physics)
    CLASS IS (physics_euler)
      c%gamma = p%gamma
    END SELECT
    c%geometry = Mesh%Geometry%GetType()
    IF (Mesh%KNUM.NE.1) CALL this%Error("timedisc_gpu","2-D meshes only (knum = 1)")
    c%inum = Mesh%INUM
    c%jnum = Mesh%JNUM
    c%xmin = Mesh%xmin
    c%xmax = Mesh%xmax
    c%ymin = Mesh%ymin
    c%ymax = Mesh%ymax
    c%gparam = Mesh%Geometry%GetScale()
    DO d = 1, 4
      c%bc(d) = this%Boundary%Boundary(d)%p%GetType()
    END DO
    SELECT TYPE(r => Fluxes%Reconstruction)
    CLASS IS (reconstruction_linear)
      c%limiter = r%GetType()
      IF (c%limiter.EQ.SWEBY .OR. c%limiter.EQ.MONOCENT) c%theta = r%limiter_param
    CLASS DEFAULT
      CALL this%Error("timedisc_gpu","linear reconstruction only")
    END SELECT
    c%recvars = MERGE(PRIMITIVE, CONSERVATIVE, Fluxes%Reconstruction%PrimRecon())
    c%method = this%scheme
    c%order = this%order
    c%cfl = this%cfl
    c%tol_rel = this%tol_rel
    c%tol_abs(:) = 0.001
    c%tol_abs(1:Physics%VNUM) = this%tol_abs(1:Physics%VNUM)
    c%stoptime = this%stoptime
    c%dtlimit = this%dtlimit
    c%beta = this%beta
    c%pmin = this%pmin
    c%rhomin = this%rhomin
    c%starttime = this%time
    c%count = 1  ! output times are fosite's Datafile's (fv_integration_step takes its dt)
    c%fargo = Mesh%fargo%GetType()
    c%dtype = this%dtype
    c%gravity = 0
    c%outbound = 0
    c%acclimit = -1.0
    c%GN = Physics%constants%GN
    c%omega = Mesh%omega
    c%rotframe = 0
    c%vismodel = 0
    vis_seen = .FALSE.
    ! the device evaluates one point mass at the origin and the rotating
    ! frame about the origin, in that order (gravity heads the list,
    ! sources_generic.f90:133-199): the source terms it accepts
    IF (ALLOCATED(Sources)) THEN
      sp => Sources%next
      DO WHILE (ASSOCIATED(sp))
        SELECT TYPE(sg => sp)
        TYPE IS (sources_gravity)
          SELECT TYPE(gp => sg%glist)
          TYPE IS (gravity_pointmass)
            IF (ASSOCIATED(sg%glist%next)) &
              CALL this%Error("timedisc_gpu","one point mass is the only gravity term supported")
            c%gravity = 1
            c%mass = gp%mass
            c%outbound = gp%outbound
            c%acclimit = gp%acclimit
          CLASS DEFAULT
            CALL this%Error("timedisc_gpu","gravity other than a point mass")
          END SELECT
        TYPE IS (sources_rotframe)
          IF (ANY(Mesh%rotcent(:).NE.0.0)) &
            CALL this%Error("timedisc_gpu","rotating frame about the origin only")
          ! the device adds the terms as gravity, frame, viscosity
          IF (vis_seen) &
            CALL this%Error("timedisc_gpu","the viscosity before the rotating frame in the source list")
          c%rotframe = MERGE(2, 1, sg%disable_centaccel)
        TYPE IS (sources_viscosity)
          ! MOLECULAR (1), ALPHA (2) and BETA (3) (sources_viscosity.f90:75-77)
          c%vismodel = sg%viscosity%GetType()
          IF (c%vismodel.LT.1 .OR. c%vismodel.GT.3) &
            CALL this%Error("timedisc_gpu","viscosity models other than MOLECULAR, ALPHA and BETA")
          c%dynconst = sg%dynconst
          c%bulkconst = sg%bulkconst
          c%cvis = sg%cvis
          vis_seen = .TRUE.
        CLASS DEFAULT
          CALL this%Error("timedisc_gpu", &
            "source terms other than point-mass gravity, the rotating frame and viscosity")
        END SELECT
        sp => sp%next
      END DO
    END IF
  END SUBROUTINE DeviceConfig

  !> the host objects from the device: Timedisc%cvar, Timedisc%pvar, their
  !! FARGO frame flags, Timedisc%w and Fluxes%bxflux/byflux
  SUBROUTINE SyncHost(this,Mesh,Fluxes)
    CLASS(timedisc_gpu), INTENT(INOUT) :: this
    CLASS(mesh_base),    INTENT(IN)    :: Mesh
    CLASS(fluxes_base),  INTENT(INOUT) :: Fluxes
    INTEGER(c_int)                     :: rc
    rc = fv_get_state(this%ctx, C_LOC(this%cvar%data1d(1)), C_LOC(this%pvar%data1d(1)))
    IF (rc.NE.0) CALL this%Error("timedisc_gpu::SyncHost", c_string(fv_last_error(this%ctx)))
    IF (ASSOCIATED(Fluxes%bxflux) .AND. ASSOCIATED(Fluxes%byflux)) &
      rc = fv_get_boundary_flux(this%ctx, C_LOC(Fluxes%bxflux), C_LOC(Fluxes%byflux))
    IF (Mesh%fargo%GetType().GT.0) THEN
      CALL SetFrame(this%cvar, .TRUE.)
      CALL SetFrame(this%pvar, .TRUE.)
      rc = fv_get_background_velocity(this%ctx, this%w(:,Mesh%KGMIN))
    END IF
  END SUBROUTINE SyncHost

  SUBROUTINE SetFrame(sv, comoving)
    CLASS(marray_compound), INTENT(INOUT) :: sv
    LOGICAL,                INTENT(IN)    :: comoving
    SELECT TYPE(s => sv)
    CLASS IS (statevector_eulerisotherm)
      s%fargo_transformation_applied = comoving
    END SELECT
  END SUBROUTINE SetFrame

  LOGICAL FUNCTION InRealFrame(sv) RESULT(r)
    CLASS(marray_compound), INTENT(IN) :: sv
    r = .FALSE.
    SELECT TYPE(s => sv)
    CLASS IS (statevector_eulerisotherm)
      r = .NOT.s%fargo_transformation_applied
    END SELECT
  END FUNCTION InRealFrame

  !> ComputeRHS: fosite calls it once, in FirstStep (fosite.f90:315-317),
  !! where the device context is created from fosite's objects and the
  !! initial data, and its first evaluation runs
  SUBROUTINE ComputeRHS_gpu(this,Mesh,Physics,Sources,Fluxes,time,dt,pvar,cvar,checkdatabm,rhs)
    CLASS(timedisc_gpu),  INTENT(INOUT) :: this
    CLASS(mesh_base),     INTENT(IN)    :: Mesh
    CLASS(physics_base),  INTENT(INOUT) :: Physics
    CLASS(sources_list), ALLOCATABLE, INTENT(INOUT) :: Sources
    CLASS(fluxes_base),   INTENT(INOUT) :: Fluxes
    REAL,                 INTENT(IN)    :: time, dt
    CLASS(marray_compound),INTENT(INOUT):: pvar,cvar,rhs
    INTEGER,              INTENT(IN)    :: checkdatabm
    TYPE(fv_config)                     :: c
    INTEGER(c_int)                      :: rc, brk
    IF (C_ASSOCIATED(this%ctx)) &
      CALL this%Error("timedisc_gpu::ComputeRHS","the device integrates the whole step; " // &
                      "ComputeRHS runs once, in FirstStep")
    IF (checkdatabm.NE.CHECK_ALL .OR. dt.NE.0.0 .OR. time.NE.this%time) &
      CALL this%Error("timedisc_gpu::ComputeRHS","expected FirstStep's evaluation")
    CALL DeviceConfig(this,Mesh,Physics,Sources,Fluxes,c)
    rc = fv_create(this%ctx, c)
    IF (rc.NE.0) CALL this%Error("timedisc_gpu::ComputeRHS","fv_create failed (see stderr)")
    SELECT TYPE(p => Physics)
    CLASS IS (physics_euler)
    CLASS IS (physics_eulerisotherm)
      rc = fv_set_sound_speeds(this%ctx, p%bccsound%data1d, p%fcsound%data1d)
      CALL Check(rc, "fv_set_sound_speeds")
    END SELECT
    IF (Mesh%fargo%GetType().EQ.2) THEN
      rc = fv_set_background_velocity(this%ctx, this%w(:,Mesh%KGMIN))
      CALL Check(rc, "fv_set_background_velocity")
    END IF
    ! the state itself: FirstStep's CenterBoundary has already replaced pvar
    ! with Convert2Primitive(cvar) under FARGO type 1 (fosite.f90:307-313)
    rc = fv_set_cvar(this%ctx, cvar%data1d)
    CALL Check(rc, "fv_set_cvar")
    rc = fv_first_rhs(this%ctx, brk)
    CALL Check(rc, "fv_first_rhs")
    IF (brk.NE.0) this%break = .TRUE.
    CALL this%SyncHost(Mesh,Fluxes)
  CONTAINS
    SUBROUTINE Check(code, what)
      INTEGER(c_int), INTENT(IN)   :: code
      CHARACTER(LEN=*), INTENT(IN) :: what
      IF (code.NE.0) CALL this%Error("timedisc_gpu::ComputeRHS", what // ": " // &
                                     c_string(fv_last_error(this%ctx)))
    END SUBROUTINE Check
  END SUBROUTINE ComputeRHS_gpu

  !> IntegrationStep (timedisc/timedisc_base.f90:649-750) on the device with
  !! the time step fosite's Datafile%AdjustTimestep left
  SUBROUTINE IntegrationStep_gpu(this,Mesh,Physics,Sources,Fluxes,iter,config,IO)
    CLASS(timedisc_gpu),  INTENT(INOUT) :: this
    CLASS(mesh_base),     INTENT(INOUT) :: Mesh
    CLASS(physics_base),  INTENT(INOUT) :: Physics
    CLASS(sources_list), ALLOCATABLE, INTENT(INOUT) :: Sources
    CLASS(fluxes_base),   INTENT(INOUT) :: Fluxes
    INTEGER                             :: iter
    TYPE(Dict_TYP),       POINTER       :: config,IO
    INTENT(INOUT)                       :: iter
    TYPE(fv_scalars)                    :: sc
    INTEGER(c_int)                      :: rc, brk
    IF (.NOT.C_ASSOCIATED(this%ctx)) &
      CALL this%Error("timedisc_gpu::IntegrationStep","no device context: FirstStep has not run")
    ! an output since the last step moved the host state to the real frame
    IF (Mesh%fargo%GetType().GT.0 .AND. InRealFrame(this%cvar)) THEN
      rc = fv_add_background_velocity(this%ctx)
      IF (rc.NE.0) CALL this%Error("timedisc_gpu::IntegrationStep", c_string(fv_last_error(this%ctx)))
    END IF
    IF (this%dt.LT.this%dtmin .AND. this%dtcause.NE.DTCAUSE_FILEIO) THEN
      this%dtmin = this%dt
      this%dtmincause = this%dtcause
    END IF
    rc = fv_integration_step(this%ctx, REAL(this%dt, c_double), INT(this%dtcause, c_int), brk)
    IF (rc.NE.0) CALL this%Error("timedisc_gpu::IntegrationStep", c_string(fv_last_error(this%ctx)))
    rc = fv_get_scalars(this%ctx, sc)
    iter = sc%iter
    this%n_adj = this%n_adj + (sc%n_adj - this%nadj_seen)
    this%nadj_seen = sc%n_adj
    this%break = brk.NE.0
    this%time = sc%time
    this%dt = sc%dt
    this%dtold = sc%dtold
    this%maxerrold = sc%maxerrold
    IF (this%sync.NE.0) CALL this%SyncHost(Mesh,Fluxes)
  END SUBROUTINE IntegrationStep_gpu

  !> CalcTimestep (timedisc/timedisc_base.f90:818-909): the CFL step of the
  !! last device evaluation
  REAL FUNCTION CalcTimestep_gpu(this,Mesh,Physics,Sources,Fluxes,time,dtcause) RESULT(dt)
    CLASS(timedisc_gpu),  INTENT(INOUT) :: this
    CLASS(mesh_base),     INTENT(IN)    :: Mesh
    CLASS(physics_base),  INTENT(INOUT) :: Physics
    CLASS(sources_list), ALLOCATABLE, INTENT(INOUT) :: Sources
    CLASS(fluxes_base),   INTENT(INOUT) :: Fluxes
    REAL,                 INTENT(IN)    :: time
    INTEGER,              INTENT(INOUT) :: dtcause
    REAL(c_double)                      :: d
    INTEGER(c_int)                      :: cause, rc
    ! an output has just moved the host state to the real frame, and fosite's
    ! CFL step is that of the real-frame pvar
    IF (Mesh%fargo%GetType().GT.0 .AND. InRealFrame(this%cvar)) THEN
      rc = fv_add_background_velocity(this%ctx)
      IF (rc.NE.0) CALL this%Error("timedisc_gpu::CalcTimestep", c_string(fv_last_error(this%ctx)))
    END IF
    rc = fv_calc_timestep(this%ctx, d, cause)
    IF (rc.NE.0) CALL this%Error("timedisc_gpu::CalcTimestep", c_string(fv_last_error(this%ctx)))
    dt = d
    dtcause = cause
  END FUNCTION CalcTimestep_gpu

  !> the device's IntegrationStep runs the whole embedded Runge-Kutta step
  SUBROUTINE SolveODE_gpu(this,Mesh,Physics,Sources,Fluxes,time,dt,err)
    CLASS(timedisc_gpu),  INTENT(INOUT) :: this
    CLASS(mesh_base),     INTENT(IN)    :: Mesh
    CLASS(physics_base),  INTENT(INOUT) :: Physics
    CLASS(sources_list), ALLOCATABLE, INTENT(INOUT) :: Sources
    CLASS(fluxes_base),   INTENT(INOUT) :: Fluxes
    REAL,                 INTENT(IN)    :: time
    REAL,                 INTENT(INOUT) :: dt, err
    CALL this%Error("timedisc_gpu::SolveODE","not used: IntegrationStep runs the step on the device")
  END SUBROUTINE SolveODE_gpu

  SUBROUTINE Finalize_gpu(this)
    CLASS(timedisc_gpu) :: this
    INTEGER(c_int)      :: rc
    IF (C_ASSOCIATED(this%ctx)) rc = fv_destroy(this%ctx)
    this%ctx = c_null_ptr
    CALL this%Finalize_base()
  END SUBROUTINE Finalize_gpu

  !> fosite%Setup (fosite.f90:152-255) with new_timedisc_device for the
  !! Timedisc; everything else as fosite builds it
  SUBROUTINE Setup_device(this)
    CLASS(fosite_device), INTENT(INOUT) :: this
    TYPE(Dict_TYP), POINTER             :: dir, IOdir, config_copy, boundary
    IF (.NOT.this%Initialized()) CALL this%Error("Setup","Sim is uninitialized")
    CALL DeleteDict(this%IO)
    CALL GetAttr(this%config, "mesh", dir)
    IF (ASSOCIATED(dir)) THEN
      NULLIFY(IOdir)
      CALL new_mesh(this%Mesh, dir, IOdir)
      IF (ASSOCIATED(IOdir)) CALL SetAttr(this%IO, "mesh", IOdir)
      IF (this%Mesh%shear_dir.GT.0) CALL SetAttr(this%config,"sources/shearing/stype",SHEARBOX)
    END IF
    CALL GetAttr(this%config, "physics", dir)
    IF (ASSOCIATED(dir)) THEN
      NULLIFY(IOdir)
      CALL new_physics(this%Physics, this%Mesh, dir, IOdir)
      IF (ASSOCIATED(IOdir)) CALL SetAttr(this%IO, "physics", IOdir)
    END IF
    CALL GetAttr(this%config, "fluxes", dir)
    IF (ASSOCIATED(dir)) THEN
      NULLIFY(IOdir)
      CALL new_fluxes(this%Fluxes, this%Mesh, this%Physics, dir, IOdir)
      IF (ASSOCIATED(IOdir)) CALL SetAttr(this%IO, "fluxes", IOdir)
    END IF
    CALL GetAttr(this%config, "timedisc", dir)
    IF (ASSOCIATED(dir)) THEN
      NULLIFY(IOdir)
      CALL new_timedisc_device(this%Timedisc, this%Mesh, this%Physics, dir, IOdir)
      IF (ASSOCIATED(IOdir)) CALL SetAttr(this%IO, "timedisc", IOdir)
    END IF
    CALL GetAttr(this%config, "boundary", dir)
    IF (.NOT.ASSOCIATED(dir)) THEN
      boundary => Dict("empty" / 0)
      CALL SetAttr(this%config, "boundary", boundary)
      CALL GetAttr(this%config, "boundary", dir)
    END IF
    IF (ASSOCIATED(dir)) THEN
      NULLIFY(IOdir)
      CALL new_boundary(this%Timedisc%Boundary, this%Mesh, this%Physics, dir, IOdir)
      IF (ASSOCIATED(IOdir)) CALL SetAttr(this%IO, "boundary", IOdir)
    END IF
    CALL GetAttr(this%config, "sources", dir)
    IF (ASSOCIATED(dir)) THEN
      IF (HasChild(dir)) THEN
        NULLIFY(IOdir)
        ALLOCATE(sources_list::this%Sources)
        CALL this%Sources%InitSources(this%Mesh, this%Physics, this%Fluxes, dir, IOdir)
        IF (ASSOCIATED(IOdir)) CALL SetAttr(this%IO, "sources", IOdir)
      END IF
    END IF
    CALL GetAttr(this%config, "datafile", dir)
    IF (ASSOCIATED(dir)) &
      CALL new_fileio(this%Datafile, this%Mesh, this%Physics, this%Timedisc, this%Sources, dir, this%IO)
    CALL GetAttr(this%config, "logfile", dir)
    IF (ASSOCIATED(dir)) &
      CALL new_fileio(this%Logfile, this%Mesh, this%Physics, this%Timedisc, this%Sources, dir, this%IO)
#ifdef VERSION
    CALL SetAttr(this%config, "version", TRIM(VERSION))
#endif
    CALL CopyDict(this%config, config_copy)
    CALL SetAttr(this%IO, "config", config_copy)
  END SUBROUTINE Setup_device

END MODULE timedisc_gpu_mod
