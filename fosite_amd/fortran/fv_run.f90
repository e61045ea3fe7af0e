!##############################################################################
!# fv_run: Fortran 2008 setup program for the GPU finite-volume step         #
!##############################################################################
!> A fosite-style setup program on the GPU host (fosite_gpu_mod): reads the
!! configuration from a namelist file and the initial data from a binary
!! file, runs FirstStep and then k_nsteps steps, either one Step() call at a
!! time or, with k_batch > 0, Run() in device-controlled batches of k_batch
!! steps, and writes the state records.
!!
!! Input:  namelist /fvcfg/ (below); initial data (stream, float64):
!!         pvar(IGMIN:IGMAX,JGMIN:JGMAX,VNUM), then for isothermal physics
!!         bccsound(cells) and fcsound(cells,4), then for k_fargo=2
!!         w(IGMIN:IGMAX).
!! Output: <k_prefix>_steps.bin, one record after FirstStep and after every
!!         k_dumpevery steps (Step mode) or every batch (Run mode):
!!         int32 step, iter, n_adj, break; float64 time, dt, dtold,
!!         maxerrold; cvar; pvar (all cells).
PROGRAM fv_run
  USE fosite_gpu_mod
  IMPLICIT NONE
  INTEGER, PARAMETER :: DP = KIND(1.0D0)
  INTEGER  :: k_physics = 17, k_geometry = 1, k_inum = 32, k_jnum = 32
  REAL(DP) :: k_gamma = 1.4_DP, k_xmin = 0, k_xmax = 1, k_ymin = 0, k_ymax = 1, k_gparam = 1
  INTEGER  :: k_bcw = 1, k_bce = 1, k_bcs = 1, k_bcn = 1
  INTEGER  :: k_limiter = 7, k_recvars = 2
  REAL(DP) :: k_theta = 1
  INTEGER  :: k_method = 5, k_order = 3
  REAL(DP) :: k_cfl = 0.4_DP, k_tolrel = 0.01_DP, k_tolabs(4) = 0.001_DP, k_stoptime = 1
  REAL(DP) :: k_dtlimit = -1, k_beta = 0.08_DP
  INTEGER  :: k_fargo = 0, k_gravity = 0, k_units = 0
  REAL(DP) :: k_mass = 1, k_acclimit = -1
  INTEGER  :: k_outbound = 0
  INTEGER  :: k_nsteps = 10, k_dumpevery = 1, k_batch = 0, k_count = 1
  REAL(DP) :: k_pmin = -1, k_rhomin = -1, k_starttime = 0, k_omega = 0
  INTEGER  :: k_rotframe = 0
  ! viscosity: 0 off, 1 MOLECULAR, 2 ALPHA; k_bulkconst is read only with k_setbulk = 1
  INTEGER  :: k_vismodel = 0, k_setbulk = 0
  REAL(DP) :: k_dynconst = 0.1_DP, k_bulkconst = 0, k_cvis = 0.5_DP
  CHARACTER(LEN=512) :: k_infile = "", k_prefix = "out"
  NAMELIST /fvcfg/ k_physics, k_geometry, k_inum, k_jnum, k_gamma, k_xmin, k_xmax, k_ymin, k_ymax, &
       k_gparam, k_bcw, k_bce, k_bcs, k_bcn, k_limiter, k_recvars, k_theta, k_method, k_order,   &
       k_cfl, k_tolrel, k_tolabs, k_stoptime, k_dtlimit, k_beta, k_fargo, k_gravity, k_units,   &
       k_mass, k_outbound, k_acclimit, k_nsteps, k_dumpevery, k_batch, k_infile, k_prefix, &
       k_count, k_pmin, k_rhomin, k_starttime, k_omega, k_rotframe, k_vismodel, k_setbulk, &
       k_dynconst, k_bulkconst, k_cvis

  TYPE(fosite_gpu) :: Sim
  CHARACTER(LEN=512) :: cfgname
  INTEGER :: u, o, ob, n, k
  LOGICAL :: done
  REAL(DP), ALLOCATABLE :: pv(:,:,:), cs(:,:), fcs(:,:,:), w(:), cv(:,:,:), pvo(:,:,:)
  REAL(DP), ALLOCATABLE :: bx(:,:,:,:), by(:,:,:,:)

  CALL GET_COMMAND_ARGUMENT(1, cfgname)
  OPEN(NEWUNIT=u, FILE=TRIM(cfgname), STATUS="old")
  READ(u, NML=fvcfg)
  CLOSE(u)

  Sim%Mesh%geometry = k_geometry
  Sim%Mesh%inum = k_inum
  Sim%Mesh%jnum = k_jnum
  Sim%Mesh%xmin = k_xmin
  Sim%Mesh%xmax = k_xmax
  Sim%Mesh%ymin = k_ymin
  Sim%Mesh%ymax = k_ymax
  Sim%Mesh%gparam = k_gparam
  Sim%Physics%problem = k_physics
  Sim%Physics%gamma = k_gamma
  Sim%Fluxes%limiter = k_limiter
  Sim%Fluxes%theta = k_theta
  Sim%Fluxes%variables = k_recvars
  Sim%Boundary%western = k_bcw
  Sim%Boundary%eastern = k_bce
  Sim%Boundary%southern = k_bcs
  Sim%Boundary%northern = k_bcn
  Sim%Sources%pointmass = k_gravity.GT.0
  Sim%Sources%mass = k_mass
  Sim%Sources%outbound = k_outbound
  Sim%Sources%acclimit = k_acclimit
  Sim%Sources%omega = k_omega
  Sim%Sources%rotframe = k_rotframe
  Sim%Sources%vismodel = k_vismodel
  Sim%Sources%dynconst = k_dynconst
  ! the reference's default bulk viscosity (sources_viscosity.f90:158)
  Sim%Sources%bulkconst = MERGE(k_bulkconst, -(2.0_DP/3.0_DP)*k_dynconst, k_setbulk.EQ.1)
  Sim%Sources%cvis = k_cvis
  Sim%Timedisc%method = k_method
  Sim%Timedisc%order = k_order
  Sim%Timedisc%cfl = k_cfl
  Sim%Timedisc%tol_rel = k_tolrel
  Sim%Timedisc%tol_abs = k_tolabs
  Sim%Timedisc%stoptime = k_stoptime
  Sim%Timedisc%dtlimit = k_dtlimit
  Sim%Timedisc%beta = k_beta
  Sim%Timedisc%fargo = k_fargo
  Sim%Timedisc%pmin = k_pmin
  Sim%Timedisc%rhomin = k_rhomin
  Sim%Timedisc%starttime = k_starttime
  Sim%Timedisc%count = k_count
  CALL Sim%Setup()

  ASSOCIATE(M => Sim%Mesh, V => Sim%Physics%VNUM)
    ALLOCATE(pv(M%IGMIN:M%IGMAX, M%JGMIN:M%JGMAX, V), cv(M%IGMIN:M%IGMAX, M%JGMIN:M%JGMAX, V), &
             pvo(M%IGMIN:M%IGMAX, M%JGMIN:M%JGMAX, V))
    OPEN(NEWUNIT=u, FILE=TRIM(k_infile), ACCESS="stream", FORM="unformatted", STATUS="old")
    READ(u) pv
    IF (k_physics.EQ.EULER_ISOTHERM) THEN
      ALLOCATE(cs(M%IGMIN:M%IGMAX, M%JGMIN:M%JGMAX), fcs(M%IGMIN:M%IGMAX, M%JGMIN:M%JGMAX, 4))
      READ(u) cs
      READ(u) fcs
      CALL Sim%SetSoundSpeeds(cs, fcs)
    END IF
    IF (k_fargo.EQ.2) THEN
      ALLOCATE(w(M%IGMIN:M%IGMAX))
      READ(u) w
      CALL Sim%SetBackgroundVelocity(w)
    END IF
    CLOSE(u)
  END ASSOCIATE
  CALL Sim%InitData(pv)

  OPEN(NEWUNIT=o, FILE=TRIM(k_prefix) // "_steps.bin", ACCESS="stream", FORM="unformatted", &
       STATUS="replace")
  OPEN(NEWUNIT=ob, FILE=TRIM(k_prefix) // "_bflux.bin", ACCESS="stream", FORM="unformatted", &
       STATUS="replace")
  ASSOCIATE(M => Sim%Mesh, V => Sim%Physics%VNUM)
    ALLOCATE(bx(M%JGMIN:M%JGMAX, 1, 2, V), by(1, M%IGMIN:M%IGMAX, 2, V))
  END ASSOCIATE
  CALL Sim%FirstStep()
  CALL Dump(0)
  n = 0
  done = .FALSE.
  DO WHILE (n.LT.k_nsteps .AND. .NOT.done)
    IF (k_batch.GT.0) THEN
      k = Sim%Run(MIN(k_batch, k_nsteps - n))
      done = k.LT.MIN(k_batch, k_nsteps - n)
      n = n + k
      CALL Dump(n)
    ELSE
      done = Sim%Step()
      IF (.NOT.done) n = n + 1
      IF (MOD(n, k_dumpevery).EQ.0 .OR. n.EQ.k_nsteps .OR. done) CALL Dump(n)
    END IF
  END DO
  CLOSE(o)
  CLOSE(ob)
  WRITE(*,'(A,I0,A,ES24.16)') "FV_RUN_STEPS ", n, " TIME ", Sim%Timedisc%time
  CALL Sim%Finalize()

CONTAINS

  SUBROUTINE Dump(step)
    INTEGER, INTENT(IN) :: step
    INTEGER(KIND=4) :: ibuf(4)
    REAL(KIND=8)    :: rbuf(4)
    CALL Sim%GetState(cv, pvo)
    ibuf = [step, Sim%iter, Sim%Timedisc%n_adj, MERGE(1, 0, Sim%Timedisc%break)]
    rbuf = [Sim%Timedisc%time, Sim%Timedisc%dt, Sim%Timedisc%dtold, Sim%Timedisc%maxerrold]
    WRITE(o) ibuf, rbuf
    WRITE(o) cv
    WRITE(o) pvo
    ! boundary fluxes and their totals per boundary (WEST, EAST, SOUTH, NORTH)
    CALL Sim%GetBoundaryFluxes(bx, by)
    WRITE(ob) bx, by
    WRITE(ob) Sim%GetBoundaryFlux(1), Sim%GetBoundaryFlux(2), Sim%GetBoundaryFlux(3), &
              Sim%GetBoundaryFlux(4)
    ! point mass: mass, accrate, massloss, mdot
    IF (k_gravity.GT.0) WRITE(ob) Sim%GetPointMass()
  END SUBROUTINE Dump

END PROGRAM fv_run
