!##############################################################################
!# fosite_fv_c: ISO_C_BINDING interface of libfosite_fv.so                   #
!##############################################################################
!> Fortran 2008 declarations of every entry point of include/fosite_fv.h.
!! The structs are BIND(C) mirrors of fv_config and fv_scalars; arrays cross
!! the boundary as plain pointers in the marray layout (i fastest, then j,
!! then field or face, ghost cells included).
MODULE fosite_fv_c
  USE, INTRINSIC :: iso_c_binding
  IMPLICIT NONE
  PRIVATE

  TYPE, BIND(C), PUBLIC :: fv_config
    INTEGER(c_int) :: physics = 17
    REAL(c_double) :: gamma = 1.4_c_double
    INTEGER(c_int) :: geometry = 1
    INTEGER(c_int) :: inum = 32, jnum = 32
    REAL(c_double) :: xmin = 0, xmax = 1, ymin = 0, ymax = 1, gparam = 1
    INTEGER(c_int) :: bc(4) = 1
    INTEGER(c_int) :: limiter = 1
    REAL(c_double) :: theta = 1
    INTEGER(c_int) :: recvars = 2
    INTEGER(c_int) :: method = 5, order = 5
    REAL(c_double) :: cfl = 0.4_c_double, tol_rel = 0.01_c_double
    REAL(c_double) :: tol_abs(4) = 0.001_c_double
    REAL(c_double) :: stoptime = 1, dtlimit = -1, beta = 0.08_c_double
    INTEGER(c_int) :: fargo = 0, gravity = 0
    REAL(c_double) :: mass = 1, GN = 1
    INTEGER(c_int) :: dtype = 0
    INTEGER(c_int) :: outbound = 0
    REAL(c_double) :: acclimit = -1
    REAL(c_double) :: pmin = -1, rhomin = -1
    INTEGER(c_int) :: count = 1
    REAL(c_double) :: starttime = 0
    !> Mesh%omega and the rotating-frame source: 0 off, 1 on, 2 disable_centaccel
    REAL(c_double) :: omega = 0
    INTEGER(c_int) :: rotframe = 0
    !> viscosity source: 0 off, 1 MOLECULAR, 2 ALPHA, its constants and Courant number
    INTEGER(c_int) :: vismodel = 0
    REAL(c_double) :: dynconst = 0.1_c_double, bulkconst = -2.0_c_double/3.0_c_double*0.1_c_double
    REAL(c_double) :: cvis = 0.5_c_double
  END TYPE fv_config

  TYPE, BIND(C), PUBLIC :: fv_scalars
    INTEGER(c_int) :: iter, n_adj, brk, dtcause
    REAL(c_double) :: time, dt, dtold, maxerrold
  END TYPE fv_scalars

  PUBLIC :: fv_mesh_array, fv_create, fv_destroy, fv_dims, fv_set_sound_speeds, &
       fv_set_background_velocity, fv_set_pvar, fv_set_cvar, fv_first_step, fv_step, fv_run, &
       fv_get_state, fv_get_scalars, fv_last_run_ms, fv_last_error, fv_profile, &
       fv_kernel_stats, fv_get_background_velocity, fv_get_boundary_flux, &
       fv_get_boundary_flux_total, fv_get_point_mass, c_string, &
       fv_first_rhs, fv_integration_step, fv_calc_timestep, fv_add_background_velocity

  INTERFACE
    INTEGER(c_int) FUNCTION fv_mesh_array(cfg, id, out) BIND(C, NAME="fv_mesh_array")
      IMPORT :: c_int, c_double, fv_config
      TYPE(fv_config), INTENT(IN) :: cfg
      INTEGER(c_int), VALUE :: id
      REAL(c_double), INTENT(OUT) :: out(*)
    END FUNCTION fv_mesh_array
    INTEGER(c_int) FUNCTION fv_create(ctx, cfg) BIND(C, NAME="fv_create")
      IMPORT :: c_int, c_ptr, fv_config
      TYPE(c_ptr), INTENT(OUT) :: ctx
      TYPE(fv_config), INTENT(IN) :: cfg
    END FUNCTION fv_create
    INTEGER(c_int) FUNCTION fv_destroy(ctx) BIND(C, NAME="fv_destroy")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
    END FUNCTION fv_destroy
    INTEGER(c_int) FUNCTION fv_dims(ctx, ni, nj, vnum) BIND(C, NAME="fv_dims")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), INTENT(OUT) :: ni, nj, vnum
    END FUNCTION fv_dims
    INTEGER(c_int) FUNCTION fv_set_sound_speeds(ctx, bccsound, fcsound) BIND(C, NAME="fv_set_sound_speeds")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(IN) :: bccsound(*), fcsound(*)
    END FUNCTION fv_set_sound_speeds
    INTEGER(c_int) FUNCTION fv_set_background_velocity(ctx, w) BIND(C, NAME="fv_set_background_velocity")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(IN) :: w(*)
    END FUNCTION fv_set_background_velocity
    INTEGER(c_int) FUNCTION fv_set_pvar(ctx, pvar) BIND(C, NAME="fv_set_pvar")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(IN) :: pvar(*)
    END FUNCTION fv_set_pvar
    INTEGER(c_int) FUNCTION fv_set_cvar(ctx, cvar) BIND(C, NAME="fv_set_cvar")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(IN) :: cvar(*)
    END FUNCTION fv_set_cvar
    INTEGER(c_int) FUNCTION fv_first_step(ctx, brk) BIND(C, NAME="fv_first_step")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), INTENT(OUT) :: brk
    END FUNCTION fv_first_step
    INTEGER(c_int) FUNCTION fv_step(ctx, done) BIND(C, NAME="fv_step")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), INTENT(OUT) :: done
    END FUNCTION fv_step
    INTEGER(c_int) FUNCTION fv_run(ctx, nsteps, taken) BIND(C, NAME="fv_run")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), VALUE :: nsteps
      INTEGER(c_int), INTENT(OUT) :: taken
    END FUNCTION fv_run
    INTEGER(c_int) FUNCTION fv_first_rhs(ctx, brk) BIND(C, NAME="fv_first_rhs")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), INTENT(OUT) :: brk
    END FUNCTION fv_first_rhs
    INTEGER(c_int) FUNCTION fv_integration_step(ctx, dt, dtcause, brk) BIND(C, NAME="fv_integration_step")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), VALUE :: dt
      INTEGER(c_int), VALUE :: dtcause
      INTEGER(c_int), INTENT(OUT) :: brk
    END FUNCTION fv_integration_step
    INTEGER(c_int) FUNCTION fv_calc_timestep(ctx, dt, dtcause) BIND(C, NAME="fv_calc_timestep")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(OUT) :: dt
      INTEGER(c_int), INTENT(OUT) :: dtcause
    END FUNCTION fv_calc_timestep
    INTEGER(c_int) FUNCTION fv_add_background_velocity(ctx) BIND(C, NAME="fv_add_background_velocity")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
    END FUNCTION fv_add_background_velocity
    INTEGER(c_int) FUNCTION fv_get_state(ctx, cvar, pvar) BIND(C, NAME="fv_get_state")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx, cvar, pvar
    END FUNCTION fv_get_state
    INTEGER(c_int) FUNCTION fv_get_scalars(ctx, sc) BIND(C, NAME="fv_get_scalars")
      IMPORT :: c_int, c_ptr, fv_scalars
      TYPE(c_ptr), VALUE :: ctx
      TYPE(fv_scalars), INTENT(OUT) :: sc
    END FUNCTION fv_get_scalars
    INTEGER(c_int) FUNCTION fv_get_background_velocity(ctx, w) BIND(C, NAME="fv_get_background_velocity")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(OUT) :: w(*)
    END FUNCTION fv_get_background_velocity
    INTEGER(c_int) FUNCTION fv_get_boundary_flux(ctx, bx, by) BIND(C, NAME="fv_get_boundary_flux")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx, bx, by
    END FUNCTION fv_get_boundary_flux
    INTEGER(c_int) FUNCTION fv_get_boundary_flux_total(ctx, direction, bflux) &
         BIND(C, NAME="fv_get_boundary_flux_total")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), VALUE :: direction
      REAL(c_double), INTENT(OUT) :: bflux(*)
    END FUNCTION fv_get_boundary_flux_total
    INTEGER(c_int) FUNCTION fv_get_point_mass(ctx, pm) BIND(C, NAME="fv_get_point_mass")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      REAL(c_double), INTENT(OUT) :: pm(*)
    END FUNCTION fv_get_point_mass
    REAL(c_double) FUNCTION fv_last_run_ms(ctx) BIND(C, NAME="fv_last_run_ms")
      IMPORT :: c_double, c_ptr
      TYPE(c_ptr), VALUE :: ctx
    END FUNCTION fv_last_run_ms
    TYPE(c_ptr) FUNCTION fv_last_error(ctx) BIND(C, NAME="fv_last_error")
      IMPORT :: c_ptr
      TYPE(c_ptr), VALUE :: ctx
    END FUNCTION fv_last_error
    INTEGER(c_int) FUNCTION fv_profile(ctx, enable) BIND(C, NAME="fv_profile")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), VALUE :: enable
    END FUNCTION fv_profile
    INTEGER(c_int) FUNCTION fv_kernel_stats(ctx, kernel, launches, ms, alg_bytes) BIND(C, NAME="fv_kernel_stats")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: ctx
      INTEGER(c_int), VALUE :: kernel
      INTEGER(c_int), INTENT(OUT) :: launches
      REAL(c_double), INTENT(OUT) :: ms, alg_bytes
    END FUNCTION fv_kernel_stats
    FUNCTION c_strlen(s) BIND(C, NAME="strlen")
      IMPORT :: c_ptr, c_size_t
      TYPE(c_ptr), VALUE :: s
      INTEGER(c_size_t) :: c_strlen
    END FUNCTION c_strlen
  END INTERFACE

CONTAINS

  !> copy of a NUL-terminated C string
  FUNCTION c_string(p) RESULT(s)
    TYPE(c_ptr), INTENT(IN) :: p
    CHARACTER(LEN=:), ALLOCATABLE :: s
    CHARACTER(KIND=c_char), POINTER :: chars(:)
    INTEGER :: n, k
    IF (.NOT.C_ASSOCIATED(p)) THEN
      s = ""
      RETURN
    END IF
    n = INT(c_strlen(p))
    CALL C_F_POINTER(p, chars, [n])
    ALLOCATE(CHARACTER(LEN=n) :: s)
    DO k = 1, n
      s(k:k) = chars(k)
    END DO
  END FUNCTION c_string

END MODULE fosite_fv_c
