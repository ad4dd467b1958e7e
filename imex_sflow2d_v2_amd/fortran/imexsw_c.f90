!> iso_c_binding interface of libimexsw (include/imexsw.h) for the Fortran
!> host of IMEX_SfloW2D.  A replacement module solver_2d keeps the public
!> names of src/solver_2d.f90 and forwards its procedures to these entry
!> points (check_solve -> imexsw_check_solve, timestep -> imexsw_timestep,
!> imex_RK_solver -> imexsw_rk_step, update_erosion_deposition_cell ->
!> imexsw_erosion_deposition, the main loop between two output times ->
!> imexsw_advance).
!>
!> Arrays cross in the host's own layout: q(n_vars, nx, ny), per-cell fields
!> (nx, ny), staggered x fields (nx+1, ny), y fields (nx, ny+1).  Optional
!> arrays are passed as C_LOC(array) or C_NULL_PTR.  LOGICAL flags of the
!> parameter modules become 0/1 integers in imexsw_params.
MODULE imexsw_c

  USE, INTRINSIC :: iso_c_binding

  IMPLICIT NONE

  INTEGER, PARAMETER :: IMEXSW_MAX_SOLID = 6
  INTEGER, PARAMETER :: IMEXSW_MAX_VARS = 4 + IMEXSW_MAX_SOLID
  INTEGER, PARAMETER :: IMEXSW_MAX_LEVELS = 8
  INTEGER, PARAMETER :: IMEXSW_MAX_QUAD = 64

  !> sides of bc_flag / bc_value
  INTEGER, PARAMETER :: IMEXSW_W = 1, IMEXSW_E = 2, IMEXSW_S = 3, IMEXSW_N = 4

  !> source-ring direction bits (geometry_2d sourceE/W/S/N)
  INTEGER(c_int8_t), PARAMETER :: IMEXSW_SRC_E = 1_c_int8_t, IMEXSW_SRC_W = 2_c_int8_t,  &
       IMEXSW_SRC_S = 4_c_int8_t, IMEXSW_SRC_N = 8_c_int8_t

  !> struct imexsw_params, field for field.  The C array bc_flag[4][MAX_VARS]
  !> is bc_flag(MAX_VARS, 4) here: bc_flag(i, side) is variable i on side.
  TYPE, BIND(C) :: imexsw_params
     INTEGER(c_int32_t) :: nx, ny
     REAL(c_double) :: x0, y0, cell_size
     INTEGER(c_int32_t) :: n_solid, n_rk
     INTEGER(c_int32_t) :: energy_flag, alpha_flag, slope_correction_flag
     INTEGER(c_int32_t) :: curvature_term_flag, topo_change_flag, radial_source_flag
     INTEGER(c_int32_t) :: rheology_flag, rheology_model, settling_flag
     INTEGER(c_int32_t) :: entrainment_flag, erodible_deposit_flag
     REAL(c_double) :: cfl, max_dt, reconstr_coeff, theta
     INTEGER(c_int32_t) :: limiter(IMEXSW_MAX_VARS + 2)
     INTEGER(c_int32_t) :: bc_flag(IMEXSW_MAX_VARS, 4)
     REAL(c_double) :: bc_value(IMEXSW_MAX_VARS, 4)
     REAL(c_double) :: grav, pres, T_ambient, sp_heat_a, sp_gas_const_a, kin_visc_c
     REAL(c_double) :: rho_s(IMEXSW_MAX_SOLID), diam_s(IMEXSW_MAX_SOLID)
     REAL(c_double) :: sp_heat_s(IMEXSW_MAX_SOLID)
     REAL(c_double) :: mu, xi
     REAL(c_double) :: erosion_coeff, erodible_porosity
     REAL(c_double) :: erodible_fract(IMEXSW_MAX_SOLID)
     REAL(c_double) :: T_erodible, alphastot_min, maximum_solid_packing
     REAL(c_double) :: x_source, y_source, r_source, h_source, vel_source, T_source
     REAL(c_double) :: alphas_source(IMEXSW_MAX_SOLID), time_param(4)
     INTEGER(c_int32_t) :: n_thickness_levels, n_dyn_pres_levels
     REAL(c_double) :: thickness_levels(IMEXSW_MAX_LEVELS)
     REAL(c_double) :: dyn_pres_levels(IMEXSW_MAX_LEVELS)
     INTEGER(c_int32_t) :: vertical_profiles_flag, bottom_conc_flag, n_quad
     REAL(c_double) :: vonK, k_s, Sc, z_dyn, friction_factor
  END TYPE imexsw_params

  !> struct imexsw_stats: cumulative counters of a context
  TYPE, BIND(C) :: imexsw_stats
     INTEGER(c_int64_t) :: steps, cell_updates_active, cell_updates_grid
     INTEGER(c_int64_t) :: cell_updates_computed, newton_cells, newton_iters
     INTEGER(c_int64_t) :: newton_nonconverged, nan_events, neg_h_events
     INTEGER(c_int64_t) :: neg_solid_events, sum_solid_events, low_T_events
     INTEGER(c_int64_t) :: line_search_evals
     INTEGER(c_int64_t) :: newton_iter_hist(8)
     INTEGER(c_int64_t) :: newton_iter_max, warm_cells, warm_fallbacks, sparse_steps
     INTEGER(c_int32_t) :: solve_cells, compute_cells
  END TYPE imexsw_stats

  INTERFACE

     TYPE(c_ptr) FUNCTION imexsw_version() BIND(C, name='imexsw_version')
       IMPORT :: c_ptr
     END FUNCTION imexsw_version

     TYPE(c_ptr) FUNCTION imexsw_last_error(ctx) BIND(C, name='imexsw_last_error')
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION imexsw_last_error

     !> allocate_solver_variables (src/solver_2d.f90:263-529)
     INTEGER(c_int) FUNCTION imexsw_create(p, ctx) BIND(C, name='imexsw_create')
       IMPORT :: c_int, c_ptr, imexsw_params
       TYPE(imexsw_params), INTENT(IN) :: p
       TYPE(c_ptr), INTENT(OUT) :: ctx
     END FUNCTION imexsw_create

     !> deallocate_solver_variables (src/solver_2d.f90:543-657)
     SUBROUTINE imexsw_destroy(ctx) BIND(C, name='imexsw_destroy')
       IMPORT :: c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END SUBROUTINE imexsw_destroy

     !> topography_reconstruction from B_cent (src/geometry_2d.f90:535-860);
     !> nodata: C_LOC of an (nx, ny) c_int8_t array or C_NULL_PTR
     INTEGER(c_int) FUNCTION imexsw_set_topography(ctx, B_cent, nodata)          &
          BIND(C, name='imexsw_set_topography')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), INTENT(IN) :: B_cent(*)
       TYPE(c_ptr), VALUE :: nodata
     END FUNCTION imexsw_set_topography

     !> the host's own geometry_2d fields (any but B_cent may be C_NULL_PTR)
     INTEGER(c_int) FUNCTION imexsw_set_geometry(ctx, B_cent, Bx, By, Bxx, Bxy,  &
          Byy, grav_coeff, grav_stag_x, grav_stag_y, dgrav_dx, nodata)          &
          BIND(C, name='imexsw_set_geometry')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), INTENT(IN) :: B_cent(*)
       TYPE(c_ptr), VALUE :: Bx, By, Bxx, Bxy, Byy, grav_coeff, grav_stag_x
       TYPE(c_ptr), VALUE :: grav_stag_y, dgrav_dx, nodata
     END FUNCTION imexsw_set_geometry

     !> init_source's ring (src/geometry_2d.f90:872-1102): source_cell(nx,ny),
     !> direction bits (nx,ny), vect(8,nx,ny) = (E_x,E_y,W_x,W_y,S_x,S_y,N_x,N_y)
     INTEGER(c_int) FUNCTION imexsw_set_source(ctx, source_cell, dir_bits, vect)  &
          BIND(C, name='imexsw_set_source')
       IMPORT :: c_int, c_ptr, c_int32_t, c_int8_t, c_double
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int32_t), INTENT(IN) :: source_cell(*)
       INTEGER(c_int8_t), INTENT(IN) :: dir_bits(*)
       REAL(c_double), INTENT(IN) :: vect(*)
     END FUNCTION imexsw_set_source

     !> the module array q (read_solution, init_2d); starts a new run
     INTEGER(c_int) FUNCTION imexsw_upload_q(ctx, q) BIND(C, name='imexsw_upload_q')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), INTENT(IN) :: q(*)
     END FUNCTION imexsw_upload_q

     !> erodible(n_solid,nx,ny), deposit and erosion (nx,ny,n_solid) or C_NULL_PTR
     INTEGER(c_int) FUNCTION imexsw_upload_erosion(ctx, erodible, deposit, erosion) &
          BIND(C, name='imexsw_upload_erosion')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, erodible, deposit, erosion
     END FUNCTION imexsw_upload_erosion

     !> check_solve(solve_all) (src/solver_2d.f90:666-857)
     INTEGER(c_int) FUNCTION imexsw_check_solve(ctx, solve_all, t, solve_cells)  &
          BIND(C, name='imexsw_check_solve')
       IMPORT :: c_int, c_ptr, c_double, c_int32_t
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: solve_all
       REAL(c_double), VALUE :: t
       INTEGER(c_int32_t), INTENT(OUT) :: solve_cells
     END FUNCTION imexsw_check_solve

     !> timestep (src/solver_2d.f90:872-1005)
     INTEGER(c_int) FUNCTION imexsw_timestep(ctx, dt) BIND(C, name='imexsw_timestep')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), INTENT(OUT) :: dt
     END FUNCTION imexsw_timestep

     !> imex_RK_solver (src/solver_2d.f90:1021-1632)
     INTEGER(c_int) FUNCTION imexsw_rk_step(ctx, t, dt) BIND(C, name='imexsw_rk_step')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), VALUE :: t, dt
     END FUNCTION imexsw_rk_step

     !> update_erosion_deposition_cell(dt) (src/solver_2d.f90:2380-2595)
     INTEGER(c_int) FUNCTION imexsw_erosion_deposition(ctx, dt)                   &
          BIND(C, name='imexsw_erosion_deposition')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), VALUE :: dt
     END FUNCTION imexsw_erosion_deposition

     !> the per-step topography_reconstruction (src/IMEX_SfloW2D.f90:387)
     INTEGER(c_int) FUNCTION imexsw_topography_reconstruction(ctx)                &
          BIND(C, name='imexsw_topography_reconstruction')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION imexsw_topography_reconstruction

     !> the post-step loop and sums (src/IMEX_SfloW2D.f90:391-483): sums =
     !> SUM(q(1,:,:)), SUM(qp(1,:,:)), COUNT(q(1,:,:) > 1e-7), solid mass,
     !> each still to be multiplied by dx*dy; C_NULL_PTR for no sums
     INTEGER(c_int) FUNCTION imexsw_post_step(ctx, sums) BIND(C, name='imexsw_post_step')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, sums
     END FUNCTION imexsw_post_step

     !> the main loop (src/IMEX_SfloW2D.f90:337-531) for up to nsteps steps,
     !> until t reaches t_output (then the host writes output, raises
     !> t_output and calls again) or t_end
     INTEGER(c_int) FUNCTION imexsw_advance(ctx, nsteps, dt0, t, t_end, t_output, &
          steps_done) BIND(C, name='imexsw_advance')
       IMPORT :: c_int, c_ptr, c_int32_t, c_double
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int32_t), VALUE :: nsteps
       REAL(c_double), VALUE :: dt0
       REAL(c_double), INTENT(INOUT) :: t
       REAL(c_double), VALUE :: t_end, t_output
       INTEGER(c_int32_t), INTENT(OUT) :: steps_done
     END FUNCTION imexsw_advance

     !> q(n_vars,nx,ny), qp(n_vars+2,nx,ny), hmax, pdynmax, mod_vel_max
     !> (nx,ny), vuln(n_thk*n_pdyn,nx,ny) as 0/1 bytes; each C_LOC or null
     INTEGER(c_int) FUNCTION imexsw_download(ctx, q, qp, hmax, pdynmax,          &
          mod_vel_max, vuln) BIND(C, name='imexsw_download')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, q, qp, hmax, pdynmax, mod_vel_max, vuln
     END FUNCTION imexsw_download

     !> erodible(n_solid,nx,ny), deposit, erosion (nx,ny,n_solid), B_cent
     INTEGER(c_int) FUNCTION imexsw_download_erosion(ctx, erodible, deposit,     &
          erosion, B_cent) BIND(C, name='imexsw_download_erosion')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx, erodible, deposit, erosion, B_cent
     END FUNCTION imexsw_download_erosion

     !> solve_mask (src/solver_2d.f90:159) as 0/1 bytes
     INTEGER(c_int) FUNCTION imexsw_download_mask(ctx, mask)                      &
          BIND(C, name='imexsw_download_mask')
       IMPORT :: c_int, c_ptr, c_int8_t
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int8_t), INTENT(OUT) :: mask(*)
     END FUNCTION imexsw_download_mask

     !> the module variable dt after the last step
     INTEGER(c_int) FUNCTION imexsw_get_dt(ctx, dt) BIND(C, name='imexsw_get_dt')
       IMPORT :: c_int, c_ptr, c_double
       TYPE(c_ptr), VALUE :: ctx
       REAL(c_double), INTENT(OUT) :: dt
     END FUNCTION imexsw_get_dt

     INTEGER(c_int) FUNCTION imexsw_stats_get(ctx, stats, reset)                  &
          BIND(C, name='imexsw_stats_get')
       IMPORT :: c_int, c_ptr, imexsw_stats
       TYPE(c_ptr), VALUE :: ctx
       TYPE(imexsw_stats), INTENT(OUT) :: stats
       INTEGER(c_int), VALUE :: reset
     END FUNCTION imexsw_stats_get

     INTEGER(c_int) FUNCTION imexsw_set_tuning(ctx, max_strip_rows, target_strips) &
          BIND(C, name='imexsw_set_tuning')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int), VALUE :: max_strip_rows, target_strips
     END FUNCTION imexsw_set_tuning

     INTEGER(c_int) FUNCTION imexsw_sync(ctx) BIND(C, name='imexsw_sync')
       IMPORT :: c_int, c_ptr
       TYPE(c_ptr), VALUE :: ctx
     END FUNCTION imexsw_sync

     !> output_solution's .q_2d file (src/inpout_2d.f90:5766-5891); the path
     !> is TRIM(name)//C_NULL_CHAR
     INTEGER(c_int) FUNCTION imexsw_q2d_write(path, nv, nx, ny, x_comp, y_comp, q) &
          BIND(C, name='imexsw_q2d_write')
       IMPORT :: c_int, c_char, c_int32_t, c_double
       CHARACTER(KIND=c_char), INTENT(IN) :: path(*)
       INTEGER(c_int32_t), VALUE :: nv, nx, ny
       REAL(c_double), INTENT(IN) :: x_comp(*), y_comp(*), q(*)
     END FUNCTION imexsw_q2d_write

     !> read_solution's .q_2d branch (src/inpout_2d.f90:5432-5453)
     INTEGER(c_int) FUNCTION imexsw_q2d_read(path, nv, n_solid, nx, ny, q)         &
          BIND(C, name='imexsw_q2d_read')
       IMPORT :: c_int, c_char, c_int32_t, c_double
       CHARACTER(KIND=c_char), INTENT(IN) :: path(*)
       INTEGER(c_int32_t), VALUE :: nv, n_solid, nx, ny
       REAL(c_double), INTENT(OUT) :: q(*)
     END FUNCTION imexsw_q2d_read

     !> qp_to_qc / qc_to_qp of n packed states with the context's model
     !> (init_2d's conversion of the initial physical states)
     INTEGER(c_int) FUNCTION imexsw_qp_to_qc(ctx, n, qp, qc)                     &
          BIND(C, name='imexsw_qp_to_qc')
       IMPORT :: c_int, c_ptr, c_int64_t, c_double
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int64_t), VALUE :: n
       REAL(c_double), INTENT(IN) :: qp(*)
       REAL(c_double), INTENT(OUT) :: qc(*)
     END FUNCTION imexsw_qp_to_qc

     INTEGER(c_int) FUNCTION imexsw_qc_to_qp(ctx, n, qc, qp)                     &
          BIND(C, name='imexsw_qc_to_qp')
       IMPORT :: c_int, c_ptr, c_int64_t, c_double
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int64_t), VALUE :: n
       REAL(c_double), INTENT(IN) :: qc(*)
       REAL(c_double), INTENT(OUT) :: qp(*)
     END FUNCTION imexsw_qc_to_qp

     !> eval_flux_coeffs of n physical states (read_param's MFR_SOURCE to
     !> H_SOURCE with vertical profiles)
     INTEGER(c_int) FUNCTION imexsw_flux_coeffs(ctx, n, qp, Bx, By, shape)      &
          BIND(C, name='imexsw_flux_coeffs')
       IMPORT :: c_int, c_ptr, c_int64_t, c_double
       TYPE(c_ptr), VALUE :: ctx
       INTEGER(c_int64_t), VALUE :: n
       REAL(c_double), INTENT(IN) :: qp(*), Bx(*), By(*)
       REAL(c_double), INTENT(OUT) :: shape(*)
     END FUNCTION imexsw_flux_coeffs

     !> lambertw0 (branch 0) / lambertwm1 (branch -1) of n arguments
     INTEGER(c_int) FUNCTION imexsw_lambertw(branch, n, x, w)                   &
          BIND(C, name='imexsw_lambertw')
       IMPORT :: c_int, c_int32_t, c_int64_t, c_double
       INTEGER(c_int32_t), VALUE :: branch
       INTEGER(c_int64_t), VALUE :: n
       REAL(c_double), INTENT(IN) :: x(*)
       REAL(c_double), INTENT(OUT) :: w(*)
     END FUNCTION imexsw_lambertw

     !> gaulegf(-1, 1, x, w, n)
     INTEGER(c_int) FUNCTION imexsw_gauss_legendre(n, x, w)                      &
          BIND(C, name='imexsw_gauss_legendre')
       IMPORT :: c_int, c_int32_t, c_double
       INTEGER(c_int32_t), VALUE :: n
       REAL(c_double), INTENT(OUT) :: x(*), w(*)
     END FUNCTION imexsw_gauss_legendre

  END INTERFACE

CONTAINS

  !> A C string returned by the library as a Fortran string.
  FUNCTION imexsw_string(p) RESULT(s)
    TYPE(c_ptr), INTENT(IN) :: p
    CHARACTER(len=:), ALLOCATABLE :: s
    CHARACTER(KIND=c_char), POINTER :: chars(:)
    INTEGER :: n
    s = ''
    IF ( .NOT. C_ASSOCIATED(p) ) RETURN
    CALL C_F_POINTER(p, chars, [4096])
    n = 0
    DO WHILE ( n < 4096 )
       IF ( chars(n + 1) == C_NULL_CHAR ) EXIT
       n = n + 1
    END DO
    ALLOCATE(CHARACTER(len=n) :: s)
    s = TRANSFER(chars(1:n), s)
  END FUNCTION imexsw_string

END MODULE imexsw_c
