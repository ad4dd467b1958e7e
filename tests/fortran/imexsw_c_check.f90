!> Exercises the Fortran interface module imexsw_c without a GPU: prints the
!> size and byte offset of every field of the BIND(C) types (the Python test
!> compares them with the C header's layout), calls the host-only entry
!> points (version string, a create that fails validation and its message,
!> a .q_2d write and read back) and reports what it found.
PROGRAM imexsw_c_check

  USE, INTRINSIC :: iso_c_binding
  USE imexsw_c

  IMPLICIT NONE

  TYPE(imexsw_params), TARGET :: p
  TYPE(imexsw_stats), TARGET :: st
  TYPE(c_ptr) :: ctx
  INTEGER(c_int) :: rc
  REAL(c_double) :: q(5, 3, 2), q2(5, 3, 2), x(3), y(2), gx(5), gw(5)
  INTEGER :: i, j, k
  CHARACTER(len=*), PARAMETER :: path = 'imexsw_c_check.q_2d'

  WRITE(*, '(A,I0)') 'sizeof imexsw_params ', C_SIZEOF(p)
  CALL off('nx', C_LOC(p%nx))
  CALL off('ny', C_LOC(p%ny))
  CALL off('x0', C_LOC(p%x0))
  CALL off('cell_size', C_LOC(p%cell_size))
  CALL off('n_solid', C_LOC(p%n_solid))
  CALL off('n_rk', C_LOC(p%n_rk))
  CALL off('energy_flag', C_LOC(p%energy_flag))
  CALL off('erodible_deposit_flag', C_LOC(p%erodible_deposit_flag))
  CALL off('cfl', C_LOC(p%cfl))
  CALL off('theta', C_LOC(p%theta))
  CALL off('limiter', C_LOC(p%limiter))
  CALL off('bc_flag', C_LOC(p%bc_flag))
  CALL off('bc_value', C_LOC(p%bc_value))
  CALL off('grav', C_LOC(p%grav))
  CALL off('kin_visc_c', C_LOC(p%kin_visc_c))
  CALL off('rho_s', C_LOC(p%rho_s))
  CALL off('sp_heat_s', C_LOC(p%sp_heat_s))
  CALL off('mu', C_LOC(p%mu))
  CALL off('xi', C_LOC(p%xi))
  CALL off('erosion_coeff', C_LOC(p%erosion_coeff))
  CALL off('erodible_fract', C_LOC(p%erodible_fract))
  CALL off('maximum_solid_packing', C_LOC(p%maximum_solid_packing))
  CALL off('x_source', C_LOC(p%x_source))
  CALL off('alphas_source', C_LOC(p%alphas_source))
  CALL off('time_param', C_LOC(p%time_param))
  CALL off('n_thickness_levels', C_LOC(p%n_thickness_levels))
  CALL off('n_dyn_pres_levels', C_LOC(p%n_dyn_pres_levels))
  CALL off('thickness_levels', C_LOC(p%thickness_levels))
  CALL off('dyn_pres_levels', C_LOC(p%dyn_pres_levels))
  CALL off('vertical_profiles_flag', C_LOC(p%vertical_profiles_flag))
  CALL off('n_quad', C_LOC(p%n_quad))
  CALL off('vonK', C_LOC(p%vonK))
  CALL off('z_dyn', C_LOC(p%z_dyn))
  CALL off('friction_factor', C_LOC(p%friction_factor))
  ! bc_flag(i, side) is C bc_flag[side][i]: side 2 (east), variable 3
  CALL off('bc_flag[1][2]', C_LOC(p%bc_flag(3, 2)))
  CALL off('bc_value[3][0]', C_LOC(p%bc_value(1, 4)))

  WRITE(*, '(A,I0)') 'sizeof imexsw_stats ', C_SIZEOF(st)
  CALL soff('steps', C_LOC(st%steps))
  CALL soff('line_search_evals', C_LOC(st%line_search_evals))
  CALL soff('newton_iter_hist', C_LOC(st%newton_iter_hist))
  CALL soff('newton_iter_max', C_LOC(st%newton_iter_max))
  CALL soff('sparse_steps', C_LOC(st%sparse_steps))
  CALL soff('solve_cells', C_LOC(st%solve_cells))
  CALL soff('compute_cells', C_LOC(st%compute_cells))

  WRITE(*, '(A,A)') 'version ', imexsw_string(imexsw_version())

  ! validation happens on the host: a one-cell grid is refused
  p%nx = 1
  p%ny = 1
  p%n_solid = 1
  p%n_rk = 2
  rc = imexsw_create(p, ctx)
  WRITE(*, '(A,I0,A,A)') 'create rc ', rc, ' message ', imexsw_string(imexsw_last_error(C_NULL_PTR))

  DO k = 1, 2
     DO j = 1, 3
        DO i = 1, 5
           q(i, j, k) = REAL(i, c_double) * 1.25_c_double + REAL(j, c_double)          &
                + 10.0_c_double * REAL(k, c_double)
        END DO
     END DO
  END DO
  q(1, 2, 2) = -1.0_c_double      ! read_solution empties a cell with q(1) <= 0
  x = [0.5_c_double, 1.5_c_double, 2.5_c_double]
  y = [0.5_c_double, 1.5_c_double]
  rc = imexsw_q2d_write(TRIM(path) // C_NULL_CHAR, 5, 3, 2, x, y, q)
  WRITE(*, '(A,I0)') 'q2d_write rc ', rc
  q2 = -7.0_c_double
  rc = imexsw_q2d_read(TRIM(path) // C_NULL_CHAR, 5, 1, 3, 2, q2)
  WRITE(*, '(A,I0)') 'q2d_read rc ', rc
  q(:, 2, 2) = 0.0_c_double
  WRITE(*, '(A,ES12.4)') 'q2d max abs diff ', MAXVAL(ABS(q2 - q))

  ! the quadrature rule of the vertical profiles (host only), printed to the
  ! last bit
  rc = imexsw_gauss_legendre(5, gx, gw)
  WRITE(*, '(A,I0)') 'gauss_legendre rc ', rc
  DO i = 1, 5
     WRITE(*, '(A,I0,A,Z16.16,A,Z16.16)') 'gl ', i, ' ', gx(i), ' ', gw(i)
  END DO

CONTAINS

  SUBROUTINE off(name, addr)
    CHARACTER(*), INTENT(IN) :: name
    TYPE(c_ptr), INTENT(IN) :: addr
    WRITE(*, '(A,A,A,I0)') 'offset params ', name, ' ',                             &
         TRANSFER(addr, 0_c_intptr_t) - TRANSFER(C_LOC(p), 0_c_intptr_t)
  END SUBROUTINE off

  SUBROUTINE soff(name, addr)
    CHARACTER(*), INTENT(IN) :: name
    TYPE(c_ptr), INTENT(IN) :: addr
    WRITE(*, '(A,A,A,I0)') 'offset stats ', name, ' ',                              &
         TRANSFER(addr, 0_c_intptr_t) - TRANSFER(C_LOC(st), 0_c_intptr_t)
  END SUBROUTINE soff

END PROGRAM imexsw_c_check
