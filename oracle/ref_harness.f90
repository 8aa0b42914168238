! ref_harness.f90 -- C-callable entry points over the reference's own
! MPI-free Fortran modules (compiled from /root/reference/src by the
! Makefile's `ref` target into oracle/_ref/libwql_ref.so).
!
! TEST INFRASTRUCTURE ONLY: the library this builds is a second checker of the
! CPU restatement (wql_oracle.c); only tests/ load it.  Nothing in the product
! links it.
!
! The modules used are the reference's, unchanged:
!   metrics            (metrics.f90:16-488, D6x / D6y / D6z)
!   anelastic_q4_model (anelastic_q4_model.f90:97-149)
!   anelastic_q8_model (anelastic_q8_model.f90:100-156)
!   rk_vector_kernels  (rk_vector_cpu.f90:7-19)
!   initial_stress_condition (initial_stress_condition.f90:8-254)
! Arrays cross the C boundary as plain arrays, x fastest; the harness copies
! them into arrays with the reference's bounds where a routine needs them.
module wql_ref_harness
  use iso_c_binding
  use common, only : wp
  use metrics, only : D6x, D6y, D6z
  use anelastic_q4_model, only : q4_parameters, build_q4_fixed_coefficients, &
       q4_max_relative_error, q4_relaxation_dt_limit
  use anelastic_q8_model, only : q8_parameters, build_q8_fixed_coefficients, &
       q8_max_relative_error, q8_relaxation_dt_limit
  use rk_vector_kernels, only : scale_rate_array, update_state_array
  use initial_stress_condition, only : initial_stress_tensor
  implicit none

contains

  ! D6x / D6y / D6z (metrics.f90) at every node of u(nx, ny, nz, 3):
  ! du(:, :, :, :, a) = derivative along axis a (1 = q, 2 = r, 3 = s),
  ! each of the three components, with spacing h(a)
  subroutine ref_d6(nx, ny, nz, h, u, du) bind(C, name='ref_d6')
    integer(c_int), value :: nx, ny, nz
    real(c_double), intent(in) :: h(3)
    real(c_double), intent(in) :: u(nx, ny, nz, 3)
    real(c_double), intent(out) :: du(nx, ny, nz, 3, 3)
    integer :: q, r, s
    do s = 1, nz
       do r = 1, ny
          do q = 1, nx
             du(q, r, s, :, 1) = D6x(h(1), u, q, r, s, 1, nx, 1, ny, 1, nz, nx)
             du(q, r, s, :, 2) = D6y(h(2), u, q, r, s, 1, nx, 1, ny, 1, nz, ny)
             du(q, r, s, :, 3) = D6z(h(3), u, q, r, s, 1, nx, 1, ny, 1, nz, nz)
          end do
       end do
    end do
  end subroutine ref_d6

  ! derivatives() (grid.f90:398-471): the coordinate derivatives from the
  ! reference's D6x / D6y / D6z, then J and the metric coefficients by the
  ! cofactor formulas of grid.f90:436-466 (that routine's module needs MPI,
  ! so its dozen lines of algebra are written out here)
  subroutine ref_metrics(nx, ny, nz, X, metx, mety, metz, J) bind(C, name='ref_metrics')
    integer(c_int), value :: nx, ny, nz
    real(c_double), intent(in) :: X(nx, ny, nz, 3)
    real(c_double), intent(out) :: metx(nx, ny, nz, 3), mety(nx, ny, nz, 3), metz(nx, ny, nz, 3)
    real(c_double), intent(out) :: J(nx, ny, nz)
    real(wp) :: X_q(3), X_r(3), X_s(3), hq, hr, hs
    integer :: q, r, s
    hq = 1.0_wp/real(nx-1, wp)
    hr = 1.0_wp/real(ny-1, wp)
    hs = 1.0_wp/real(nz-1, wp)
    do s = 1, nz
       do r = 1, ny
          do q = 1, nx
             X_q = D6x(hq, X, q, r, s, 1, nx, 1, ny, 1, nz, nx)
             X_r = D6y(hr, X, q, r, s, 1, nx, 1, ny, 1, nz, ny)
             X_s = D6z(hs, X, q, r, s, 1, nx, 1, ny, 1, nz, nz)
             J(q, r, s) = X_q(1)*(X_r(2)*X_s(3) - X_r(3)*X_s(2)) &
                  - X_q(2)*(X_r(1)*X_s(3) - X_r(3)*X_s(1)) &
                  + X_q(3)*(X_r(1)*X_s(2) - X_r(2)*X_s(1))
             metx(q, r, s, 1) = (1.0_wp/J(q, r, s))*(X_r(2)*X_s(3) - X_r(3)*X_s(2))
             metx(q, r, s, 2) = (1.0_wp/J(q, r, s))*(X_q(3)*X_s(2) - X_q(2)*X_s(3))
             metx(q, r, s, 3) = (1.0_wp/J(q, r, s))*(X_q(2)*X_r(3) - X_q(3)*X_r(2))
             mety(q, r, s, 1) = (1.0_wp/J(q, r, s))*(X_r(3)*X_s(1) - X_r(1)*X_s(3))
             mety(q, r, s, 2) = (1.0_wp/J(q, r, s))*(X_q(1)*X_s(3) - X_q(3)*X_s(1))
             mety(q, r, s, 3) = (1.0_wp/J(q, r, s))*(X_q(3)*X_r(1) - X_q(1)*X_r(3))
             metz(q, r, s, 1) = (1.0_wp/J(q, r, s))*(X_r(1)*X_s(2) - X_r(2)*X_s(1))
             metz(q, r, s, 2) = (1.0_wp/J(q, r, s))*(X_q(2)*X_s(1) - X_q(1)*X_s(2))
             metz(q, r, s, 3) = (1.0_wp/J(q, r, s))*(X_q(1)*X_r(2) - X_q(2)*X_r(1))
          end do
       end do
    end do
  end subroutine ref_metrics

  ! build_q{4,8}_fixed_coefficients with the default (fixed-q50) parameters;
  ! returns 0, or 1 for an nmech other than 4 / 8
  integer(c_int) function ref_q_coefficients(nmech, tau, weight) bind(C, name='ref_q_coefficients')
    integer(c_int), value :: nmech
    real(c_double), intent(out) :: tau(8), weight(8)
    type(q4_parameters) :: p4
    type(q8_parameters) :: p8
    tau = 0.0_wp
    weight = 0.0_wp
    ref_q_coefficients = 0
    if (nmech == 4) then
       call build_q4_fixed_coefficients(p4, tau(1:4), weight(1:4))
    else if (nmech == 8) then
       call build_q8_fixed_coefficients(p8, tau, weight)
    else
       ref_q_coefficients = 1
    end if
  end function ref_q_coefficients

  ! q{4,8}_max_relative_error: the largest |Q(f)/target - 1| over 256
  ! log-spaced frequencies in [fmin, fmax] of the fixed coefficients
  real(c_double) function ref_q_max_relative_error(nmech, target_Q, fmin, fmax) &
       bind(C, name='ref_q_max_relative_error')
    integer(c_int), value :: nmech
    real(c_double), value :: target_Q, fmin, fmax
    real(wp) :: tau(8), weight(8), err
    integer(c_int) :: rc
    rc = ref_q_coefficients(nmech, tau, weight)
    err = -1.0_wp
    if (nmech == 4) call q4_max_relative_error(target_Q, tau(1:4), weight(1:4), fmin, fmax, err)
    if (nmech == 8) call q8_max_relative_error(target_Q, tau, weight, fmin, fmax, err)
    ref_q_max_relative_error = err
  end function ref_q_max_relative_error

  ! q{4,8}_relaxation_dt_limit with the default parameters
  real(c_double) function ref_q_dt_limit(nmech) bind(C, name='ref_q_dt_limit')
    integer(c_int), value :: nmech
    type(q4_parameters) :: p4
    type(q8_parameters) :: p8
    ref_q_dt_limit = -1.0_wp
    if (nmech == 4) ref_q_dt_limit = q4_relaxation_dt_limit(p4)
    if (nmech == 8) ref_q_dt_limit = q8_relaxation_dt_limit(p8)
  end function ref_q_dt_limit

  ! scale_rate_array / update_state_array (rk_vector_cpu.f90:7-19) on arrays
  ! with the reference's ghost bounds (mb:pb per axis, 9 components): the
  ! values are copied in and out of allocatables with those bounds
  subroutine ref_rk_vector(n1, n2, n3, lo1, lo2, lo3, state, rate, A, inc) &
       bind(C, name='ref_rk_vector')
    integer(c_int), value :: n1, n2, n3, lo1, lo2, lo3
    real(c_double), intent(inout) :: state(n1, n2, n3, 9), rate(n1, n2, n3, 9)
    real(c_double), value :: A, inc
    real(wp), allocatable :: s(:, :, :, :), r(:, :, :, :)
    allocate(s(lo1:lo1+n1-1, lo2:lo2+n2-1, lo3:lo3+n3-1, 1:9))
    allocate(r(lo1:lo1+n1-1, lo2:lo2+n2-1, lo3:lo3+n3-1, 1:9))
    s = state
    r = rate
    call scale_rate_array(r, A)
    call update_state_array(s, r, inc)
    state = s
    rate = r
    deallocate(s, r)
  end subroutine ref_rk_vector

  ! initial_stress_tensor (initial_stress_condition.f90:8-254) at point x for
  ! a problem name of nchar characters
  subroutine ref_initial_stress(nchar, name, x, s0) bind(C, name='ref_initial_stress')
    integer(c_int), value :: nchar
    character(kind=c_char), intent(in) :: name(nchar)
    real(c_double), intent(in) :: x(3)
    real(c_double), intent(out) :: s0(6)
    character(len=64) :: problem
    integer :: k
    problem = ' '
    do k = 1, min(nchar, 64)
       problem(k:k) = name(k)
    end do
    s0 = 0.0_wp
    call initial_stress_tensor(s0, x, trim(problem))
  end subroutine ref_initial_stress

end module wql_ref_harness
