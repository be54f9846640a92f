! ref_bind.f90 -- bind(C) entry points onto the reference's own pure Fortran
! modules (precision, mathconstants, matrixoperations, su3, su2,
! ranlxd_generator), compiled from /root/reference/src by oracle/build_ref.sh
! into oracle/_ref/libnrqcd_ref.so.
!
! TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's RNG stream,
! generator table, algebra maps and group exponential against the
! reference's compiled code.  Nothing in the product loads it.
module ref_bind
  use, intrinsic :: iso_c_binding
  use, intrinsic :: iso_fortran_env
  use precision, only: fp
  implicit none
contains
  ! ranlxd_generator.f90:251-313
  subroutine ref_ranlxd_init(level, seed) bind(C, name='ref_ranlxd_init')
    use ranlxd_generator, only: ranlxd_init
    integer(c_int64_t), value :: level, seed
    integer(int64) :: l, s
    l = level
    s = seed
    call ranlxd_init(l, s)
  end subroutine ref_ranlxd_init

  ! ranlxd_generator.f90:319-349
  subroutine ref_ranlxd(r, n) bind(C, name='ref_ranlxd')
    use ranlxd_generator, only: ranlxd
    integer(c_int64_t), value :: n
    real(c_double), intent(out) :: r(n)
    call ranlxd(r)
  end subroutine ref_ranlxd

  ! ranlxd_generator.f90:355-382
  subroutine ref_ranlxd_get(state) bind(C, name='ref_ranlxd_get')
    use ranlxd_generator, only: ranlxd_get
    integer(c_int64_t), intent(out) :: state(25)
    integer(int64) :: st(25)
    call ranlxd_get(st)
    state = st
  end subroutine ref_ranlxd_get

  ! su3.f90:97-131 via GetGenerator (su3.f90:105-120)
  subroutine ref_su3_generators(t) bind(C, name='ref_su3_generators')
    use su3, only: GetGenerator
    complex(c_double_complex), intent(out) :: t(3, 3, 8)
    integer(int8) :: a
    do a = 1, 8
       t(:, :, a) = GetGenerator(a)
    end do
  end subroutine ref_su3_generators

  ! su2.f90:68-88 via GetGenerator
  subroutine ref_su2_generators(t) bind(C, name='ref_su2_generators')
    use su2, only: GetGenerator
    complex(c_double_complex), intent(out) :: t(2, 2, 3)
    integer(int8) :: a
    do a = 1, 3
       t(:, :, a) = GetGenerator(a)
    end do
  end subroutine ref_su2_generators

  ! su3.f90:235-251 (GetGroupExp_fromAlgebraCoordinates)
  subroutine ref_group_exp(alpha, u) bind(C, name='ref_group_exp')
    use su3, only: GetGroupExp
    real(c_double), intent(in) :: alpha(8)
    complex(c_double_complex), intent(out) :: u(3, 3)
    real(fp) :: al(8)
    al = alpha
    u = GetGroupExp(al)
  end subroutine ref_group_exp

  ! su3.f90:287-301
  subroutine ref_algebra_matrix(alpha, m) bind(C, name='ref_algebra_matrix')
    use su3, only: GetAlgebraMatrix
    real(c_double), intent(in) :: alpha(8)
    complex(c_double_complex), intent(out) :: m(3, 3)
    real(fp) :: al(8)
    al = alpha
    m = GetAlgebraMatrix(al)
  end subroutine ref_algebra_matrix

  ! su3.f90:349-357
  function ref_algebra_coordinate(a, m) result(res) bind(C, name='ref_algebra_coordinate')
    use su3, only: GetAlgebraCoordinate
    integer(c_int), value :: a
    complex(c_double_complex), intent(in) :: m(3, 3)
    real(c_double) :: res
    complex(fp) :: mm(3, 3)
    integer(int8) :: ai
    mm = m
    ai = int(a, int8)
    res = GetAlgebraCoordinate(ai, mm)
  end function ref_algebra_coordinate

  ! su3.f90:369-381
  subroutine ref_trace_with_generator(a, m, res) bind(C, name='ref_trace_with_generator')
    use su3, only: GetTraceWithGenerator
    integer(c_int), value :: a
    complex(c_double_complex), intent(in) :: m(3, 3)
    complex(c_double_complex), intent(out) :: res
    complex(fp) :: mm(3, 3)
    integer(int8) :: ai
    mm = m
    ai = int(a, int8)
    res = GetTraceWithGenerator(ai, mm)
  end subroutine ref_trace_with_generator

  ! matrixoperations.f90:547-567
  subroutine ref_expah(a, res) bind(C, name='ref_expah')
    use matrixoperations, only: ExpAH
    complex(c_double_complex), intent(in) :: a(3, 3)
    complex(c_double_complex), intent(out) :: res(3, 3)
    complex(fp) :: aa(3, 3)
    aa = a
    res = ExpAH(aa)
  end subroutine ref_expah

  ! su3.f90:318-334 (ArgLogU, matrixoperations.f90:474-495)
  subroutine ref_group_log(u, res) bind(C, name='ref_group_log')
    use su3, only: GetGroupLog
    complex(c_double_complex), intent(in) :: u(3, 3)
    real(c_double), intent(out) :: res(8)
    complex(fp) :: uu(3, 3)
    uu = u
    res = GetGroupLog(uu)
  end subroutine ref_group_log

  ! matrixoperations.f90:474-495 (the full matrix, trace included)
  subroutine ref_arglogu(u, res) bind(C, name='ref_arglogu')
    use matrixoperations, only: ArgLogU
    complex(c_double_complex), intent(in) :: u(3, 3)
    complex(c_double_complex), intent(out) :: res(3, 3)
    complex(fp) :: uu(3, 3)
    uu = u
    res = ArgLogU(uu)
  end subroutine ref_arglogu

  ! matrixoperations.f90:81-104
  subroutine ref_kron(a, ra, ca, b, rb, cb, ab) bind(C, name='ref_kron')
    use matrixoperations, only: GetKronProd
    integer(c_int), value :: ra, ca, rb, cb
    complex(c_double_complex), intent(in) :: a(ra, ca), b(rb, cb)
    complex(c_double_complex), intent(out) :: ab(ra*rb, ca*cb)
    complex(fp) :: aa(ra, ca), bb(rb, cb)
    aa = a
    bb = b
    ab = GetKronProd(aa, bb)
  end subroutine ref_kron
end module ref_bind
