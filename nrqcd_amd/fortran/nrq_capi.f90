!> ISO_C_BINDING interfaces to libnrqcd_amd.so (include/nrqcd_amd.h).
!! The Fortran modules lattice, random, gaugeconfiguration_su3 and nrqcd are
!! written on top of these.
module nrq_capi
  use, intrinsic :: iso_c_binding
  implicit none
  public

  integer(c_int), parameter :: NRQ_OK = 0
  integer(c_int), parameter :: NRQ_PRECISION_FP64 = 64, NRQ_PRECISION_FP32 = 32

  type, bind(C) :: nrq_solve_stats
     integer(c_int64_t) :: iterations = 0
     integer(c_int64_t) :: solves = 0
     integer(c_int64_t) :: restarts = 0
     real(c_double) :: max_true_relres = 0
  end type nrq_solve_stats

  interface
     integer(c_int) function nrq_last_error(buf, n) bind(C, name='nrq_last_error')
       import :: c_int, c_char, c_size_t
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_size_t), value :: n
     end function nrq_last_error
     integer(c_int) function nrq_synchronize() bind(C, name='nrq_synchronize')
       import :: c_int
     end function nrq_synchronize

     integer(c_int) function nrq_lattice_init(ext, spacing) bind(C, name='nrq_lattice_init')
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), intent(in) :: ext(3)
       real(c_double), intent(in) :: spacing(4)
     end function nrq_lattice_init
     integer(c_int) function nrq_lattice_finalize() bind(C, name='nrq_lattice_finalize')
       import :: c_int
     end function nrq_lattice_finalize
     integer(c_int) function nrq_lattice_index(pos, idx) bind(C, name='nrq_lattice_index')
       import :: c_int, c_int64_t
       integer(c_int64_t), intent(in) :: pos(3)
       integer(c_int64_t), intent(out) :: idx
     end function nrq_lattice_index
     integer(c_int) function nrq_lattice_position(idx, pos) bind(C, name='nrq_lattice_position')
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: idx
       integer(c_int64_t), intent(out) :: pos(3)
     end function nrq_lattice_position
     integer(c_int) function nrq_lattice_neib(dir, idx, out) bind(C, name='nrq_lattice_neib')
       import :: c_int, c_int64_t
       integer(c_int), value :: dir
       integer(c_int64_t), value :: idx
       integer(c_int64_t), intent(out) :: out
     end function nrq_lattice_neib
     integer(c_int) function nrq_lattice_negative_index(idx, out) bind(C, name='nrq_lattice_negative_index')
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: idx
       integer(c_int64_t), intent(out) :: out
     end function nrq_lattice_negative_index
     integer(c_int) function nrq_lattice_momentum(idx, p) bind(C, name='nrq_lattice_momentum')
       import :: c_int, c_int64_t, c_double_complex
       integer(c_int64_t), value :: idx
       complex(c_double_complex), intent(out) :: p(3)
     end function nrq_lattice_momentum

     integer(c_int) function nrq_random_init(seed) bind(C, name='nrq_random_init')
       import :: c_int, c_int64_t
       integer(c_int64_t), value :: seed
     end function nrq_random_init
     integer(c_int) function nrq_random_uniform(r, n) bind(C, name='nrq_random_uniform')
       import :: c_int, c_int64_t, c_double
       integer(c_int64_t), value :: n
       real(c_double), intent(out) :: r(n)
     end function nrq_random_uniform

     integer(c_int) function nrq_gauge_create(precision, h) bind(C, name='nrq_gauge_create')
       import :: c_int, c_ptr
       integer(c_int), value :: precision
       type(c_ptr), intent(out) :: h
     end function nrq_gauge_create
     integer(c_int) function nrq_gauge_destroy(h) bind(C, name='nrq_gauge_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function nrq_gauge_destroy
     integer(c_int) function nrq_gauge_copy(dst, src) bind(C, name='nrq_gauge_copy')
       import :: c_int, c_ptr
       type(c_ptr), value :: dst, src
     end function nrq_gauge_copy
     integer(c_int) function nrq_gauge_cold(h) bind(C, name='nrq_gauge_cold')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function nrq_gauge_cold
     integer(c_int) function nrq_gauge_box_ic(h, qs, amp, g) bind(C, name='nrq_gauge_box_ic')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), value :: qs, amp, g
     end function nrq_gauge_box_ic
     integer(c_int) function nrq_gauge_box_ic_correlators(h, qs, amp, g, aa, ee) &
          bind(C, name='nrq_gauge_box_ic_correlators')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h, aa, ee
       real(c_double), value :: qs, amp, g
     end function nrq_gauge_box_ic_correlators
     integer(c_int) function nrq_gauge_transverse_init(h, occupation, specific_sender, aa, ee) &
          bind(C, name='nrq_gauge_transverse_init')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, occupation, aa, ee
       integer(c_int), value :: specific_sender
     end function nrq_gauge_transverse_init
     integer(c_int) function nrq_gauge_equilibrium_init(h, beta, nefieldinit, nequilibrium, tol, rho, energies, &
          deviations, projections) bind(C, name='nrq_gauge_equilibrium_init')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h, rho, energies, deviations, projections
       real(c_double), value :: beta, tol
       integer(c_int64_t), value :: nefieldinit, nequilibrium
     end function nrq_gauge_equilibrium_init
     integer(c_int) function nrq_gauge_upload(h, links, efield) bind(C, name='nrq_gauge_upload')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, links, efield
     end function nrq_gauge_upload
     integer(c_int) function nrq_gauge_download(h, links, efield) bind(C, name='nrq_gauge_download')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, links, efield
     end function nrq_gauge_download
     integer(c_int) function nrq_gauge_get_link(h, dir, idx, link) bind(C, name='nrq_gauge_get_link')
       import :: c_int, c_ptr, c_int64_t, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: dir
       integer(c_int64_t), value :: idx
       complex(c_double_complex), intent(out) :: link(3, 3)
     end function nrq_gauge_get_link
     integer(c_int) function nrq_gauge_get_efield(h, dir, idx, e) bind(C, name='nrq_gauge_get_efield')
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: h
       integer(c_int), value :: dir
       integer(c_int64_t), value :: idx
       real(c_double), intent(out) :: e(8)
     end function nrq_gauge_get_efield
     integer(c_int) function nrq_gauge_leapfrog(h, sw, n) bind(C, name='nrq_gauge_leapfrog')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), value :: h
       real(c_double), value :: sw
       integer(c_int64_t), value :: n
     end function nrq_gauge_leapfrog
     integer(c_int) function nrq_gauge_leapfrog_batch(h, n, sw, nsteps) bind(C, name='nrq_gauge_leapfrog_batch')
       import :: c_int, c_ptr, c_double, c_int64_t
       type(c_ptr), intent(in) :: h(*)
       integer(c_int), value :: n
       real(c_double), value :: sw
       integer(c_int64_t), value :: nsteps
     end function nrq_gauge_leapfrog_batch
     integer(c_int) function nrq_gauge_observables(h, out) bind(C, name='nrq_gauge_observables')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: out(4)
     end function nrq_gauge_observables
     integer(c_int) function nrq_gauge_plaquette(h, i, j, idx, p) bind(C, name='nrq_gauge_plaquette')
       import :: c_int, c_ptr, c_int64_t, c_double_complex
       type(c_ptr), value :: h
       integer(c_int), value :: i, j
       integer(c_int64_t), value :: idx
       complex(c_double_complex), intent(out) :: p(3, 3)
     end function nrq_gauge_plaquette
     integer(c_int) function nrq_gauge_reunitarise(h) bind(C, name='nrq_gauge_reunitarise')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function nrq_gauge_reunitarise

     integer(c_int) function nrq_gauge_set_reunitarise_interval(h, interval) &
          bind(C, name='nrq_gauge_set_reunitarise_interval')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int64_t), value :: interval
     end function nrq_gauge_set_reunitarise_interval
     integer(c_int) function nrq_gauge_coulomb_gaugefix(h, tol, alpha, maxit, iterations, dev) &
          bind(C, name='nrq_gauge_coulomb_gaugefix')
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: h
       real(c_double), value :: tol, alpha
       integer(c_int64_t), value :: maxit
       integer(c_int64_t), intent(out) :: iterations
       real(c_double), intent(out) :: dev
     end function nrq_gauge_coulomb_gaugefix
     integer(c_int) function nrq_gauge_transverse_correlators(h, aa, ee) &
          bind(C, name='nrq_gauge_transverse_correlators')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, aa, ee
     end function nrq_gauge_transverse_correlators
     integer(c_int) function nrq_lattice_max_norm2_momentum(out) bind(C, name='nrq_lattice_max_norm2_momentum')
       import :: c_int, c_double
       real(c_double), intent(out) :: out
     end function nrq_lattice_max_norm2_momentum

     integer(c_int) function nrq_cn_init(mass, c, sw, tol) bind(C, name='nrq_cn_init')
       import :: c_int, c_double, c_double_complex
       real(c_double), value :: mass, sw, tol
       complex(c_double_complex), intent(in) :: c(4)
     end function nrq_cn_init
     integer(c_int) function nrq_cn_finalize() bind(C, name='nrq_cn_finalize')
       import :: c_int
     end function nrq_cn_finalize
     integer(c_int) function nrq_heavy_create(h) bind(C, name='nrq_heavy_create')
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: h
     end function nrq_heavy_create
     integer(c_int) function nrq_heavy_destroy(h) bind(C, name='nrq_heavy_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function nrq_heavy_destroy
     integer(c_int) function nrq_heavy_copy(dst, src) bind(C, name='nrq_heavy_copy')
       import :: c_int, c_ptr
       type(c_ptr), value :: dst, src
     end function nrq_heavy_copy
     integer(c_int) function nrq_heavy_init_point(h, iq, ia) bind(C, name='nrq_heavy_init_point')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int64_t), value :: iq, ia
     end function nrq_heavy_init_point
     integer(c_int) function nrq_heavy_init_point_dof(h, sq, cq, iq, sa, ca, ia) &
          bind(C, name='nrq_heavy_init_point_dof')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: h
       integer(c_int), value :: sq, cq, sa, ca
       integer(c_int64_t), value :: iq, ia
     end function nrq_heavy_init_point_dof
     integer(c_int) function nrq_heavy_upload(h, q, a) bind(C, name='nrq_heavy_upload')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, q, a
     end function nrq_heavy_upload
     integer(c_int) function nrq_heavy_download(h, q, a) bind(C, name='nrq_heavy_download')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, q, a
     end function nrq_heavy_download
     integer(c_int) function nrq_heavy_update(h, g, mass, c, sw, stats) bind(C, name='nrq_heavy_update')
       import :: c_int, c_ptr, c_double, c_double_complex, nrq_solve_stats
       type(c_ptr), value :: h, g
       real(c_double), value :: mass, sw
       complex(c_double_complex), intent(in) :: c(4)
       type(nrq_solve_stats), intent(out) :: stats
     end function nrq_heavy_update
     integer(c_int) function nrq_heavy_update_batch(h, mass, n, g, c, sw, stats) bind(C, name='nrq_heavy_update_batch')
       import :: c_int, c_ptr, c_double, c_double_complex, nrq_solve_stats
       type(c_ptr), intent(in) :: h(*)
       real(c_double), intent(in) :: mass(*)
       integer(c_int), value :: n
       type(c_ptr), value :: g
       real(c_double), value :: sw
       complex(c_double_complex), intent(in) :: c(4)
       type(nrq_solve_stats), intent(out) :: stats
     end function nrq_heavy_update_batch
     integer(c_int) function nrq_heavy_norms(h, nq, na) bind(C, name='nrq_heavy_norms')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: nq, na
     end function nrq_heavy_norms
     integer(c_int) function nrq_heavy_correlators(h, out) bind(C, name='nrq_heavy_correlators')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(out) :: out(4)
     end function nrq_heavy_correlators
     integer(c_int) function nrq_heavy_correlators_gconjg(h, out) bind(C, name='nrq_heavy_correlators_gconjg')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(out) :: out(4)
     end function nrq_heavy_correlators_gconjg

     integer(c_int) function nrq_spectral_functions(signals, n, batch, dt, spectra, stages) &
          bind(C, name='nrq_spectral_functions')
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: signals, spectra, stages
       integer(c_int64_t), value :: n, batch
       real(c_double), value :: dt
     end function nrq_spectral_functions
     integer(c_int) function nrq_t2w(data, n, dt) bind(C, name='nrq_t2w')
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: data
       integer(c_int64_t), value :: n
       real(c_double), value :: dt
     end function nrq_t2w
     integer(c_int) function nrq_w2t(data, n, dt) bind(C, name='nrq_w2t')
       import :: c_int, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: data
       integer(c_int64_t), value :: n
       real(c_double), value :: dt
     end function nrq_w2t
  end interface

contains

  !> Stops the program with the library's last error message when rc /= 0
  !! (the reference aborts through MPIstop, mpiinterface.f90:148-173).
  subroutine nrq_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: k, code
    if (rc == NRQ_OK) return
    code = nrq_last_error(buf, int(512, c_size_t))
    msg = ''
    do k = 1, 511
       if (buf(k) == c_null_char) exit
       msg(k:k) = buf(k)
    end do
    write(*, '(A,A,A,I0,A,A)') 'Error in ', where, ' (code ', rc, '): ', trim(msg)
    error stop 1
  end subroutine nrq_check
end module nrq_capi
