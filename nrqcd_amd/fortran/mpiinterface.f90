!> Single-process stand-in for the reference's mpiinterface module
!! (mpiinterface.f90:68-262): the whole lattice lives on one GPU, so there is
!! one "rank" and no MPI.  The names the drivers use are kept.
module mpiinterface
  use, intrinsic :: iso_fortran_env, only: int32, output_unit
  use nrq_capi, only: nrq_synchronize, nrq_check
  implicit none
  private
  public :: InitModule, FinalizeModule, ThisProc, NumProcs, SyncAll, MPIstop, intmpi

  integer, parameter :: intmpi = int32
  logical :: initialised = .false.
contains
  subroutine InitModule()
    initialised = .true.
  end subroutine InitModule

  subroutine FinalizeModule()
    initialised = .false.
  end subroutine FinalizeModule

  pure integer(intmpi) function ThisProc()
    ThisProc = 0
  end function ThisProc

  pure integer(intmpi) function NumProcs()
    NumProcs = 1
  end function NumProcs

  !> Waits for all queued device work (the reference's mpi_barrier).
  subroutine SyncAll()
    call nrq_check(nrq_synchronize(), 'SyncAll')
  end subroutine SyncAll

  !> Prints the message and stops (mpiinterface.f90:148-173).
  subroutine MPIstop(errormessage, errorcode)
    character(len=*), intent(in), optional :: errormessage
    integer, intent(in), optional :: errorcode
    if (present(errormessage)) write(output_unit, '(A)') trim(errormessage)
    if (present(errorcode)) then
       write(output_unit, '(A,I0)') 'Error code: ', errorcode
    end if
    error stop 1
  end subroutine MPIstop
end module mpiinterface
