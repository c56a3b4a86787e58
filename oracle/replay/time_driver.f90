! time_driver.f90 -- TEST INFRASTRUCTURE ONLY (CPU-baseline calibration).
!
! This project's own timing program around the unmodified P-TRANS module RTMC
! (src/RTMCphon.f90): loadMCP, then TransmissionFunction (RT:1544-1604) for one sampling
! mean free path, timed with system_clock, with the reference's own RANDOM_NUMBER (no
! interposer) and its OpenMP loop on OMP_NUM_THREADS threads.  Built and run in the
! development container only (oracle/Makefile, target `timing`).
!
! Usage (in a directory holding Materials/<name>.anphon):
!   time_ref <file.mcp> <mfp_nm>      prints "TIME seconds PATHS n TRANSMITTED t"
program time_driver
  use RTMC
  implicit none
  character(len=256) :: mcp, arg
  integer :: status
  integer(8) :: t0, t1, rate, npaths
  real(8) :: mfp

  verbose = 1   ! main.f90:14
  call getarg(1, mcp)
  call getarg(2, arg); read(arg,*) mfp
  call init_random_seed()
  call loadMCP(trim(mcp), status)
  if (status /= 0) stop 3
  param%MeanFreePath_current = mfp
  interval_trace = huge(1)
  open(fd_trace, file='trace.out')
  call system_clock(t0, rate)
  call TransmissionFunction
  call system_clock(t1)
  close(fd_trace)
  npaths = int(anphon%nbranch, 8) * anphon%nkpoint * (param%nangle - 1) * param%nphonon
  write(*,'(a,f14.6,a,i14,a,i14)') 'TIME ', dble(t1 - t0) / dble(rate), ' PATHS ', npaths, &
    ' TRANSMITTED ', nint(sum(transmit) * param%nphonon, 8)
end program time_driver
