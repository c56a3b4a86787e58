! driver.f90 -- TEST INFRASTRUCTURE ONLY (replay harness for the reference).
!
! A small program of this project's own that USEs the unmodified P-TRANS module RTMC
! (src/RTMCphon.f90) and calls its public procedures, to write golden vectors.  It is
! built and run in the development container only (oracle/Makefile), linked with
! interpose.c, which serves Philox random numbers to the reference's RANDOM_NUMBER.
!
! Usage (run in a directory holding Materials/<name>.anphon, RT:671):
!   replay grid  <file.mcp> <mfp_nm> <seed>
!       TransmissionFunction (RT:1544) once; prints "C ib ik ia count" per cell.
!   In every mode, the draws of loadMCP (FluctuationSpecularity, RT:1151-1157) come from
!   the Philox stream (seed, 2^62), the host's set-up stream.
!   replay paths <file.mcp> <mfp_nm> <seed> <first_path> <n>
!       calcTransmit (RT:1608) one particle at a time for paths first..first+n-1 of the
!       numbering path = cell*Np + p, cell = ((ik-1)*nb + (ib-1))*(nangle-1) + (ia-1);
!       prints "P path trans draws"; trajectories go to trace.out (RT:1661-1665,
!       1818-1831), random-number call sites to draws.log.
!   replay post  <file.mcp> <transmit.txt>
!       effectiveMFP (RT:2418) on transmit(nb,nk,nangle) blocks read from the file (one
!       per sampling MFP of the .mcp), then fileoutput (RT:2601); prints SEMFP.
!   replay geom  <cases.txt>
!       runs setupSurfaceGeometry, isInsidePolygon, calcSurfaceDistance*,
!       calcSurfaceNormal* and calcSurfaceRotation* on the cases in the file.
!   replay laws  <file.mcp> <seed>
!       loadMCP with the set-up draws (FluctuationSpecularity, RT:1151-1157) served from
!       the Philox stream (seed, 2^62); prints "T i2 i1 kind value" for every material
!       pair (GammaDeterm's gamma for TBC rows, RT:858-870) and "S ig is p" for every
!       surface of every grain.
!   replay dump  <file.mcp>
!       loadMCP only; prints the parameters the trace reads ("PARAM ...", "MFP s value",
!       "CELLSPEC ..."), then per grain "G ig incident nsurface generatrix(3)" and per
!       surface "S ig is kindShape internal incident grain_inside grain_outside
!       imaterial1 imaterial2 kindTrans nvertex", "R specularity radius law", "C center",
!       "N normal", "T tangent" and "V vertex" lines (law: param%transmissivity of the
!       surface's material pair, 0 for the box grain).
program replay_driver
  use RTMC
  use iso_c_binding
  implicit none

  interface
    subroutine replay_configure(seed, first_path) bind(C, name='replay_configure')
      import :: c_int64_t
      integer(c_int64_t), value :: seed, first_path
    end subroutine replay_configure
    function replay_draws() bind(C, name='replay_draws') result(n)
      import :: c_int64_t
      integer(c_int64_t) :: n
    end function replay_draws
    subroutine replay_log_open(path, n) bind(C, name='replay_log_open')
      import :: c_char, c_int
      character(kind=c_char), dimension(*) :: path
      integer(c_int), value :: n
    end subroutine replay_log_open
    subroutine replay_log_close() bind(C, name='replay_log_close')
    end subroutine replay_log_close
    subroutine replay_log_enable(on) bind(C, name='replay_log_enable')
      import :: c_int
      integer(c_int), value :: on
    end subroutine replay_log_enable
  end interface

  character(len=256) :: mode, arg
  integer :: status

  verbose = 1   ! main.f90:14
  call getarg(1, mode)
  select case (trim(mode))
  case ('grid')
    call run_grid()
  case ('paths')
    call run_paths()
  case ('geom')
    call run_geom()
  case ('post')
    call run_post()
  case ('laws')
    call run_laws()
  case ('dump')
    call run_dump()
  case default
    write(*,*) 'usage: replay grid|paths|geom ...'
    stop 2
  end select

contains

  subroutine load(status)
    integer, intent(out) :: status
    character(len=256) :: mcp
    call getarg(2, mcp)
    call loadMCP(trim(mcp), status)
    if (status /= 0) then
      write(*,*) 'replay: loadMCP failed'
      stop 3
    end if
  end subroutine load

  subroutine run_grid()
    integer :: ib, ik, ia
    real(8) :: mfp
    integer(c_int64_t) :: seed
    call getarg(3, arg); read(arg,*) mfp
    call getarg(4, arg); read(arg,*) seed
    call replay_configure(seed, 4611686018427387904_c_int64_t)   ! set-up draws
    call load(status)
    param%MeanFreePath_current = mfp
    interval_trace = huge(1)
    open(fd_trace, file='trace.out')
    call replay_configure(seed, 0_c_int64_t)
    call TransmissionFunction
    do ia = 1, param%nangle
      do ik = 1, anphon%nkpoint
        do ib = 1, anphon%nbranch
          write(*,'(a,3i8,i12)') 'C ', ib, ik, ia, nint(transmit(ib,ik,ia)*param%nphonon, kind=8)
        end do
      end do
    end do
    close(fd_trace)
  end subroutine run_grid

  subroutine run_paths()
    integer :: ib, ik, ia, np, na, nb
    integer(c_int64_t) :: seed, first, n, path, cell, draws
    real(8) :: mfp, trans, dummy
    call getarg(3, arg); read(arg,*) mfp
    call getarg(4, arg); read(arg,*) seed
    call replay_configure(seed, 4611686018427387904_c_int64_t)   ! set-up draws
    call load(status)
    call getarg(5, arg); read(arg,*) first
    call getarg(6, arg); read(arg,*) n
    param%MeanFreePath_current = mfp
    do ia = 1, param%nangle          ! as TransmissionFunction, RT:1554-1556
      param%vangle(ia) = 0.0d0 + dble(ia-1)/dble(param%nangle-1)* M_PI_2
    end do
    np = param%nphonon
    na = param%nangle - 1
    nb = anphon%nbranch
    ! only particles with mod(id,interval_trace)==1 are traced (RT:1662): with an
    ! interval of 2 and one untraced dummy particle after each real one, every real
    ! particle is traced.
    interval_trace = 2
    open(fd_trace, file='trace.out')
    call replay_log_open('draws.log', 9)
    param%nphonon = 1
    do path = first, first + n - 1
      cell = path / np
      ia = int(mod(cell, int(na, c_int64_t))) + 1
      ib = int(mod(cell / na, int(nb, c_int64_t))) + 1
      ik = int(cell / (int(na, c_int64_t) * nb)) + 1
      call replay_configure(seed, path)
      call calcTransmit(trans, ib, ik, ia)
      draws = replay_draws()
      write(*,'(a,i12,i3,i10)') 'P ', path, nint(trans), draws
      call replay_log_enable(0)
      call replay_configure(seed + 1_c_int64_t, path)
      call calcTransmit(dummy, ib, ik, ia)
      call replay_log_enable(1)
    end do
    call replay_log_close()
    close(fd_trace)
  end subroutine run_paths

  subroutine run_laws()
    integer(c_int64_t) :: seed
    integer :: i1, i2, ig, is
    call getarg(3, arg); read(arg,*) seed
    call replay_configure(seed, 4611686018427387904_c_int64_t)
    call load(status)
    do i1 = 1, size(param%kindTrans, 2)
      do i2 = 1, size(param%kindTrans, 1)
        write(*,'(a,2i6,i4,es27.17e3)') 'T ', i2, i1, param%kindTrans(i2,i1), &
          param%transmissivity(i2,i1)
      end do
    end do
    do ig = lbound(vgrain, 1), ubound(vgrain, 1)
      do is = 1, vgrain(ig)%nsurface
        write(*,'(a,2i8,es27.17e3)') 'S ', ig, is, vgrain(ig)%vsurface(is)%specularity
      end do
    end do
  end subroutine run_laws

  subroutine run_dump()
    integer :: ig, is, k
    real(8) :: law
    type(Surface_type), pointer :: sf
    call load(status)
    write(*,'(a,5es27.17e3,es27.17e3,3i12,2i6,es27.17e3)') 'PARAM ', param%boxx, param%boxy, &
      param%boxz, param%scopey, param%scopez, param%temperature, param%nphonon, param%nangle, &
      param%nMeanFreePath, param%kindGrain, param%nmaterial, param%specfluc
    do is = 1, param%nMeanFreePath
      write(*,'(a,i6,es27.17e3)') 'MFP ', is, param%vMeanFreePath(is)
    end do
    write(*,'(a,6es27.17e3)') 'CELLSPEC ', param%cellspec
    write(*,'(a,4i8)') 'GRAINS ', lbound(vgrain, 1), ubound(vgrain, 1), igrain_box, igrain_fill
    do ig = lbound(vgrain, 1), ubound(vgrain, 1)
      write(*,'(a,i8,i3,i8,3es27.17e3)') 'G ', ig, merge(1, 0, vgrain(ig)%incident), &
        vgrain(ig)%nsurface, vgrain(ig)%generatrix
      do is = 1, vgrain(ig)%nsurface
        sf => vgrain(ig)%vsurface(is)
        law = 0.0d0
        if (ig /= igrain_box) law = param%transmissivity(sf%imaterial2, sf%imaterial1)
        write(*,'(a,11(1x,i0))') 'S ', ig, is, sf%kindShape, merge(1, 0, sf%internal), &
          merge(1, 0, sf%incident), sf%grain_inside, sf%grain_outside, sf%imaterial1, &
          sf%imaterial2, sf%kindTrans, sf%nvertex
        write(*,'(a,3es27.17e3)') 'R ', sf%specularity, sf%radius, law
        write(*,'(a,3es27.17e3)') 'C ', sf%center
        write(*,'(a,3es27.17e3)') 'N ', sf%normal
        write(*,'(a,9es27.17e3)') 'T ', sf%tangent
        if (sf%kindShape == ID_SHAPE_POLYGON) then
          do k = 1, sf%nvertex
            write(*,'(a,3es27.17e3)') 'V ', sf%vertex(:, k)
          end do
        end if
      end do
    end do
  end subroutine run_dump

  subroutine run_post()
    character(len=256) :: tfile, mcp
    integer :: u, ia, is
    real(8), allocatable :: semfp(:,:,:)
    call load(status)
    call getarg(2, mcp)
    call getarg(3, tfile)
    do ia = 1, param%nangle          ! as TransmissionFunction, RT:1554-1556
      param%vangle(ia) = 0.0d0 + dble(ia-1)/dble(param%nangle-1)* M_PI_2
    end do
    allocate(semfp(anphon%nkpoint, anphon%nbranch, param%nMeanFreePath))
    open(newunit=u, file=trim(tfile), action='read')
    do is = 1, param%nMeanFreePath
      read(u,*) transmit
      call effectiveMFP(semfp(:,:,is))
    end do
    close(u)
    write(*,'(a)') 'SEMFP'
    write(*,*) semfp
    call fileoutput(semfp, trim(mcp))
  end subroutine run_post

  subroutine run_geom()
    character(len=256) :: casefile
    character(len=8) :: tag
    integer :: u, ios, nv, i
    real(8) :: rs(3), nd(3), c(3), r, boxz, l, nrm(3), rot(3,3), p(3)
    real(8), allocatable :: v(:,:)
    logical :: hit
    type(Surface_type), target :: sf
    call getarg(2, casefile)
    open(newunit=u, file=trim(casefile), action='read')
    do
      read(u,*,iostat=ios) tag
      if (ios /= 0) exit
      select case (trim(tag))
      case ('POLY')
        read(u,*) nv
        allocate(v(3, nv+1))
        read(u,*) v(:, 1:nv)
        v(:, nv+1) = v(:, 1)
        read(u,*) rs, nd, p
        sf%kindShape = ID_SHAPE_POLYGON
        sf%nvertex = nv
        allocate(sf%vertex(3, nv+1))
        sf%vertex = v
        call setupSurfaceGeometry(sf)
        call calcSurfaceDistancePolygon(hit, l, rs, nd, sf%vertex, sf%normal, nv)
        write(*,*) 'POLY'
        write(*,*) sf%center
        write(*,*) sf%normal
        write(*,*) sf%tangent
        write(*,*) merge(1, 0, hit), l
        write(*,*) merge(1, 0, isInsidePolygon(sf%vertex, nv, p))
        deallocate(sf%vertex)
        deallocate(v)
      case ('SPH', 'CYL')
        read(u,*) rs, nd, c, r, boxz
        param%boxz = boxz
        if (trim(tag) == 'SPH') then
          call calcSurfaceDistanceSphere(hit, l, rs, nd, c, r)
        else
          call calcSurfaceDistanceCylinder(hit, l, rs, nd, c, r)
        end if
        write(*,*) trim(tag)
        write(*,*) merge(1, 0, hit), l
        ! normal and frame at the hit point (or at the ray origin for a miss)
        if (hit) then
          p = rs + nd*l
        else
          p = rs
        end if
        if (trim(tag) == 'SPH') then
          call calcSurfaceNormalSphere(nrm, p, c, r)
          call calcSurfaceRotationSphere(rot, p, c)
        else
          call calcSurfaceNormalCylinder(nrm, p, c, r)
          call calcSurfaceRotationCylinder(rot, p, c)
        end if
        write(*,*) p
        write(*,*) nrm
        write(*,*) rot
      case default
        write(*,*) 'replay: unknown case ', tag
        stop 4
      end select
    end do
    close(u)
    i = 0
  end subroutine run_geom

end program replay_driver
