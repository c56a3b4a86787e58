! ft_trace.f90 -- a Fortran host calling libptrans through module ptrans (iso_c_binding).
!
!   ft_trace <geom.bin> <modes.bin> <counts.bin> <seed> <mfp_nm> <nangle> <np> <precision>
!            [<trace.out> <npaths> <stride>]
!
! With the optional arguments it also records paths 0, stride, ... with pt_trace_log and
! writes them as the reference writes trace.out (write(u,*) real(position), two empty
! records after each path).
!
! Reads a flattened geometry and a mode table written by the Python test (native
! byte order, stream access; surface records read straight into type(pt_surface), so
! the binding's layout is checked as well), runs pt_trace_grid on device 0 and writes
! the int64 counts.  Prints "ft_trace ok <paths> <events>" on success.
program ft_trace
  use iso_c_binding
  use ptrans
  implicit none
  character(len=512) :: fgeom, fmodes, fout, arg
  integer(c_int32_t) :: hdr(8), nb, nk, nangle, prec
  integer(c_int64_t) :: nvertex, seed, np
  real(c_double) :: mfp, max_omega
  real(c_double), target :: box(5)
  integer(c_int32_t), allocatable, target :: goff(:), ginc(:)
  real(c_double), allocatable, target :: gen(:), vtx(:), omega(:)
  type(pt_surface), allocatable, target :: surf(:)
  integer(c_int64_t), allocatable, target :: counts(:), npts(:)
  real(c_double), allocatable, target :: pts(:,:,:)
  character(len=512) :: ftrace
  integer(c_int64_t) :: ntr, stride, i, k, maxp
  type(pt_geom_desc) :: gd
  type(pt_modes_desc) :: md
  type(pt_stats), target :: st
  type(c_ptr) :: ctx, geom, modes
  integer(c_int) :: rc
  integer :: u, ngt

  call get_command_argument(1, fgeom)
  call get_command_argument(2, fmodes)
  call get_command_argument(3, fout)
  call get_command_argument(4, arg); read(arg, *) seed
  call get_command_argument(5, arg); read(arg, *) mfp
  call get_command_argument(6, arg); read(arg, *) nangle
  call get_command_argument(7, arg); read(arg, *) np
  call get_command_argument(8, arg); read(arg, *) prec

  open(newunit=u, file=trim(fgeom), access='stream', form='unformatted', status='old')
  read(u) hdr, nvertex, box
  ngt = hdr(5)
  allocate(goff(ngt + 1), ginc(ngt), gen(3 * ngt), surf(hdr(6)), vtx(max(3_c_int64_t, 3 * nvertex)))
  read(u) goff, ginc, gen, surf
  if (nvertex > 0) read(u) vtx(1:3 * nvertex)
  close(u)
  gd%box = box(1:3)
  gd%scope = box(4:5)
  gd%kind_grain = hdr(1)
  gd%ngrain = hdr(2)
  gd%igrain_fill = hdr(3)
  gd%nmaterial = hdr(4)
  gd%ngrain_total = hdr(5)
  gd%nsurface = hdr(6)
  gd%nvertex = nvertex
  gd%grain_offset = c_loc(goff)
  gd%grain_incident = c_loc(ginc)
  gd%generatrix = c_loc(gen)
  gd%surfaces = c_loc(surf)
  gd%vertices = c_loc(vtx)

  open(newunit=u, file=trim(fmodes), access='stream', form='unformatted', status='old')
  read(u) nb, nk, max_omega
  allocate(omega(nb * nk))
  read(u) omega
  close(u)
  md%nbranch = nb
  md%nkpoint = nk
  md%omega = c_loc(omega)
  md%max_omega = max_omega

  allocate(counts(int(nb, 8) * nk * nangle))
  counts = 0
  rc = pt_ctx_create(0_c_int, seed, ctx)
  if (rc == PT_OK) rc = pt_geom_create(ctx, gd, geom)
  if (rc == PT_OK) rc = pt_modes_create(ctx, md, modes)
  if (rc == PT_OK) rc = pt_trace_grid(ctx, geom, modes, mfp, nangle, c_null_ptr, np, prec, &
                                      c_loc(counts), c_loc(st))
  if (rc /= PT_OK) then
    write(*, '(a, i0, a, a)') 'ft_trace error ', rc, ': ', ptrans_error_message()
    stop 1
  end if
  if (command_argument_count() >= 11) then
    call get_command_argument(9, ftrace)
    call get_command_argument(10, arg); read(arg, *) ntr
    call get_command_argument(11, arg); read(arg, *) stride
    maxp = 4096
    allocate(pts(3, maxp, ntr), npts(ntr))
    rc = pt_trace_log(ctx, geom, modes, mfp, nangle, c_null_ptr, np, 0_c_int64_t, stride, ntr, &
                      prec, maxp, c_loc(pts), c_loc(npts), c_null_ptr)
    if (rc /= PT_OK) then
      write(*, '(a, i0, a, a)') 'ft_trace error ', rc, ': ', ptrans_error_message()
      stop 1
    end if
    open(newunit=u, file=trim(ftrace), status='replace')
    do i = 1, ntr
      do k = 1, min(npts(i), maxp)
        write(u, *) real(pts(:, k, i))
      end do
      write(u, *)
      write(u, *)
    end do
    close(u)
  end if
  rc = pt_modes_destroy(modes)
  rc = pt_geom_destroy(geom)
  rc = pt_ctx_destroy(ctx)
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) counts
  close(u)
  write(*, '(a, i0, 1x, i0)') 'ft_trace ok ', st%paths, st%events
end program ft_trace
