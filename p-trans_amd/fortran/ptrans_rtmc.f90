! ptrans_rtmc.f90 -- the P-TRANS side of the drop-in: replaces the body of
! TransmissionFunction (src/RTMCphon.f90:1544-1604) with one libptrans call.
!
! Compiled together with the P-TRANS module RTMC (it USEs RTMC's public globals: param,
! vgrain, ngrain, igrain_box, igrain_fill, anphon, transmit) and module ptrans.
!
!   call loadMCP(...)                 ! unchanged (RT:252-1251), also reads the .anphon
!   call ptrans_rtmc_init(0, seed, st)        ! upload vgrain and anphon once
!   do s = 1, param%nMeanFreePath             ! RayTrace's loop (RT:3388-3398)
!     param%MeanFreePath_current = param%vMeanFreePath(s)
!     call ptrans_transmission_function(st)   ! was: call TransmissionFunction
!     call effectiveMFP(SEMFP(:,:,s))          ! unchanged
!   end do
!   call ptrans_rtmc_finalize()
!
! The geometry handed over is vgrain exactly as loadMCP built it: per surface the
! reference's normal, tangent, centre and specularity (including any specularity
! fluctuation drawn at load time, RT:1151-1157) and the transmission law
! param%transmissivity(imaterial2, imaterial1) of kindTrans (RT:1161-1162, 1392-1393);
! box walls have law VALUE with t = 0 (RT:1327-1330).
module ptrans_rtmc
  use iso_c_binding
  use RTMC
  use ptrans
  implicit none
  private
  public :: ptrans_rtmc_init, ptrans_transmission_function, ptrans_rtmc_finalize, &
       ptrans_precision, ptrans_write_trace

  type(c_ptr), save :: ctx = c_null_ptr, geom = c_null_ptr, modes = c_null_ptr
  integer(c_int32_t), save :: ptrans_precision = PT_PREC_FP64

contains

  subroutine ptrans_rtmc_init(device, seed, status)
    integer, intent(in) :: device
    integer(c_int64_t), intent(in) :: seed
    integer, intent(out) :: status

    type(pt_geom_desc) :: gd
    type(pt_modes_desc) :: md
    type(pt_surface), allocatable, target :: surf(:)
    integer(c_int32_t), allocatable, target :: goff(:), ginc(:)
    real(c_double), allocatable, target :: gen(:), vtx(:), omega(:,:)
    integer :: g0, g1, ng, ig, is, ns, nv, k, iv
    type(Surface_type), pointer :: s

    g0 = lbound(vgrain, 1)
    g1 = ubound(vgrain, 1)
    ng = g1 - g0 + 1
    ns = 0
    nv = 0
    do ig = g0, g1
      ns = ns + vgrain(ig)%nsurface
      do is = 1, vgrain(ig)%nsurface
        if (vgrain(ig)%vsurface(is)%kindShape == PT_SHAPE_POLYGON) &
          nv = nv + vgrain(ig)%vsurface(is)%nvertex
      end do
    end do
    allocate(surf(ns), goff(ng + 1), ginc(ng), gen(3 * ng), vtx(max(3, 3 * nv)))

    k = 0
    iv = 0
    goff(1) = 0
    do ig = g0, g1
      ginc(ig - g0 + 1) = merge(1, 0, vgrain(ig)%incident)
      gen(3 * (ig - g0) + 1:3 * (ig - g0) + 3) = vgrain(ig)%generatrix
      do is = 1, vgrain(ig)%nsurface
        s => vgrain(ig)%vsurface(is)
        k = k + 1
        surf(k)%kind = s%kindShape
        surf(k)%internal = merge(1, 0, s%internal)
        surf(k)%incident = merge(1, 0, s%incident)
        surf(k)%grain_inside = s%grain_inside
        surf(k)%grain_outside = s%grain_outside
        surf(k)%trans_kind = s%kindTrans
        if (ig == igrain_box) then
          surf(k)%trans_param = 0.0d0
        else
          surf(k)%trans_param = param%transmissivity(s%imaterial2, s%imaterial1)
        end if
        surf(k)%specularity = s%specularity
        surf(k)%center = s%center
        surf(k)%radius = s%radius
        surf(k)%normal = s%normal
        surf(k)%tangent = reshape(s%tangent, [9])
        surf(k)%vertex_offset = iv
        if (s%kindShape == PT_SHAPE_POLYGON) then
          surf(k)%nvertex = s%nvertex
          vtx(3 * iv + 1:3 * (iv + s%nvertex)) = reshape(s%vertex(:, 1:s%nvertex), [3 * s%nvertex])
          iv = iv + s%nvertex
        else
          surf(k)%nvertex = 0
        end if
      end do
      goff(ig - g0 + 2) = k
    end do

    gd%box = [param%boxx, param%boxy, param%boxz]
    gd%scope = [param%scopey, param%scopez]
    gd%kind_grain = param%kindGrain
    gd%ngrain = ngrain
    gd%igrain_fill = merge(igrain_fill, -1, param%kindGrain == ID_GRAIN_BRICKS)
    gd%nmaterial = param%nmaterial
    gd%ngrain_total = ng
    gd%nsurface = ns
    gd%nvertex = nv
    gd%grain_offset = c_loc(goff)
    gd%grain_incident = c_loc(ginc)
    gd%generatrix = c_loc(gen)
    gd%surfaces = c_loc(surf)
    gd%vertices = c_loc(vtx)

    allocate(omega(anphon%nbranch, anphon%nkpoint))
    omega = anphon%omega
    md%nbranch = anphon%nbranch
    md%nkpoint = anphon%nkpoint
    md%omega = c_loc(omega)
    md%max_omega = anphon%maxOmega
    md%temperature = param%temperature

    status = pt_ctx_create(int(device, c_int), seed, ctx)
    if (status == PT_OK) status = pt_geom_create(ctx, gd, geom)
    if (status == PT_OK) status = pt_modes_create(ctx, md, modes)
    if (status /= PT_OK) write(fd_stdout, '(" libptrans: ", a)') ptrans_error_message()
  end subroutine ptrans_rtmc_init

  ! TransmissionFunction on one GPU: transmit(ib,ik,ia) = transmitted/nphonon for every
  ! cell with ia < nangle, the same numbers calcTransmit produces (RT:1833-1838).
  subroutine ptrans_transmission_function(status)
    integer, intent(out) :: status
    integer(c_int64_t), allocatable, target :: counts(:,:,:)
    real(c_double), allocatable, target :: va(:)
    type(pt_stats), target :: stats
    integer :: ia

    do ia = 1, param%nangle                                   ! RT:1555-1557
      param%vangle(ia) = 0.0d0 + dble(ia - 1) / dble(param%nangle - 1) * M_PI_2
    end do
    allocate(va(param%nangle))
    va = param%vangle
    allocate(counts(anphon%nbranch, anphon%nkpoint, param%nangle))
    counts = 0
    status = pt_trace_grid(ctx, geom, modes, param%MeanFreePath_current, &
         int(param%nangle, c_int32_t), c_loc(va), int(param%nphonon, c_int64_t), &
         ptrans_precision, c_loc(counts), c_loc(stats))
    if (status /= PT_OK) then
      write(fd_stdout, '(" libptrans: ", a)') ptrans_error_message()
      return
    end if
    ! paths stopped by the library's event safety limit are counted as not transmitted
    ! (the reference's event loop has no bound, RT:1693): say so, they bias T low
    if (stats%event_limited > 0) write(fd_stdout, '(" libptrans: warning: ", i0, &
         & " paths stopped at the event limit (counted as not transmitted)")') stats%event_limited
    transmit(:,:,:) = dble(counts) / dble(param%nphonon)
  end subroutine ptrans_transmission_function

  ! trace.out for the current sampling mean free path (RT:1661-1665, 1818-1831): the
  ! paths first, first + stride, ... (n of them) of TransmissionFunction's numbering,
  ! each as its start position and the position after every event, written with the
  ! reference's own statement (write(fd_trace,*) real(Position)) and two empty records.
  ! The reference traces every interval_trace-th particle of its OpenMP master thread;
  ! stride = interval_trace gives the same density on the deterministic numbering.
  subroutine ptrans_write_trace(first, stride, n, max_points, status)
    integer(c_int64_t), intent(in) :: first, stride, n, max_points
    integer, intent(out) :: status
    real(c_double), allocatable, target :: pts(:,:,:), va(:)
    integer(c_int64_t), allocatable, target :: npts(:)
    integer(c_int64_t) :: i, k

    allocate(va(param%nangle))
    va = param%vangle
    allocate(pts(3, max_points, n), npts(n))
    status = pt_trace_log(ctx, geom, modes, param%MeanFreePath_current, &
         int(param%nangle, c_int32_t), c_loc(va), int(param%nphonon, c_int64_t), first, &
         stride, n, ptrans_precision, max_points, c_loc(pts), c_loc(npts), c_null_ptr)
    if (status /= PT_OK) then
      write(fd_stdout, '(" libptrans: ", a)') ptrans_error_message()
      return
    end if
    if (fd_trace <= 0) return
    do i = 1, n
      do k = 1, min(npts(i), max_points)
        write(fd_trace, *) real(pts(:, k, i))
      end do
      write(fd_trace, *)
      write(fd_trace, *)
    end do
  end subroutine ptrans_write_trace

  subroutine ptrans_rtmc_finalize()
    integer(c_int) :: st
    if (c_associated(modes)) st = pt_modes_destroy(modes)
    if (c_associated(geom)) st = pt_geom_destroy(geom)
    if (c_associated(ctx)) st = pt_ctx_destroy(ctx)
    modes = c_null_ptr
    geom = c_null_ptr
    ctx = c_null_ptr
  end subroutine ptrans_rtmc_finalize

end module ptrans_rtmc
