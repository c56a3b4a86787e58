! ptrans.f90 -- Fortran binding of libptrans (include/ptrans.h) through iso_c_binding.
!
! The interoperable types below have the layout of the C structs (sizes 184, 112, 56,
! 96 and 80 bytes); every entry point is a bind(C) interface to the function of the
! same name.  Status codes are returned as integer(c_int); ptrans_error_message()
! fetches the thread's last message.  No CUDA, no wrappers around HIP: the Fortran host
! only sees plain pointers and sizes.
module ptrans
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PT_OK = 0, PT_E_INVALID = -1, PT_E_HIP = -2, PT_E_NOMEM = -3, &
       PT_E_GEOMETRY = -4, PT_E_UNSUPPORTED = -5, PT_E_NODEVICE = -6
  integer(c_int32_t), parameter :: PT_SHAPE_POLYGON = 1, PT_SHAPE_SPHERE = 2, &
       PT_SHAPE_CYLINDER = 3
  integer(c_int32_t), parameter :: PT_TRANS_VALUE = 1, PT_TRANS_GAMMA = 2, PT_TRANS_TBC = 3
  integer(c_int32_t), parameter :: PT_GRAIN_VORONOI = 1, PT_GRAIN_BRICKS = 2
  integer(c_int32_t), parameter :: PT_PREC_FP64 = 0, PT_PREC_FP32_RETEST = 1, &
       PT_PREC_FP32_FILTER = 2

  ! Surface_type (src/RTMCphon.f90:94-111) as the trace reads it
  type, bind(C) :: pt_surface
    integer(c_int32_t) :: kind = 0
    integer(c_int32_t) :: nvertex = 0
    integer(c_int64_t) :: vertex_offset = 0
    integer(c_int32_t) :: internal = 0
    integer(c_int32_t) :: incident = 0
    integer(c_int32_t) :: grain_inside = 0
    integer(c_int32_t) :: grain_outside = 0
    integer(c_int32_t) :: trans_kind = 0
    integer(c_int32_t) :: reserved = 0
    real(c_double) :: trans_param = 0.0d0
    real(c_double) :: specularity = 0.0d0
    real(c_double) :: center(3) = 0.0d0
    real(c_double) :: radius = 0.0d0
    real(c_double) :: normal(3) = 0.0d0
    real(c_double) :: tangent(9) = 0.0d0      ! tangent(3,3) column-major
  end type pt_surface

  type, bind(C) :: pt_geom_desc
    real(c_double) :: box(3) = 0.0d0
    real(c_double) :: scope(2) = 0.0d0
    integer(c_int32_t) :: kind_grain = 0
    integer(c_int32_t) :: ngrain = 0
    integer(c_int32_t) :: igrain_fill = -1
    integer(c_int32_t) :: nmaterial = 1
    integer(c_int32_t) :: ngrain_total = 0
    integer(c_int32_t) :: nsurface = 0
    integer(c_int64_t) :: nvertex = 0
    type(c_ptr) :: grain_offset = c_null_ptr      ! int32 [ngrain_total+1]
    type(c_ptr) :: grain_incident = c_null_ptr    ! int32 [ngrain_total]
    type(c_ptr) :: generatrix = c_null_ptr        ! double [3*ngrain_total]
    type(c_ptr) :: surfaces = c_null_ptr          ! pt_surface [nsurface]
    type(c_ptr) :: vertices = c_null_ptr          ! double [3*nvertex]
  end type pt_geom_desc

  ! Anphon_type (src/RTMCphon.f90:161-174), arrays column-major (nbranch, nkpoint)
  type, bind(C) :: pt_modes_desc
    integer(c_int32_t) :: nbranch = 0
    integer(c_int32_t) :: nkpoint = 0
    type(c_ptr) :: omega = c_null_ptr
    type(c_ptr) :: vel = c_null_ptr
    type(c_ptr) :: mfp = c_null_ptr
    type(c_ptr) :: weight = c_null_ptr
    real(c_double) :: max_omega = 0.0d0
    real(c_double) :: temperature = 300.0d0
  end type pt_modes_desc

  type, bind(C) :: pt_stats
    integer(c_int64_t) :: paths = 0, events = 0, draws = 0, lost = 0, start_mismatch = 0, &
         skipped = 0
    real(c_double) :: kernel_ms = 0.0d0, reduce_ms = 0.0d0
    integer(c_int64_t) :: fp64_ops = 0, philox_blocks = 0, retests = 0, fallbacks = 0
    integer(c_int64_t) :: bvh_nodes = 0, bvh_tests = 0, event_limited = 0
  end type pt_stats

  type, bind(C) :: pt_path_record
    integer(c_int32_t) :: end_side = 0, flags = 0
    integer(c_int64_t) :: events = 0, draws = 0
    integer(c_int32_t) :: which_grain = 0, start_grain = 0
    real(c_double) :: pos(3) = 0.0d0, dir(3) = 0.0d0
  end type pt_path_record

  interface
    function pt_version() bind(C, name='pt_version') result(v)
      import :: c_int32_t
      integer(c_int32_t) :: v
    end function pt_version

    function pt_last_error(buf, len) bind(C, name='pt_last_error') result(n)
      import :: c_char, c_size_t
      character(kind=c_char), dimension(*) :: buf
      integer(c_size_t), value :: len
      integer(c_size_t) :: n
    end function pt_last_error

    function pt_ctx_create(device, seed, ctx) bind(C, name='pt_ctx_create') result(st)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int), value :: device
      integer(c_int64_t), value :: seed
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: st
    end function pt_ctx_create

    function pt_ctx_destroy(ctx) bind(C, name='pt_ctx_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: st
    end function pt_ctx_destroy

    function pt_ctx_set_seed(ctx, seed) bind(C, name='pt_ctx_set_seed') result(st)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed
      integer(c_int) :: st
    end function pt_ctx_set_seed

    function pt_ctx_set_event_limit(ctx, limit) bind(C, name='pt_ctx_set_event_limit') result(st)
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: limit
      integer(c_int) :: st
    end function pt_ctx_set_event_limit

    function pt_geom_create(ctx, desc, geom) bind(C, name='pt_geom_create') result(st)
      import :: c_int, c_ptr, pt_geom_desc
      type(c_ptr), value :: ctx
      type(pt_geom_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: geom
      integer(c_int) :: st
    end function pt_geom_create

    function pt_geom_destroy(geom) bind(C, name='pt_geom_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: geom
      integer(c_int) :: st
    end function pt_geom_destroy

    function pt_surface_setup_polygon(nvertex, vertices, center, normal, tangent) &
         bind(C, name='pt_surface_setup_polygon') result(st)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: nvertex
      real(c_double), intent(in) :: vertices(*)
      real(c_double), intent(out) :: center(3), normal(3), tangent(9)
      integer(c_int) :: st
    end function pt_surface_setup_polygon

    function pt_modes_create(ctx, desc, modes) bind(C, name='pt_modes_create') result(st)
      import :: c_int, c_ptr, pt_modes_desc
      type(c_ptr), value :: ctx
      type(pt_modes_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: modes
      integer(c_int) :: st
    end function pt_modes_create

    function pt_modes_destroy(modes) bind(C, name='pt_modes_destroy') result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: modes
      integer(c_int) :: st
    end function pt_modes_destroy

    function pt_trace_grid(ctx, geom, modes, mfp_nm, nangle, vangle, np, precision, counts, &
         stats) bind(C, name='pt_trace_grid') result(st)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, geom, modes
      real(c_double), value :: mfp_nm
      integer(c_int32_t), value :: nangle
      type(c_ptr), value :: vangle                 ! double [nangle] or c_null_ptr
      integer(c_int64_t), value :: np
      integer(c_int32_t), value :: precision
      type(c_ptr), value :: counts                 ! int64 [nbranch*nkpoint*nangle]
      type(c_ptr), value :: stats                  ! pt_stats or c_null_ptr
      integer(c_int) :: st
    end function pt_trace_grid

    function pt_trace_log(ctx, geom, modes, mfp_nm, nangle, vangle, np, first_path, stride, n, &
         precision, max_points, points, npoints, records) bind(C, name='pt_trace_log') result(st)
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, geom, modes
      real(c_double), value :: mfp_nm
      integer(c_int32_t), value :: nangle
      type(c_ptr), value :: vangle                 ! double [nangle] or c_null_ptr
      integer(c_int64_t), value :: np, first_path, stride, n
      integer(c_int32_t), value :: precision
      integer(c_int64_t), value :: max_points
      type(c_ptr), value :: points                 ! double [3, max_points, n]
      type(c_ptr), value :: npoints                ! int64 [n]
      type(c_ptr), value :: records                ! pt_path_record [n] or c_null_ptr
      integer(c_int) :: st
    end function pt_trace_log

    function pt_trace_alias(ctx, geom, modes, npaths, precision, branch_scale, mode_paths, &
         mode_trans, stats) bind(C, name='pt_trace_alias') result(st)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx, geom, modes
      integer(c_int64_t), value :: npaths
      integer(c_int32_t), value :: precision
      type(c_ptr), value :: branch_scale           ! double [nbranch] or c_null_ptr
      type(c_ptr), value :: mode_paths, mode_trans ! int64 [nbranch*nkpoint], accumulated
      type(c_ptr), value :: stats
      integer(c_int) :: st
    end function pt_trace_alias
  end interface

contains

  ! The calling thread's last libptrans error message.
  function ptrans_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char) :: buf(512)
    integer(c_size_t) :: n
    integer :: i
    n = pt_last_error(buf, int(size(buf), c_size_t))
    n = min(n, int(size(buf) - 1, c_size_t))
    allocate(character(len=n) :: msg)
    do i = 1, int(n)
      msg(i:i) = buf(i)
    end do
  end function ptrans_error_message

end module ptrans
