! ft_sizes.f90 -- prints c_sizeof of module ptrans's interoperable types, which must
! equal sizeof of the C structs of include/ptrans.h.
program ft_sizes
  use iso_c_binding
  use ptrans
  implicit none
  type(pt_surface) :: s
  type(pt_geom_desc) :: g
  type(pt_modes_desc) :: m
  type(pt_stats) :: st
  type(pt_path_record) :: r
  write(*, '(5(i0, 1x))') c_sizeof(s), c_sizeof(g), c_sizeof(m), c_sizeof(st), c_sizeof(r)
end program ft_sizes
