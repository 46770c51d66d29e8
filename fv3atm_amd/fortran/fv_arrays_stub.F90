! fv_arrays_stub.F90 -- stand-in modules with the names, types and
! components the tracer transport + remap path reads from the dycore's
! fv_arrays_mod and from FMS mpp_domains_mod.  Linked into FV3, the real
! modules take their place; only the components below are touched.
!
! fv_grid_bounds_type: moving_nest/fv_moving_nest.F90:177-185 (Atm(n)%bd%isd ...)
! fv_grid_type:        moving_nest/fv_moving_nest.F90:81 (gridstruct)
! kind R_GRID (8-byte metrics): moving_nest/bounding_box.F90:32

module mpp_domains_mod
  implicit none
  private
  ! opaque FMS handles: the device path refreshes halos itself
  type, public :: domain2d
     integer :: id = 0
  end type domain2d
  type, public :: group_halo_update_type
     integer :: id = 0
  end type group_halo_update_type
end module mpp_domains_mod

module fv_arrays_mod
  implicit none
  private
  integer, public, parameter :: R_GRID = selected_real_kind(12, 100)
  ! kind of the dycore's default real on this path: 8 bytes, or 4 in the
  ! 32BIT build (OVERLOAD_R4, ci/CMakeLists.txt:20-22,35,59-63); the grid
  ! metrics stay R_GRID either way
#ifdef OVERLOAD_R4
  integer, public, parameter :: FVPRC = selected_real_kind(6, 37)
#else
  integer, public, parameter :: FVPRC = R_GRID
#endif

  type, public :: fv_grid_bounds_type
     integer :: is, ie, js, je
     integer :: isd, ied, jsd, jed
  end type fv_grid_bounds_type

  type, public :: fv_grid_type
     ! (isd:ied, jsd:jed)
     real(kind=R_GRID), allocatable :: area(:,:), rarea(:,:), dxa(:,:), dya(:,:)
     ! dx, del6_u (isd:ied, jsd:jed+1); dy, del6_v (isd:ied+1, jsd:jed)
     real(kind=R_GRID), allocatable :: dx(:,:), dy(:,:), del6_u(:,:), del6_v(:,:)
     ! (isd:ied, jsd:jed, 9)
     real(kind=R_GRID), allocatable :: sin_sg(:,:,:)
     real(kind=R_GRID) :: da_min = 0.0
     logical :: sw_corner = .false., se_corner = .false., ne_corner = .false., nw_corner = .false.
     logical :: bounded_domain = .false., nested = .false.
     integer :: grid_type = 0
  end type fv_grid_type
end module fv_arrays_mod
