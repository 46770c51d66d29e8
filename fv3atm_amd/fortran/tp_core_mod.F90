! tp_core_mod.F90 -- fv_tp_2d with the dycore's argument list, computed by
! the MI355X library (fv3_fv_tp_2d, include/fv3atm_amd.h).  Replaces
! tp_core.F90 fv_tp_2d of GFDL_atmos_cubed_sphere (the submodule left empty at
! .gitmodules:1-4); the caller is tracer_2d / tracer_2d_1L.

module tp_core_mod
  use iso_c_binding
  use fv_arrays_mod, only: fv_grid_bounds_type, fv_grid_type, FVPRC
  use fv3atm_amd_cabi, only: fv3_check, fv3_context, fv3_fv_tp_2d
  implicit none
  private
  public :: fv_tp_2d

contains

  subroutine fv_tp_2d(q, crx, cry, npx, npy, hord, fx, fy, xfx, yfx, gridstruct, bd, ra_x, ra_y, &
       lim_fac, mfx, mfy, mass, nord, damp_c)
    type(fv_grid_bounds_type), intent(in) :: bd
    integer, intent(in) :: npx, npy, hord
    real(FVPRC), intent(inout), target :: q(bd%isd:bd%ied, bd%jsd:bd%jed)
    real(FVPRC), intent(in), target :: crx(bd%is:bd%ie+1, bd%jsd:bd%jed)
    real(FVPRC), intent(in), target :: xfx(bd%is:bd%ie+1, bd%jsd:bd%jed)
    real(FVPRC), intent(in), target :: cry(bd%isd:bd%ied, bd%js:bd%je+1)
    real(FVPRC), intent(in), target :: yfx(bd%isd:bd%ied, bd%js:bd%je+1)
    real(FVPRC), intent(in), target :: ra_x(bd%is:bd%ie, bd%jsd:bd%jed)
    real(FVPRC), intent(in), target :: ra_y(bd%isd:bd%ied, bd%js:bd%je)
    real(FVPRC), intent(out), target :: fx(bd%is:bd%ie+1, bd%js:bd%je)
    real(FVPRC), intent(out), target :: fy(bd%is:bd%ie, bd%js:bd%je+1)
    type(fv_grid_type), intent(in), target :: gridstruct
    real(FVPRC), intent(in) :: lim_fac
    real(FVPRC), optional, intent(in), target :: mfx(bd%is:bd%ie+1, bd%js:bd%je)
    real(FVPRC), optional, intent(in), target :: mfy(bd%is:bd%ie, bd%js:bd%je+1)
    real(FVPRC), optional, intent(in), target :: mass(bd%isd:bd%ied, bd%jsd:bd%jed)
    integer, optional, intent(in), target :: nord
    real(FVPRC), optional, intent(in), target :: damp_c
    type(c_ptr) :: ctx, p_mfx, p_mfy, p_mass, p_nord, p_damp
    integer(c_int), target :: nord_c
    real(c_double), target :: damp_cc
    ctx = fv3_context(gridstruct, bd, npx, npy, 1)
    p_mfx = c_null_ptr
    p_mfy = c_null_ptr
    p_mass = c_null_ptr
    p_nord = c_null_ptr
    p_damp = c_null_ptr
    if (present(mfx)) p_mfx = c_loc(mfx)
    if (present(mfy)) p_mfy = c_loc(mfy)
    if (present(mass)) p_mass = c_loc(mass)
    if (present(nord)) then
       nord_c = nord
       p_nord = c_loc(nord_c)
    end if
    if (present(damp_c)) then
       damp_cc = damp_c
       p_damp = c_loc(damp_cc)
    end if
#ifdef OVERLOAD_R4
    block
      ! fv_tp_2d computes in FP64: 8-byte copies of the field and its inputs
      real(c_double), allocatable, target :: q8(:,:), crx8(:,:), cry8(:,:), fx8(:,:), fy8(:,:), &
           xfx8(:,:), yfx8(:,:), rax8(:,:), ray8(:,:), mfx8(:,:), mfy8(:,:), mass8(:,:)
      q8 = q; crx8 = crx; cry8 = cry; xfx8 = xfx; yfx8 = yfx; rax8 = ra_x; ray8 = ra_y
      allocate(fx8(size(fx, 1), size(fx, 2)), fy8(size(fy, 1), size(fy, 2)))
      if (present(mfx)) then
         mfx8 = mfx
         p_mfx = c_loc(mfx8)
      end if
      if (present(mfy)) then
         mfy8 = mfy
         p_mfy = c_loc(mfy8)
      end if
      if (present(mass)) then
         mass8 = mass
         p_mass = c_loc(mass8)
      end if
      call fv3_check(fv3_fv_tp_2d(ctx, c_loc(q8), c_loc(crx8), c_loc(cry8), npx, npy, hord, c_loc(fx8), &
           c_loc(fy8), c_loc(xfx8), c_loc(yfx8), c_loc(rax8), c_loc(ray8), real(lim_fac, c_double), &
           p_mfx, p_mfy, p_mass, p_nord, p_damp), 'fv_tp_2d')
      q = real(q8, FVPRC)
      fx = real(fx8, FVPRC)
      fy = real(fy8, FVPRC)
    end block
#else
    call fv3_check(fv3_fv_tp_2d(ctx, c_loc(q), c_loc(crx), c_loc(cry), npx, npy, hord, c_loc(fx), &
         c_loc(fy), c_loc(xfx), c_loc(yfx), c_loc(ra_x), c_loc(ra_y), real(lim_fac, c_double), &
         p_mfx, p_mfy, p_mass, p_nord, p_damp), 'fv_tp_2d')
#endif
  end subroutine fv_tp_2d

end module tp_core_mod
