! fv_tracer2d_mod.F90 -- tracer_2d_1L and tracer_2d with the dycore's
! argument lists, computed by the MI355X library (fv3_tracer_2d_1l /
! fv3_tracer_2d, include/fv3atm_amd.h).  Replaces fv_tracer2d.F90 of
! GFDL_atmos_cubed_sphere (submodule empty at .gitmodules:1-4), called from
! fv_dynamics, reached from atmosphere_dynamics (atmos_model.F90:879).
!
! Semantics as upstream: q, dp1 updated in place, cx/cy/mfx/mfy scaled by
! 1/nsplit in place.  domain, q_pack and dp1_pack are FMS halo handles the
! device path does not need: the halo of a single tile stays as passed in.
! Built with OVERLOAD_R4 (the 32-bit dycore) the same routines take 4-byte
! arrays and run the FP32 fused kernels (fv3_tracer_2d_1l_r4 / _r4).

module fv_tracer2d_mod
  use iso_c_binding
  use fv_arrays_mod, only: fv_grid_bounds_type, fv_grid_type, FVPRC
  use mpp_domains_mod, only: domain2d, group_halo_update_type
#ifdef OVERLOAD_R4
  use fv3atm_amd_cabi, only: fv3_check, fv3_context, fv3_tracer_2d_1l => fv3_tracer_2d_1l_r4, &
       fv3_tracer_2d => fv3_tracer_2d_r4
#else
  use fv3atm_amd_cabi, only: fv3_check, fv3_context, fv3_tracer_2d_1l, fv3_tracer_2d
#endif
  implicit none
  private
  public :: tracer_2d_1L, tracer_2d
  ! sub-cycle counts of the last call, per level (diagnostic)
  integer, allocatable, public, save :: last_nsplit(:)

contains

  subroutine tracer_2d_1L(q, dp1, mfx, mfy, cx, cy, gridstruct, bd, domain, npx, npy, npz, nq, &
       hord, q_split, dt, id_divg, q_pack, dp1_pack, nord_tr, trdm, lim_fac)
    type(fv_grid_bounds_type), intent(in) :: bd
    integer, intent(in) :: npx, npy, npz, nq, hord, q_split, id_divg, nord_tr
    real(FVPRC), intent(in) :: dt, trdm, lim_fac
    real(FVPRC), intent(inout), target :: q(bd%isd:bd%ied, bd%jsd:bd%jed, npz, nq)
    real(FVPRC), intent(inout), target :: dp1(bd%isd:bd%ied, bd%jsd:bd%jed, npz)
    real(FVPRC), intent(inout), target :: mfx(bd%is:bd%ie+1, bd%js:bd%je, npz)
    real(FVPRC), intent(inout), target :: mfy(bd%is:bd%ie, bd%js:bd%je+1, npz)
    real(FVPRC), intent(inout), target :: cx(bd%is:bd%ie+1, bd%jsd:bd%jed, npz)
    real(FVPRC), intent(inout), target :: cy(bd%isd:bd%ied, bd%js:bd%je+1, npz)
    type(fv_grid_type), intent(in), target :: gridstruct
    type(domain2d), intent(inout) :: domain
    type(group_halo_update_type), intent(inout) :: q_pack, dp1_pack
    type(c_ptr) :: ctx
    integer(c_int), allocatable, target :: ns(:)
    ctx = fv3_context(gridstruct, bd, npx, npy, npz)
    allocate(ns(npz))
    call fv3_check(fv3_tracer_2d_1l(ctx, c_loc(q), c_loc(dp1), c_loc(mfx), c_loc(mfy), c_loc(cx), &
         c_loc(cy), npx, npy, npz, nq, hord, q_split, real(dt, c_double), id_divg, nord_tr, &
         real(trdm, c_double), real(lim_fac, c_double), c_loc(ns)), 'tracer_2d_1L')
    last_nsplit = ns
  end subroutine tracer_2d_1L

  subroutine tracer_2d(q, dp1, mfx, mfy, cx, cy, gridstruct, bd, domain, npx, npy, npz, nq, &
       hord, q_split, dt, id_divg, q_pack, dp1_pack, nord_tr, trdm, lim_fac)
    type(fv_grid_bounds_type), intent(in) :: bd
    integer, intent(in) :: npx, npy, npz, nq, hord, q_split, id_divg, nord_tr
    real(FVPRC), intent(in) :: dt, trdm, lim_fac
    real(FVPRC), intent(inout), target :: q(bd%isd:bd%ied, bd%jsd:bd%jed, npz, nq)
    real(FVPRC), intent(inout), target :: dp1(bd%isd:bd%ied, bd%jsd:bd%jed, npz)
    real(FVPRC), intent(inout), target :: mfx(bd%is:bd%ie+1, bd%js:bd%je, npz)
    real(FVPRC), intent(inout), target :: mfy(bd%is:bd%ie, bd%js:bd%je+1, npz)
    real(FVPRC), intent(inout), target :: cx(bd%is:bd%ie+1, bd%jsd:bd%jed, npz)
    real(FVPRC), intent(inout), target :: cy(bd%isd:bd%ied, bd%js:bd%je+1, npz)
    type(fv_grid_type), intent(in), target :: gridstruct
    type(domain2d), intent(inout) :: domain
    type(group_halo_update_type), intent(inout) :: q_pack, dp1_pack
    type(c_ptr) :: ctx
    integer(c_int), allocatable, target :: ns(:)
    ctx = fv3_context(gridstruct, bd, npx, npy, npz)
    allocate(ns(npz))
    call fv3_check(fv3_tracer_2d(ctx, c_loc(q), c_loc(dp1), c_loc(mfx), c_loc(mfy), c_loc(cx), &
         c_loc(cy), npx, npy, npz, nq, hord, q_split, real(dt, c_double), id_divg, nord_tr, &
         real(trdm, c_double), real(lim_fac, c_double), c_loc(ns)), 'tracer_2d')
    last_nsplit = ns
  end subroutine tracer_2d

end module fv_tracer2d_mod
