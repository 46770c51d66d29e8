! fv_mapz_mod.F90 -- the vertical remap routines of fv_mapz.F90 on this path
! (map1_ppm, map_scalar, mapn_tracer, cs_profile) with the dycore's argument
! lists, and the tracer/pt subset of Lagrangian_to_Eulerian, computed by the
! MI355X library (include/fv3atm_amd.h).  fv_mapz_mod is imported by the
! reference at moving_nest/fv_moving_nest.F90:85; its source sits in the
! empty GFDL_atmos_cubed_sphere submodule (.gitmodules:1-4).

module fv_mapz_mod
  use iso_c_binding
  use fv_arrays_mod, only: fv_grid_bounds_type, fv_grid_type, FVPRC
  use fv3atm_amd_cabi, only: fv3_check, fv3_context, fv3_map1_ppm, fv3_map_scalar, fv3_mapn_tracer, &
       fv3_cs_profile, fv3_lagrangian_to_eulerian_tracers, fv3_lagrangian_to_eulerian_tracers_r4
  implicit none
  private
  public :: map1_ppm, map_scalar, mapn_tracer, cs_profile, lagrangian_to_eulerian_tracers

contains

  subroutine map1_ppm(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord)
    integer, intent(in) :: km, kn, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord
    real(FVPRC), intent(in), target :: pe1(i1:i2, km+1), pe2(i1:i2, kn+1), qs(i1:i2)
    real(FVPRC), intent(in), target :: q1(ibeg:iend, jbeg:jend, km)
    real(FVPRC), intent(inout), target :: q2(ibeg:iend, jbeg:jend, kn)
#ifdef OVERLOAD_R4
    ! the row entry points compute in FP64: 8-byte copies of the row
    real(c_double), allocatable, target :: p1(:,:), p2(:,:), s1(:), a1(:,:,:), a2(:,:,:)
    p1 = pe1; p2 = pe2; s1 = qs; a1 = q1; a2 = q2
    call fv3_check(fv3_map1_ppm(km, c_loc(p1), c_loc(a1), c_loc(s1), kn, c_loc(p2), c_loc(a2), &
         i1, i2, j, ibeg, iend, jbeg, jend, iv, kord), 'map1_ppm')
    q2(i1:i2, j, :) = real(a2(i1:i2, j, :), FVPRC)  ! a2 keeps q2's bounds
#else
    call fv3_check(fv3_map1_ppm(km, c_loc(pe1), c_loc(q1), c_loc(qs), kn, c_loc(pe2), c_loc(q2), &
         i1, i2, j, ibeg, iend, jbeg, jend, iv, kord), 'map1_ppm')
#endif
  end subroutine map1_ppm

  subroutine map_scalar(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord, q_min)
    integer, intent(in) :: km, kn, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord
    real(FVPRC), intent(in), target :: pe1(i1:i2, km+1), pe2(i1:i2, kn+1), qs(i1:i2)
    real(FVPRC), intent(in), target :: q1(ibeg:iend, jbeg:jend, km)
    real(FVPRC), intent(inout), target :: q2(ibeg:iend, jbeg:jend, kn)
    real(FVPRC), intent(in) :: q_min
#ifdef OVERLOAD_R4
    real(c_double), allocatable, target :: p1(:,:), p2(:,:), s1(:), a1(:,:,:), a2(:,:,:)
    p1 = pe1; p2 = pe2; s1 = qs; a1 = q1; a2 = q2
    call fv3_check(fv3_map_scalar(km, c_loc(p1), c_loc(a1), c_loc(s1), kn, c_loc(p2), c_loc(a2), &
         i1, i2, j, ibeg, iend, jbeg, jend, iv, kord, real(q_min, c_double)), 'map_scalar')
    q2(i1:i2, j, :) = real(a2(i1:i2, j, :), FVPRC)  ! a2 keeps q2's bounds
#else
    call fv3_check(fv3_map_scalar(km, c_loc(pe1), c_loc(q1), c_loc(qs), kn, c_loc(pe2), c_loc(q2), &
         i1, i2, j, ibeg, iend, jbeg, jend, iv, kord, real(q_min, c_double)), 'map_scalar')
#endif
  end subroutine map_scalar

  subroutine mapn_tracer(nq, km, pe1, pe2, q1, dp2, kord, j, i1, i2, isd, ied, jsd, jed, q_min, fill)
    integer, intent(in) :: nq, km, j, i1, i2, isd, ied, jsd, jed
    integer, intent(in), target :: kord(nq)
    real(FVPRC), intent(in), target :: pe1(i1:i2, km+1), pe2(i1:i2, km+1), dp2(i1:i2, km)
    real(FVPRC), intent(inout), target :: q1(isd:ied, jsd:jed, km, nq)
    real(FVPRC), intent(in) :: q_min
    logical, intent(in) :: fill
    integer(c_int), target :: kord_c(nq)
#ifdef OVERLOAD_R4
    real(c_double), allocatable, target :: p1(:,:), p2(:,:), d2(:,:), a1(:,:,:,:)
    p1 = pe1; p2 = pe2; d2 = dp2; a1 = q1
    kord_c = kord
    call fv3_check(fv3_mapn_tracer(nq, km, c_loc(p1), c_loc(p2), c_loc(a1), c_loc(d2), c_loc(kord_c), &
         j, i1, i2, isd, ied, jsd, jed, real(q_min, c_double), merge(1, 0, fill)), 'mapn_tracer')
    q1(i1:i2, j, :, :) = real(a1(i1:i2, j, :, :), FVPRC)  ! a1 keeps q1's bounds
#else
    kord_c = kord
    call fv3_check(fv3_mapn_tracer(nq, km, c_loc(pe1), c_loc(pe2), c_loc(q1), c_loc(dp2), c_loc(kord_c), &
         j, i1, i2, isd, ied, jsd, jed, real(q_min, c_double), merge(1, 0, fill)), 'mapn_tracer')
#endif
  end subroutine mapn_tracer

  subroutine cs_profile(qs, a4, delp, km, i1, i2, iv, kord)
    integer, intent(in) :: km, i1, i2, iv, kord
    real(FVPRC), intent(in), target :: qs(i1:i2), delp(i1:i2, km)
    real(FVPRC), intent(inout), target :: a4(4, i1:i2, km)
#ifdef OVERLOAD_R4
    real(c_double), allocatable, target :: s1(:), d1(:,:), b4(:,:,:)
    s1 = qs; d1 = delp; b4 = a4
    call fv3_check(fv3_cs_profile(c_loc(s1), c_loc(b4), c_loc(d1), km, i1, i2, iv, kord), 'cs_profile')
    a4 = real(b4, FVPRC)
#else
    call fv3_check(fv3_cs_profile(c_loc(qs), c_loc(a4), c_loc(delp), km, i1, i2, iv, kord), 'cs_profile')
#endif
  end subroutine cs_profile

  ! Tracer/pt part of Lagrangian_to_Eulerian for one tile: pe1 = ptop +
  ! cumsum(delp), pe2 = ak + bk*ps; tracers by mapn_tracer (nq > 5) or map1_q2
  ! (iv = 0, kord_tr), pt by map_scalar (iv = 1, kord_tm, floor t_min);
  ! delp <- dp2; fillz when fill.
  subroutine lagrangian_to_eulerian_tracers(km, nq, ak, bk, ptop, delp, q, pt, kord_tr, kord_tm, &
       t_min, fill, gridstruct, bd, npx, npy)
    type(fv_grid_bounds_type), intent(in) :: bd
    integer, intent(in) :: km, nq, kord_tm, npx, npy
    integer, intent(in) :: kord_tr(nq)
    real(FVPRC), intent(in), target :: ak(km+1), bk(km+1)
    real(FVPRC), intent(in) :: ptop, t_min
    real(FVPRC), intent(inout), target :: delp(bd%isd:bd%ied, bd%jsd:bd%jed, km)
    real(FVPRC), intent(inout), target :: q(bd%isd:bd%ied, bd%jsd:bd%jed, km, nq)
    real(FVPRC), intent(inout), target :: pt(bd%isd:bd%ied, bd%jsd:bd%jed, km)
    logical, intent(in) :: fill
    type(fv_grid_type), intent(in), target :: gridstruct
    type(c_ptr) :: ctx
    integer(c_int), target :: kt(nq)
#ifdef OVERLOAD_R4
    real(c_double), target :: ak8(km+1), bk8(km+1)
#endif
    ctx = fv3_context(gridstruct, bd, npx, npy, km)
    kt = kord_tr
#ifdef OVERLOAD_R4
    ! the 32-bit build: 4-byte delp, q, pt through the FP32 column-shared remap
    ak8 = ak
    bk8 = bk
    call fv3_check(fv3_lagrangian_to_eulerian_tracers_r4(ctx, km, nq, c_loc(ak8), c_loc(bk8), &
         real(ptop, c_double), c_loc(delp), c_loc(q), c_loc(pt), c_loc(kt), kord_tm, &
         real(t_min, c_double), merge(1, 0, fill)), 'lagrangian_to_eulerian_tracers')
#else
    call fv3_check(fv3_lagrangian_to_eulerian_tracers(ctx, km, nq, c_loc(ak), c_loc(bk), &
         real(ptop, c_double), c_loc(delp), c_loc(q), c_loc(pt), c_loc(kt), kord_tm, &
         real(t_min, c_double), merge(1, 0, fill)), 'lagrangian_to_eulerian_tracers')
#endif
  end subroutine lagrangian_to_eulerian_tracers

end module fv_mapz_mod
