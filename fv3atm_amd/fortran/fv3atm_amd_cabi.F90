! fv3atm_amd_cabi.F90 -- ISO_C_BINDING declarations of libfv3atm_amd.so
! (include/fv3atm_amd.h) and the context the Fortran interfaces share.
!
! Every entry point returns an integer status; fv3_check turns a nonzero
! status into the dycore's fatal-error behaviour (mpp_error(FATAL) aborts,
! atmos_model.F90:657-658): print the library's last-error text, then stop.

module fv3atm_amd_cabi
  use iso_c_binding
  use fv_arrays_mod, only: fv_grid_bounds_type, fv_grid_type, R_GRID
  implicit none
  private

  integer, public, parameter :: FV3_FORM_PLAIN64 = 0, FV3_FORM_FUSED64 = 1, FV3_FORM_FUSED32 = 2

  type, public, bind(C) :: fv3_bounds_c
     integer(c_int) :: is, ie, js, je, isd, ied, jsd, jed
  end type fv3_bounds_c

  type, public, bind(C) :: fv3_grid_c
     type(c_ptr) :: area, rarea, dx, dy, dxa, dya, sin_sg, del6_u, del6_v
     real(c_double) :: da_min
     integer(c_int) :: sw_corner, se_corner, nw_corner, ne_corner, bounded_domain, grid_type
  end type fv3_grid_c

  public :: fv3_check, fv3_context, fv3_context_release, fv3_error_text

  interface
     function fv3_last_error() bind(C, name="fv3_last_error")
       import :: c_ptr
       type(c_ptr) :: fv3_last_error
     end function fv3_last_error

     integer(c_int) function fv3_ctx_create(bd, npx, npy, npz, ntiles, ctx) bind(C, name="fv3_ctx_create")
       import :: c_int, c_ptr, fv3_bounds_c
       type(fv3_bounds_c), intent(in) :: bd
       integer(c_int), value :: npx, npy, npz, ntiles
       type(c_ptr), intent(out) :: ctx
     end function fv3_ctx_create

     integer(c_int) function fv3_ctx_destroy(ctx) bind(C, name="fv3_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fv3_ctx_destroy

     integer(c_int) function fv3_ctx_set_grid(ctx, tile, g) bind(C, name="fv3_ctx_set_grid")
       import :: c_int, c_ptr, fv3_grid_c
       type(c_ptr), value :: ctx
       integer(c_int), value :: tile
       type(fv3_grid_c), intent(in) :: g
     end function fv3_ctx_set_grid

     integer(c_int) function fv3_fv_tp_2d(ctx, q, crx, cry, npx, npy, hord, fx, fy, xfx, yfx, &
          ra_x, ra_y, lim_fac, mfx, mfy, mass, nord, damp_c) bind(C, name="fv3_fv_tp_2d")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, crx, cry, fx, fy, xfx, yfx, ra_x, ra_y, mfx, mfy, mass, nord, damp_c
       integer(c_int), value :: npx, npy, hord
       real(c_double), value :: lim_fac
     end function fv3_fv_tp_2d

     integer(c_int) function fv3_tracer_2d_1l(ctx, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, &
          hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac, nsplit_out) bind(C, name="fv3_tracer_2d_1l")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, dp1, mfx, mfy, cx, cy, nsplit_out
       integer(c_int), value :: npx, npy, npz, nq, hord, q_split, id_divg, nord_tr
       real(c_double), value :: dt, trdm, lim_fac
     end function fv3_tracer_2d_1l

     integer(c_int) function fv3_tracer_2d(ctx, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, &
          hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac, nsplit_out) bind(C, name="fv3_tracer_2d")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, dp1, mfx, mfy, cx, cy, nsplit_out
       integer(c_int), value :: npx, npy, npz, nq, hord, q_split, id_divg, nord_tr
       real(c_double), value :: dt, trdm, lim_fac
     end function fv3_tracer_2d

     integer(c_int) function fv3_map1_ppm(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, &
          jbeg, jend, iv, kord) bind(C, name="fv3_map1_ppm")
       import :: c_int, c_ptr
       integer(c_int), value :: km, kn, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord
       type(c_ptr), value :: pe1, q1, qs, pe2, q2
     end function fv3_map1_ppm

     integer(c_int) function fv3_map_scalar(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, &
          jbeg, jend, iv, kord, q_min) bind(C, name="fv3_map_scalar")
       import :: c_int, c_ptr, c_double
       integer(c_int), value :: km, kn, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord
       type(c_ptr), value :: pe1, q1, qs, pe2, q2
       real(c_double), value :: q_min
     end function fv3_map_scalar

     integer(c_int) function fv3_mapn_tracer(nq, km, pe1, pe2, q1, dp2, kord, j, i1, i2, isd, ied, &
          jsd, jed, q_min, fill) bind(C, name="fv3_mapn_tracer")
       import :: c_int, c_ptr, c_double
       integer(c_int), value :: nq, km, j, i1, i2, isd, ied, jsd, jed, fill
       type(c_ptr), value :: pe1, pe2, q1, dp2, kord
       real(c_double), value :: q_min
     end function fv3_mapn_tracer

     integer(c_int) function fv3_cs_profile(qs, a4, delp, km, i1, i2, iv, kord) bind(C, name="fv3_cs_profile")
       import :: c_int, c_ptr
       type(c_ptr), value :: qs, a4, delp
       integer(c_int), value :: km, i1, i2, iv, kord
     end function fv3_cs_profile

     integer(c_int) function fv3_lagrangian_to_eulerian_tracers(ctx, km, nq, ak, bk, ptop, delp, q, &
          pt, kord_tr, kord_tm, t_min, fill) bind(C, name="fv3_lagrangian_to_eulerian_tracers")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, ak, bk, delp, q, pt, kord_tr
       integer(c_int), value :: km, nq, kord_tm, fill
       real(c_double), value :: ptop, t_min
     end function fv3_lagrangian_to_eulerian_tracers

     integer(c_int) function fv3_step_dev(ctx, form, nq, q_in, q_out, pt, dp1, delp_out, cx, cy, &
          mfx, mfy, ak, bk, ptop, hord, kord_tr, kord_tm, t_min, nsplit_cap, stream) bind(C, name="fv3_step_dev")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk, stream
       integer(c_int), value :: form, nq, hord, kord_tr, kord_tm, nsplit_cap
       real(c_double), value :: ptop, t_min
     end function fv3_step_dev

     integer(c_int) function fv3_ctx_set_timing(ctx, on) bind(C, name="fv3_ctx_set_timing")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function fv3_ctx_set_timing

     integer(c_int) function fv3_ctx_set_launch(ctx, cfg, rows, remap) bind(C, name="fv3_ctx_set_launch")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: cfg, rows, remap
     end function fv3_ctx_set_launch

     integer(c_int) function fv3_ctx_set_concurrency(ctx, on) bind(C, name="fv3_ctx_set_concurrency")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function fv3_ctx_set_concurrency

     integer(c_int) function fv3_ctx_set_pipeline(ctx, tracers_per_chunk) bind(C, name="fv3_ctx_set_pipeline")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tracers_per_chunk
     end function fv3_ctx_set_pipeline

     integer(c_int) function fv3_ctx_set_transport(ctx, per_level, q_split) bind(C, name="fv3_ctx_set_transport")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: per_level, q_split
     end function fv3_ctx_set_transport

     integer(c_int) function fv3_step_timing(ctx, transport_ms, remap_ms, passes) bind(C, name="fv3_step_timing")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: transport_ms, remap_ms
       integer(c_int), intent(out) :: passes
     end function fv3_step_timing

     integer(c_int) function fv3_tracer_2d_1l_r4(ctx, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, &
          hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac, nsplit_out) bind(C, name="fv3_tracer_2d_1l_r4")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, dp1, mfx, mfy, cx, cy, nsplit_out
       integer(c_int), value :: npx, npy, npz, nq, hord, q_split, id_divg, nord_tr
       real(c_double), value :: dt, trdm, lim_fac
     end function fv3_tracer_2d_1l_r4

     integer(c_int) function fv3_tracer_2d_r4(ctx, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, &
          hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac, nsplit_out) bind(C, name="fv3_tracer_2d_r4")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q, dp1, mfx, mfy, cx, cy, nsplit_out
       integer(c_int), value :: npx, npy, npz, nq, hord, q_split, id_divg, nord_tr
       real(c_double), value :: dt, trdm, lim_fac
     end function fv3_tracer_2d_r4

     integer(c_int) function fv3_lagrangian_to_eulerian_tracers_r4(ctx, km, nq, ak, bk, ptop, delp, q, &
          pt, kord_tr, kord_tm, t_min, fill) bind(C, name="fv3_lagrangian_to_eulerian_tracers_r4")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, ak, bk, delp, q, pt, kord_tr
       integer(c_int), value :: km, nq, kord_tm, fill
       real(c_double), value :: ptop, t_min
     end function fv3_lagrangian_to_eulerian_tracers_r4

     integer(c_int) function fv3_step_dev_noalloc(ctx, form, nq, q_in, q_out, pt, dp1, delp_out, cx, cy, &
          mfx, mfy, ak, bk, ptop, hord, kord_tr, kord_tm, t_min, nsplit_cap, stream) &
          bind(C, name="fv3_step_dev_noalloc")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk, stream
       integer(c_int), value :: form, nq, hord, kord_tr, kord_tm, nsplit_cap
       real(c_double), value :: ptop, t_min
     end function fv3_step_dev_noalloc

     integer(c_int) function fv3_ctx_set_graph_safe(ctx, on) bind(C, name="fv3_ctx_set_graph_safe")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function fv3_ctx_set_graph_safe

     integer(c_int) function fv3_step_status(ctx, nsplit_overflow) bind(C, name="fv3_step_status")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: nsplit_overflow
     end function fv3_step_status

     integer(c_int) function fv3_step_tl_dev(ctx, per_level, q_split, nq, q0, dq0, q1, dq1, dp1, &
          delp_out, cx, cy, mfx, mfy, ak, bk, ptop, hord, kord_tr, stream) bind(C, name="fv3_step_tl_dev")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q0, dq0, q1, dq1, dp1, delp_out, cx, cy, mfx, mfy, ak, bk, stream
       integer(c_int), value :: per_level, q_split, nq, hord, kord_tr
       real(c_double), value :: ptop
     end function fv3_step_tl_dev

     integer(c_int) function fv3_step_ad_dev(ctx, per_level, q_split, nq, q0, qbar1, qbar0, dp1, cx, &
          cy, mfx, mfy, ak, bk, ptop, hord, kord_tr, stream) bind(C, name="fv3_step_ad_dev")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, q0, qbar1, qbar0, dp1, cx, cy, mfx, mfy, ak, bk, stream
       integer(c_int), value :: per_level, q_split, nq, hord, kord_tr
       real(c_double), value :: ptop
     end function fv3_step_ad_dev

     integer(c_int) function fv3_halo_update(ctx, field, nslab, elem_bytes, stream) bind(C, name="fv3_halo_update")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, field, stream
       integer(c_int), value :: nslab, elem_bytes
     end function fv3_halo_update

     integer(c_int) function fv3_halo_update_vector(ctx, u, v, nslab, grid, elem_bytes, stream) &
          bind(C, name="fv3_halo_update_vector")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, u, v, stream
       integer(c_int), value :: nslab, grid, elem_bytes
     end function fv3_halo_update_vector

     integer(c_int) function fv3_kmp(npz, p_ref, ak, bk, kmp) bind(C, name="fv3_kmp")
       import :: c_int, c_ptr, c_double
       integer(c_int), value :: npz
       real(c_double), value :: p_ref
       type(c_ptr), value :: ak, bk
       integer(c_int), intent(out) :: kmp
     end function fv3_kmp

     function c_strlen(s) bind(C, name="strlen")
       import :: c_ptr, c_size_t
       type(c_ptr), value :: s
       integer(c_size_t) :: c_strlen
     end function c_strlen
  end interface

  public :: fv3_last_error, fv3_ctx_create, fv3_ctx_destroy, fv3_ctx_set_grid, fv3_fv_tp_2d
  public :: fv3_tracer_2d_1l, fv3_tracer_2d, fv3_map1_ppm, fv3_map_scalar, fv3_mapn_tracer
  public :: fv3_cs_profile, fv3_lagrangian_to_eulerian_tracers, fv3_step_dev
  public :: fv3_ctx_set_timing, fv3_ctx_set_launch, fv3_ctx_set_concurrency, fv3_ctx_set_transport, &
            fv3_ctx_set_pipeline
  public :: fv3_step_timing
  public :: fv3_tracer_2d_1l_r4, fv3_tracer_2d_r4, fv3_lagrangian_to_eulerian_tracers_r4
  public :: fv3_step_dev_noalloc, fv3_ctx_set_graph_safe, fv3_step_status, fv3_step_tl_dev, &
            fv3_step_ad_dev, fv3_halo_update, fv3_halo_update_vector, fv3_kmp

  ! Device contexts, one per (bounds, npx, npy, npz): fv_tp_2d (npz = 1) and
  ! tracer_2d (npz levels) alternate without re-creating them.  The metrics
  ! are uploaded on every call, so a gridstruct recomputed in place (a moving
  ! nest, fv_moving_nest.F90:1607-1613) is always the one the kernels see.
  integer, parameter :: NCTX = 4
  type(c_ptr), save :: ctx_cached(NCTX) = c_null_ptr
  integer, save :: key(11, NCTX) = -huge(1)
  integer, save :: ctx_next = 1

contains

  function fv3_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: n, i
    p = fv3_last_error()
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = s(i)
    end do
  end function fv3_error_text

  subroutine fv3_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
       write(0, '(a)') 'FATAL from ' // where // ': ' // fv3_error_text()
       error stop 1
    end if
  end subroutine fv3_check

  ! The device context for one tile of gridstruct: found (or created) by its
  ! shape, then given the current metrics.
  function fv3_context(gridstruct, bd, npx, npy, npz) result(ctx)
    type(fv_grid_type), intent(in), target :: gridstruct
    type(fv_grid_bounds_type), intent(in) :: bd
    integer, intent(in) :: npx, npy, npz
    type(c_ptr) :: ctx
    integer :: k(11), slot, i
    type(fv3_bounds_c) :: b
    type(fv3_grid_c) :: g
    k = [bd%is, bd%ie, bd%js, bd%je, bd%isd, bd%ied, bd%jsd, bd%jed, npx, npy, npz]
    slot = 0
    do i = 1, NCTX
       if (c_associated(ctx_cached(i)) .and. all(k == key(:, i))) slot = i
    end do
    if (slot == 0) then
       slot = ctx_next
       ctx_next = mod(ctx_next, NCTX) + 1
       if (c_associated(ctx_cached(slot))) &
            call fv3_check(fv3_ctx_destroy(ctx_cached(slot)), 'fv3_ctx_destroy')
       b = fv3_bounds_c(bd%is, bd%ie, bd%js, bd%je, bd%isd, bd%ied, bd%jsd, bd%jed)
       call fv3_check(fv3_ctx_create(b, npx, npy, npz, 1, ctx_cached(slot)), 'fv3_ctx_create')
       key(:, slot) = k
    end if
    g%area = c_loc(gridstruct%area)
    g%rarea = c_loc(gridstruct%rarea)
    g%dx = c_loc(gridstruct%dx)
    g%dy = c_loc(gridstruct%dy)
    g%dxa = c_loc(gridstruct%dxa)
    g%dya = c_loc(gridstruct%dya)
    g%sin_sg = c_loc(gridstruct%sin_sg)
    g%del6_u = c_null_ptr
    g%del6_v = c_null_ptr
    if (allocated(gridstruct%del6_u)) g%del6_u = c_loc(gridstruct%del6_u)
    if (allocated(gridstruct%del6_v)) g%del6_v = c_loc(gridstruct%del6_v)
    g%da_min = gridstruct%da_min
    g%sw_corner = merge(1, 0, gridstruct%sw_corner)
    g%se_corner = merge(1, 0, gridstruct%se_corner)
    g%nw_corner = merge(1, 0, gridstruct%nw_corner)
    g%ne_corner = merge(1, 0, gridstruct%ne_corner)
    g%bounded_domain = merge(1, 0, gridstruct%bounded_domain)
    g%grid_type = gridstruct%grid_type
    call fv3_check(fv3_ctx_set_grid(ctx_cached(slot), 1, g), 'fv3_ctx_set_grid')
    ctx = ctx_cached(slot)
  end function fv3_context

  subroutine fv3_context_release()
    integer :: i
    do i = 1, NCTX
       if (c_associated(ctx_cached(i))) call fv3_check(fv3_ctx_destroy(ctx_cached(i)), 'fv3_ctx_destroy')
       ctx_cached(i) = c_null_ptr
       key(:, i) = -huge(1)
    end do
    ctx_next = 1
  end subroutine fv3_context_release

end module fv3atm_amd_cabi
