! harness.F90 -- a Fortran caller of the path through the dycore-shaped
! interfaces: reads one tile's inputs from a stream file, runs
!   tracer_2d_1L (fv_tracer2d_mod) and then the tracer/pt remap
!   (fv_mapz_mod lagrangian_to_eulerian_tracers)
! the way fv_dynamics sequences them, and writes the results.
!
!   fv3_harness IN OUT            run the step
!   fv3_harness IN OUT --regrid   run it once, then change the metrics in place
!                                 (area x 1.1, rarea = 1/area, dxa x 1.05) with
!                                 the same bounds and run it again from the
!                                 same state: the output is the second run's
!   fv3_harness --host            argument checks that fail before any device work
! Built with OVERLOAD_R4 (fv3_harness_r4) the arrays are the 32-bit build's
! 4-byte reals; the files stay float64.
!
! IN (little-endian stream): int32 n, npz, nq, hord, kord_tr, kord_tm, fill,
! sw, se, ne, nw, grid_type; float64 ptop, t_min, dt, trdm, lim_fac, da_min;
! then float64 arrays in column-major order: area, rarea, dxa, dya (nd,nd),
! dx (nd,nd+1), dy (nd+1,nd), sin_sg (nd,nd,9), del6_u (nd,nd+1),
! del6_v (nd+1,nd), q (nd,nd,npz,nq),
! dp1 (nd,nd,npz), mfx (n+1,n,npz), mfy (n,n+1,npz), cx (n+1,nd,npz),
! cy (nd,n+1,npz), ak, bk (npz+1), pt (nd,nd,npz); nd = n + 6.
! OUT: q, dp1 after transport, delp and pt after the remap, int32 nsplit(npz).

program fv3_harness
  use iso_c_binding
  use fv_arrays_mod
  use mpp_domains_mod
  use fv3atm_amd_cabi, only: fv3_map1_ppm, fv3_error_text, fv3_context_release
  use fv_tracer2d_mod, only: tracer_2d_1L, last_nsplit
  use fv_mapz_mod, only: lagrangian_to_eulerian_tracers
  implicit none
  character(len=4096) :: fin, fout, opt
  integer :: hdr(12), n, npz, nq, hord, kord_tr, kord_tm, nd, u
  real(R_GRID) :: rhdr(6)
  type(fv_grid_bounds_type) :: bd
  type(fv_grid_type), target :: gs
  type(domain2d) :: domain
  type(group_halo_update_type) :: q_pack, dp1_pack
  real(FVPRC), allocatable, target :: q(:,:,:,:), dp1(:,:,:), mfx(:,:,:), mfy(:,:,:), cx(:,:,:), &
       cy(:,:,:), ak(:), bk(:), pt(:,:,:), delp(:,:,:)
  ! the inputs as read (float64)
  real(R_GRID), allocatable :: q8(:,:,:,:), dp18(:,:,:), mfx8(:,:,:), mfy8(:,:,:), cx8(:,:,:), &
       cy8(:,:,:), ak8(:), bk8(:), pt8(:,:,:)
  integer, allocatable :: kord(:)

  call get_command_argument(1, fin)
  if (trim(fin) == '--host') then
     call host_checks()
     stop
  end if
  call get_command_argument(2, fout)
  call get_command_argument(3, opt)

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) hdr
  read(u) rhdr
  n = hdr(1); npz = hdr(2); nq = hdr(3); hord = hdr(4); kord_tr = hdr(5); kord_tm = hdr(6)
  nd = n + 6
  bd = fv_grid_bounds_type(1, n, 1, n, -2, n + 3, -2, n + 3)
  allocate(gs%area(-2:n+3, -2:n+3), gs%rarea(-2:n+3, -2:n+3), gs%dxa(-2:n+3, -2:n+3), &
       gs%dya(-2:n+3, -2:n+3), gs%dx(-2:n+3, -2:n+4), gs%dy(-2:n+4, -2:n+3), &
       gs%sin_sg(-2:n+3, -2:n+3, 9), gs%del6_u(-2:n+3, -2:n+4), gs%del6_v(-2:n+4, -2:n+3))
  read(u) gs%area, gs%rarea, gs%dxa, gs%dya, gs%dx, gs%dy, gs%sin_sg, gs%del6_u, gs%del6_v
  gs%sw_corner = hdr(8) /= 0
  gs%se_corner = hdr(9) /= 0
  gs%ne_corner = hdr(10) /= 0
  gs%nw_corner = hdr(11) /= 0
  gs%grid_type = hdr(12)
  gs%da_min = rhdr(6)
  allocate(q8(-2:n+3, -2:n+3, npz, nq), dp18(-2:n+3, -2:n+3, npz), mfx8(1:n+1, 1:n, npz), &
       mfy8(1:n, 1:n+1, npz), cx8(1:n+1, -2:n+3, npz), cy8(-2:n+3, 1:n+1, npz), ak8(npz+1), bk8(npz+1), &
       pt8(-2:n+3, -2:n+3, npz))
  read(u) q8, dp18, mfx8, mfy8, cx8, cy8, ak8, bk8, pt8
  close(u)
  allocate(kord(nq))
  kord = kord_tr

  call run_step()
  if (trim(opt) == '--regrid') then
     ! a gridstruct recomputed in place with the same bounds (a moving nest)
     gs%area = gs%area * 1.1_R_GRID
     gs%rarea = 1.0_R_GRID / gs%area
     gs%dxa = gs%dxa * 1.05_R_GRID
     call run_step()
  end if
  call fv3_context_release()

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) real(q, R_GRID), real(dp1, R_GRID), real(delp, R_GRID), real(pt, R_GRID), &
       int(last_nsplit, c_int32_t)
  close(u)

contains

  ! tracer_2d_1L, then the tracer/pt remap from the Lagrangian thickness the
  ! step starts from, as fv_dynamics sequences them
  subroutine run_step()
    q = real(q8, FVPRC)
    dp1 = real(dp18, FVPRC)
    mfx = real(mfx8, FVPRC)
    mfy = real(mfy8, FVPRC)
    cx = real(cx8, FVPRC)
    cy = real(cy8, FVPRC)
    ak = real(ak8, FVPRC)
    bk = real(bk8, FVPRC)
    pt = real(pt8, FVPRC)
    delp = dp1
    call tracer_2d_1L(q, dp1, mfx, mfy, cx, cy, gs, bd, domain, n + 1, n + 1, npz, nq, hord, 0, &
         real(rhdr(3), FVPRC), 0, q_pack, dp1_pack, 0, real(rhdr(4), FVPRC), real(rhdr(5), FVPRC))
    call lagrangian_to_eulerian_tracers(npz, nq, ak, bk, real(rhdr(1), FVPRC), delp, q, pt, kord, &
         kord_tm, real(rhdr(2), FVPRC), hdr(7) /= 0, gs, bd, n + 1, n + 1)
  end subroutine run_step

  subroutine host_checks()
    real(c_double), target :: a(64)
    integer(c_int) :: rc
    a = 0.0_c_double
    ! kord 7 (ppm_profile) is outside this path: refused before any device work
    rc = fv3_map1_ppm(8, c_loc(a), c_loc(a), c_loc(a), 8, c_loc(a), c_loc(a), 1, 2, 1, 1, 2, 1, 1, 0, 7)
    if (rc == 0 .or. index(fv3_error_text(), 'kord') == 0) error stop 2
    rc = fv3_map1_ppm(8, c_null_ptr, c_loc(a), c_null_ptr, 8, c_loc(a), c_loc(a), 1, 2, 1, 1, 2, 1, 1, 0, 9)
    if (rc == 0 .or. index(fv3_error_text(), 'NULL') == 0) error stop 3
    print '(a)', 'host checks passed: ' // fv3_error_text()
  end subroutine host_checks

end program fv3_harness
