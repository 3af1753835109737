! Host program: the Taylor-Green case (periodic 64^3 D3Q27, Ma 0.1, Re 1600,
! seeded 1e-4 noise) advanced through the Fortran time loop and the C ABI,
! then rho, u, v, w written as raw FP64 (x fastest) to the file named on the
! command line, followed by the step statistics on stdout.
!   uclbs_driver <out.bin> [n] [steps] [precision 0|1] [levels 1|2]
! With two levels, level 1 refines the middle of the box (node indices n/2 ..
! 3n/2 of the fine index space) and the loop advances both with subcycling.
program uclbs_driver
  use iso_c_binding
  use uclbs_hip
  use uclbs_timestep
  implicit none
  type(uclbs_config) :: cfg
  type(uclbs_stats) :: st
  type(c_ptr) :: ctx
  integer(c_int32_t) :: lo(3), hi(3), per(3)
  integer(c_int64_t) :: timestep(0:0), nn
  real(c_double) :: time(0:0), dt(0:0), u0, pi
  real(c_double), allocatable :: rho(:), u(:), v(:), w(:)
  character(len=256) :: outname, arg
  integer :: n, steps, prec, unit_out, nlev, k

  call get_command_argument(1, outname)
  n = 64
  steps = 200
  prec = 0
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) n
  end if
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    read(arg, *) steps
  end if
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) prec
  end if
  nlev = 1
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    read(arg, *) nlev
  end if
  pi = 4.0_c_double * atan(1.0_c_double)
  u0 = 0.1_c_double / sqrt(3.0_c_double)
  cfg%lattice = UCLBS_LATTICE_D3Q27
  cfg%precision = prec
  cfg%alpha_mode = UCLBS_ALPHA_NEWTON
  cfg%alpha_tol = 1.0e-10_c_double
  cfg%nu = u0 * (real(n, c_double) / (2.0_c_double * pi)) / 1600.0_c_double
  call uclbs_check(uclbs_create(cfg, ctx), 'uclbs_create')
  lo = 0
  hi = n - 1
  per = 1
  call uclbs_check(uclbs_define_level(ctx, 0_c_int, lo, hi, per), 'uclbs_define_level')
  call uclbs_check(uclbs_init_tgv(ctx, 0_c_int, u0, 1.0e-4_c_double, 20260101_c_int64_t), &
                   'uclbs_init_tgv')
  timestep = 0
  time = 0.0_c_double
  dt = 1.0_c_double
  if (nlev == 2) then
    lo = n / 2
    hi = n / 2 + n
    per = 0
    call uclbs_check(uclbs_define_level(ctx, 1_c_int, lo, hi, per), 'uclbs_define_level')
    call uclbs_check(uclbs_init_from_coarse(ctx, 1_c_int), 'uclbs_init_from_coarse')
    do k = 1, steps
      call amr_advance(ctx, 0, 1)
    end do
    timestep(0) = steps
    call uclbs_check(uclbs_sync(ctx), 'uclbs_sync')
  else
    ! steps coarse steps: time(0) runs 0 .. steps-1
    call run_loop(ctx, 0, dt, real(steps - 1, c_double), timestep, time)
  end if
  nn = int(n, c_int64_t)**3
  allocate(rho(nn), u(nn), v(nn), w(nn))
  call uclbs_check(uclbs_download_macro(ctx, 0_c_int, rho, u, v, w), 'uclbs_download_macro')
  call uclbs_check(uclbs_get_stats(ctx, st), 'uclbs_get_stats')
  open(newunit=unit_out, file=trim(outname), access='stream', form='unformatted', status='replace')
  write(unit_out) rho, u, v, w
  close(unit_out)
  write(*, '(a,i0,a,i0,a,i0,a,i0)') 'steps ', timestep(0), ' node_updates ', st%node_updates, &
      ' newton_iters ', st%newton_iters, ' noconv ', st%n_noconv
  call uclbs_check(uclbs_destroy(ctx), 'uclbs_destroy')
end program uclbs_driver
