! Host program: the compressible D3Q39 two-population model advanced through
! the Fortran time loop and the C ABI -- a box at Mach 1.5 (shifted lattice)
! with the supersonic face set (freestream inflow x-, pure outflow x+,
! inviscid y and z sides) around a perturbed freestream.  Writes the state
! after initialisation and after the steps (f, g, multipliers, FP64, the
! library's host layout) to the file named on the command line, then the step
! statistics on stdout.  With an output directory it also writes the final
! state as Plot3D files (domain placed at [0, n dx], dx = 1e-3) and one
! microphone's sample, as the reference's lb_output would (lb_output.f95:44,71).
!   uclbs_driver39 <out.bin> [nx ny nz] [steps] [precision 0|1] [output directory]
program uclbs_driver39
  use iso_c_binding
  use uclbs_hip
  use uclbs_timestep
  implicit none
  type(uclbs_config) :: cfg
  type(uclbs_gas) :: gas
  type(uclbs_stats) :: st
  type(uclbs_plot3d_header) :: hdr
  real(c_double) :: mloc(3, 1)
  integer(c_int32_t) :: mtype(1), mlvl(1), mvalid(1)
  character(len=256) :: outdir
  type(c_ptr) :: ctx
  integer(c_int32_t) :: lo(3), hi(3), per(3), faces(6)
  integer(c_int64_t) :: timestep(0:0), nn
  real(c_double) :: time(0:0), dt(0:0), pi, t0, x, y, z
  real(c_double), allocatable :: rho(:), u(:), v(:), w(:), temp(:), f(:), g(:), lm(:)
  character(len=256) :: outname, arg
  integer :: n(3), steps, prec, unit_out, a, b, c, k, d

  call get_command_argument(1, outname)
  n = (/24, 16, 12/)
  steps = 20
  prec = 1
  if (command_argument_count() >= 4) then
    do d = 1, 3
      call get_command_argument(1 + d, arg)
      read(arg, *) n(d)
    end do
  end if
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    read(arg, *) steps
  end if
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg)
    read(arg, *) prec
  end if
  pi = 4.0_c_double * atan(1.0_c_double)
  cfg%lattice = UCLBS_LATTICE_D3Q39
  cfg%precision = prec
  call uclbs_check(uclbs_create(cfg, ctx), 'uclbs_create')
  ! basic_fluids_calcs.f95 freestream: Mach 1.5, 300 K, 1 atm; the shifted
  ! lattice follows from the Mach number (directions_and_weights.f95:477)
  gas%mach = 1.5_c_double
  call uclbs_check(uclbs_set_gas(ctx, gas), 'uclbs_set_gas')
  lo = 0
  hi = n - 1
  per = 0
  call uclbs_check(uclbs_define_level(ctx, 0_c_int, lo, hi, per), 'uclbs_define_level')
  faces = (/-1000, UCLBS_FACE_INVISCID, UCLBS_FACE_INVISCID, UCLBS_FACE_OUTFLOW, &
            UCLBS_FACE_INVISCID, UCLBS_FACE_INVISCID/)
  call uclbs_check(uclbs_set_faces(ctx, 0_c_int, faces), 'uclbs_set_faces')
  nn = int(n(1), c_int64_t) * n(2) * n(3)
  allocate(rho(nn), u(nn), v(nn), w(nn), temp(nn))
  allocate(f(39 * nn), g(39 * nn), lm(13 * nn))
  t0 = 1.0_c_double / 1.4_c_double
  do c = 0, n(3) - 1
    do b = 0, n(2) - 1
      do a = 0, n(1) - 1
        k = 1 + a + n(1) * (b + n(2) * c)
        x = real(a, c_double) / n(1)
        y = real(b, c_double) / n(2)
        z = real(c, c_double) / n(3)
        rho(k) = 1.0_c_double + 0.01_c_double * sin(2 * pi * (x + y))
        u(k) = 1.5_c_double + 0.02_c_double * sin(2 * pi * x)
        v(k) = 0.01_c_double * sin(2 * pi * y)
        w(k) = 0.01_c_double * cos(2 * pi * (x + z))
        temp(k) = t0 * (1.0_c_double + 0.01_c_double * cos(2 * pi * z))
      end do
    end do
  end do
  call uclbs_check(uclbs_init_fields(ctx, 0_c_int, rho, u, v, w, temp), 'uclbs_init_fields')
  call uclbs_check(uclbs_download_state(ctx, 0_c_int, f, g, lm, nn), 'uclbs_download_state')
  open(newunit=unit_out, file=trim(outname), access='stream', form='unformatted', status='replace')
  write(unit_out) f, g, lm
  timestep = 0
  time = 0.0_c_double
  dt = 1.0_c_double
  call run_loop(ctx, 0, dt, real(steps - 1, c_double), timestep, time)
  call uclbs_check(uclbs_download_state(ctx, 0_c_int, f, g, lm, nn), 'uclbs_download_state')
  write(unit_out) f, g, lm
  close(unit_out)
  if (command_argument_count() >= 7) then
    call get_command_argument(7, outdir)
    call uclbs_check(uclbs_set_placement(ctx, (/0.0_c_double, 0.0_c_double, 0.0_c_double/), &
                     real(n, c_double) * 1.0e-3_c_double), 'uclbs_set_placement')
    hdr%mach = 1.5_c_double
    hdr%gamma = 1.4_c_double
    hdr%time = time(0)
    hdr%timestep = int(timestep(0), c_int32_t)
    call uclbs_check(uclbs_plot3d_write(ctx, trim(outdir)//c_null_char, hdr), 'uclbs_plot3d_write')
    mloc(:, 1) = (/5.5e-3_c_double, 4.25e-3_c_double, 3.1e-3_c_double/)
    mtype = 63
    mlvl = 0
    call uclbs_check(uclbs_mic_setup(ctx, trim(outdir)//c_null_char, 1_c_int, mloc, mtype, mlvl, &
                     mvalid), 'uclbs_mic_setup')
    call uclbs_check(uclbs_mic_sample(ctx), 'uclbs_mic_sample')
    call uclbs_check(uclbs_mic_flush(ctx), 'uclbs_mic_flush')
    call uclbs_check(uclbs_output_wait(ctx), 'uclbs_output_wait')
  end if
  call uclbs_check(uclbs_get_stats(ctx, st), 'uclbs_get_stats')
  write(*, '(a,i0,a,i0,a,i0,a,i0)') 'steps ', timestep(0), ' node_updates ', st%node_updates, &
      ' newton_iters ', st%newton_iters, ' noconv ', st%n_noconv
  call uclbs_check(uclbs_destroy(ctx), 'uclbs_destroy')
end program uclbs_driver39
