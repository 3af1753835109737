! Drop-in replacement of the per-level phase calls of the reference time loop.
!
! looping/timestepping/loop.f95:94-147 runs, per coarse step,
!   valid_levels -> streaming -> calculate_flow_variables -> interlevel_interpolation
!   -> fine_to_coarse -> collision -> damp_it -> lb_output -> steps_counter,
! each phase looping over the levels selected by valid_levels.  Here the
! level scheduling (valid_levels / steps_counter) is kept in Fortran with the
! reference's rules, and the fused device step replaces streaming +
! calculate_flow_variables + collision for a level: one call per level step.
module uclbs_timestep
  use iso_c_binding
  use uclbs_hip
  implicit none
  private
  public :: valid_levels, steps_counter, level_advance, run_loop, amr_advance

contains

  ! component_subroutines/valid_levels.f95:21-59: the finest level always
  ! steps; a coarser level steps when the next finer one has just completed an
  ! odd step and stepped itself.
  subroutine valid_levels(level_step, fine, timestep, coarsest)
    integer, intent(in) :: fine
    integer(c_int64_t), intent(in) :: timestep(0:)
    logical, intent(out) :: level_step(0:)
    integer, intent(out) :: coarsest
    integer :: lvl
    level_step = .false.
    coarsest = fine
    do lvl = fine, 0, -1
      if (lvl == fine) then
        level_step(lvl) = .true.
      else if (modulo(timestep(lvl + 1), 2_c_int64_t) == 1 .and. level_step(lvl + 1)) then
        level_step(lvl) = .true.
        coarsest = lvl
      end if
    end do
  end subroutine

  ! component_subroutines/steps_counter.f95:18-31
  subroutine steps_counter(level_step, fine, timestep, time, time_prev, dt)
    logical, intent(in) :: level_step(0:)
    integer, intent(in) :: fine
    integer(c_int64_t), intent(inout) :: timestep(0:)
    real(c_double), intent(inout) :: time(0:), time_prev(0:)
    real(c_double), intent(in) :: dt(0:)
    integer :: lvl
    do lvl = 0, ubound(timestep, 1)
      if (level_step(lvl)) then
        timestep(lvl) = timestep(lvl) + 1
        time_prev(lvl) = time(lvl)
        time(lvl) = time(lvl) + dt(lvl)
      else if (lvl > fine) then
        timestep(lvl) = timestep(lvl) + 2**(lvl - fine)
        time_prev(lvl) = time(lvl)
        time(lvl) = time(lvl) + dt(lvl) * 2**(lvl - fine)
      end if
    end do
  end subroutine

  ! streaming(level_step,fine,lvl) + calculate_flow_variables(...) +
  ! collision(...) of loop.f95:105,112,133 for one level, fused on the device
  subroutine level_advance(ctx, level_step, lvl)
    type(c_ptr), intent(in) :: ctx
    logical, intent(in) :: level_step(0:)
    integer, intent(in) :: lvl
    if (.not. level_step(lvl)) return
    call uclbs_check(uclbs_step_level(ctx, int(lvl, c_int), 1_c_int), 'uclbs_step_level')
  end subroutine

  ! The coarse-step loop of loop.f95:94-147 for levels 0..fine until
  ! time(0) > total_time.
  subroutine run_loop(ctx, fine, dt, total_time, timestep, time)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: fine
    real(c_double), intent(in) :: dt(0:), total_time
    integer(c_int64_t), intent(inout) :: timestep(0:)
    real(c_double), intent(inout) :: time(0:)
    logical :: level_step(0:ubound(dt, 1))
    real(c_double) :: time_prev(0:ubound(dt, 1))
    integer :: coarsest, lvl
    time_prev = time
    do while (time(0) <= total_time)
      call valid_levels(level_step, fine, timestep, coarsest)
      do lvl = fine, coarsest, -1
        call level_advance(ctx, level_step, lvl)
      end do
      call steps_counter(level_step, fine, timestep, time, time_prev, dt)
    end do
    call uclbs_check(uclbs_sync(ctx), 'uclbs_sync')
  end subroutine

  ! One step of level lvl with the levels above it subcycled at ratio 2: the
  ! finer level's ghost shell is filled from this level's previous state
  ! (theta = 0) and then half way (theta = 1/2) before its two substeps, and
  ! the finer level is injected back once it has caught up
  ! (interlevel_interpolation.f95 / fine_to_coarse.f95 between the level
  ! steps of loop.f95:104-139).
  recursive subroutine amr_advance(ctx, lvl, finest)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: lvl, finest
    integer :: k
    call uclbs_check(uclbs_step_level(ctx, int(lvl, c_int), 1_c_int), 'uclbs_step_level')
    if (lvl < finest) then
      do k = 0, 1
        call uclbs_check(uclbs_interlevel(ctx, int(lvl + 1, c_int), 0.5_c_double * k), &
                         'uclbs_interlevel')
        call amr_advance(ctx, lvl + 1, finest)
      end do
      call uclbs_check(uclbs_fine_to_coarse(ctx, int(lvl + 1, c_int)), 'uclbs_fine_to_coarse')
    end if
  end subroutine

end module uclbs_timestep
