! Fortran binding of libuclbs_hip.so (include/uclbs_hip.h) through iso_c_binding.
!
! This is the host side of the drop-in: the reference's time loop
! (looping/timestepping/loop.f95:104-139) calls six per-level phase
! subroutines that sweep AMReX MultiFabs; with this module the level's state
! lives in HBM inside the library and one call advances it.  Every interface
! below is the C entry point of the same name; a nonzero return code is turned
! into the reference's error_out + STOP behaviour by uclbs_check
! (general/background_calcs/error_out.f95:11-14).
module uclbs_hip
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: UCLBS_LATTICE_D3Q27 = 27, UCLBS_LATTICE_D3Q39 = 39
  integer(c_int32_t), parameter, public :: UCLBS_PREC_FP32_SHIFTED = 0, UCLBS_PREC_FP64 = 1
  integer(c_int32_t), parameter, public :: UCLBS_ALPHA_NEWTON = 0, UCLBS_ALPHA_FIXED2 = 1, &
                                           UCLBS_ALPHA_LEARNED = 2
  integer(c_int32_t), parameter, public :: UCLBS_FACE_PERIODIC = 2147483647
  integer(c_int32_t), parameter, public :: UCLBS_FACE_OUTFLOW = -6, UCLBS_FACE_INVISCID = -11

  type, bind(c), public :: uclbs_config
    integer(c_int32_t) :: lattice = UCLBS_LATTICE_D3Q27
    integer(c_int32_t) :: precision = UCLBS_PREC_FP32_SHIFTED
    integer(c_int32_t) :: alpha_mode = UCLBS_ALPHA_NEWTON
    integer(c_int32_t) :: alpha_maxit = 0
    real(c_double) :: alpha_tol = 0.0_c_double
    real(c_double) :: nu = 1.0e-3_c_double
    integer(c_int32_t) :: reserved(8) = 0
  end type

  ! D3Q39 gas and freestream (include/uclbs_hip.h uclbs_gas)
  type, bind(c), public :: uclbs_gas
    real(c_double) :: gama = 1.4_c_double, temperature = 300.0_c_double, &
                      pressure = 101325.0_c_double, mach = 0.3_c_double, dx = 1.0e-3_c_double
    integer(c_int32_t) :: startup_timesteps = 25, turbulent = 1, shifted = -1, lm_maxit = 0
    integer(c_int32_t) :: reserved(4) = 0
  end type

  type, bind(c), public :: uclbs_stats
    integer(c_int64_t) :: node_updates, newton_iters, n_trivial, n_bound, n_nonpos, &
                          n_noconv, n_fallback, n_logpath, n_ibb_links, n_repair
  end type

  ! Plot3D header values (include/uclbs_hip.h uclbs_plot3d_header)
  type, bind(c), public :: uclbs_plot3d_header
    real(c_double) :: mach = 0, reynolds = 0, time = 0, gamma = 0, t_ref = 0, &
                      stag_temperature = 0, cv = 0, cp = 0, gas_r = 0
    integer(c_int32_t) :: timestep = 0, max_level = 31
    integer(c_int64_t) :: max_subrecord = 0
  end type

  public :: uclbs_set_placement, uclbs_plot3d_write, uclbs_output_wait, uclbs_mic_setup, &
            uclbs_mic_sample, uclbs_mic_flush, uclbs_regrid_level, uclbs_clear_level, &
            uclbs_tag_ggm, uclbs_cluster_tags
  public :: uclbs_create, uclbs_destroy, uclbs_define_level, uclbs_upload_f, &
            uclbs_download_f, uclbs_download_macro, uclbs_set_keep_alpha, &
            uclbs_download_alpha, uclbs_init_tgv, uclbs_init_uniform, uclbs_set_faces, &
            uclbs_set_bc_states, uclbs_build_links, uclbs_set_links, uclbs_download_links, &
            uclbs_step_level, uclbs_sync, uclbs_get_stats, uclbs_reset_stats, &
            uclbs_geom_icosphere, uclbs_check, uclbs_set_gas, uclbs_init_fields, &
            uclbs_upload_state, uclbs_download_state, uclbs_set_timestep, uclbs_interlevel, &
            uclbs_init_from_coarse, uclbs_fine_to_coarse, uclbs_mlp_init, uclbs_mlp_set_params, &
            uclbs_mlp_get_params, uclbs_mlp_reserve, uclbs_mlp_collect, uclbs_mlp_train

  interface
    function c_last_error() bind(c, name='uclbs_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function uclbs_create(cfg, ctx) bind(c, name='uclbs_create') result(rc)
      import :: uclbs_config, c_ptr, c_int
      type(uclbs_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function
    function uclbs_destroy(ctx) bind(c, name='uclbs_destroy') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function uclbs_define_level(ctx, lvl, lo, hi, periodic) bind(c, name='uclbs_define_level') result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int32_t), intent(in) :: lo(3), hi(3), periodic(3)
      integer(c_int) :: rc
    end function
    function uclbs_upload_f(ctx, lvl, f, n) bind(c, name='uclbs_upload_f') result(rc)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(in) :: f(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function uclbs_download_f(ctx, lvl, f, n) bind(c, name='uclbs_download_f') result(rc)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(out) :: f(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function uclbs_download_macro(ctx, lvl, rho, u, v, w) bind(c, name='uclbs_download_macro') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(out) :: rho(*), u(*), v(*), w(*)
      integer(c_int) :: rc
    end function
    function uclbs_set_keep_alpha(ctx, keep) bind(c, name='uclbs_set_keep_alpha') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: keep
      integer(c_int) :: rc
    end function
    function uclbs_download_alpha(ctx, lvl, alpha) bind(c, name='uclbs_download_alpha') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(out) :: alpha(*)
      integer(c_int) :: rc
    end function
    function uclbs_init_tgv(ctx, lvl, u0, amp, seed) bind(c, name='uclbs_init_tgv') result(rc)
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), value :: u0, amp
      integer(c_int64_t), value :: seed
      integer(c_int) :: rc
    end function
    function uclbs_init_uniform(ctx, lvl, rho, u, v, w) bind(c, name='uclbs_init_uniform') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), value :: rho, u, v, w
      integer(c_int) :: rc
    end function
    function uclbs_set_faces(ctx, lvl, codes) bind(c, name='uclbs_set_faces') result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int32_t), intent(in) :: codes(6)
      integer(c_int) :: rc
    end function
    function uclbs_set_bc_states(ctx, n_in, in_states, n_out, out_states, fs_state) &
        bind(c, name='uclbs_set_bc_states') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_in, n_out
      real(c_double), intent(in) :: in_states(*), out_states(*)
      type(c_ptr), value :: fs_state
      integer(c_int) :: rc
    end function
    function uclbs_build_links(ctx, lvl, origin, dx, ntri, tri, nrm, n_solid, n_tie) &
        bind(c, name='uclbs_build_links') result(rc)
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(in) :: origin(3)
      real(c_double), value :: dx
      integer(c_int64_t), value :: ntri
      real(c_double), intent(in) :: tri(*), nrm(*)
      integer(c_int64_t), intent(out) :: n_solid, n_tie
      integer(c_int) :: rc
    end function
    function uclbs_set_links(ctx, lvl, state, n) bind(c, name='uclbs_set_links') result(rc)
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int32_t), intent(in) :: state(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function uclbs_download_links(ctx, lvl, state, n) bind(c, name='uclbs_download_links') result(rc)
      import :: c_ptr, c_int, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int32_t), intent(out) :: state(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function uclbs_step_level(ctx, lvl, nsteps) bind(c, name='uclbs_step_level') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl, nsteps
      integer(c_int) :: rc
    end function
    function uclbs_sync(ctx) bind(c, name='uclbs_sync') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function uclbs_get_stats(ctx, st) bind(c, name='uclbs_get_stats') result(rc)
      import :: c_ptr, c_int, uclbs_stats
      type(c_ptr), value :: ctx
      type(uclbs_stats), intent(out) :: st
      integer(c_int) :: rc
    end function
    function uclbs_reset_stats(ctx) bind(c, name='uclbs_reset_stats') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function uclbs_geom_icosphere(level, radius, center, seed, tri, nrm, cap, ntri) &
        bind(c, name='uclbs_geom_icosphere') result(rc)
      import :: c_int, c_double, c_int64_t
      integer(c_int), value :: level
      real(c_double), value :: radius
      real(c_double), intent(in) :: center(3)
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: tri(*), nrm(*)
      integer(c_int64_t), value :: cap
      integer(c_int64_t), intent(out) :: ntri
      integer(c_int) :: rc
    end function
    function uclbs_set_gas(ctx, gas) bind(c, name='uclbs_set_gas') result(rc)
      import :: c_ptr, c_int, uclbs_gas
      type(c_ptr), value :: ctx
      type(uclbs_gas), intent(in) :: gas
      integer(c_int) :: rc
    end function
    function uclbs_init_fields(ctx, lvl, rho, u, v, w, temp) bind(c, name='uclbs_init_fields') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(in) :: rho(*), u(*), v(*), w(*), temp(*)
      integer(c_int) :: rc
    end function
    function uclbs_upload_state(ctx, lvl, f, g, lm, n) bind(c, name='uclbs_upload_state') result(rc)
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(in) :: f(*), g(*), lm(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function uclbs_download_state(ctx, lvl, f, g, lm, n) bind(c, name='uclbs_download_state') result(rc)
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), intent(out) :: f(*), g(*), lm(*)
      integer(c_int64_t), value :: n
      integer(c_int) :: rc
    end function
    function uclbs_set_timestep(ctx, lvl, step) bind(c, name='uclbs_set_timestep') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int64_t), value :: step
      integer(c_int) :: rc
    end function
    function uclbs_interlevel(ctx, lvl, theta) bind(c, name='uclbs_interlevel') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      real(c_double), value :: theta
      integer(c_int) :: rc
    end function
    function uclbs_init_from_coarse(ctx, lvl) bind(c, name='uclbs_init_from_coarse') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int) :: rc
    end function
    function uclbs_fine_to_coarse(ctx, lvl) bind(c, name='uclbs_fine_to_coarse') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int) :: rc
    end function
    function uclbs_mlp_init(ctx, seed) bind(c, name='uclbs_mlp_init') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: seed
      integer(c_int) :: rc
    end function
    function uclbs_mlp_set_params(ctx, p, ymean, ystd) bind(c, name='uclbs_mlp_set_params') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: p(*)
      real(c_float), value :: ymean, ystd
      integer(c_int) :: rc
    end function
    function uclbs_mlp_get_params(ctx, p, ymean, ystd) bind(c, name='uclbs_mlp_get_params') result(rc)
      import :: c_ptr, c_int, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(out) :: p(*), ymean, ystd
      integer(c_int) :: rc
    end function
    function uclbs_mlp_reserve(ctx, cap) bind(c, name='uclbs_mlp_reserve') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: cap
      integer(c_int) :: rc
    end function
    function uclbs_mlp_collect(ctx, lvl, stride, n) bind(c, name='uclbs_mlp_collect') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl, stride
      integer(c_int64_t), intent(out) :: n
      integer(c_int) :: rc
    end function
    function uclbs_mlp_train(ctx, steps, batch, lr, seed, l0, l1) bind(c, name='uclbs_mlp_train') result(rc)
      import :: c_ptr, c_int, c_int64_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: steps, seed
      integer(c_int), value :: batch
      real(c_double), value :: lr
      real(c_double), intent(out) :: l0, l1
      integer(c_int) :: rc
    end function
    function c_strlen(p) bind(c, name='strlen') result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: p
      integer(c_size_t) :: n
    end function
    ! output: the plot3d_output / mic_setup / mic_write calls of lb_output.f95:44,71
    function uclbs_set_placement(ctx, problo, probhi) bind(c, name='uclbs_set_placement') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: problo(3), probhi(3)
      integer(c_int) :: rc
    end function
    function uclbs_plot3d_write(ctx, dir, h) bind(c, name='uclbs_plot3d_write') result(rc)
      import :: c_ptr, c_int, c_char, uclbs_plot3d_header
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: dir(*)
      type(uclbs_plot3d_header), intent(in) :: h
      integer(c_int) :: rc
    end function
    function uclbs_output_wait(ctx) bind(c, name='uclbs_output_wait') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function uclbs_mic_setup(ctx, dir, n, loc, mtype, lvl, valid) &
        bind(c, name='uclbs_mic_setup') result(rc)
      import :: c_ptr, c_int, c_char, c_double, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int), value :: n
      real(c_double), intent(in) :: loc(3, *)
      integer(c_int32_t), intent(in) :: mtype(*), lvl(*)
      integer(c_int32_t), intent(out) :: valid(*)
      integer(c_int) :: rc
    end function
    function uclbs_mic_sample(ctx) bind(c, name='uclbs_mic_sample') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function uclbs_mic_flush(ctx) bind(c, name='uclbs_mic_flush') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    ! remake (amr_processes.f95:670-1080): lo, hi (3, nbox) level indices
    function uclbs_regrid_level(ctx, lvl, nbox, lo, hi) bind(c, name='uclbs_regrid_level') result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl, nbox
      integer(c_int32_t), intent(in) :: lo(*), hi(*)
      integer(c_int) :: rc
    end function
    ! clear_lvl (amr_processes.f95:478-545)
    function uclbs_clear_level(ctx, lvl) bind(c, name='uclbs_clear_level') result(rc)
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl
      integer(c_int) :: rc
    end function
    ! ggm_calculator (ggm_calculator.f95:47-68); ggm, tags over the stored nodes
    function uclbs_tag_ggm(ctx, lvl, quantity, scheme, limit, ggm, tags, ntagged) &
        bind(c, name='uclbs_tag_ggm') result(rc)
      import :: c_ptr, c_int, c_double, c_int8_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: lvl, quantity, scheme
      real(c_double), value :: limit
      real(c_double), intent(out) :: ggm(*)
      integer(c_int8_t), intent(out) :: tags(*)
      integer(c_int64_t), intent(out) :: ntagged
      integer(c_int) :: rc
    end function
    ! boxes over tags(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3)) (host only)
    function uclbs_cluster_tags(tags, lo, hi, efficiency, blocking, buffer, max_size, max_boxes, &
        box_lo, box_hi, nbox) bind(c, name='uclbs_cluster_tags') result(rc)
      import :: c_int, c_double, c_int8_t, c_int32_t
      integer(c_int8_t), intent(in) :: tags(*)
      integer(c_int32_t), intent(in) :: lo(3), hi(3)
      real(c_double), value :: efficiency
      integer(c_int), value :: blocking, buffer, max_size, max_boxes
      integer(c_int32_t), intent(out) :: box_lo(*), box_hi(*)
      integer(c_int), intent(out) :: nbox
      integer(c_int) :: rc
    end function
  end interface

contains

  ! error_out + STOP semantics of the reference for a failed library call
  subroutine uclbs_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, i
    character(len=512) :: text
    if (rc == 0) return
    p = c_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    text = ''
    do i = 1, min(n, 512)
      text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'uclbs: ', where, ' failed (', rc, '): ', trim(text)
    stop 1
  end subroutine

end module uclbs_hip
