!--------------------------------------------------------------------
! poongback_gpu -- iso_c_binding interface to the gfx950 time-step
! library declared in include/pb_capi.h.
!
! A Fortran driver built on SOURCE_PASSOPB (main.f90, time_march.f90)
! uses this module to replace the body of time_march (time_march.f90:
! 63-120: y_to_x x4, dft_x_realspace, x_to_y x8, rescale_data, wavespace)
! by one pb_substep call on device-resident state (pb_set_state once,
! pb_substep per substep, pb_get_state when the driver needs the fields).
! Sizes and layouts are those of the reference: U(ny,nx/2,nz,slot) and
! RK_STORE(ny,nx/2,nz,3) complex(c_double_complex), y fastest.
!--------------------------------------------------------------------
module poongback_gpu
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: PB_OK = 0, PB_E_ARG = -1, PB_E_HIP = -2, &
       PB_E_NUMERIC = -3, PB_E_STATE = -4, PB_E_NONFINITE = -5
  integer(c_int32_t), parameter :: PB_KNOTS_REFERENCE = 0, PB_KNOTS_TANH = 1
  integer(c_int32_t), parameter :: PB_SCALAR_SOURCE = 0, PB_SCALAR_MIXED = 1

  type, bind(C) :: pb_config
     integer(c_int64_t) :: nx, ny, nz
     integer(c_int32_t) :: labs_k, knot_map
     real(c_double) :: stretch
     real(c_double) :: lmfx, lmfy, lmfz
     real(c_double) :: nu, pr
     real(c_double) :: bulk_u, bulk_w, bulk_t
     real(c_double) :: u_gal
     real(c_double) :: cfl, dt_factor, dy_factor
     real(c_double) :: opp_alpha, opp_ydist
     integer(c_int32_t) :: rk_step, scalar_bc
  end type pb_config

  type, bind(C) :: pb_substep_out
     real(c_double) :: dt, velmax, phi_u, phi_t, tau_w, q_w
  end type pb_substep_out

  interface
     integer(c_int) function pb_create(cfg, ctx) bind(C, name="pb_create")
       import :: c_int, c_ptr, pb_config
       type(pb_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: ctx
     end function pb_create

     integer(c_int) function pb_destroy(ctx) bind(C, name="pb_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function pb_destroy

     type(c_ptr) function pb_last_error() bind(C, name="pb_last_error")
       import :: c_ptr
     end function pb_last_error

     integer(c_int) function pb_get_info(ctx, bw, device_bytes) bind(C, name="pb_get_info")
       import :: c_int, c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: bw
       integer(c_int64_t), intent(out) :: device_bytes
     end function pb_get_info

     ! pointers may be c_null_ptr (pb_get_labs skips them)
     integer(c_int) function pb_get_labs(ctx, y, d0, d1, d2, d0_lu, integ) bind(C, name="pb_get_labs")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, y, d0, d1, d2, d0_lu, integ
     end function pb_get_labs

     ! U(ny,nx/2,nz,4) = u, v, w, theta; rk = RK_STORE or c_null_ptr
     integer(c_int) function pb_set_state(ctx, U, rk) bind(C, name="pb_set_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, U, rk
     end function pb_set_state

     integer(c_int) function pb_get_state(ctx, U, rk) bind(C, name="pb_get_state")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, U, rk
     end function pb_get_state

     ! rk: c_loc(RK_STORE) or c_null_ptr (then pb_init_stores forms the stores)
     integer(c_int) function pb_set_state_slots(ctx, u, v, w, theta, rk) bind(C, name="pb_set_state_slots")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: u(*), v(*), w(*), theta(*)
       type(c_ptr), value :: rk
     end function pb_set_state_slots

     integer(c_int) function pb_get_state_slots(ctx, u, v, w, theta, rk) bind(C, name="pb_get_state_slots")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(out) :: u(*), v(*), w(*), theta(*)
       type(c_ptr), value :: rk
     end function pb_get_state_slots

     integer(c_int) function pb_init_stores(ctx) bind(C, name="pb_init_stores")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function pb_init_stores

     ! restart kinds v/g and g/phi (initial_step.f90:20-45); means: ny doubles
     integer(c_int) function pb_set_state_vg(ctx, v, g, theta, u_mean, w_mean, t_mean) &
          bind(C, name="pb_set_state_vg")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: v(*), g(*), theta(*)
       real(c_double), intent(in) :: u_mean(*), w_mean(*), t_mean(*)
     end function pb_set_state_vg

     integer(c_int) function pb_set_state_gphi(ctx, g, phi, theta, u_mean, w_mean, t_mean) &
          bind(C, name="pb_set_state_gphi")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: g(*), phi(*), theta(*)
       real(c_double), intent(in) :: u_mean(*), w_mean(*), t_mean(*)
     end function pb_set_state_gphi

     ! what data_store writes; every argument c_loc(...) or c_null_ptr
     integer(c_int) function pb_get_checkpoint(ctx, u, v, w, theta, g, phi, u_mean, w_mean, t_mean) &
          bind(C, name="pb_get_checkpoint")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, u, v, w, theta, g, phi, u_mean, w_mean, t_mean
     end function pb_get_checkpoint

     integer(c_int) function pb_substep(ctx, s, dt, out) bind(C, name="pb_substep")
       import :: c_int, c_int32_t, c_double, c_ptr, pb_substep_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: s
       real(c_double), intent(inout) :: dt
       type(pb_substep_out), intent(out) :: out
     end function pb_substep

     ! p: c_loc of a (ny,nx/2,nz) complex array, vv: c_loc of ny doubles; either may be c_null_ptr
     integer(c_int) function pb_substep_stats(ctx, s, dt, out, p, vv) bind(C, name="pb_substep_stats")
       import :: c_int, c_int32_t, c_double, c_ptr, pb_substep_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: s
       real(c_double), intent(inout) :: dt
       type(pb_substep_out), intent(out) :: out
       type(c_ptr), value :: p, vv
     end function pb_substep_stats

     ! pb_substep_stats + get_statistics_2 (29 lines of ny); p, vv, stat2: c_loc or c_null_ptr
     integer(c_int) function pb_substep_stats2(ctx, s, dt, out, p, vv, stat2) bind(C, name="pb_substep_stats2")
       import :: c_int, c_int32_t, c_double, c_ptr, pb_substep_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: s
       real(c_double), intent(inout) :: dt
       type(pb_substep_out), intent(out) :: out
       type(c_ptr), value :: p, vv, stat2
     end function pb_substep_stats2

     ! the three substeps of main.f90:134-156; out(rk_step)
     integer(c_int) function pb_step(ctx, dt, out) bind(C, name="pb_step")
       import :: c_int, c_double, c_ptr, pb_substep_out
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: dt
       type(pb_substep_out), intent(out) :: out(*)
     end function pb_step

     integer(c_int) function pb_timing(ctx, enable) bind(C, name="pb_timing")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: enable
     end function pb_timing

     integer(c_int) function pb_timing_read(ctx, ms, launches) bind(C, name="pb_timing_read")
       import :: c_int, c_double, c_int64_t, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms(*)
       integer(c_int64_t), intent(out) :: launches(*)
     end function pb_timing_read

     integer(c_int) function pb_statistics(ctx, out) bind(C, name="pb_statistics")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: out(*)
     end function pb_statistics

     integer(c_int) function pb_spectra(ctx, ex, ez) bind(C, name="pb_spectra")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ex(*), ez(*)
     end function pb_spectra

     ! pb_step as a captured hipGraph (captured on the first call)
     integer(c_int) function pb_step_graph(ctx, dt, out) bind(C, name="pb_step_graph")
       import :: c_int, c_double, c_ptr, pb_substep_out
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: dt
       type(pb_substep_out), intent(out) :: out(*)
     end function pb_step_graph

     ! ensembles: n channels in one captured graph per step
     integer(c_int) function pb_ensemble_create(n, cfgs, ens) bind(C, name="pb_ensemble_create")
       import :: c_int, c_int32_t, c_ptr, pb_config
       integer(c_int32_t), value :: n
       type(pb_config), intent(in) :: cfgs(*)
       type(c_ptr), intent(out) :: ens
     end function pb_ensemble_create

     integer(c_int) function pb_ensemble_channel(ens, i, ctx) bind(C, name="pb_ensemble_channel")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ens
       integer(c_int32_t), value :: i
       type(c_ptr), intent(out) :: ctx
     end function pb_ensemble_channel

     integer(c_int) function pb_ensemble_step(ens, dt, out) bind(C, name="pb_ensemble_step")
       import :: c_int, c_double, c_ptr, pb_substep_out
       type(c_ptr), value :: ens
       real(c_double), intent(inout) :: dt(*)
       type(pb_substep_out), intent(out) :: out(*)
     end function pb_ensemble_step

     integer(c_int) function pb_ensemble_tune(ens, reps, ms_out) bind(C, name="pb_ensemble_tune")
       import :: c_int, c_int32_t, c_ptr
       type(c_ptr), value :: ens, ms_out
       integer(c_int32_t), value :: reps
     end function pb_ensemble_tune

     integer(c_int) function pb_ensemble_destroy(ens) bind(C, name="pb_ensemble_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ens
     end function pb_ensemble_destroy

     integer(c_int) function pb_sync(ctx) bind(C, name="pb_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function pb_sync

     ! v_old: c_loc(U(:,10)) with opposition control, else c_null_ptr
     integer(c_int) function pb_wavespace_host(ctx, s, dt, U, rk, v_old, out) bind(C, name="pb_wavespace_host")
       import :: c_int, c_int32_t, c_double, c_ptr, c_double_complex, pb_substep_out
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: s
       real(c_double), value :: dt
       complex(c_double_complex), intent(inout) :: U(*), rk(*)
       type(c_ptr), value :: v_old
       type(pb_substep_out), intent(out) :: out
     end function pb_wavespace_host

     integer(c_int) function pb_nonlinear_host(ctx, U, P, want_vv, velmax) bind(C, name="pb_nonlinear_host")
       import :: c_int, c_int32_t, c_double, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: U(*)
       complex(c_double_complex), intent(inout) :: P(*)
       integer(c_int32_t), value :: want_vv
       real(c_double), intent(out) :: velmax
     end function pb_nonlinear_host

     ! y_to_x of u, v, w, theta: U(ny,nx/2,nz,4) -> Upad(nx/2,3nz/2,ny,4)
     integer(c_int) function pb_y_to_x_host(ctx, U, Upad) bind(C, name="pb_y_to_x_host")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: U(*)
       complex(c_double_complex), intent(out) :: Upad(*)
     end function pb_y_to_x_host

     ! dft_x_realspace: Upad_in(nx/2,3nz/2,ny,4) -> Upad_out(nx/2,3nz/2,ny,10)
     integer(c_int) function pb_realspace_host(ctx, Upad_in, Upad_out, want_vv, velmax) &
          bind(C, name="pb_realspace_host")
       import :: c_int, c_int32_t, c_double, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: Upad_in(*)
       complex(c_double_complex), intent(out) :: Upad_out(*)
       integer(c_int32_t), value :: want_vv
       real(c_double), intent(out) :: velmax
     end function pb_realspace_host

     ! x_to_y of nfield fields, times scale (1/(nx32*nz32) adds rescale_data)
     integer(c_int) function pb_x_to_y_host(ctx, nfield, Upad, U, scale) bind(C, name="pb_x_to_y_host")
       import :: c_int, c_int32_t, c_double, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nfield
       complex(c_double_complex), intent(in) :: Upad(*)
       complex(c_double_complex), intent(out) :: U(*)
       real(c_double), value :: scale
     end function pb_x_to_y_host

     ! get_pressure: state U(ny,nx/2,nz,4), products P(ny,nx/2,nz,10) with v*v in slot 6
     integer(c_int) function pb_pressure_host(ctx, U, P, p_out, vv_out) bind(C, name="pb_pressure_host")
       import :: c_int, c_ptr, c_double, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: U(*), P(*)
       complex(c_double_complex), intent(out) :: p_out(*)
       real(c_double), intent(out) :: vv_out(*)
     end function pb_pressure_host

     ! LABS_lu / LABS_bs / LABS_mat_mul (module_labs.f90:462, 499, 570) on the device
     integer(c_int) function pb_labs_lu(n, bw, nmat, A) bind(C, name="pb_labs_lu")
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: n, bw, nmat
       real(c_double), intent(inout) :: A(*)
     end function pb_labs_lu

     integer(c_int) function pb_labs_bs(n, bw, A_lu, nrhs, b) bind(C, name="pb_labs_bs")
       import :: c_int, c_int32_t, c_double, c_double_complex
       integer(c_int32_t), value :: n, bw, nrhs
       real(c_double), intent(in) :: A_lu(*)
       complex(c_double_complex), intent(inout) :: b(*)
     end function pb_labs_bs

     ! pb_labs_bs staged through the library's device scratch (8 MiB): no allocation
     integer(c_int) function pb_labs_bs_noalloc(n, bw, A_lu, nrhs, b) bind(C, name="pb_labs_bs_noalloc")
       import :: c_int, c_int32_t, c_double, c_double_complex
       integer(c_int32_t), value :: n, bw, nrhs
       real(c_double), intent(in) :: A_lu(*)
       complex(c_double_complex), intent(inout) :: b(*)
     end function pb_labs_bs_noalloc

     ! opposition gain and wall actuation (act: nx/2*nz complex, or c_null_ptr)
     integer(c_int) function pb_set_control(ctx, alpha, act) bind(C, name="pb_set_control")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: alpha
       type(c_ptr), value :: act
     end function pb_set_control

     ! discrete adjoint: the transposed wall-normal stage, nonlinear stage and substep
     integer(c_int) function pb_adj_wavespace_host(ctx, s, dt, AU, rk, vd, avd, act_bar, alpha_bar, tau_bar, q_bar) &
         bind(C, name="pb_adj_wavespace_host")
       import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: s
       real(c_double), value :: dt, tau_bar, q_bar
       complex(c_double_complex), intent(inout) :: AU(*), rk(*)
       type(c_ptr), value :: vd, avd, act_bar, alpha_bar
     end function pb_adj_wavespace_host

     integer(c_int) function pb_adj_nonlinear_host(ctx, U, aprod, astate) bind(C, name="pb_adj_nonlinear_host")
       import :: c_int, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: U(*), aprod(*)
       complex(c_double_complex), intent(out) :: astate(*)
     end function pb_adj_nonlinear_host

     integer(c_int) function pb_adj_substep(ctx, s, dt, state, astate, rk, tau_bar, q_bar, alpha_bar, act_bar) &
         bind(C, name="pb_adj_substep")
       import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: s
       real(c_double), value :: dt, tau_bar, q_bar
       complex(c_double_complex), intent(in) :: state(*)
       complex(c_double_complex), intent(inout) :: astate(*), rk(*)
       type(c_ptr), value :: alpha_bar, act_bar
     end function pb_adj_substep

     ! gradient of the window-averaged tau_w + wq q_w w.r.t. alpha and the actuation
     integer(c_int) function pb_gradient(ctx, nsteps, ckpt_every, dt, wq, J, dJ_dalpha, dJ_da) &
         bind(C, name="pb_gradient")
       import :: c_int, c_int32_t, c_double, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps, ckpt_every
       real(c_double), value :: dt, wq
       real(c_double), intent(out) :: J, dJ_dalpha
       complex(c_double_complex), intent(out) :: dJ_da(*)
     end function pb_gradient

     integer(c_int) function pb_labs_mat_mul(n, bw, A, nrhs, x, b) bind(C, name="pb_labs_mat_mul")
       import :: c_int, c_int32_t, c_double, c_double_complex
       integer(c_int32_t), value :: n, bw, nrhs
       real(c_double), intent(in) :: A(*)
       complex(c_double_complex), intent(in) :: x(*)
       complex(c_double_complex), intent(out) :: b(*)
     end function pb_labs_mat_mul
  end interface
end module poongback_gpu
