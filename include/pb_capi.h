/*
 * pb_capi.h -- C ABI of the MI355X (gfx950) RK3-CN channel time step.
 *
 * The library replaces the per-substep call chain of the PoongBack channel
 * DNS (reference: SOURCE_PASSOPB/time_march.f90:36-128) on one GPU.  Every
 * entry point takes plain pointers and sizes, returns 0 on success and a
 * negative PB_E* code on failure (pb_last_error() then describes it); no
 * entry point aborts the process.
 *
 * Spectral layout at the boundary (identical to the reference y-pencil):
 *   U(ny, nx/2, nz, nslot)   Fortran order, complex(c_double_complex)
 *   i.e. element (y, kx, kz, slot) at ((slot*nz + kz)*(nx/2) + kx)*ny + y,
 *   interleaved (re, im) doubles.  kx = 0 .. nx/2-1; kz index j maps to
 *   0 .. nz/2, -(nz/2-1) .. -1 (module_ppdfftw.f90:360-374).  Values are
 *   collocation-point values, not B-spline coefficients (hvhg.f90:68-69).
 *   RK_STORE(ny, nx/2, nz, 3): slot 0 = g (mean U at kx=kz=0), slot 1 = phi
 *   (mean W), slot 2 = theta store (mean Theta)  (module_ppdfftw.f90:304).
 *
 * Inside the library the state is device resident and laid out
 * (slot, y, kz, kx) with kx fastest; conversions happen only in
 * pb_set_state / pb_get_state and the host-array drop-ins below.
 *
 * One context must not be used from two threads at once; all device work
 * of a context runs on one HIP stream.
 */
#ifndef PB_CAPI_H
#define PB_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define PB_OK            0
#define PB_E_ARG        -1   /* invalid argument / size */
#define PB_E_HIP        -2   /* HIP runtime error (no device, OOM, launch) */
#define PB_E_NUMERIC    -3   /* dt out of range (dft_x_realspace.f90:494) */
#define PB_E_STATE      -4   /* call out of order (e.g. substep before set_state) */
#define PB_E_NONFINITE  -5   /* the state holds a NaN or Inf (a CFL reduction or a mean-line
                                output of the substep was not finite); the reference has no
                                check and would only trip dt_factor's range test one step later */

/* knot maps for the B-spline breakpoints (labs_fill_full.cpp:80-101) */
#define PB_KNOTS_REFERENCE 0 /* uniform if |stretch| <= 1e-16, else sin map (line 92) */
#define PB_KNOTS_TANH      1 /* 0.5*Ly*tanh(s*xi)/tanh(s)  (line 89, commented out in the reference) */

/* Scalar forcing and wall condition.
 * PB_SCALAR_SOURCE: the reference's: theta = 0 at both walls and a
 *   y-uniform source whose strength phi_t is chosen every substep to hold the
 *   bulk scalar at bulk_t (wave00.f90:473-667, calculate_phi 704-714).
 * PB_SCALAR_MIXED: uniform wall heat flux with the mixed-type condition
 *   (Kasagi, Tomita & Kuroda 1992): the wall temperature rises linearly in x,
 *   the scalar theta = T - T_w obeys the transport equation with the source
 *   (nu/Pr) u / bulk_u on every mode (u the streamwise velocity of the
 *   substep's input state, treated explicitly like the convective terms)
 *   and theta = 0 at both walls; the wall flux per wall is nu/Pr in steady
 *   state, the bulk scalar is free (phi_t = 0).  Laminar Poiseuille flow has
 *   theta = (5 - 6 y^2 + y^4)/8 and a mixed-mean Nusselt number 70/17
 *   (based on 2h).  Not in the reference (its scalar has the SOURCE form). */
#define PB_SCALAR_SOURCE 0
#define PB_SCALAR_MIXED  1

/* Physical and numerical parameters.  Mirrors the `input` file records that
 * reach the time step (read_input_file.f90:39-116) plus the opposition
 * control the reference leaves commented out (solve_v_f.f90:253-257). */
typedef struct pb_config {
  int64_t nx, ny, nz;        /* grid; nx, nz even; ny = number of B-splines */
  int32_t labs_k;            /* LABS_K; GSL order passed to the generator is labs_k+1 (module_labs.f90:103) */
  int32_t knot_map;          /* PB_KNOTS_* */
  double  stretch;           /* knot stretching factor */
  double  lmfx, lmfy, lmfz;  /* Lx = 2*pi*lmfx, Ly = 2*lmfy, Lz = 2*pi*lmfz (read_input_file.f90:100-102) */
  double  nu, pr;            /* viscosity, Prandtl number */
  double  bulk_u, bulk_w, bulk_t; /* held bulk values (wave00.f90:202, 395, 612) */
  double  u_gal;             /* Galilean shift (wave00.f90:171, 227) */
  double  cfl, dt_factor, dy_factor; /* dt = dt_factor*dt + (1-dt_factor)*CFL/pi/velmax (dft_x_realspace.f90:492) */
  double  opp_alpha;         /* opposition-control gain alpha; 0 = reference behaviour (wall v = 0) */
  double  opp_ydist;         /* detection plane distance from each wall in outer units (y+_d / Re_tau) */
  int32_t rk_step;           /* number of RK substeps per step; 3 (read_input_file.f90:121) */
  int32_t scalar_bc;         /* PB_SCALAR_* */
} pb_config;

typedef struct pb_ctx pb_ctx;

/* Per-substep scalar outputs that the reference returns through module
 * globals (wave00.f90:208-211, 619-622; dft_x_realspace.f90:491-492). */
typedef struct pb_substep_out {
  double dt;        /* time step after the CFL update (substep 1 only changes it) */
  double velmax;    /* max(|u|dxr + |v|dyr + |w|dzr); computed on substep 1 */
  double phi_u;     /* mean pressure-gradient forcing coefficient, phi_dpdx(s) */
  double phi_t;     /* scalar source coefficient, phi_dtdx(s) */
  double tau_w;     /* wall_shear_stress (set when s == rk_step) */
  double q_w;       /* wall_theta_stress (set when s == rk_step) */
} pb_substep_out;

/* ---- lifetime ---------------------------------------------------------- */

/* Build the B-spline operators (LABS_initialize, module_labs.f90:77-238),
 * the RK3 coefficients (RK3_initialize, module_runge_kutta.f90:30-51), the
 * wavenumber tables and dyr (module_ppdfftw.f90:347-384, 572-583) and
 * allocate the device state on the current HIP device. */
int pb_create(const pb_config *cfg, pb_ctx **out);
int pb_destroy(pb_ctx *ctx);
const char *pb_last_error(void);

/* Grid facts computed at create time: bw = LABS_BW (module_labs.f90:245-293). */
int pb_get_info(const pb_ctx *ctx, int32_t *bw, int64_t *device_bytes);

/* Host copies of the collocation operators: y (ny), D0/D1/D2 bands
 * (2bw+1, ny) in the reference's near-boundary packed format
 * (LABS_full_to_band, module_labs.f90:301-328), D0_LU and the integration
 * weights.  Any pointer may be NULL. */
int pb_get_labs(const pb_ctx *ctx, double *y, double *d0, double *d1, double *d2,
                double *d0_lu, double *integ);

/* ---- state exchange (off the timed path) ------------------------------- */

/* U: (ny, nx/2, nz, 4) slots u, v, w, theta; rk: (ny, nx/2, nz, 3). */
int pb_set_state(pb_ctx *ctx, const double *U, const double *rk);
int pb_get_state(pb_ctx *ctx, double *U, double *rk);
/* The same with one pointer per field, as the Fortran driver holds them:
 * u = U(:,1), v = U(:,2), w = U(:,3), theta = U(:,7), each (ny, nx/2, nz)
 * complex, and rk = RK_STORE(ny, nx/2, nz, 3) (or NULL). */
int pb_set_state_slots(pb_ctx *ctx, const double *u, const double *v, const double *w, const double *theta,
                       const double *rk);
int pb_get_state_slots(pb_ctx *ctx, double *u, double *v, double *w, double *theta, double *rk);
/* Fill the RK stores from u, v, w, theta already set by pb_set_state with
 * rk == NULL (set_initial_data_uvw, initial_step.f90:139-179). */
int pb_init_stores(pb_ctx *ctx);
/* The reference's other two restart kinds (initial_step.f90:20-45): fields
 * (ny, nx/2, nz) complex as its restart file holds them, the mean lines as ny
 * doubles with u_mean in the stored frame (u_gal is subtracted, as data_read_vg
 * / data_read_gphi do).  From v, g, theta: u and w from continuity and g,
 * the phi store = (D2 LU(D0)^-1 - k^2) v (set_initial_data_vg, 301-353).
 * From g, phi, theta: v from [D2 - k^2 D0] c = phi with zero wall values,
 * then u, w (set_initial_data_gphi, 430-513).  Both leave the stores set. */
int pb_set_state_vg(pb_ctx *ctx, const double *v, const double *g, const double *theta, const double *u_mean,
                    const double *w_mean, const double *t_mean);
int pb_set_state_gphi(pb_ctx *ctx, const double *g, const double *phi, const double *theta,
                      const double *u_mean, const double *w_mean, const double *t_mean);
/* What data_store writes (data_store.f90:92-118): u, v, w, theta, g and phi
 * (the stores: after a full step they hold g and phi of the state) as
 * (ny, nx/2, nz) complex, and the mean lines U (+ u_gal), W, Theta as ny
 * doubles.  Any pointer may be NULL.  The HDF5 writing stays with the caller
 * (layout 0 of vel_step_*.h5 is this array order). */
int pb_get_checkpoint(pb_ctx *ctx, double *u, double *v, double *w, double *theta, double *g, double *phi,
                      double *u_mean, double *w_mean, double *t_mean);

/* ---- the time step (device resident; no host<->device copies) ---------- */

/* One RK3-CN substep s (1..rk_step) with time step *dt_inout:
 * y_to_x x4, dft_x_realspace, x_to_y x8, rescale_data and wavespace
 * (time_march.f90:63-120) in one call.  On substep 1 dt is updated from the
 * CFL condition when dt_factor != 1. */
int pb_substep(pb_ctx *ctx, int32_t s, double *dt_inout, pb_substep_out *out);
/* pb_substep with the statistics-time work of substep 1 (time_march.f90:
 * 47-110 with stat_store_or_not set): <vv>(y) and the pressure wall values
 * nu*D2(wall row).LU(D0)^-1 v from the substep's input state (get_uv,
 * get_statistics.f90:982-1017; get_y_dir_deriv_of_field_with_PPE_BC,
 * 689-716), the v*v product (x_to_y of field 6), and the pressure of
 * get_pressure (get_pressure.f90:17-133; p = -<vv> on the mean line).
 * p_out receives p in the reference y-pencil layout (ny, nx/2, nz) complex,
 * vv_out receives <vv> (ny doubles); either may be NULL.  The state advances
 * exactly as with pb_substep.  For s != 1 it is pb_substep. */
int pb_substep_stats(pb_ctx *ctx, int32_t s, double *dt_inout, pb_substep_out *out, double *p_out,
                     double *vv_out);
/* pb_substep_stats plus get_statistics_2 (get_statistics.f90:376-539): stat2
 * receives PB_NSTAT2_LINES lines of ny doubles, the reference's datasets in
 * its order: uuu vvv www, uuuu vvvv wwww, u6 v6 w6, u8 v8 w8, u10 v10 w10
 * (physical-space averages of the raw powers over the 3/2-padded grid,
 * zero on the wall rows; dft_x_realspace.f90:128-484), -duuvdy -dvvvdy
 * -dvwwdy -duvvdy (D1 LU(D0)^-1 of the averages of -uuv, -vvv, -vww, -uvv),
 * up, -dpudy, vp, -2dpvdy, wp, pp (get_uv with the pressure), 2pdudx,
 * 2pdvdy, 2pdwdz, pdudy+pdvdx (get_turb_transport_PI1_ij, 817-902).
 * Substep 1 only; stat2 == NULL makes it pb_substep_stats.  Allocates its
 * buffers (3 fields) on the first call. */
#define PB_NSTAT2_LINES 29
int pb_substep_stats2(pb_ctx *ctx, int32_t s, double *dt_inout, pb_substep_out *out, double *p_out,
                      double *vv_out, double *stat2);
/* Profiles of get_statistics_1 (get_statistics.f90:54-155) from the state
 * held between steps: out receives 16 lines of ny doubles, in this order:
 * U (+u_gal), W, Theta (get_u, 1024-1048); dU/dy, dW/dy, dTheta/dy
 * (D1 LU(D0)^-1 of the mean column, 546-652); <uu>, <vv>, <ww>, <tt>,
 * <uv>, <uw>, <vw>, <ut>, <wt>, <vt> (get_uv, 982-1017: all modes, kx = 0
 * once, kx > 0 twice); then u_center and t_center (2 doubles). */
#define PB_NSTAT_LINES 16
int pb_statistics(pb_ctx *ctx, double *out);
/* 1-D energy (co)spectra of the state held between steps
 * (write_energy_spectra.f90:9-300, replacing its get_E_uu_x/get_E_uu_z
 * datasets; the HDF5 output stays with the caller).  du/dy and dw/dy are
 * D1 LU(D0)^-1 of every line (get_y_dir_deriv_of_field,
 * get_statistics.f90:658-681), the vorticity follows get_vorticity_field
 * (912-948).  ex receives PB_NSPECTRA planes (nx/2, ny) of sums over kz
 * (get_E_uu_x, write_energy_spectra.f90:326-359), ez PB_NSPECTRA planes
 * (nz, ny) of sums over kx (get_E_uu_z, 366-399), y fastest, kx = 0 once and
 * kx > 0 twice, in the reference's dataset order: uu vv ww tt uv vw uw ut vt
 * wt oxox oyoy ozoz oxoy oxoz oyoz voz woy (its kz "ozoz" is the (ox, oz)
 * co-spectrum, as in write_energy_spectra.f90:218-219).  Either output may be
 * NULL.  Uses product slots 3-5 and the scratch columns; the state is kept. */
#define PB_NSPECTRA 18
int pb_spectra(pb_ctx *ctx, double *ex, double *ez);
/* rk_step substeps; out may be NULL or point to rk_step records. */
int pb_step(pb_ctx *ctx, double *dt_inout, pb_substep_out *out);
/* pb_step with the step's device work captured in a hipGraph on the first
 * call (dt upload, the rk_step substeps with their side-stream forks and
 * joins, the scalars back) and replayed by every later call: one launch and
 * one synchronisation per step.  Same results as pb_step, bit for bit.  Not
 * while stage timing is on (PB_E_STATE).  No counterpart in the reference
 * (its time_march is host-driven). */
int pb_step_graph(pb_ctx *ctx, double *dt_inout, pb_substep_out *out);
/* Wait for the context's stream. */
int pb_sync(pb_ctx *ctx);

/* ---- ensembles of independent channels ---------------------------------- */
/* n channels of uneven size and parameters (one pb_config each) advanced
 * together: every channel's step is a captured graph (as pb_step_graph),
 * either all of them forked from one stream and joined back inside one
 * graph (one launch per ensemble step) or each channel's graph launched on
 * its own stream, whichever pb_ensemble_tune found faster; one wait per
 * ensemble step.  Each channel gives bit-identical
 * results to stepping it alone.  Channels are pb_ctx handles owned by the
 * ensemble (pb_ensemble_channel) for the state exchange; do not destroy
 * them.  No counterpart in the reference (one channel per run). */
typedef struct pb_ensemble pb_ensemble;
int pb_ensemble_create(int32_t n, const pb_config *cfgs, pb_ensemble **out);
int pb_ensemble_channel(pb_ensemble *e, int32_t i, pb_ctx **ctx);
/* dt[n] in/out (one per channel); out: rk_step records per channel, channel
 * after channel (or NULL).  The first error of any channel is returned. */
int pb_ensemble_step(pb_ensemble *e, double *dt, pb_substep_out *out);
/* The autotuner: times candidate schedules as captured steps (reps launches
 * each) and keeps the fastest -- per channel the tile width of the per-block
 * z transforms; for the ensemble one graph holding every channel versus each
 * channel's own graph on its own stream, each with the channels as given,
 * largest first or smallest first.  No candidate changes any
 * arithmetic.  Call it before pb_set_state: it steps the channels on zeroed
 * fields and leaves them without a state.  ms_out[n+1] (or NULL): the tuned
 * step time of each channel, then of the ensemble. */
int pb_ensemble_tune(pb_ensemble *e, int32_t reps, double *ms_out);
int pb_ensemble_destroy(pb_ensemble *e);

/* Per-stage device timing with HIP events recorded on the context's stream
 * around every kernel launch (no counterpart in the reference, whose phase
 * timers are commented out: time_march.f90:62-124).  enable != 0 starts
 * recording and clears the accumulators. */
#define PB_STAGE_OPPVD     0  /* detection-plane v for opposition control */
#define PB_STAGE_ZINV      1  /* y_to_x: padded inverse z DFT */
#define PB_STAGE_XPHYS     2  /* dft_x_realspace: x DFTs + products */
#define PB_STAGE_VELMAX    3  /* CFL maximum / dt */
#define PB_STAGE_ZFWD      4  /* x_to_y + rescale_data */
#define PB_STAGE_WAVESPACE 5  /* wavespace: all (kx,kz) systems */
#define PB_NSTAGE          6
int pb_timing(pb_ctx *ctx, int32_t enable);
/* Sum of event times (ms) and launch counts per stage; synchronises. */
int pb_timing_read(pb_ctx *ctx, double *ms, int64_t *launches);

/* ---- host-array drop-ins for the reference routines -------------------- */
/* These copy host arrays in the reference layout to the device, run the
 * corresponding kernels and copy back.  They exist for incremental adoption
 * and parity testing; the timed path is pb_substep. */

/* wavespace(U, RK_STORE, ...)  (wavespace.f90:22): U(ny,nx/2,nz,10) holds the
 * rescaled product spectra in slots 1-5, 7-9 (0-based 0-4, 6-8) and slot 10
 * (0-based 9) is vold, which the library reads only under PB_SCALAR_MIXED,
 * where it must hold u of the substep's input state (the scalar source);
 * on return slots 0,1,2,6 hold u, v, w, theta.
 * rk(ny,nx/2,nz,3) is updated in place.  v_old (ny,nx/2,nz) is the v of
 * the substep's input state (used by opposition control only; may be NULL
 * when opp_alpha == 0). */
int pb_wavespace_host(pb_ctx *ctx, int32_t s, double dt, double *U, double *rk,
                      const double *v_old, pb_substep_out *out);

/* y_to_x + dft_x_realspace + x_to_y + rescale_data for the four state
 * fields u, v, w, theta in U(ny,nx/2,nz,4): writes the product spectra
 * P(ny,nx/2,nz,10) in the reference slot order, slots 0-8 (slot 5 = v*v
 * only when want_vv != 0; slot 9, the reference's copy of v, is not
 * written).  velmax receives the CFL maximum. */
int pb_nonlinear_host(pb_ctx *ctx, const double *U, double *P, int32_t want_vv,
                      double *velmax);

/* The stages of pb_nonlinear_host one by one, in the layouts the reference
 * hands between them.  The x-pencil layout of the padded fields is
 * Upad(nx/2, 3nz/2, ny, slot) complex, kx fastest (module_ppdfftw.f90:
 * 1019-1146 with one rank): the z direction is in physical space on the
 * 3/2-padded grid, x still spectral.
 *
 * y_to_x (y_to_x.f90:21-49, module_ppdfftw.f90:830-861) of u, v, w, theta:
 * U(ny,nx/2,nz,4) -> Upad(nx/2,3nz/2,ny,4): kz padded to 3nz/2 (Nyquist slot
 * and gap zero), unnormalised backward DFT in z. */
int pb_y_to_x_host(pb_ctx *ctx, const double *U, double *Upad);
/* dft_x_realspace (dft_x_realspace.f90:24-213): Upad_in(nx/2,3nz/2,ny,4)
 * holds u, v, w, theta after y_to_x; Upad_out(nx/2,3nz/2,ny,10) receives the
 * product spectra in the reference slots 1-5, 7-9 (0-based 0-4, 6-8), v*v in
 * slot 6 (0-based 5) when want_vv (else zero) and the input v in slot 10
 * (0-based 9), as the reference keeps it for vold.  velmax receives the CFL
 * maximum max(|u| dxr + |v| dyr + |w| dzr) (lines 125-127). */
int pb_realspace_host(pb_ctx *ctx, const double *Upad_in, double *Upad_out, int32_t want_vv,
                      double *velmax);
/* x_to_y (x_to_y.f90:21-45, module_ppdfftw.f90:794-823) of nfield (1..9)
 * padded fields Upad(nx/2,3nz/2,ny,nfield) -> U(ny,nx/2,nz,nfield):
 * unnormalised forward DFT in z, kz truncated to nz with the Nyquist slot
 * zero, every value times scale (1 = x_to_y alone; 1/(3nx/2 * 3nz/2) =
 * x_to_y followed by rescale_data, rescale_data.f90:10-39). */
int pb_x_to_y_host(pb_ctx *ctx, int32_t nfield, const double *Upad, double *U, double scale);
/* get_pressure (get_pressure.f90:1-138) with the statistics-time inputs of
 * get_statistics_1: U(ny,nx/2,nz,4) the substep's input state (<vv>(y) and
 * the PPE wall values nu D2(wall row) LU(D0)^-1 v, get_statistics.f90:
 * 689-716, 982-1017) and P(ny,nx/2,nz,10) its rescaled product spectra with
 * v*v in slot 6 (0-based 5).  p_out (ny,nx/2,nz) receives p; vv_out (ny
 * doubles) <vv> when not NULL. */
int pb_pressure_host(pb_ctx *ctx, const double *U, const double *P, double *p_out, double *vv_out);

/* ---- LABS band operations (module_labs.f90) on host arrays -------------- */
/* ---- opposition gain, wall actuation and the discrete adjoint ----------------
 * Beside the drop-in: the reference has one opposition gain read at start-up
 * (opp_alpha) and no actuation or adjoint.  pb_set_control sets the gain and
 * a wall-actuation field act (nx/2 * nz complex, [kz][kx], the same value
 * added to the wall v of both walls for (kx,kz) != (0,0): v_wall = -alpha
 * v(y_d) + act), or none (act = NULL); a captured step is re-captured.  A
 * non-zero gain needs the detection plane (opp_ydist in (0,1)). */
int pb_set_control(pb_ctx *ctx, double alpha, const double *act);
/* Transpose of the wall-normal stage of substep s at a fixed dt (the
 * transpose of pb_wavespace_host's map, a linear map for fixed inputs):
 * AU (9 fields, layout of pb_wavespace_host's U) holds the adjoints of u, v,
 * w, theta (slots 0, 1, 2, 6) on entry and those of the products (0-4, 6-8)
 * and of the mixed-type source u (5) on return; rk (3 fields) the adjoints of
 * the new stores -> of the old ones.  vd: the forward's detection-plane v
 * (2 planes [kz][kx] complex) or NULL; avd (2 planes) receives its adjoint,
 * act_bar (1 plane) accumulates the actuation's, alpha_bar (nx/2 * nz
 * doubles) receives the per-system terms of d/d alpha (all may be NULL).
 * tau_bar, q_bar: adjoints of tau_w, q_w (last substep). */
int pb_adj_wavespace_host(pb_ctx *ctx, int32_t s, double dt, double *AU, double *rk, const double *vd, double *avd,
                          double *act_bar, double *alpha_bar, double tau_bar, double q_bar);
/* Transpose of the nonlinear stage (y_to_x, products, x_to_y) at state U
 * (4 fields, pb_set_state's layout): aprod (9 fields, pb_nonlinear_host's
 * layout; slots 0-4, 6-8 read) -> astate (4 fields). */
int pb_adj_nonlinear_host(pb_ctx *ctx, const double *U, const double *aprod, double *astate);
/* Transpose of one substep at a fixed dt around the input state `state` (4
 * fields): astate (4 fields) and rk (3 fields) hold the adjoints of the
 * substep's output state and stores on entry, of its input state and stores
 * on return; tau_bar, q_bar as above; *alpha_bar and act_bar (1 plane)
 * accumulate (may be NULL).  The context's u, v, w, theta become `state`. */
int pb_adj_substep(pb_ctx *ctx, int32_t s, double dt, const double *state, double *astate, double *rk,
                   double tau_bar, double q_bar, double *alpha_bar, double *act_bar);
/* Gradient of J = (1/nsteps) sum over steps of (tau_w + wq q_w) (the values
 * after each step's last substep) over nsteps steps of fixed dt from the
 * context's state, with respect to the opposition gain (*dJ_dalpha) and the
 * actuation field (dJ_da, 1 plane complex): forward with a checkpoint every
 * ckpt_every steps, then each segment recomputed from its checkpoint and
 * reversed substep by substep.  *J receives J; the context ends at the
 * state after the window.  Needs dt_factor = 1.  Allocates on first use
 * (nsteps/ckpt_every + 1) * 7 fields of checkpoints and, per substep of one
 * segment, the z inverse of its input state and its detection planes (4
 * padded fields + 2 planes: no transform is repeated in the reverse pass),
 * or, when that exceeds a quarter of the free device memory or the
 * environment sets PB_ADJ_SAVE=state, its input state (4 fields). */
int pb_gradient(pb_ctx *ctx, int32_t nsteps, int32_t ckpt_every, double dt, double wq, double *J, double *dJ_dalpha,
                double *dJ_da);

/* Packed bands A(2bw+1, n) in the reference's near-boundary format
 * (LABS_full_to_band, module_labs.f90:301-328), one after the other; vectors
 * are n complex values, one after the other.  They run the device solver's
 * band kernels (one thread per matrix or vector) on the current device and
 * synchronise; count = 1 is the reference's single call.
 * LABS_lu (module_labs.f90:462; no-pivot Crout, fixed_bandec*): in place. */
int pb_labs_lu(int32_t n, int32_t bw, int32_t nmat, double *A);
/* LABS_bs (module_labs.f90:499; fixed_banbks*): b <- A^-1 b for nrhs columns
 * with the factor A_lu of pb_labs_lu. */
int pb_labs_bs(int32_t n, int32_t bw, const double *A_lu, int32_t nrhs, double *b);
/* LABS_mat_mul (module_labs.f90:570): b = A x for nrhs columns. */
int pb_labs_mat_mul(int32_t n, int32_t bw, const double *A, int32_t nrhs, const double *x, double *b);
/* pb_labs_bs without any allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight loops that call
 * LABS_bs per line, harnesses that forbid hipMalloc).  Same arguments and the
 * same kernel as pb_labs_bs (bit-identical results); the factor and the
 * right-hand sides are staged through a library-owned device array of
 * PB_NOALLOC_SCRATCH_BYTES on the current GPU (one per GPU, held until the
 * process ends), so n*(2bw+1) doubles, rounded up to 32, plus 2*nrhs*n
 * doubles must fit in it (PB_E_ARG otherwise).  Calls that overlap on one
 * GPU (other threads) share that array and race. */
#define PB_NOALLOC_SCRATCH_BYTES (8u << 20)
int pb_labs_bs_noalloc(int32_t n, int32_t bw, const double *A_lu, int32_t nrhs, double *b);

#ifdef __cplusplus
}
#endif
#endif /* PB_CAPI_H */
