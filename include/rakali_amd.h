/*
 * rakali_amd.h -- C ABI of the MI355X (gfx950) structured 2-D shallow-water
 * step: the kernel-level boundary ("rkg_*") that replaces the calls
 * solver_step_impl makes into the structured barotropic kernels.
 *
 * Reference interfaces replaced (file:line in JorgeG94/rakali):
 *   rkg_step / rkg_advance  -> the fused SSP-RK2 branch of solver_step_impl
 *                              (src/core/coastal/solver/rki_solver.F90:692-815)
 *                              as driven by rki_step / rki_advance
 *                              (src/api/rki_api.F90:358-410), i.e. per step:
 *                              apply_boundary_conditions (rki_boundary.F90:40),
 *                              compute_timestep (rki_kernel_timestep.F90:16),
 *                              compute_flux_hll[c] (rki_kernel_flux.F90:17,62),
 *                              fused_update_stage1/2 (rki_kernel_fused.F90:164,305)
 *   rkg_create / rkg_destroy -> barotropic_state_init + solver_enter_data /
 *                              solver_exit_data (rki_barotropic_state.F90:114,
 *                              rki_solver.F90:131,281)
 *   rkg_upload / rkg_download-> rki_set_bathymetry / rki_set_initial_h /
 *                              rki_set_h/hu/hv and the rki_get_*_ptr getters
 *                              (rki_api.F90:502-686, 836-1123) at the level of
 *                              device buffers
 *   rkg_total_mass          -> compute_total_mass (rki_init.F90:287)
 *
 * All functions return 0 on success and a negative code on failure, in the
 * reference's convention (rki_api.F90:22-27): -1 not ready, -2 shape or
 * argument out of range, -100 invalid handle; plus -4 for a HIP runtime
 * error and -5 for a configuration this build does not implement.
 * Arrays on the host side are Fortran-ordered FP64 (x fastest) with a
 * leading dimension `ld` >= nx, so a caller may pass the interior of a
 * ghosted (nx_total, ny_total) array as (&a[ng + ng*nx_total], ld=nx_total).
 */
#ifndef RAKALI_AMD_H
#define RAKALI_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define RKG_OK 0
#define RKG_ERR_NOT_READY (-1)
#define RKG_ERR_RANGE (-2)
#define RKG_ERR_HIP (-4)
#define RKG_ERR_UNSUPPORTED (-5)
#define RKG_ERR_BAD_HANDLE (-100)

/* Boundary tags: rki_boundary_types.F90:18-32. */
#define RKG_BC_WALL 1
#define RKG_BC_OPEN 2
#define RKG_BC_TIDAL 3
#define RKG_BC_INFLOW 5
#define RKG_BC_DISCHARGE 6
#define RKG_BC_CLAMPED 7
#define RKG_BC_SPONGE 8
/* Chapman / Orlanski radiation (bc_chapman_edge, rki_boundary.F90:471-552):
 * the ghost surface moves with the phase speed estimated from the interior
 * eta tendency since the previous boundary call and the along-normal
 * gradient; the normal momentum is the Flather radiation about it.  The
 * adjoint and the ensembles do not support it. */
#define RKG_BC_CHAPMAN 9

/* Field selectors.  H0/HU0/HV0 are U^n of the last step (the reference's
 * RK save h0/hu0/hv0, rki_kernel_fused.F90:224-226), kept only while
 * rkg_set_rk_save is on; HS/HUS/HVS are U*, the stage-1 result of the last
 * step.  Fields above B are read-only. */
#define RKG_FIELD_H 0
#define RKG_FIELD_HU 1
#define RKG_FIELD_HV 2
#define RKG_FIELD_B 3
#define RKG_FIELD_H0 4
#define RKG_FIELD_HU0 5
#define RKG_FIELD_HV0 6
#define RKG_FIELD_HS 7
#define RKG_FIELD_HUS 8
#define RKG_FIELD_HVS 9
/* Per-cell Manning n (writable once rkg_use_manning_field has enabled it).
 * The reference's manning_n is a scalar (rki_barotropic_state.F90:18-110);
 * the field is an addition for gradients with respect to roughness. */
#define RKG_FIELD_N 10
/* Surface forcing planes (state%forcing, rki_surface_forcing.F90), allocated
 * by rkg_set_forcing and zero until uploaded: wind stress (Pa, the x and y
 * components), atmospheric pressure (Pa) and precipitation rate (m/s).
 * The pressure gradient is central, with pressure 0 outside the grid (the
 * reference's ghost cells, which its setters never write,
 * rki_api.F90:926-980). */
#define RKG_FIELD_WIND_X 11
#define RKG_FIELD_WIND_Y 12
#define RKG_FIELD_PRESSURE 13
#define RKG_FIELD_PRECIP 14
/* Read-only: the state of the last step's last boundary call (U* of
 * SSP-RK2 stage 2, the last stage's input of a generic scheme, U^n for
 * euler) and h of the boundary call before it, for rebuilding the ghost
 * cells the reference leaves after a step.  Need rkg_set_rk_save on. */
#define RKG_FIELD_BC_H 15
#define RKG_FIELD_BC_HU 16
#define RKG_FIELD_BC_HV 17
#define RKG_FIELD_BC_PREV_H 18

#define RKG_MAX_TIDAL_CONSTITUENTS 10

/* Explicit time schemes, tags as TIME_* in rki_timestepper.F90:27-33. */
#define RKG_TIME_EULER 1
#define RKG_TIME_RK2 2
#define RKG_TIME_RK3 3
#define RKG_TIME_RK43 4

/* Scalar configuration of one grid; the namelist keys the step reads
 * (config_t, rki_config.F90:37-260; bc_params_t, rki_boundary_types.F90:37). */
typedef struct rkg_config {
    int nx, ny;                 /* physical cells (>= 2 each) */
    double dx, dy;              /* cell sizes (m) */
    double g;                   /* gravity; reference GRAVITY = 9.80665 */
    double cfl;                 /* CFL number (0.45) */
    double manning_n;           /* scalar Manning n (0 = no friction) */
    double coriolis_f;          /* Coriolis parameter (1/s) */
    double dt_fixed;            /* > 0: fixed dt */
    double t_end;               /* bounds dt in rkg_step (dt_max = t_end - t) */
    int cfl_interval;           /* dt recomputed every cfl_interval steps */
    int use_hllc;               /* 1 = HLLC, 0 = HLL (reference default HLL) */
    int bc[4];                  /* west, east, south, north tags */
    double tidal_amplitude;     /* single-constituent tide (m) */
    double tidal_period;        /* (s), default 44712 */
    int n_tidal_constituents;   /* > 0: constituent sum replaces the above */
    double tidal_amp[RKG_MAX_TIDAL_CONSTITUENTS];
    double tidal_phase[RKG_MAX_TIDAL_CONSTITUENTS];
    double tidal_omega[RKG_MAX_TIDAL_CONSTITUENTS];
    double inflow_velocity;     /* BC_INFLOW normal velocity (m/s) */
    double discharge_flux;      /* BC_DISCHARGE normal flux (m^2/s) */
    double clamped_elevation;   /* BC_CLAMPED surface (m) */
} rkg_config;

typedef struct rkg_grid rkg_grid;   /* opaque handle */

/* Reference defaults (rki_config.F90, rki_constants.F90). */
int rkg_config_default(rkg_config *cfg);

/* Allocates device state for one grid on the current HIP device.  `stream`
 * is a hipStream_t to launch on, or NULL for a library-owned stream.  A
 * stream obtained from another grid with rkg_get_stream is shared: the grids
 * may be destroyed in any order and the stream goes with the last of them.
 * Any other caller-provided stream is borrowed and must outlive every grid
 * that uses it. */
int rkg_create(const rkg_config *cfg, void *stream, rkg_grid **out);
/* Frees the grid and nulls *g; a NULL or already-destroyed handle is a no-op.
 * If adjoint sessions of the grid are still alive, its device state is kept
 * until the last of them is destroyed; those sessions then accept only
 * rkg_adjoint_destroy (every other call returns RKG_ERR_BAD_HANDLE). */
void rkg_destroy(rkg_grid **g);

/* Host <-> device copies of one interior field (see the header comment for
 * `ld`).  Uploading h, hu or hv (or b) marks the CFL state stale; the next
 * step recomputes it. */
int rkg_upload(rkg_grid *g, int field, const double *src, int ld);
int rkg_download(rkg_grid *g, int field, double *dst, int ld);
/* Device-to-device variants: `src`/`dst` is a device pointer with leading
 * dimension `ld` elements. */
int rkg_upload_device(rkg_grid *g, int field, const double *src, int ld);
int rkg_download_device(rkg_grid *g, int field, double *dst, int ld);
/* Zero-copy device view of a field: pointer and row pitch (elements). */
int rkg_field_device_ptr(rkg_grid *g, int field, double **ptr, long long *pitch);

/* Advances n SSP-RK2 steps (rki_step semantics: dt_max = t_end - t).
 * Asynchronous: returns once the launches are queued. */
int rkg_step(rkg_grid *g, int n);
/* Steps while t < target - 1e-10 with dt_max = target - t (rki_advance
 * semantics).  Synchronous; *steps_taken may be NULL. */
int rkg_advance(rkg_grid *g, double target, long long *steps_taken);

/* rkg_step without a grid and without allocating: n SSP-RK2 steps on
 * state planes the caller owns.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  h, hu, hv (in
 * place) and b are device planes of the cfg->nx x cfg->ny interior, row
 * pitch `pitch` >= nx elements; *t, *dt, *step_count hold the clock on
 * entry (as rkg_get_clock returns it) and the clock after the steps on
 * return.  Results equal rkg_step's on a grid holding the same state and
 * clock, bit for bit (no mass budget, RK save, forcing, sponge or Manning
 * field).  The intermediates (U* and the control slots) live in a fixed
 * device array of 256 B + 24 B x RKG_NOALLOC_MAX_CELLS per GPU, held until
 * the process ends; pitch x ny above RKG_NOALLOC_MAX_CELLS returns -2.
 * Calls on one GPU share that array: calls that overlap on one GPU race.
 * Launches on `stream` (NULL: the null stream) and synchronizes it to
 * return the clock. */
#define RKG_NOALLOC_MAX_CELLS 2097152
int rkg_step_noalloc(const rkg_config *cfg, double *h, double *hu, double *hv, const double *b,
                     long long pitch, int n, double *t, double *dt, long long *step_count, void *stream);

/* Clock: time, dt of the last step, and step count (synchronizes). */
int rkg_get_clock(rkg_grid *g, double *t, double *dt, long long *step_count);
/* Sets the time and step count (e.g. after a restart). */
int rkg_set_clock(rkg_grid *g, double t, long long step_count);
/* Total mass, sum of h*dx*dy over interior cells with 0 < h < 1e6, in a
 * fixed summation order (deterministic). */
int rkg_total_mass(rkg_grid *g, double *mass);
/* Waits for all queued work of this grid. */
int rkg_sync(rkg_grid *g);

/* Rows per tile segment of the stage kernels (1..64; 0, the default,
 * chooses 16 for grids of at least 8192 such tiles and 8 below).  The get
 * call returns the rows in use. */
int rkg_set_tile_rows(rkg_grid *g, int rows);
int rkg_get_tile_rows(rkg_grid *g, int *rows);
/* Whether the stage kernels try their fast path (wavefronts whose faces
 * and cells are all wet and subsonic skip the dry and shock branches and
 * use cheaper exact sequences; any other wavefront takes the general code,
 * so results are the same bits either way): 0 (default) when the launch
 * has at least 6144 tiles, where it pays; 1 always; -1 never.  A tuning and
 * testing knob with no counterpart in the reference. */
int rkg_set_fast_path(rkg_grid *g, int mode);
/* Keeps U^n of every step in three extra planes (24 B per cell per step of
 * extra traffic) so that RKG_FIELD_H0..HV0 can be read; off by default. */
int rkg_set_rk_save(rkg_grid *g, int enable);
/* enable != 0: friction uses a per-cell Manning n (field RKG_FIELD_N,
 * initialised to the scalar manning_n; friction where n > 0); 0: back to
 * the scalar.  No counterpart in the reference (scalar n only). */
int rkg_use_manning_field(rkg_grid *g, int enable);
/* Time scheme of rkg_step / rkg_advance (state%time_scheme, set from the
 * namelist key time_scheme, rki_config.F90).  RKG_TIME_RK2 (default) is
 * the fused SSP-RK2 path (rki_solver.F90:760-815); euler, rk3 and rk43 run
 * the generic SSP stage loop (solver_advance_2d_generic,
 * rki_solver.F90:861-925, coefficients rki_timestepper.F90:60-100; rk43
 * doubles the CFL dt).  rk3 and rk43 allocate three more planes. */
int rkg_set_time_scheme(rkg_grid *g, int scheme);
int rkg_get_time_scheme(rkg_grid *g, int *scheme);
/* Boundary mass budget state%barotropic%mass_out
 * (accumulate_boundary_mass_out, rki_solver.F90:507-537): while enabled,
 * every SSP-RK2 step subtracts 0.5*dt*sum(flux_h)*dx*dy per stage, the
 * volume that left through the boundary.  Enabling starts it at 0; the
 * generic schemes do not accumulate it (as in the reference).  The sum is
 * in a fixed order (per tile, then over tiles). */
int rkg_set_mass_budget(rkg_grid *g, int enable);
int rkg_get_mass_budget(rkg_grid *g, double *mass_out);
/* Surface forcing on (active != 0) or off: state%forcing%active, which the
 * reference's setters rki_set_wind / rki_set_atm_pressure /
 * rki_set_precipitation switch on (rki_api.F90:898-980).  While on, each
 * stage adds the wind stress and the pressure gradient to hu, hv of wet
 * cells and the precipitation to h of every cell (rki_kernel_fused.F90:
 * 273-289).  The first call allocates the four RKG_FIELD_WIND_X..PRECIP
 * planes and two for the pressure gradient (48 B per cell); they stay until
 * rkg_destroy.  The gradient is taken from the pressure plane before the
 * next step after an upload of it (every step once its device pointer has
 * been requested).  The adjoint does not support forcing. */
/* bc%time and bc%dt of the last step's last boundary call (stage 2 at
 * t + dt for SSP-RK2, t + c_last dt for a generic scheme). */
int rkg_get_last_bc(rkg_grid *g, double *time, double *dt);
int rkg_set_forcing(rkg_grid *g, int active);
int rkg_get_forcing(rkg_grid *g, int *active);
/* Surface mass source state%barotropic%mass_src (accumulate_mass_src,
 * rki_solver.F90:539-606): while the mass budget is enabled and the
 * forcing is on, each SSP-RK2 step adds 0.5*dt*sum(precipitation)*dx*dy
 * per stage, and the sponge's change of volume (rkg_set_sponge).
 * rkg_set_mass_budget(g, 1) starts it at 0. */
int rkg_get_mass_source(rkg_grid *g, double *mass_src);
/* Gauge stations (rki_gauges.F90): points (x, y) in metres from the grid's
 * lower-left corner.  rkg_gauges_set keeps the stations inside the domain
 * (compute_local_gauges, :208-243; n_local of them, in input order,
 * rkg_gauges_local gives their input indices) with their bilinear
 * interpolation cells and weights (compute_interp_weights, :245-336: the
 * two cell centres around the point, clamped to the edge cell in the
 * outer half cells); n = 0 removes them.  A sample is eta = h + b and
 * u, v = hu, hv / max(h, 1e-6), each interpolated in the reference's
 * operation order (gauges_write, :490-575). */
#define RKG_MAX_GAUGES 1048576
int rkg_gauges_set(rkg_grid *g, int n, const double *x, const double *y, int *n_local);
int rkg_gauges_local(rkg_grid *g, int *local_idx);
/* One sample of the current state (n_local values each). */
int rkg_gauges_sample(rkg_grid *g, double *eta, double *u, double *v);
/* Records a sample after every step on the device, at the driver's cadence
 * (rki_driver.F90:345-352,470-474): with interval > 0, when the step's end
 * time reaches next = t_start + interval, next += interval; with
 * interval <= 0, after every step.  `capacity` records fit in the device
 * buffer, which is drained to host memory when it may fill (a pause of the
 * step loop); capacity 0 stops recording.  Starting clears the series. */
int rkg_gauges_record(rkg_grid *g, double interval, int capacity);
/* Takes up to max_records of the oldest recorded samples: times t
 * [n_records], eta, u, v [n_records][n_local] (station fastest), and
 * removes them from the series.  With t == NULL or max_records == 0,
 * returns the count only. */
int rkg_gauges_series(rkg_grid *g, long long max_records, long long *n_records, double *t, double *eta,
                      double *u, double *v);
/* Sponge relaxation (apply_sponge, rki_kernel_sponge.F90:16-88; namelist
 * sponge_width, sponge_strength, sponge_h_ref, bc_params_t
 * rki_boundary_types.F90:73-77).  Interior cells closer than `width`
 * cells to an RKG_BC_SPONGE edge relax toward (h_ref, 0, 0) after every
 * stage: h -= dt*alpha*(h - h_ref), hu -= dt*alpha*hu, hv likewise, with
 * alpha = strength*((width - d)/width)^2 and d the distance in cells from
 * the nearest sponge edge.  Off for width 0 (the default) or strength <= 0.
 * With the mass budget on, an SSP-RK2 step adds 0.5*(sum of the change of
 * h)*dx*dy per stage to the mass source (accumulate_sponge_mass_src,
 * rki_solver.F90:608-657).  The adjoint does not support the sponge. */
int rkg_set_sponge(rkg_grid *g, int width, double strength, double h_ref);
int rkg_get_sponge(rkg_grid *g, int *width, double *strength, double *h_ref);
/* Reads / replaces the scalar configuration (nx and ny cannot change):
 * boundary values, tide, friction, CFL, t_end, HLL/HLLC. */
int rkg_get_config(rkg_grid *g, rkg_config *cfg);
int rkg_set_config(rkg_grid *g, const rkg_config *cfg);

/* Kernel timing: when enabled, rkg_step brackets each stage launch with
 * HIP events on the grid's stream and waits for every step, accumulating
 * the stage-1 and stage-2 kernel times (ms).  Enabling resets the sums. */
int rkg_profile(rkg_grid *g, int enable);
int rkg_profile_read(rkg_grid *g, double *stage1_ms, double *stage2_ms, long long *steps);
/* The hipStream_t the grid launches on. */
int rkg_get_stream(rkg_grid *g, void **stream);
/* Last-error text of this thread. */
const char *rkg_last_error(void);

/* HBM bandwidth reference: times `reps` launches of a plain copy kernel
 * (k_copy) of `bytes` bytes on the current device; *gbs counts read plus
 * written bytes.  Used to calibrate traffic counters and state the
 * achievable copy bandwidth next to the 8 TB/s nominal peak. */
int rkg_copy_bandwidth(long long bytes, int reps, double *gbs, double *ms_per_copy);

/* Evaluates the step's transcendental functions on n inputs, on the host
 * (on_device = 0) or the current GPU: fn 0 = x**(1.0/3.0) (Manning,
 * rki_kernel_fused.F90:268), 1 = sin, 2 = cos (tidal surface
 * rki_boundary.F90:96-100, Coriolis rki_kernel_fused.F90:212-213).  Both
 * sides return identical bits, correctly rounded except for inputs whose
 * exact result lies within about 1e-26 (x**(1/3), for x in [2^-900,
 * 2^1000): 4e-24) relative of a rounding midpoint. */
int rkg_math_eval(int fn, const double *x, double *y, int n, int on_device);

/* ------------------------------------------------------------------------
 * Ensembles of independent grids of uneven sizes, stepped together with
 * one launch per RK stage.  No counterpart in the reference, which steps
 * one grid per handle; these are additive entry points for batched runs.
 * Each member keeps its own clock (t, dt, step count), bed, physics and
 * boundaries; results per member are bit-identical to a single grid with
 * the same configuration.  All members share use_hllc.
 * ------------------------------------------------------------------------ */
typedef struct rkg_ensemble rkg_ensemble;   /* opaque handle */

int rkg_ensemble_create(const rkg_config *cfgs, int n_members, void *stream, rkg_ensemble **out);
void rkg_ensemble_destroy(rkg_ensemble **e);
/* member count, total interior cells and tiles in the table */
int rkg_ensemble_size(rkg_ensemble *e, int *n_members, long long *cells, int *n_tiles);
/* fields as for rkg_upload / rkg_download (H, HU, HV, B writable; HS.. readable) */
int rkg_ensemble_upload(rkg_ensemble *e, int member, int field, const double *src, int ld);
int rkg_ensemble_download(rkg_ensemble *e, int member, int field, double *dst, int ld);
/* n steps of every member (each bounded by its own t_end, as rkg_step) */
int rkg_ensemble_step(rkg_ensemble *e, int n);
int rkg_ensemble_sync(rkg_ensemble *e);
int rkg_ensemble_get_clock(rkg_ensemble *e, int member, double *t, double *dt, long long *step_count);
int rkg_ensemble_total_mass(rkg_ensemble *e, int member, double *mass);
/* schedule: 0 = one workgroup per tile, 1 = persistent workgroups taking
 * tiles from an atomic queue; tile_rows in [4, 64] (0 = keep);
 * persistent_blocks (0 = keep); use_graph: replay steps from a hipGraph. */
int rkg_ensemble_set_schedule(rkg_ensemble *e, int schedule, int tile_rows, int persistent_blocks, int use_graph);
int rkg_ensemble_get_schedule(rkg_ensemble *e, int *schedule, int *tile_rows, int *persistent_blocks, int *use_graph);
/* Times `steps` steps of each candidate schedule and tile height on the
 * live ensemble, keeps the fastest, restores the state. */
int rkg_ensemble_autotune(rkg_ensemble *e, int steps, double *best_ms_per_step);

/* ------------------------------------------------------------------------
 * Discrete adjoint of the step: gradients of a gauge misfit
 *   L = sum_{steps s} sum_{gauges k} (h + b - target_{s,k})^2
 * (h + b at the gauge cell after step s) with respect to the bed, the
 * per-cell Manning n and the initial h, hu, hv.  No counterpart in the
 * reference; for callers that fit bathymetry or roughness.  The forward
 * run keeps a checkpoint every `checkpoint_interval` steps; the reverse
 * sweep replays each segment from its checkpoint and reverses it.  dt and
 * the boundary times are those of the forward run (exact at fixed dt).
 * Creating an adjoint switches the grid to a per-cell Manning field.
 * ------------------------------------------------------------------------ */
typedef struct rkg_adjoint rkg_adjoint;     /* opaque handle */

/* gauge cells (gauge_i[k], gauge_j[k]), 0-based interior indices;
 * targets: n_steps x n_gauges, step-major (may be NULL: zeros). */
int rkg_adjoint_create(rkg_grid *g, int n_steps, int checkpoint_interval, int n_gauges, const int *gauge_i,
                       const int *gauge_j, const double *targets, rkg_adjoint **out);
/* The same loss over stations at (x[k], y[k]) metres, eta sampled as
 * rkg_gauges_sample does (bilinear, rki_gauges.F90:245-336,516-550); every
 * station must lie inside the domain.  No counterpart in the reference,
 * whose gauge series is output only. */
int rkg_adjoint_create_stations(rkg_grid *g, int n_steps, int checkpoint_interval, int n_stations,
                                const double *x, const double *y, const double *targets, rkg_adjoint **out);
/* Frees the session and nulls *a; releases the session's hold on its grid. */
void rkg_adjoint_destroy(rkg_adjoint **a);
int rkg_adjoint_set_targets(rkg_adjoint *a, const double *targets);
/* n_steps steps from the grid's current state; returns L; the grid is
 * left at the final state. */
int rkg_adjoint_forward(rkg_adjoint *a, double *loss);
/* gauge eta of every step of the last forward run (n_steps x n_gauges) */
int rkg_adjoint_gauge_eta(rkg_adjoint *a, double *eta);
/* gradients of loss_scale * L; tile_rows of the reverse y-face kernel (0:
 * chosen from the grid size so that about 2048 tiles run at once, 4 to 32
 * rows).  With two or more segments, the first call also allocates a
 * second segment workspace ((12 K + 3) planes) and a stream of its own when
 * that takes at most half of the free device memory: segment s - 1 is then
 * replayed while segment s is reversed (the same gradients, bit for bit). */
int rkg_adjoint_backward(rkg_adjoint *a, double loss_scale, int tile_rows);
/* adds the bed part of a still-water initial state h = max(eta0 - b, 0) */
int rkg_adjoint_fold_still_water(rkg_adjoint *a, double eta0);
/* field: RKG_FIELD_B, RKG_FIELD_N, or RKG_FIELD_H/HU/HV (initial state) */
int rkg_adjoint_gradient(rkg_adjoint *a, int field, double *dst, int ld, int on_device);
/* HIP-event time of the last forward and backward calls (ms) */
int rkg_adjoint_timing(rkg_adjoint *a, double *forward_ms, double *backward_ms);
/* Batches of sessions: the forward runs / reverse sweeps of n sessions whose
 * grids have their own streams (and share n_steps and the checkpoint
 * interval), queued step by step across the sessions so that their kernels
 * overlap on the GPU.  Results per session equal rkg_adjoint_forward /
 * rkg_adjoint_backward (automatic tile rows) bit for bit.  losses may be
 * NULL.  No counterpart in the reference (ensembles of gradients). */
int rkg_adjoint_forward_many(int n, rkg_adjoint *const *sessions, double *losses);
int rkg_adjoint_backward_many(int n, rkg_adjoint *const *sessions, double loss_scale);

#ifdef __cplusplus
}
#endif
#endif
