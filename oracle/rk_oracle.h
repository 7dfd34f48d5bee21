/*
 * rk_oracle.h -- CPU restatement of Rakali's structured 2D shallow-water
 * step.  TEST INFRASTRUCTURE ONLY: this is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (rakali_amd/) never does.
 *
 * Arrays follow the reference layout: Fortran column-major
 * (nx_total, ny_total) FP64 with i (x) fastest and `ng` ghost cells on each
 * side (rki_state.F90:5-7).  Index (i, j) below is 1-based like the
 * reference.
 */
#ifndef RK_ORACLE_H
#define RK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Boundary tags, values as rki_boundary_types.F90:18-32. */
enum { ORC_BC_WALL = 1, ORC_BC_OPEN = 2, ORC_BC_TIDAL = 3, ORC_BC_NESTED = 4,
       ORC_BC_INFLOW = 5, ORC_BC_DISCHARGE = 6, ORC_BC_CLAMPED = 7,
       ORC_BC_SPONGE = 8, ORC_BC_CHAPMAN = 9 };

#define ORC_MAX_TC 10 /* MAX_TIDAL_CONSTITUENTS, rki_config.F90 */

typedef struct {
    int nx, ny, ng;          /* physical sizes and ghost width (ng >= 2) */
    double dx, dy;
    double g;                /* GRAVITY, rki_constants.F90:38 (9.80665) */
    double cfl;              /* config cfl (0.45 default) */
    double manning_n;        /* scalar Manning n */
    double coriolis_f;
    double dt_fixed;         /* > 0 -> fixed dt (rki_solver.F90:728) */
    double t_end;            /* config t_end, bounds rki_step's dt */
    int cfl_interval;        /* dt recomputed every cfl_interval steps */
    int use_hllc;            /* 1 = HLLC, 0 = HLL */
    int bc[4];               /* west, east, south, north tags */
    double tidal_amplitude, tidal_period;
    int n_tc;
    double tc_amp[ORC_MAX_TC], tc_phase[ORC_MAX_TC], tc_omega[ORC_MAX_TC];
    double inflow_velocity, discharge_flux, clamped_elevation;
    int time_scheme;         /* ORC_TIME_* (rki_timestepper.F90:27-36) */
    /* Surface forcing (rki_surface_forcing.F90): ghosted arrays like the
     * state, or NULL (zero); applied when forcing_active
     * (rki_kernel_fused.F90:268-283, :286-289). */
    const double *wind_x, *wind_y, *atm_pressure, *precipitation;
    int forcing_active;
    /* Sponge relaxation next to BC_SPONGE edges (bc_params_t,
     * rki_boundary_types.F90:73-77); off when width <= 0 or strength <= 0. */
    int sponge_width;
    double sponge_strength, sponge_h_ref;
    /* Chapman radiation edges: each edge's eta_old (bc_params_t
     * eta_old_west/east/south/north, rki_boundary_types.F90; W/E of length
     * ny_total, S/N nx_total), updated by every boundary call, or NULL. */
    double *chap_eta_old[4];
} orc_params;

/* Explicit SSP schemes, tags as rki_timestepper.F90:27-33.  ORC_TIME_RK2
 * runs the fused RK2 sequence; ORC_TIME_RK2_GENERIC runs the same Heun
 * coefficients through the generic stage loop (a check of that loop). */
enum { ORC_TIME_EULER = 1, ORC_TIME_RK2 = 2, ORC_TIME_RK3 = 3, ORC_TIME_RK43 = 4,
       ORC_TIME_RK2_GENERIC = 102 };

typedef struct {
    double t;            /* handle t_current */
    double dt;           /* state%dt of the last step */
    double mass_out;     /* boundary mass budget (rki_solver.F90:507-537) */
    long long step_count;
    double mass_src;     /* surface source budget (rki_solver.F90:539-606) */
} orc_clock;

/* Fills default parameters (rki_config.F90 defaults, rki_constants.F90). */
void orc_default_params(orc_params *p);

/* Primitives (each mirrors one reference routine). */
void orc_compute_flux(const orc_params *p, const double *h, const double *hu,
                      const double *hv, const double *b, double *flux_h,
                      double *flux_hu, double *flux_hv, int use_hllc);
void orc_fused_stage1(const orc_params *p, double *h, double *hu, double *hv,
                      double *h0, double *hu0, double *hv0,
                      const double *fh, const double *fhu, const double *fhv,
                      double dt);
void orc_fused_stage2(const orc_params *p, double *h, double *hu, double *hv,
                      const double *h0, const double *hu0, const double *hv0,
                      const double *fh, const double *fhu, const double *fhv,
                      double dt);
/* fused_update_stage_generic (rki_kernel_fused.F90:22-162) and
 * compute_ssp_combine (rki_kernel_update.F90:152-208). */
void orc_fused_stage_generic(const orc_params *p, double *h, double *hu, double *hv,
                             const double *fh, const double *fhu, const double *fhv,
                             double dt);
void orc_ssp_combine(const orc_params *p, double *h, double *hu, double *hv,
                     const double *h0, const double *hu0, const double *hv0,
                     double a_n);
/* apply_sponge (rki_kernel_sponge.F90:16-88). */
void orc_apply_sponge(const orc_params *p, double *h, double *hu, double *hv, double dt);
/* Gauges (rki_gauges.F90): compute_local_gauges + compute_interp_weights
 * (:208-336) for n stations at (x, y) in metres.  Returns n_local; writes
 * local_idx (0-based station ids inside the domain), ix0 / iy0 (1-based
 * ghosted array indices of the lower-left interpolation cell, as the
 * reference) and the weights wx, wy. */
int orc_gauges_setup(const orc_params *p, int n, const double *x, const double *y,
                     int *local_idx, int *ix0, int *iy0, double *wx, double *wy);
/* gauges_write's bilinear interpolation (:516-550): eta = h + b and the
 * depth-averaged velocities with the 1e-6 m depth guard. */
void orc_gauges_sample(const orc_params *p, const double *h, const double *hu,
                       const double *hv, const double *b, int n_local, const int *ix0,
                       const int *iy0, const double *wx, const double *wy,
                       double *eta, double *u, double *v);
double orc_compute_timestep(const orc_params *p, const double *h,
                            const double *hu, const double *hv);
double orc_tidal_eta(const orc_params *p, double time);
/* apply_boundary_conditions at bc%time = time with bc%dt = 0 (its value
 * before the first step); orc_apply_bc_dt with the step's dt (bc%dt =
 * state%dt, rki_solver.F90:748), which the Chapman edges read. */
void orc_apply_bc(const orc_params *p, double time, double *h, double *hu,
                  double *hv, double *b);
void orc_apply_bc_dt(const orc_params *p, double time, double bc_dt, double *h,
                     double *hu, double *hv, double *b);
/* bc_init_chapman (rki_boundary_types.F90:155-202): seeds eta_old of the
 * Chapman edges from h + b of the first interior column / row. */
void orc_chapman_seed(const orc_params *p, const double *h, const double *b);
void orc_fill_bathymetry_ghosts(const orc_params *p, double *b);
void orc_init_still_water(const orc_params *p, double eta0, const double *b,
                          double *h, double *hu, double *hv);
double orc_total_mass(const orc_params *p, const double *h);

/* The step sequence of rki_solver.F90:692-815 (fused SSP-RK2) or, for the
 * other time schemes, :861-925 (generic SSP stage loop), driven by
 * rki_step (mode 0, n steps, dt_max = t_end - t; rki_api.F90:358-383) or
 * rki_advance (mode 1, loop while t < target - 1e-10; rki_api.F90:385-410).
 * Returns the number of steps taken, or a negative error. */
long long orc_run(const orc_params *p, double *h, double *hu, double *hv,
                  double *b, orc_clock *clk, int mode, long long nsteps,
                  double target);

/* Transcendentals: 0 (default) evaluates pow(h, 1/3), sin and cos correctly
 * rounded (rk_crmath.h); 1 calls the platform libm as the reference does. */
void orc_set_libm(int use_platform_libm);
int orc_get_libm(void);
/* fn 0: x**(1/3), 1: sin, 2: cos -- in the current mode. */
double orc_eval(int fn, double x);

/* Worker threads used by the loops (OpenMP); 0 = library default. */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
