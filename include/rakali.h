/*
 * rakali.h -- the reference's handle API ("rki_*", librakali.so) for the
 * structured 2-D barotropic path, implemented over the MI355X step
 * (rakali_amd.h).  Signatures, argument meaning, return codes and the
 * host/device state machine follow src/api/rki_api.F90; each declaration
 * names the reference entry point it replaces.
 *
 * Return codes (rki_api.F90:22-27): 0 ok; -1 not initialised / not
 * finalised; -2 shape mismatch, already finalised or out of range; -3
 * multilayer inactive or restart mismatch; -100 invalid handle.  Two codes
 * are additions: -4 a HIP runtime error, -5 a configuration this build
 * does not implement (multilayer, non-hydrostatic, nested edges, the
 * semi-implicit scheme, NetCDF bathymetry files).  Where the reference ends the process with
 * ERROR STOP (an invalid namelist), these functions return -2 and
 * rki_last_error() describes the problem.
 *
 * Field arrays are Fortran-ordered FP64 with ghost cells,
 * (nx_total, ny_total) = (nx + 2 ng, ny + 2 ng), x fastest; the pointers
 * returned by the getters stay owned by the handle.
 */
#ifndef RAKALI_H
#define RAKALI_H

#ifdef __cplusplus
extern "C" {
#endif

/* rki_api.F90:157 -- bytes of the working real kind (8). */
int rki_working_precision(void);

/* rki_api.F90:169 -- phase 1 from a namelist file. */
int rki_create(const char *filename, int len, void **handle);
/* rki_api.F90:219 -- phase 2: bed ghosts, initial condition, device upload. */
int rki_finalize(void *handle);
/* rki_api.F90:261 -- rki_create + rki_finalize. */
int rki_create_from_namelist(const char *filename, int len, void **handle);
/* rki_api.F90:274 -- phase 1 from an in-memory namelist. */
int rki_create_from_string(const char *text, int len, void **handle);
/* rki_api.F90:315 -- rki_create_from_string + rki_finalize. */
int rki_create_from_namelist_string(const char *text, int len, void **handle);
/* rki_api.F90:328 -- frees the handle and nulls *handle; idempotent. */
void rki_destroy(void **handle);

/* rki_api.F90:358 -- n steps, each bounded by t_end - t. */
int rki_step(void *handle, int n_steps);
/* rki_api.F90:385 -- steps while t < target - 1e-10. */
int rki_advance(void *handle, double target_time);

/* rki_api.F90:416-498 -- queries (0 on an invalid handle). */
double rki_get_time(void *handle);
double rki_get_dt(void *handle);
void rki_get_grid_info(void *handle, int *nx, int *ny, int *nghost, double *dx, double *dy);
int rki_get_step_count(void *handle);
int rki_get_nz_ml(void *handle);

/* rki_api.F90:502-686 -- zero-copy host views with ghosts (device -> host
 * first if the device holds the newer state). */
int rki_get_h_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
int rki_get_hu_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
int rki_get_hv_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
int rki_get_b_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
int rki_get_h0_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
int rki_get_hu0_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
int rki_get_hv0_ptr(void *handle, double **ptr, int *nx_total, int *ny_total);
/* rki_api.F90:690-830 -- multilayer views: -1 (multilayer is inactive). */
int rki_get_h_layer_ptr(void *handle, double **ptr, int *nx_total, int *ny_total, int *nz);
int rki_get_hu_layer_ptr(void *handle, double **ptr, int *nx_total, int *ny_total, int *nz);
int rki_get_hv_layer_ptr(void *handle, double **ptr, int *nx_total, int *ny_total, int *nz);
int rki_get_salinity_ptr(void *handle, double **ptr, int *nx_total, int *ny_total, int *nz);
int rki_get_temperature_ptr(void *handle, double **ptr, int *nx_total, int *ny_total, int *nz);

/* rki_api.F90:836,865 -- interior (nx, ny) Fortran-ordered fields. */
int rki_set_bathymetry(void *handle, const double *b, int nx, int ny);
int rki_set_initial_h(void *handle, const double *h, int nx, int ny);
/* rki_api.F90:1040-1123 -- mid-run state setters. */
int rki_set_h(void *handle, const double *h, int nx, int ny);
int rki_set_hu(void *handle, const double *hu, int nx, int ny);
int rki_set_hv(void *handle, const double *hv, int nx, int ny);
/* rki_api.F90:896-1034 -- surface forcing fields, interior (nx, ny).  Wind
 * stress (Pa), atmospheric pressure (Pa) and precipitation (m/s) switch
 * the forcing on (state%forcing%active) and enter every stage; until the
 * wind is set it is the namelist's uniform wind_stress_x/y.  Heat flux and
 * E-P are stored: they feed only the multilayer path, which this build
 * does not run. */
int rki_set_wind(void *handle, const double *wx, const double *wy, int nx, int ny);
int rki_set_atm_pressure(void *handle, const double *p, int nx, int ny);
int rki_set_precipitation(void *handle, const double *p, int nx, int ny);
int rki_set_heat_flux(void *handle, const double *q, int nx, int ny);
int rki_set_evap_minus_precip(void *handle, const double *ep, int nx, int ny);
/* rki_api.F90:1125-1207 -- multilayer tracers: -3 (multilayer inactive). */
int rki_set_salinity(void *handle, const double *s, int nx, int ny, int nz);
int rki_set_temperature(void *handle, const double *t, int nx, int ny, int nz);

/* rki_api.F90:1209-1337 -- boundary scalars and physics queries. */
int rki_set_inflow_velocity(void *handle, double v);
int rki_set_inflow_salinity(void *handle, double s);
int rki_set_inflow_temperature(void *handle, double t);
int rki_set_discharge_flux(void *handle, double q);
int rki_set_clamped_elevation(void *handle, double eta);
double rki_get_clamped_elevation(void *handle);
double rki_get_manning_n(void *handle);
double rki_get_coriolis_f(void *handle);
/* rki_api.F90:1339 -- n <= 10 constituents. */
int rki_set_tidal_constituents(void *handle, const double *amp, const double *phase,
                               const double *omega, int n);

/* rki_api.F90:1376,1390 -- device -> host, keeping / dropping the device copy. */
void rki_sync_host(void *handle);
void rki_sync_from_device(void *handle);

/* rki_api.F90:1410,1436 -- restart files in restart_write's layout
 * (rki_output.F90:740-848): classic 64-bit-offset NetCDF with dimensions
 * x, y, FP64 h, hu, hv, b on (y, x) (interior cells), a scalar time and the
 * reference's units and global attributes, written and read without the
 * NetCDF library.  Load also accepts the binary "RKRS" files of earlier
 * versions of this library.  Load is a phase-1 call (-2 after finalize)
 * and returns -3 on a missing or unreadable file or a grid mismatch. */
int rki_save_restart(void *handle, const char *filename, int len);
int rki_load_restart(void *handle, const char *filename, int len);

/* rki_api.F90:1469,1491 -- solver switches. */
int rki_set_use_ppm_tracer(void *handle, int flag);
int rki_set_use_hllc(void *handle, int flag);

/* Additions with no counterpart in the reference. */
/* Text of the last error of this thread (namelist errors, HIP errors). */
const char *rki_last_error(void);
/* The kernel-level handle (rkg_grid*, rakali_amd.h) of a finalised
 * simulation, for device-side access; null before rki_finalize. */
void *rki_get_device_grid(void *handle);

#ifdef __cplusplus
}
#endif

#endif /* RAKALI_H */
