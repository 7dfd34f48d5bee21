// rkg_internal.hpp -- error reporting shared by the translation units of
// librakali_amd.so (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/rakali_amd.h"
#include "swe_kernels.hpp"

namespace rk {

// Records the message returned by rkg_last_error() on this thread.
__attribute__((visibility("hidden"), format(printf, 1, 2))) void set_error(const char* fmt, ...);

}  // namespace rk

namespace rk {
struct GaugeState;   // gauge stations and their time series (rkg_gauges.hip)
}

// One grid: device buffers, stream and control slots (rkg_grid of the ABI).
struct rkg_grid {
    unsigned magic;
    rkg_config cfg;
    int device;
    hipStream_t stream;
    bool own_stream;     // the stream was made by rkg_create (this grid holds a reference to it)
    int refs;            // 1 for the caller's handle + 1 per attached adjoint session
    long long pitch;
    size_t plane;        // elements per field plane
    double* buf;         // 7 planes: A(h,hu,hv) B(h,hu,hv) bed
    double* save;        // 3 planes: U^n of the last step (RK save), or null
    double* nfield;      // per-cell Manning n (1 plane), or null
    double* cbuf;        // 3 planes: third stage buffer of rk3 / rk43, or null
    double* mpart;       // per-tile sums of both stages (mass budget): flux_h, then the sponge's
                         // change of h (2 x mpart_cap each), or null
    long long mpart_cap; // tiles per stage that mpart holds
    int scheme;          // RKG_TIME_*
    double* frc;         // 6 planes: wind stress x, y, atmospheric pressure, precipitation, and
                         // the pressure gradient x, y (from the pressure plane), or null
    bool forcing;        // forcing%active: the stage kernels apply the planes in frc
    bool frc_grad_stale; // the pressure plane may have changed since the gradient was taken
    bool frc_pres_out;   // a device pointer to the pressure plane was handed out
    int sponge_width;    // sponge relaxation next to RKG_BC_SPONGE edges (bc_params_t)
    double sponge_strength, sponge_h_ref;
    rk::CtrlSlot* ctrl;      // device [2]
    rk::CtrlSlot* ctrl_host; // pinned [2]
    double* mass_buf;    // device [MASS_BLOCKS + 2]: partial sums, total mass, precipitation sum
    double* mass_host;   // pinned [1]
    int parity;
    int ty;              // rows per stage tile, 0 = chosen from the grid size
    int fast_mode;       // fast path: 0 by the launch's tile count, 1 always, -1 never
    bool speed_stale;
    bool prof;             // time each stage launch with events
    hipEvent_t ev[3];
    double prof_ms[2];     // accumulated stage-1 / stage-2 kernel time
    long long prof_n;
    rk::GaugeState* gauges;  // stations, or null
};

namespace rk {

constexpr unsigned GRID_MAGIC_LIVE = 0x72616B47u;  // 'rakG', rkg_grid::magic of a live grid
inline bool grid_valid(const rkg_grid* g) { return g != nullptr && g->magic == GRID_MAGIC_LIVE; }

// Holds / drops a reference to the grid's device state: an adjoint session
// retains its grid, so destroying the grid first leaves the session's
// buffers valid until the session is destroyed too (the grid is freed with
// its last reference).
void grid_retain(rkg_grid* g);
void grid_release(rkg_grid* g);

// The checks rkg_create applies to a configuration (sizes, spacing, tags);
// RKG_OK or an error code with the message set.
int validate_config(const rkg_config* c);

inline double* plane_ptr(rkg_grid* g, int k) { return g->buf + (size_t)k * g->plane; }

// Gauge recording around the step launches (rkg_gauges.hip): the steps one
// launch may take before the device series must be drained, the drain, the
// record after a step (slot p holds the state after it), and the release.
// Station (x, y) in metres: false when it lies outside the domain
// (compute_local_gauges, rki_gauges.F90:208-243); else its lower-left
// interpolation cell (c0, j0, 0-based interior) and weights
// (compute_interp_weights, :245-336).
bool station_weights(const rkg_config& c, double x, double y, int* c0, int* j0, double* wx, double* wy);

long long gauges_room(const rkg_grid* g);
int gauges_drain(rkg_grid* g);
int gauges_after_step(rkg_grid* g, int p);
void gauges_free(rkg_grid* g);

inline double* field_ptr(rkg_grid* g, int field) {
    switch (field) {
    case RKG_FIELD_H: return plane_ptr(g, 0);
    case RKG_FIELD_HU: return plane_ptr(g, 1);
    case RKG_FIELD_HV: return plane_ptr(g, 2);
    case RKG_FIELD_B: return plane_ptr(g, 6);
    case RKG_FIELD_H0: return g->save ? g->save : nullptr;
    case RKG_FIELD_HU0: return g->save ? g->save + g->plane : nullptr;
    case RKG_FIELD_HV0: return g->save ? g->save + 2 * g->plane : nullptr;
    case RKG_FIELD_HS: return plane_ptr(g, 3);
    case RKG_FIELD_HUS: return plane_ptr(g, 4);
    case RKG_FIELD_HVS: return plane_ptr(g, 5);
    case RKG_FIELD_N: return g->nfield;
    case RKG_FIELD_WIND_X: return g->frc;
    case RKG_FIELD_WIND_Y: return g->frc ? g->frc + g->plane : nullptr;
    case RKG_FIELD_PRESSURE: return g->frc ? g->frc + 2 * g->plane : nullptr;
    case RKG_FIELD_PRECIP: return g->frc ? g->frc + 3 * g->plane : nullptr;
    case RKG_FIELD_BC_H:
    case RKG_FIELD_BC_HU:
    case RKG_FIELD_BC_HV:
    case RKG_FIELD_BC_PREV_H: {
        // the last step's last boundary call: its state and the previous
        // call's (rki_solver.F90:760-815 for SSP-RK2, :861-925 for the
        // generic stage loop); stage buffers B = planes 3-5, C = cbuf
        double* B = plane_ptr(g, 3);
        double* cur = nullptr;
        double* prev = nullptr;
        switch (g->scheme) {
        case RKG_TIME_EULER: cur = g->save; prev = g->save; break;   // U^n, after the pre-dt call on U^n
        case RKG_TIME_RK3: cur = g->cbuf; prev = B; break;           // stage 3 input, stage 2 input
        case RKG_TIME_RK43: cur = B; prev = g->cbuf; break;          // stage 4 input, stage 3 input
        default: cur = B; prev = g->save; break;                     // U*, U^n
        }
        if (!cur || !prev) return nullptr;
        if (field == RKG_FIELD_BC_PREV_H) return prev;
        return cur + (size_t)(field - RKG_FIELD_BC_H) * g->plane;
    }
    default: return nullptr;
    }
}

// Rows per stage tile: the set value, or (0) 16 rows when the grid has
// at least 8192 tiles of 16 rows (several waves of tiles per wave slot, so
// longer tiles waste less on halos), else 8 (smaller grids run one tile
// per slot, and shorter tiles finish sooner).
// Rows per stage tile by grid size: 16 once a grid has 8192 tiles of 16
// rows, else 8 (a small grid needs more, shorter tiles to fill the chip).
inline int default_tile_rows(int nx, int ny) {
    const long long strips = (nx + 63) / 64;
    return strips * ((ny + 15) / 16) >= 8192 ? 16 : 8;
}

inline int tile_rows_of(const rkg_grid* g) {
    if (g->ty > 0) return g->ty;
    return default_tile_rows(g->cfg.nx, g->cfg.ny);
}

// apply_sponge does something (rki_kernel_sponge.F90:42-47 and a sponge edge).
inline bool sponge_on(const rkg_grid* g) {
    if (g->sponge_width <= 0 || !(g->sponge_strength > 0.0)) return false;
    for (int s = 0; s < 4; ++s)
        if (g->cfg.bc[s] == RKG_BC_SPONGE) return true;
    return false;
}

// The configuration-derived part of a descriptor for a grid of row pitch
// `pitch` and `ty` rows per tile; the buffer pointers are left null.
inline GridDesc desc_of_config(const rkg_config& c, long long pitch, int ty) {
    GridDesc d;
    memset(&d, 0, sizeof d);
    d.nx = c.nx; d.ny = c.ny; d.pitch = pitch;
    d.dx = c.dx; d.dy = c.dy;
    d.plane = pitch * (long long)c.ny;
    d.g = c.g; d.cfl = c.cfl; d.manning_n = c.manning_n; d.coriolis_f = c.coriolis_f;
    d.use_hllc = c.use_hllc;
    for (int s = 0; s < 4; ++s) d.bc[s] = c.bc[s];
    d.tidal_amplitude = c.tidal_amplitude; d.tidal_period = c.tidal_period;
    d.n_tc = c.n_tidal_constituents;
    for (int n = 0; n < RKG_MAX_TIDAL_CONSTITUENTS; ++n) {
        d.tc_amp[n] = c.tidal_amp[n]; d.tc_phase[n] = c.tidal_phase[n]; d.tc_omega[n] = c.tidal_omega[n];
    }
    d.inflow_velocity = c.inflow_velocity; d.discharge_flux = c.discharge_flux;
    d.clamped_elevation = c.clamped_elevation;
    d.dt_fixed = c.dt_fixed; d.t_end = c.t_end;
    d.cfl_interval = c.cfl_interval;
    d.ty = ty;
    d.nstrips = (c.nx + WAVE - 1) / WAVE;
    d.nsegs = (c.ny + d.ty - 1) / d.ty;
    d.fast = (long long)d.nstrips * d.nsegs >= FAST_MIN_TILES ? 1 : 0;
    d.ssp_coeff = 1.0;
    return d;
}

inline GridDesc make_desc(rkg_grid* g) {
    GridDesc d = desc_of_config(g->cfg, g->pitch, tile_rows_of(g));
    d.ah = plane_ptr(g, 0); d.ahu = plane_ptr(g, 1); d.ahv = plane_ptr(g, 2);
    d.bh = plane_ptr(g, 3); d.bhu = plane_ptr(g, 4); d.bhv = plane_ptr(g, 5);
    d.bed = plane_ptr(g, 6);
    d.o2h = d.ah; d.o2hu = d.ahu; d.o2hv = d.ahv;
    d.nfield = g->nfield;
    d.rec = nullptr;
    d.sh = g->save;
    d.shu = g->save ? g->save + g->plane : nullptr;
    d.shv = g->save ? g->save + 2 * g->plane : nullptr;
    d.ctrl = g->ctrl;
    d.mpart = g->mpart;
    if (sponge_on(g)) {
        d.sponge_width = g->sponge_width;
        d.sponge_strength = g->sponge_strength;
        d.sponge_h_ref = g->sponge_h_ref;
        d.spart = g->mpart ? g->mpart + 2 * g->mpart_cap : nullptr;
    }
    if (g->forcing) {
        d.fwx = g->frc;
        d.fwy = g->frc + g->plane;
        d.frain = g->frc + 3 * g->plane;
        d.fdpx = g->frc + 4 * g->plane;
        d.fdpy = g->frc + 5 * g->plane;
    }
    d.ssp_coeff = g->scheme == RKG_TIME_RK43 ? 2.0 : 1.0;  // rki_timestepper.F90:60-100
    if (g->fast_mode != 0) d.fast = g->fast_mode > 0 ? 1 : 0;
    return d;
}


}  // namespace rk

#define HIP_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            rk::set_error("%s failed: %s", #expr, hipGetErrorString(e_));      \
            return RKG_ERR_HIP;                                                \
        }                                                                      \
    } while (0)
