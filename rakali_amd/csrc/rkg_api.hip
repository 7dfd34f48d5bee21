// rkg_api.hip -- stage kernels, CFL/mass reductions and the rkg_* C ABI.
//
// One SSP-RK2 step is two launches:
//   stage 1: dt from the CFL speed of U^n (carried in the control slot),
//            U* = clampfriction(U^n + dt L(U^n))                -> B
//   stage 2: U^{n+1} = clamp(0.5 (U^n + clampfriction(U* + dt L(U*)))) -> A,
//            plus the max CFL speed of U^{n+1} for the next step's dt
//            (wave max + one atomic per workgroup).
// This is the data flow of rki_solver.F90:718-815 with the boundary fill
// folded into the tile loader and compute_timestep folded into the
// stage-2 epilogue.  dt and t never leave the device between steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/rakali_amd.h"
#include "rkg_internal.hpp"
#include "stage_kernels.hpp"
#include "stage_launch.hpp"

using namespace rk;

namespace {
thread_local std::string g_last_error;
}

void rk::set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

namespace {

constexpr unsigned GRID_MAGIC = GRID_MAGIC_LIVE;
constexpr unsigned GRID_CLOSED = 0x72616B43u;  // 'rakC': destroyed by its caller, still held by a session
constexpr unsigned GRID_DEAD = 0xDEADBEEFu;

// Streams that rkg_create made, with the number of grids launching on each.
// A grid created with such a stream (from rkg_get_stream of another grid)
// shares its ownership, so the grids may be destroyed in any order and the
// stream is destroyed with the last of them.  Streams the caller made are
// not in the table and are never destroyed by the library.
std::mutex g_stream_mu;
std::unordered_map<hipStream_t, int> g_stream_refs;

bool stream_retain(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    auto it = g_stream_refs.find(s);
    if (it == g_stream_refs.end()) return false;
    ++it->second;
    return true;
}

void stream_register(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    g_stream_refs[s] = 1;
}

// True when the caller held the last reference and must destroy the stream.
bool stream_release(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_stream_mu);
    auto it = g_stream_refs.find(s);
    if (it == g_stream_refs.end()) return false;
    if (--it->second > 0) return false;
    g_stream_refs.erase(it);
    return true;
}

// Plain copy, 16 bytes per lane per iteration: the bandwidth reference
// for the stage kernels and the calibration of the HBM traffic counters.
__global__ __launch_bounds__(256) void k_copy(const double2* __restrict__ src, double2* __restrict__ dst,
                                              long long n2) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n2; i += (long long)gridDim.x * 256)
        dst[i] = src[i];
}

__global__ void k_math_eval(int fn, const double* x, double* y, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i] = fn == 0 ? crm::pow13(x[i]) : (fn == 1 ? crm::sin(x[i]) : crm::cos(x[i]));
}

}  // namespace


namespace {

bool valid(const rkg_grid* g) { return g != nullptr && g->magic == GRID_MAGIC; }

int refresh_speed(rkg_grid* g) {
    GridDesc d = make_desc(g);
    const CtrlSlot& cur = g->ctrl_host[g->parity];
    (void)cur;
    // keep t / dt / step_count of the current slot, clear its speed max
    HIP_TRY(hipMemcpyAsync(g->ctrl_host, g->ctrl, 2 * sizeof(CtrlSlot), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    const CtrlSlot c = g->ctrl_host[g->parity];
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, g->stream, g->ctrl, g->parity, c.t, c.dt,
                       c.step_count, c.mass_out, c.mass_src);
    const long long n = (long long)g->cfg.nx * g->cfg.ny;
    int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_init_speed, dim3(blocks), dim3(256), 0, g->stream, d, g->parity);
    HIP_TRY(hipGetLastError());
    g->speed_stale = false;
    return RKG_OK;
}

// Source terms a descriptor asks the stage kernels for.
int feat_of(const GridDesc& d) {
    bool chapman = false;
    for (int s = 0; s < 4; ++s) chapman = chapman || d.bc[s] == RKG_BC_CHAPMAN;
    return (d.fwx ? FEAT_FORCING : 0) | (d.sponge_width > 0 ? FEAT_SPONGE : 0) | (chapman ? FEAT_CHAPMAN : 0);
}

template <int STAGE>
void launch_stage(rkg_grid* g, const GridDesc& d, int ntiles, int mode, double tlim) {
    const bool hllc = g->cfg.use_hllc != 0, nf = d.nfield != nullptr;
    void (*const table[8])(int, hipStream_t, const GridDesc&, int, int, int, double, bool, bool) = {
        launch_fused_stage<0>, launch_fused_stage<1>, launch_fused_stage<2>, launch_fused_stage<3>,
        launch_fused_stage<4>, launch_fused_stage<5>, launch_fused_stage<6>, launch_fused_stage<7>};
    table[feat_of(d)](STAGE, g->stream, d, ntiles, g->parity, mode, tlim, hllc, nf);
}

// Shu-Osher table of a generic scheme, make_timestepper
// (rki_timestepper.F90:60-100); ssp_coeff is in make_desc.
struct SchemeTab {
    int n;
    double alpha[4], beta[4], c[4];
};

SchemeTab scheme_tab(int scheme) {
    SchemeTab t{};
    if (scheme == RKG_TIME_EULER) {
        t.n = 1; t.alpha[0] = 0.0; t.beta[0] = 1.0; t.c[0] = 0.0;
    } else if (scheme == RKG_TIME_RK3) {
        t.n = 3;
        t.alpha[0] = 0.0; t.alpha[1] = 0.75; t.alpha[2] = 1.0 / 3.0;
        t.beta[0] = 1.0; t.beta[1] = 1.0; t.beta[2] = 1.0;
        t.c[0] = 0.0; t.c[1] = 1.0; t.c[2] = 0.5;
    } else {  // RKG_TIME_RK43
        t.n = 4;
        t.alpha[0] = 0.0; t.alpha[1] = 0.0; t.alpha[2] = 2.0 / 3.0; t.alpha[3] = 0.0;
        t.beta[0] = 0.5; t.beta[1] = 0.5; t.beta[2] = 0.5; t.beta[3] = 0.5;
        t.c[0] = 0.0; t.c[1] = 0.5; t.c[2] = 1.0; t.c[3] = 0.5;
    }
    return t;
}

void launch_gen(rkg_grid* g, const GridDesc& d, const GenStage& gs, int ntiles, int mode, double tlim) {
    const bool hllc = g->cfg.use_hllc != 0, nf = d.nfield != nullptr;
    void (*const table[8])(hipStream_t, const GridDesc&, const GenStage&, int, int, int, double, bool, bool) = {
        launch_generic_stage<0>, launch_generic_stage<1>, launch_generic_stage<2>, launch_generic_stage<3>,
        launch_generic_stage<4>, launch_generic_stage<5>, launch_generic_stage<6>, launch_generic_stage<7>};
    table[feat_of(d)](g->stream, d, gs, ntiles, g->parity, mode, tlim, hllc, nf);
}

// Generic explicit SSP steps (euler / rk3 / rk43), solver_advance_2d_generic
// (rki_solver.F90:861-925).  Stage i reads the previous iterate and writes
// B or C in turn; the last stage writes A, which holds U^n until then.  A
// one-stage scheme writes B and k_euler_finish copies it to A.
int launch_generic(rkg_grid* g, int n, int mode, double tlim) {
    GridDesc d = make_desc(g);
    d.mpart = nullptr;  // the generic path keeps no mass budget (as the reference)
    d.spart = nullptr;
    const int ntiles = d.nstrips * d.nsegs;
    const SchemeTab T = scheme_tab(g->scheme);
    double* A = plane_ptr(g, 0);
    double* B = plane_ptr(g, 3);
    double* C = g->cbuf;
    const size_t P = g->plane;
    for (int k = 0; k < n; ++k) {
        const double* in = A;
        const double* prev = nullptr;
        for (int st = 0; st < T.n; ++st) {
            double* out = (T.n > 1 && st == T.n - 1) ? A : ((st % 2 == 0) ? B : C);
            GenStage gs;
            gs.ph = prev;   // the previous boundary call's state (Chapman eta_old)
            gs.ih = in; gs.ihu = in + P; gs.ihv = in + 2 * P;
            gs.oh = out; gs.ohu = out + P; gs.ohv = out + 2 * P;
            gs.alpha = T.alpha[st]; gs.beta = T.beta[st]; gs.c = T.c[st];
            gs.c_last = T.c[T.n - 1];
            gs.first = st == 0;
            gs.last = T.n > 1 && st == T.n - 1;
            launch_gen(g, d, gs, ntiles, mode, tlim);
            prev = in;
            in = out;
        }
        if (T.n == 1) {
            const long long cells = (long long)d.nx * d.ny;
            const int blocks = (int)std::min<long long>((cells + 255) / 256, 4096);
            hipLaunchKernelGGL(k_euler_finish, dim3(blocks), dim3(256), 0, g->stream, d, g->parity);
        }
        g->parity ^= 1;
        if (g->gauges) {
            int rc = gauges_after_step(g, g->parity);
            if (rc) return rc;
        }
    }
    HIP_TRY(hipGetLastError());
    return RKG_OK;
}

int launch_steps_now(rkg_grid* g, int n, int mode, double tlim);

// Steps in pieces that fit the gauge series' device buffer, drained between
// pieces (each step records at most one gauge record).  n = 0 still makes
// the control slot's CFL speed current (callers rely on rkg_step(g, 0)).
int launch_steps(rkg_grid* g, int n, int mode, double tlim) {
    if (g->speed_stale) {
        int rc = refresh_speed(g);
        if (rc) return rc;
    }
    while (n > 0) {
        long long room = gauges_room(g);
        if (room == 0) {
            int rc = gauges_drain(g);
            if (rc) return rc;
            room = gauges_room(g);
        }
        const int piece = (int)std::min<long long>(n, room);
        int rc = launch_steps_now(g, piece, mode, tlim);
        if (rc) return rc;
        n -= piece;
    }
    return RKG_OK;
}

int launch_steps_now(rkg_grid* g, int n, int mode, double tlim) {
    if (g->speed_stale) {
        int rc = refresh_speed(g);
        if (rc) return rc;
    }
    if (g->forcing && (g->frc_grad_stale || g->frc_pres_out)) {
        const long long cells = (long long)g->cfg.nx * g->cfg.ny;
        const int blocks = (int)std::min<long long>((cells + 255) / 256, 4096);
        hipLaunchKernelGGL(k_pressure_gradient, dim3(blocks), dim3(256), 0, g->stream, g->frc + 2 * g->plane,
                           g->frc + 4 * g->plane, g->frc + 5 * g->plane, g->cfg.nx, g->cfg.ny, (long long)g->pitch,
                           g->cfg.dx, g->cfg.dy);
        HIP_TRY(hipGetLastError());
        g->frc_grad_stale = false;
    }
    if (g->scheme != RKG_TIME_RK2) return launch_generic(g, n, mode, tlim);
    GridDesc d = make_desc(g);
    const int ntiles = d.nstrips * d.nsegs;
    const double area = d.dx * d.dy;
    // Interior sum of the precipitation rate for the surface mass source
    // (fixed over the call: the planes change only between calls).
    const double* rain = nullptr;
    if (d.mpart && d.frain) {
        const int blocks = std::min(MASS_BLOCKS, g->cfg.ny);
        hipLaunchKernelGGL(k_plane_sum_partial, dim3(blocks), dim3(256), 0, g->stream, d.frain, d.pitch, d.nx, d.ny,
                           g->mass_buf);
        hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, g->stream, g->mass_buf, blocks,
                           g->mass_buf + MASS_BLOCKS + 1);
        rain = g->mass_buf + MASS_BLOCKS + 1;
    }
    for (int k = 0; k < n; ++k) {
        if (g->prof) HIP_TRY(hipEventRecord(g->ev[0], g->stream));
        launch_stage<1>(g, d, ntiles, mode, tlim);
        if (g->prof) HIP_TRY(hipEventRecord(g->ev[1], g->stream));
        launch_stage<2>(g, d, ntiles, mode, tlim);
        if (d.mpart)
            hipLaunchKernelGGL(k_budget, dim3(1), dim3(256), 0, g->stream, g->ctrl, g->parity, d.mpart, ntiles, area,
                               rain, d.spart);
        g->parity ^= 1;
        if (g->gauges) {
            int rc = gauges_after_step(g, g->parity);
            if (rc) return rc;
        }
        if (g->prof) {
            float a = 0.f, b = 0.f;
            HIP_TRY(hipEventRecord(g->ev[2], g->stream));
            HIP_TRY(hipEventSynchronize(g->ev[2]));
            HIP_TRY(hipEventElapsedTime(&a, g->ev[0], g->ev[1]));
            HIP_TRY(hipEventElapsedTime(&b, g->ev[1], g->ev[2]));
            g->prof_ms[0] += a;
            g->prof_ms[1] += b;
            g->prof_n += 1;
        }
    }
    HIP_TRY(hipGetLastError());
    return RKG_OK;
}

int read_clock(rkg_grid* g, CtrlSlot* out) {
    HIP_TRY(hipMemcpyAsync(g->ctrl_host, g->ctrl, 2 * sizeof(CtrlSlot), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    *out = g->ctrl_host[g->parity];
    return RKG_OK;
}

int check_config(const rkg_config* c) {
    if (c->nx < 2 || c->ny < 2) { set_error("nx and ny must be >= 2 (got %d, %d)", c->nx, c->ny); return RKG_ERR_RANGE; }
    if (!(c->dx >= 1e-6 && c->dx <= 1e9) || !(c->dy >= 1e-6 && c->dy <= 1e9)) {
        set_error("dx and dy must lie in [1e-6, 1e9]");
        return RKG_ERR_RANGE;
    }
    if (c->cfl_interval < 1) { set_error("cfl_interval must be >= 1"); return RKG_ERR_RANGE; }
    if (c->n_tidal_constituents < 0 || c->n_tidal_constituents > RKG_MAX_TIDAL_CONSTITUENTS) {
        set_error("n_tidal_constituents out of range"); return RKG_ERR_RANGE;
    }
    for (int s = 0; s < 4; ++s) {
        switch (c->bc[s]) {
        case RKG_BC_WALL: case RKG_BC_OPEN: case RKG_BC_TIDAL: case RKG_BC_INFLOW:
        case RKG_BC_DISCHARGE: case RKG_BC_CLAMPED: case RKG_BC_SPONGE: case RKG_BC_CHAPMAN: break;
        default:
            set_error("boundary tag %d on edge %d is not implemented (nested)", c->bc[s], s);
            return RKG_ERR_UNSUPPORTED;
        }
    }
    return RKG_OK;
}

}  // namespace

int rk::validate_config(const rkg_config* c) { return check_config(c); }

void rk::grid_retain(rkg_grid* g) { ++g->refs; }

void rk::grid_release(rkg_grid* g) {
    if (--g->refs > 0) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->buf) (void)hipFree(g->buf);
    if (g->save) (void)hipFree(g->save);
    if (g->nfield) (void)hipFree(g->nfield);
    if (g->cbuf) (void)hipFree(g->cbuf);
    if (g->mpart) (void)hipFree(g->mpart);
    if (g->frc) (void)hipFree(g->frc);
    gauges_free(g);
    if (g->ctrl) (void)hipFree(g->ctrl);
    if (g->mass_buf) (void)hipFree(g->mass_buf);
    if (g->ctrl_host) (void)hipHostFree(g->ctrl_host);
    if (g->mass_host) (void)hipHostFree(g->mass_host);
    for (int k = 0; k < 3; ++k)
        if (g->ev[k]) (void)hipEventDestroy(g->ev[k]);
    if (g->own_stream && g->stream && stream_release(g->stream)) (void)hipStreamDestroy(g->stream);
    g->magic = GRID_DEAD;
    delete g;
}

extern "C" {

int rkg_config_default(rkg_config* c) {
    if (!c) return RKG_ERR_RANGE;
    memset(c, 0, sizeof *c);
    c->nx = 200; c->ny = 1;
    c->dx = 1.0; c->dy = 1.0;
    c->g = 9.80665;
    c->cfl = 0.45;
    c->t_end = 1.0;
    c->cfl_interval = 1;
    c->use_hllc = 0;
    for (int s = 0; s < 4; ++s) c->bc[s] = RKG_BC_WALL;
    c->tidal_period = 44712.0;
    return RKG_OK;
}

int rkg_create(const rkg_config* cfg, void* stream, rkg_grid** out) {
    if (!cfg || !out) { set_error("null argument"); return RKG_ERR_RANGE; }
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    rkg_grid* g = new rkg_grid();
    memset(g, 0, sizeof *g);
    g->cfg = *cfg;
    g->ty = 0;    // chosen from the grid size (tile_rows_of)
    g->scheme = RKG_TIME_RK2;
    HIP_TRY(hipGetDevice(&g->device));
    g->refs = 1;
    if (stream) {
        g->stream = (hipStream_t)stream;
        g->own_stream = stream_retain(g->stream);   // shared with the grid that made it
    } else {
        hipError_t e = hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { set_error("hipStreamCreate: %s", hipGetErrorString(e)); delete g; return RKG_ERR_HIP; }
        stream_register(g->stream);
        g->own_stream = true;
    }
    long long pitch = ((long long)cfg->nx + 15) / 16 * 16;
    if (pitch % 1024 == 0) pitch += 32;  // break power-of-two row strides
    g->pitch = pitch;
    g->plane = (size_t)pitch * (size_t)cfg->ny;
    hipError_t e = hipMalloc((void**)&g->buf, 7 * g->plane * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&g->ctrl, 2 * sizeof(CtrlSlot));
    if (e == hipSuccess) e = hipHostMalloc((void**)&g->ctrl_host, 2 * sizeof(CtrlSlot), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&g->mass_buf, (MASS_BLOCKS + 2) * sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void**)&g->mass_host, sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemsetAsync(g->buf, 0, 7 * g->plane * sizeof(double), g->stream);
    if (e == hipSuccess) e = hipMemsetAsync(g->ctrl, 0, 2 * sizeof(CtrlSlot), g->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) {
        set_error("device allocation failed: %s", hipGetErrorString(e));
        g->magic = GRID_MAGIC;
        rkg_destroy(&g);
        return RKG_ERR_HIP;
    }
    memset(g->ctrl_host, 0, 2 * sizeof(CtrlSlot));
    g->parity = 0;
    g->speed_stale = true;
    g->magic = GRID_MAGIC;
    *out = g;
    return RKG_OK;
}

void rkg_destroy(rkg_grid** pg) {
    if (!pg || !*pg) return;
    rkg_grid* g = *pg;
    *pg = nullptr;
    if (g->magic != GRID_MAGIC) return;
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    // API calls on the handle fail from here on; an adjoint session still
    // attached keeps the device state until it is destroyed itself
    g->magic = GRID_CLOSED;
    grid_release(g);
}


static int copy_field(rkg_grid* g, int field, double* host_or_dev, int ld, bool upload, hipMemcpyKind kind) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    double* f = field_ptr(g, field);
    if (!f) { set_error("unknown or unallocated field %d", field); return RKG_ERR_RANGE; }
    if (upload && ((field > RKG_FIELD_B && field < RKG_FIELD_N) || field >= RKG_FIELD_BC_H)) {
        set_error("field %d is read-only", field);
        return RKG_ERR_RANGE;
    }
    if (!host_or_dev || ld < g->cfg.nx) { set_error("bad buffer or leading dimension"); return RKG_ERR_RANGE; }
    const size_t w = (size_t)g->cfg.nx * sizeof(double);
    if (upload) {
        HIP_TRY(hipMemcpy2DAsync(f, (size_t)g->pitch * sizeof(double), host_or_dev, (size_t)ld * sizeof(double), w,
                                 (size_t)g->cfg.ny, kind, g->stream));
        g->speed_stale = true;
        if (field == RKG_FIELD_PRESSURE) g->frc_grad_stale = true;
    } else {
        HIP_TRY(hipMemcpy2DAsync(host_or_dev, (size_t)ld * sizeof(double), f, (size_t)g->pitch * sizeof(double), w,
                                 (size_t)g->cfg.ny, kind, g->stream));
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    return RKG_OK;
}

int rkg_upload(rkg_grid* g, int field, const double* src, int ld) {
    return copy_field(g, field, const_cast<double*>(src), ld, true, hipMemcpyHostToDevice);
}
int rkg_download(rkg_grid* g, int field, double* dst, int ld) {
    return copy_field(g, field, dst, ld, false, hipMemcpyDeviceToHost);
}
int rkg_upload_device(rkg_grid* g, int field, const double* src, int ld) {
    return copy_field(g, field, const_cast<double*>(src), ld, true, hipMemcpyDeviceToDevice);
}
int rkg_download_device(rkg_grid* g, int field, double* dst, int ld) {
    return copy_field(g, field, dst, ld, false, hipMemcpyDeviceToDevice);
}

int rkg_field_device_ptr(rkg_grid* g, int field, double** ptr, long long* pitch) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    double* f = field_ptr(g, field);
    if (!f || !ptr || !pitch) return RKG_ERR_RANGE;
    // the caller may write the pressure at any time from now on: the
    // gradient planes are then refreshed before every launch
    if (field == RKG_FIELD_PRESSURE) g->frc_pres_out = true;
    *ptr = f;
    *pitch = g->pitch;
    return RKG_OK;
}

int rkg_step(rkg_grid* g, int n) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (n < 0) return RKG_ERR_RANGE;
    return launch_steps(g, n, MODE_STEP, g->cfg.t_end);
}

int rkg_advance(rkg_grid* g, double target, long long* steps_taken) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    CtrlSlot c;
    int rc = read_clock(g, &c);
    if (rc) return rc;
    const long long start = c.step_count;
    // Launch in chunks; steps past the target are no-ops on the device.
    int chunk = 8;
    while (c.t < target - 1.0e-10) {
        rc = launch_steps(g, chunk, MODE_ADVANCE, target);
        if (rc) return rc;
        rc = read_clock(g, &c);
        if (rc) return rc;
        if (c.t < target - 1.0e-10 && c.dt > 0.0) {
            const double need = (target - c.t) / c.dt + 2.0;
            chunk = (int)std::min(4096.0, std::max(8.0, need));
        } else {
            chunk = std::min(chunk * 2, 4096);
        }
    }
    if (steps_taken) *steps_taken = c.step_count - start;
    return RKG_OK;
}

int rkg_get_clock(rkg_grid* g, double* t, double* dt, long long* step_count) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    CtrlSlot c;
    int rc = read_clock(g, &c);
    if (rc) return rc;
    if (t) *t = c.t;
    if (dt) *dt = c.dt;
    if (step_count) *step_count = c.step_count;
    return RKG_OK;
}

int rkg_set_clock(rkg_grid* g, double t, long long step_count) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    CtrlSlot c;
    int rc = read_clock(g, &c);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, g->stream, g->ctrl, g->parity, t, c.dt, step_count,
                       c.mass_out, c.mass_src);
    HIP_TRY(hipGetLastError());
    g->speed_stale = true;
    return RKG_OK;
}

int rkg_total_mass(rkg_grid* g, double* mass) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!mass) return RKG_ERR_RANGE;
    GridDesc d = make_desc(g);
    const int blocks = std::min(MASS_BLOCKS, g->cfg.ny);
    hipLaunchKernelGGL(k_mass_partial, dim3(blocks), dim3(256), 0, g->stream, d, g->mass_buf);
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, g->stream, g->mass_buf, blocks,
                       g->mass_buf + MASS_BLOCKS);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(g->mass_host, g->mass_buf + MASS_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    *mass = *g->mass_host;
    return RKG_OK;
}

int rkg_sync(rkg_grid* g) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    HIP_TRY(hipStreamSynchronize(g->stream));
    return RKG_OK;
}

int rkg_set_tile_rows(rkg_grid* g, int rows) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (rows < 0 || rows > WAVE) { set_error("tile rows must be in 1..64, or 0 for the size-based choice"); return RKG_ERR_RANGE; }
    g->ty = rows;
    return RKG_OK;
}

int rkg_set_time_scheme(rkg_grid* g, int scheme) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (scheme < RKG_TIME_EULER || scheme > RKG_TIME_RK43) {
        set_error("time scheme must be 1 (euler), 2 (rk2), 3 (rk3) or 4 (rk43); got %d", scheme);
        return RKG_ERR_RANGE;
    }
    if ((scheme == RKG_TIME_RK3 || scheme == RKG_TIME_RK43) && !g->cbuf) {
        HIP_TRY(hipMalloc((void**)&g->cbuf, 3 * g->plane * sizeof(double)));
        HIP_TRY(hipMemsetAsync(g->cbuf, 0, 3 * g->plane * sizeof(double), g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    g->scheme = scheme;
    return RKG_OK;
}

int rkg_get_time_scheme(rkg_grid* g, int* scheme) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!scheme) return RKG_ERR_RANGE;
    *scheme = g->scheme;
    return RKG_OK;
}

int rkg_set_mass_budget(rkg_grid* g, int enable) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (enable && !g->mpart) {
        // room for the tiles of both stages at any tile height, for the
        // flux and the sponge sums
        const long long cap = (long long)((g->cfg.nx + WAVE - 1) / WAVE) * g->cfg.ny;
        HIP_TRY(hipMalloc((void**)&g->mpart, 4 * cap * sizeof(double)));
        g->mpart_cap = cap;
        CtrlSlot c;
        int rc = read_clock(g, &c);
        if (rc) return rc;
        hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, g->stream, g->ctrl, g->parity, c.t, c.dt,
                           c.step_count, 0.0, 0.0);
        HIP_TRY(hipGetLastError());
        g->speed_stale = true;
    } else if (!enable && g->mpart) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        HIP_TRY(hipFree(g->mpart));
        g->mpart = nullptr;
        g->mpart_cap = 0;
    }
    return RKG_OK;
}

// Fills a plane with one value.
__global__ void k_fill(double* p, long long n, double v) {
    for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}

int rkg_set_forcing(rkg_grid* g, int active) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (active && !g->frc) {
        HIP_TRY(hipMalloc((void**)&g->frc, 6 * g->plane * sizeof(double)));
        hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, g->stream, g->frc, (long long)(6 * g->plane), 0.0);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(g->stream));
    }
    g->forcing = active != 0;
    g->frc_grad_stale = true;
    return RKG_OK;
}

int rkg_set_sponge(rkg_grid* g, int width, double strength, double h_ref) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (width < 0 || !std::isfinite(strength) || !std::isfinite(h_ref)) {
        set_error("sponge width must be >= 0 and strength, h_ref finite");
        return RKG_ERR_RANGE;
    }
    g->sponge_width = width;
    g->sponge_strength = strength;
    g->sponge_h_ref = h_ref;
    return RKG_OK;
}

int rkg_get_sponge(rkg_grid* g, int* width, double* strength, double* h_ref) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!width || !strength || !h_ref) return RKG_ERR_RANGE;
    *width = g->sponge_width;
    *strength = g->sponge_strength;
    *h_ref = g->sponge_h_ref;
    return RKG_OK;
}

int rkg_get_forcing(rkg_grid* g, int* active) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!active) return RKG_ERR_RANGE;
    *active = g->forcing ? 1 : 0;
    return RKG_OK;
}

int rkg_get_last_bc(rkg_grid* g, double* time, double* dt) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!time || !dt) return RKG_ERR_RANGE;
    CtrlSlot c;
    int rc = read_clock(g, &c);
    if (rc) return rc;
    *time = c.t_bc;
    *dt = c.dt;
    return RKG_OK;
}

int rkg_get_mass_source(rkg_grid* g, double* mass_src) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!mass_src) return RKG_ERR_RANGE;
    CtrlSlot c;
    int rc = read_clock(g, &c);
    if (rc) return rc;
    *mass_src = c.mass_src;
    return RKG_OK;
}

int rkg_get_mass_budget(rkg_grid* g, double* mass_out) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!mass_out) return RKG_ERR_RANGE;
    CtrlSlot c;
    int rc = read_clock(g, &c);
    if (rc) return rc;
    *mass_out = c.mass_out;
    return RKG_OK;
}

int rkg_set_rk_save(rkg_grid* g, int enable) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (enable && !g->save) {
        HIP_TRY(hipMalloc((void**)&g->save, 3 * g->plane * sizeof(double)));
        HIP_TRY(hipMemsetAsync(g->save, 0, 3 * g->plane * sizeof(double), g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
    } else if (!enable && g->save) {
        HIP_TRY(hipStreamSynchronize(g->stream));
        HIP_TRY(hipFree(g->save));
        g->save = nullptr;
    }
    return RKG_OK;
}

int rkg_use_manning_field(rkg_grid* g, int enable) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (enable && !g->nfield) {
        HIP_TRY(hipMalloc((void**)&g->nfield, g->plane * sizeof(double)));
        hipLaunchKernelGGL(k_fill, dim3(1024), dim3(256), 0, g->stream, g->nfield, (long long)g->plane,
                           g->cfg.manning_n);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(g->stream));
    } else if (!enable && g->nfield) {
        HIP_TRY(hipFree(g->nfield));
        g->nfield = nullptr;
    }
    return RKG_OK;
}

int rkg_get_config(rkg_grid* g, rkg_config* cfg) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!cfg) return RKG_ERR_RANGE;
    *cfg = g->cfg;
    return RKG_OK;
}

int rkg_set_config(rkg_grid* g, const rkg_config* cfg) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!cfg) return RKG_ERR_RANGE;
    if (cfg->nx != g->cfg.nx || cfg->ny != g->cfg.ny) { set_error("nx and ny cannot change"); return RKG_ERR_RANGE; }
    int rc = check_config(cfg);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(g->stream));
    const bool speed = cfg->g != g->cfg.g || cfg->dx != g->cfg.dx || cfg->dy != g->cfg.dy;
    g->cfg = *cfg;
    if (speed) g->speed_stale = true;
    return RKG_OK;
}

int rkg_profile(rkg_grid* g, int enable) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    for (int k = 0; k < 3; ++k)
        if (!g->ev[k]) HIP_TRY(hipEventCreate(&g->ev[k]));
    g->prof = enable != 0;
    g->prof_ms[0] = g->prof_ms[1] = 0.0;
    g->prof_n = 0;
    return RKG_OK;
}

int rkg_profile_read(rkg_grid* g, double* stage1_ms, double* stage2_ms, long long* steps) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (stage1_ms) *stage1_ms = g->prof_ms[0];
    if (stage2_ms) *stage2_ms = g->prof_ms[1];
    if (steps) *steps = g->prof_n;
    return RKG_OK;
}

int rkg_get_stream(rkg_grid* g, void** stream) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!stream) return RKG_ERR_RANGE;
    *stream = (void*)g->stream;
    return RKG_OK;
}

int rkg_set_fast_path(rkg_grid* g, int mode) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (mode < -1 || mode > 1) { set_error("fast path mode must be -1 (never), 0 (by tile count) or 1 (always)"); return RKG_ERR_RANGE; }
    g->fast_mode = mode;
    return RKG_OK;
}

int rkg_get_tile_rows(rkg_grid* g, int* rows) {
    if (!valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!rows) return RKG_ERR_RANGE;
    *rows = tile_rows_of(g);
    return RKG_OK;
}

const char* rkg_last_error(void) { return g_last_error.c_str(); }

int rkg_copy_bandwidth(long long bytes, int reps, double* gbs, double* ms_per_copy) {
    if (bytes < 4096 || reps < 1 || !gbs) { set_error("bad arguments"); return RKG_ERR_RANGE; }
    const long long n2 = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipMalloc((void**)&a, n2 * sizeof(double2)));
    hipError_t e = hipMalloc((void**)&b, n2 * sizeof(double2));
    if (e == hipSuccess) e = hipMemset(a, 0, n2 * sizeof(double2));
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.f;
    if (e == hipSuccess) {
        const int blocks = 256 * 8 * 4;  // 8 waves per CU per XCD slot, grid-stride
        hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n2);  // warm-up
        e = hipEventRecord(e0, 0);
        for (int r = 0; r < reps && e == hipSuccess; ++r) {
            hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, 0, a, b, n2);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, 0);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) { set_error("rkg_copy_bandwidth: %s", hipGetErrorString(e)); return RKG_ERR_HIP; }
    const double per = (double)ms / reps;
    *gbs = 2.0 * (double)(n2 * 16) / (per * 1e-3) / 1e9;   // read + write
    if (ms_per_copy) *ms_per_copy = per;
    return RKG_OK;
}

int rkg_math_eval(int fn, const double* x, double* y, int n, int on_device) {
    if (fn < 0 || fn > 2 || n < 0 || (n > 0 && (!x || !y))) { set_error("bad arguments"); return RKG_ERR_RANGE; }
    if (!on_device) {
        for (int i = 0; i < n; ++i)
            y[i] = fn == 0 ? crm::pow13(x[i]) : (fn == 1 ? crm::sin(x[i]) : crm::cos(x[i]));
        return RKG_OK;
    }
    if (n == 0) return RKG_OK;
    double* d = nullptr;
    HIP_TRY(hipMalloc((void**)&d, 2 * (size_t)n * sizeof(double)));
    hipError_t e = hipMemcpy(d, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_math_eval, dim3((n + 255) / 256), dim3(256), 0, 0, fn, d, d + n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, d + n, (size_t)n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) { set_error("rkg_math_eval: %s", hipGetErrorString(e)); return RKG_ERR_HIP; }
    return RKG_OK;
}

}  // extern "C"
