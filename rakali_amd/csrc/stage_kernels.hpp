// stage_kernels.hpp -- the stage kernels and the small reductions shared by
// the single-grid ABI (rkg_api.hip) and the ensemble ABI
// (rkg_ensemble.hip).  Kernels have internal linkage: each translation unit
// that launches them holds its own copy.
#pragma once

#include "swe_kernels.hpp"

#ifndef RKG_WAVES_PER_EU
#define RKG_WAVES_PER_EU 3
#endif

namespace rk {
namespace {

constexpr int MODE_STEP = 0;
constexpr int MODE_ADVANCE = 1;
constexpr int MASS_BLOCKS = 512;

// dt of a step from the control slot of the state it starts from:
// rki_solver.F90:728-745 with rki_api.F90:377 (step) / :402 (advance).
__device__ __forceinline__ double step_dt(const GridDesc& d, const CtrlSlot& cp, double dt_max) {
    if (d.dt_fixed > 0.0) return dmin(d.dt_fixed, dt_max);
    if (cp.step_count % d.cfl_interval == 0 || cp.dt <= 0.0) {
        const double max_speed = __longlong_as_double((long long)cp.speed_bits);
        double dt;
        if (max_speed > 0.0)
            dt = d.cfl / max_speed;
        else
            dt = d.cfl * dmin(d.dx, d.dy) / sqrt(d.g * DRY_TOLERANCE);
        dt = dt * d.ssp_coeff;  // rki_solver.F90:736-741
        return dmin(dt, dt_max);
    }
    return cp.dt;
}

__device__ __forceinline__ SourceTerms make_src(const GridDesc& d, double dt) {
    SourceTerms s;
    s.g = d.g;
    s.n_sq = d.manning_n * d.manning_n;
    s.has_friction = d.manning_n > 0.0;
    s.has_coriolis = fabs(d.coriolis_f) > 1.0e-20;
    s.cosf = 0.0;
    s.sinf = 0.0;
    s.coriolis_f = d.coriolis_f;
    if (s.has_coriolis) {
        s.cosf = crm::cos(d.coriolis_f * dt);
        s.sinf = crm::sin(d.coriolis_f * dt);
    }
    return s;
}

__device__ __forceinline__ bool tile_is_interior(const GridDesc& d, int strip, int seg) {
    const int c0 = strip * WAVE;
    const int j0 = seg * d.ty;
    const int j1 = min(j0 + d.ty, d.ny);
    return c0 >= 2 && c0 + WAVE + 1 < d.nx && j0 >= 2 && j1 + 1 < d.ny;
}

// One tile of one stage of one grid.  `where(d)` gives the tile as
// (strip, segment, leader); the leader is the single tile per grid and
// launch that advances the control slot.  STAGE 1 and 2 are the fused
// SSP-RK2 stages; STAGE 3 is one stage `gs` of a generic SSP scheme.
template <int STAGE, bool HLLC, bool NF, bool REC, class Where, int FEAT = 0, bool FASTK = false>
__device__ __forceinline__ void stage_tile(const GridDesc& d, TileLds& L, int p, int mode, double tlim,
                                           Where where, const GenStage& gs = GenStage{}) {
    const int3 w = where(d);
    const bool leader = w.z != 0;
    StageCtx s;
    s.alpha = 0.0; s.a_prev = 1.0;
    s.blend = false; s.first = false; s.last = false;
    s.radiate = false; s.prev_h = nullptr; s.bc_dt = 0.0;
    s.eta_tidal = 0.0;
    double t_tide;   // bc%time of the stage; the tide is evaluated by edge tiles only
    if (STAGE == 1 || (STAGE == 3 && gs.first)) {
        const CtrlSlot cp = d.ctrl[p];
        const bool active = (mode == MODE_STEP) || (cp.t < tlim - 1.0e-10);
        const double dt_max = tlim - cp.t;
        const double dt = active ? step_dt(d, cp, dt_max) : cp.dt;
        if (leader && threadIdx.x == 0) {
            CtrlSlot* q = &d.ctrl[1 - p];
            q->t = active ? cp.t + dt : cp.t;
            q->dt = dt;
            q->step_count = active ? cp.step_count + 1 : cp.step_count;
            q->speed_bits = active ? 0ull : cp.speed_bits;
            q->mass_out = cp.mass_out;
            q->mass_src = cp.mass_src;
            q->active = active ? 1 : 0;
            // the last boundary call: stage 2 at t + dt, or the last generic
            // stage at t + c_last dt (rki_solver.F90:790-795, :895-897)
            if (!active) q->t_bc = cp.t_bc;
            else if (STAGE == 1) q->t_bc = cp.t + dt;
            else q->t_bc = cp.t + gs.c_last * dt;
            if (REC && d.hist != nullptr) {   // the step's clock for a reverse sweep (rki_solver.F90:731)
                d.hist[0] = cp.t;
                d.hist[1] = dt;
                d.hist[2] = (d.dt_fixed <= 0.0 && (cp.step_count % d.cfl_interval == 0 || cp.dt <= 0.0)) ? 1.0 : 0.0;
            }
        }
        if (!active) return;
        s.bc_dt = dt;   // bc%dt = state%dt, rki_solver.F90:748
        if (STAGE == 1) {
            s.ih = d.ah; s.ihu = d.ahu; s.ihv = d.ahv;
            s.oh = d.bh; s.ohu = d.bhu; s.ohv = d.bhv;
            s.dt = dt;
        } else {
            s.ih = gs.ih; s.ihu = gs.ihu; s.ihv = gs.ihv;
            s.oh = gs.oh; s.ohu = gs.ohu; s.ohv = gs.ohv;
            s.dt = gs.beta * dt;
        }
        t_tide = cp.t;  // stage abscissa c_1 = 0
    } else if (STAGE == 3) {
        const CtrlSlot cp = d.ctrl[p];
        const CtrlSlot cq = d.ctrl[1 - p];
        if (!cq.active) return;
        s.ih = gs.ih; s.ihu = gs.ihu; s.ihv = gs.ihv;
        s.oh = gs.oh; s.ohu = gs.ohu; s.ohv = gs.ohv;
        s.dt = gs.beta * cq.dt;
        t_tide = cp.t + gs.c * cq.dt;  // bc%time = t_base + c_i dt
        s.bc_dt = cq.dt;
        s.radiate = true;       // against the previous stage's input
        s.prev_h = gs.ph;
    } else {
        const CtrlSlot cq = d.ctrl[1 - p];
        if (!cq.active) return;
        s.ih = d.bh; s.ihu = d.bhu; s.ihv = d.bhv;
        s.oh = d.o2h; s.ohu = d.o2hu; s.ohv = d.o2hv;
        s.dt = cq.dt;
        t_tide = cq.t;  // bc%time = t + dt
        s.bc_dt = cq.dt;
        s.radiate = true;       // against U^n, the input of stage 1
        s.prev_h = d.ah;
    }
    if (STAGE == 3) {
        s.alpha = gs.alpha;
        s.a_prev = 1.0 - gs.alpha;
        s.blend = gs.alpha != 0.0;
        s.first = gs.first != 0;
        s.last = gs.last != 0;
    }
    s.src = make_src(d, s.dt);
    s.rdx = recip(d.dx);
    s.rdy = recip(d.dy);
    const int strip = w.x, seg = w.y;
    s.row0 = seg * d.ty;
    s.row1 = min(s.row0 + d.ty, d.ny);
    double smax, fsum, ssum;
    if (tile_is_interior(d, strip, seg)) {
        // no ghost cells: the tide (a correctly rounded sine, a long
        // dependent chain) is not needed
        smax = process_tile<STAGE, HLLC, false, NF, REC, FEAT, FASTK>(d, s, L, strip, seg, fsum, ssum);
    } else {
        s.eta_tidal = tidal_eta(d, t_tide);
        smax = process_tile<STAGE, HLLC, true, NF, REC, FEAT, FASTK>(d, s, L, strip, seg, fsum, ssum);
    }
    if (STAGE == 2 || STAGE == 3) {
        smax = wave_max(smax);
        if (threadIdx.x == 0 && smax > 0.0)
            atomicMax(&d.ctrl[1 - p].speed_bits, (unsigned long long)__double_as_longlong(smax));
    }
    if (STAGE != 3 && d.mpart != nullptr) {  // accumulate_boundary_mass_out's sum, per tile
        fsum = wave_sum(fsum);
        if (threadIdx.x == 0)
            d.mpart[(long long)(STAGE - 1) * d.nstrips * d.nsegs + (long long)seg * d.nstrips + strip] = fsum;
    }
    if ((FEAT & FEAT_SPONGE) && STAGE != 3 && d.spart != nullptr) {  // accumulate_sponge_mass_src
        ssum = wave_sum(ssum);
        if (threadIdx.x == 0)
            d.spart[(long long)(STAGE - 1) * d.nstrips * d.nsegs + (long long)seg * d.nstrips + strip] = ssum;
    }
}

// Single grid: one workgroup (one wavefront) per tile.
struct SingleGridTile {
    __device__ int3 operator()(const GridDesc& g) const {
        const int tile = xcd_tile(blockIdx.x, g.nstrips * g.nsegs);
        return make_int3(tile % g.nstrips, tile / g.nstrips, blockIdx.x == 0);
    }
};

// FEAT: FEAT_FORCING / FEAT_SPONGE source terms.  FASTK: the fast path
// compiled in (launches with d.fast set, FAST_MIN_TILES).
template <int STAGE, bool HLLC, bool NF, bool REC, int FEAT = 0, bool FASTK = false>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(RKG_WAVES_PER_EU, 8))) void k_stage(GridDesc d, int p, int mode, double tlim) {
    __shared__ TileLds L;
    stage_tile<STAGE, HLLC, NF, REC, SingleGridTile, FEAT, FASTK>(d, L, p, mode, tlim, SingleGridTile{});
}

// One stage of a generic SSP scheme, single grid.
template <bool HLLC, bool NF, int FEAT = 0>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(RKG_WAVES_PER_EU, 8))) void k_stage_gen(GridDesc d, GenStage gs, int p, int mode, double tlim) {
    __shared__ TileLds L;
    stage_tile<3, HLLC, NF, false, SingleGridTile, FEAT>(d, L, p, mode, tlim, SingleGridTile{}, gs);
}

// Resets control slot p to time t / step count n with an empty speed max;
// the mass budget is set to mass_out.
__global__ void k_ctrl_reset(CtrlSlot* ctrl, int p, double t, double dt, long long n, double mass_out,
                             double mass_src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctrl[p].t = t;
        ctrl[p].dt = dt;
        ctrl[p].step_count = n;
        ctrl[p].speed_bits = 0ull;
        ctrl[p].mass_out = mass_out;
        ctrl[p].mass_src = mass_src;
        ctrl[p].active = 1;
        ctrl[p].t_bc = t;
    }
}

// Boundary mass budget of one SSP-RK2 step (rki_solver.F90:507-537, called
// after each stage's flux at :769 and :796): mass_out -= 0.5 dt sum(flux_h)
// area per stage.  The per-tile sums of both stages are added in a fixed
// order (strided per thread, then a tree).  mass_src, per stage and in
// the reference's order: with the surface forcing on, `rain` holds the
// interior sum of the precipitation rate and mass_src += 0.5 dt rain area
// (accumulate_mass_src, rki_solver.F90:539-606, at :783 and :810); with the
// sponge on, `spart` holds the per-tile sums of its change of h and
// mass_src += 0.5 sum area (accumulate_sponge_mass_src, :608-657, at :787
// and :814).  Slot 1-p is the step's result.
__global__ __launch_bounds__(256) void k_budget(CtrlSlot* ctrl, int p, const double* mpart, int ntiles,
                                                double area, const double* rain, const double* spart) {
    __shared__ double red[4][256];
    if (!ctrl[1 - p].active) return;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < ntiles; i += 256) {
        a[0] = a[0] + mpart[i];
        a[1] = a[1] + mpart[ntiles + i];
        if (spart != nullptr) {
            a[2] = a[2] + spart[i];
            a[3] = a[3] + spart[ntiles + i];
        }
    }
    for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = a[q];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        CtrlSlot* q = &ctrl[1 - p];
        const double dt = q->dt;
        double m = q->mass_out;
        m = m - 0.5 * dt * red[0][0] * area;
        m = m - 0.5 * dt * red[1][0] * area;
        q->mass_out = m;
        double src = q->mass_src;
        for (int st = 0; st < 2; ++st) {
            if (rain != nullptr) src = src + 0.5 * dt * (*rain) * area;
            if (spart != nullptr) src = src + 0.5 * red[2 + st][0] * area;
        }
        q->mass_src = src;
    }
}

// End of a one-stage (Euler) step: the stage wrote B, so B is copied to A
// and the CFL max of the new state goes to slot 1-p.  (The stage kernel
// cannot take that max itself: its leader clears the slot's max while the
// other tiles of the same launch run.)
__global__ __launch_bounds__(256) void k_euler_finish(GridDesc d, int p) {
    if (!d.ctrl[1 - p].active) return;
    double m = 0.0;
    const Recip rdx = recip(d.dx), rdy = recip(d.dy);
    const long long n = (long long)d.nx * d.ny;
    for (long long k = blockIdx.x * 256ll + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / d.nx), i = (int)(k % d.nx);
        const long long o = (long long)j * d.pitch + i;
        const double h = d.bh[o], hu = d.bhu[o], hv = d.bhv[o];
        d.ah[o] = h; d.ahu[o] = hu; d.ahv[o] = hv;
        m = dmax(m, cfl_speed<true>(d.g, rdx, rdy, h, hu, hv));
    }
    m = wave_max(m);
    if ((threadIdx.x & (WAVE - 1)) == 0 && m > 0.0)
        atomicMax(&d.ctrl[1 - p].speed_bits, (unsigned long long)__double_as_longlong(m));
}

// Max CFL speed of the state A into slot p (compute_timestep's reduction,
// rki_kernel_timestep.F90:53-70).
__global__ __launch_bounds__(256) void k_init_speed(GridDesc d, int p) {
    double m = 0.0;
    const Recip rdx = recip(d.dx), rdy = recip(d.dy);
    const long long n = (long long)d.nx * d.ny;
    for (long long k = blockIdx.x * 256ll + threadIdx.x; k < n; k += (long long)gridDim.x * 256) {
        const int j = (int)(k / d.nx), i = (int)(k % d.nx);
        const long long o = (long long)j * d.pitch + i;
        m = dmax(m, cfl_speed<true>(d.g, rdx, rdy, d.ah[o], d.ahu[o], d.ahv[o]));
    }
    m = wave_max(m);
    if ((threadIdx.x & (WAVE - 1)) == 0 && m > 0.0)
        atomicMax(&d.ctrl[p].speed_bits, (unsigned long long)__double_as_longlong(m));
}

// Mass partials: block b sums rows b, b + gridDim.x, ... in a fixed order
// (compute_total_mass, rki_init.F90:287-317).
__global__ __launch_bounds__(256) void k_mass_partial(GridDesc d, double* partial) {
    __shared__ double red[256];
    const double area = d.dx * d.dy;
    double acc = 0.0;
    for (int j = blockIdx.x; j < d.ny; j += gridDim.x) {
        for (int i = threadIdx.x; i < d.nx; i += 256) {
            const double h = d.ah[(long long)j * d.pitch + i];
            if (h > 0.0 && h < 1.0e6) acc = acc + h * area;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// The central pressure gradient of the surface forcing
// (rki_kernel_fused.F90:277-278), dp/dx = (p(i+1) - p(i-1)) / (2 dx), with
// pressure 0 outside the grid (the ghost cells, which no setter writes);
// computed once per pressure field, in the same operations as per stage.
__global__ __launch_bounds__(256) void k_pressure_gradient(const double* p, double* dpx, double* dpy, int nx,
                                                           int ny, long long pitch, double dx, double dy) {
    const long long n = (long long)nx * ny;
    for (long long q = blockIdx.x * 256ll + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
        const int c = (int)(q % nx), j = (int)(q / nx);
        const long long k = (long long)j * pitch + c;
        const double pe = c + 1 < nx ? p[k + 1] : 0.0;
        const double pw = c > 0 ? p[k - 1] : 0.0;
        const double pn = j + 1 < ny ? p[k + pitch] : 0.0;
        const double ps = j > 0 ? p[k - pitch] : 0.0;
        dpx[k] = (pe - pw) / (2.0 * dx);
        dpy[k] = (pn - ps) / (2.0 * dy);
    }
}

// Per-block sums of the interior of one plane (rows strided over blocks),
// added up by k_mass_final: a fixed order, so the result is deterministic.
__global__ __launch_bounds__(256) void k_plane_sum_partial(const double* f, long long pitch, int nx, int ny,
                                                           double* partial) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int j = blockIdx.x; j < ny; j += gridDim.x)
        for (int i = threadIdx.x; i < nx; i += 256) acc = acc + f[(long long)j * pitch + i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_mass_final(const double* partial, int n, double* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) acc = acc + partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace
}  // namespace rk
