// rkg_ensemble.hip -- ensembles of independent grids of uneven sizes
// stepped together: one launch per RK stage for all members.
//
// Every member keeps the single-grid layout and control slots (its own t,
// dt and CFL speed); a flat table lists the tiles of all members, member
// by member.  Two schedules run the same tile routine as the single-grid
// kernel (stage_tile, bit-identical results per member):
//   flat       - one workgroup per table entry, the XCD-aware mapping over
//                the whole table;
//   persistent - a fixed grid of workgroups that take table entries from an
//                atomic counter until the table is exhausted.  Each stage
//                clears the other stage's counter, so back-to-back steps
//                (and hipGraph replays of them) need no host reset.
// Steps can be captured into a hipGraph of two steps (four launches, both
// parities) and replayed.  rkg_ensemble_autotune times the schedules and
// tile heights on the live ensemble and restores its state afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/rakali_amd.h"
#include "rkg_internal.hpp"
#include "stage_kernels.hpp"

using namespace rk;

namespace {

constexpr unsigned ENS_MAGIC = 0x72616B45u;  // 'rakE'
constexpr unsigned ENS_DEAD = 0xDEADBEEFu;

struct BatchTile {
    int member, strip, seg, leader;
};

// The tile of a table entry, as stage_tile asks for it.
struct TileOf {
    BatchTile bt;
    __device__ int3 operator()(const GridDesc&) const { return make_int3(bt.strip, bt.seg, bt.leader); }
};

// The tile loop of the persistent variant needs about 250 VGPRs to keep
// the tile body free of spills, so it runs at two waves per SIMD.
template <int STAGE, bool HLLC, bool PERSISTENT, bool FASTK>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(PERSISTENT ? 2 : RKG_WAVES_PER_EU, 8))) void k_stage_batch(
    const GridDesc* __restrict__ descs, const BatchTile* __restrict__ tiles, int ntiles, int p,
    unsigned* __restrict__ queue) {
    __shared__ TileLds L;
    if (!PERSISTENT) {
        const int t = xcd_tile(blockIdx.x, ntiles);
        const BatchTile bt = tiles[t];
        stage_tile<STAGE, HLLC, false, false, TileOf, 0, FASTK>(descs[bt.member], L, p, MODE_STEP,
                                                              descs[bt.member].t_end, TileOf{bt});
        return;
    }
    // The other stage's counter is idle during this launch: clear it for
    // the next one.
    if (blockIdx.x == 0 && threadIdx.x == 0) queue[STAGE == 1 ? 1 : 0] = 0u;
    for (;;) {
        __shared__ unsigned slot;
        if (threadIdx.x == 0) slot = atomicAdd(&queue[STAGE - 1], 1u);
        __syncthreads();
        const int t = __builtin_amdgcn_readfirstlane((int)slot);
        if (t >= ntiles) break;
        const BatchTile bt = tiles[t];
        stage_tile<STAGE, HLLC, false, false, TileOf, 0, FASTK>(descs[bt.member], L, p, MODE_STEP,
                                                              descs[bt.member].t_end, TileOf{bt});
        __syncthreads();   // the next tile rewrites the LDS buffers and the slot
    }
}

}  // namespace

struct rkg_ensemble {
    unsigned magic;
    int n;
    std::vector<rkg_config> cfg;
    std::vector<long long> pitch;
    std::vector<size_t> plane, offset;   // elements
    int device;
    hipStream_t stream;
    bool own_stream;
    double* buf;
    CtrlSlot* ctrl;        // device [2 n]
    CtrlSlot* ctrl_host;   // pinned [2 n]
    GridDesc* descs;       // device [n]
    BatchTile* tiles;      // device
    int ntiles;
    unsigned* queue;       // device [2]
    double* mass_buf;
    double* mass_host;
    int parity;
    int ty;
    int schedule;          // 0 flat, 1 persistent
    int persistent_blocks;
    int cus;               // compute units of the device
    int use_hllc;
    std::vector<char> speed_stale;
    bool use_graph;
    hipGraphExec_t graph;  // two steps, starting at parity 0
    bool graph_valid;
    long long steps_since_graph;
};

namespace {

bool valid(const rkg_ensemble* e) { return e != nullptr && e->magic == ENS_MAGIC; }

double* ens_plane(rkg_ensemble* e, int m, int k) { return e->buf + e->offset[m] + (size_t)k * e->plane[m]; }

GridDesc ens_desc(rkg_ensemble* e, int m) {
    GridDesc d;
    memset(&d, 0, sizeof d);
    const rkg_config& c = e->cfg[m];
    d.nx = c.nx; d.ny = c.ny; d.pitch = e->pitch[m];
    d.dx = c.dx; d.dy = c.dy;
    d.ah = ens_plane(e, m, 0); d.ahu = ens_plane(e, m, 1); d.ahv = ens_plane(e, m, 2);
    d.bh = ens_plane(e, m, 3); d.bhu = ens_plane(e, m, 4); d.bhv = ens_plane(e, m, 5);
    d.bed = ens_plane(e, m, 6);
    d.o2h = d.ah; d.o2hu = d.ahu; d.o2hv = d.ahv;
    d.plane = (long long)e->plane[m];
    d.g = c.g; d.cfl = c.cfl; d.manning_n = c.manning_n; d.coriolis_f = c.coriolis_f;
    d.use_hllc = c.use_hllc;
    for (int s = 0; s < 4; ++s) d.bc[s] = c.bc[s];
    d.tidal_amplitude = c.tidal_amplitude; d.tidal_period = c.tidal_period;
    d.n_tc = c.n_tidal_constituents;
    for (int k = 0; k < RKG_MAX_TIDAL_CONSTITUENTS; ++k) {
        d.tc_amp[k] = c.tidal_amp[k]; d.tc_phase[k] = c.tidal_phase[k]; d.tc_omega[k] = c.tidal_omega[k];
    }
    d.inflow_velocity = c.inflow_velocity; d.discharge_flux = c.discharge_flux;
    d.clamped_elevation = c.clamped_elevation;
    d.dt_fixed = c.dt_fixed; d.t_end = c.t_end;
    d.cfl_interval = c.cfl_interval;
    d.ty = e->ty;
    d.nstrips = (c.nx + WAVE - 1) / WAVE;
    d.nsegs = (c.ny + e->ty - 1) / e->ty;
    d.ctrl = e->ctrl + 2 * m;
    d.ssp_coeff = 1.0;  // ensembles run SSP-RK2
    return d;
}

void drop_graph(rkg_ensemble* e) {
    if (e->graph_valid) (void)hipGraphExecDestroy(e->graph);
    e->graph_valid = false;
}

// Descriptors and the tile table for the current tile height.
int build_tables(rkg_ensemble* e) {
    std::vector<GridDesc> d(e->n);
    std::vector<BatchTile> t;
    for (int m = 0; m < e->n; ++m) {
        d[m] = ens_desc(e, m);
        for (int seg = 0; seg < d[m].nsegs; ++seg)
            for (int strip = 0; strip < d[m].nstrips; ++strip)
                t.push_back(BatchTile{m, strip, seg, (strip == 0 && seg == 0) ? 1 : 0});
    }
    HIP_TRY(hipStreamSynchronize(e->stream));
    if (e->tiles) HIP_TRY(hipFree(e->tiles));
    e->tiles = nullptr;
    HIP_TRY(hipMalloc((void**)&e->tiles, t.size() * sizeof(BatchTile)));
    HIP_TRY(hipMemcpyAsync(e->tiles, t.data(), t.size() * sizeof(BatchTile), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(e->descs, d.data(), d.size() * sizeof(GridDesc), hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipMemsetAsync(e->queue, 0, 2 * sizeof(unsigned), e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->ntiles = (int)t.size();
    drop_graph(e);
    return RKG_OK;
}

int refresh_speeds(rkg_ensemble* e) {
    bool any = false;
    for (int m = 0; m < e->n; ++m) any = any || e->speed_stale[m];
    if (!any) return RKG_OK;
    HIP_TRY(hipMemcpyAsync(e->ctrl_host, e->ctrl, 2 * e->n * sizeof(CtrlSlot), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    for (int m = 0; m < e->n; ++m) {
        if (!e->speed_stale[m]) continue;
        const GridDesc d = ens_desc(e, m);
        const CtrlSlot c = e->ctrl_host[2 * m + e->parity];
        hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, e->stream, d.ctrl, e->parity, c.t, c.dt,
                           c.step_count, c.mass_out, c.mass_src);
        const long long cells = (long long)d.nx * d.ny;
        const int blocks = (int)std::min<long long>((cells + 255) / 256, 4096);
        hipLaunchKernelGGL(k_init_speed, dim3(blocks), dim3(256), 0, e->stream, d, e->parity);
        e->speed_stale[m] = 0;
    }
    HIP_TRY(hipGetLastError());
    return RKG_OK;
}

template <int STAGE, bool HLLC, bool FASTK>
void launch_one_k(rkg_ensemble* e, int p) {
    if (e->schedule == 1)
        hipLaunchKernelGGL((k_stage_batch<STAGE, HLLC, true, FASTK>), dim3(e->persistent_blocks), dim3(WAVE), 0,
                           e->stream, e->descs, e->tiles, e->ntiles, p, e->queue);
    else
        hipLaunchKernelGGL((k_stage_batch<STAGE, HLLC, false, FASTK>), dim3(e->ntiles), dim3(WAVE), 0, e->stream,
                           e->descs, e->tiles, e->ntiles, p, e->queue);
}

// The kernels with the fast path when the launch holds enough tiles.
template <int STAGE, bool HLLC>
void launch_one(rkg_ensemble* e, int p) {
    if ((long long)e->ntiles >= FAST_MIN_TILES) launch_one_k<STAGE, HLLC, true>(e, p);
    else launch_one_k<STAGE, HLLC, false>(e, p);
}

void launch_stage(rkg_ensemble* e, int stage, int p) {
    if (stage == 1) {
        if (e->use_hllc) launch_one<1, true>(e, p);
        else launch_one<1, false>(e, p);
    } else {
        if (e->use_hllc) launch_one<2, true>(e, p);
        else launch_one<2, false>(e, p);
    }
}

int capture_graph(rkg_ensemble* e) {
    drop_graph(e);
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    for (int p = 0; p < 2; ++p) {
        launch_stage(e, 1, p);
        launch_stage(e, 2, p);
    }
    hipError_t err = hipStreamEndCapture(e->stream, &g);
    if (err != hipSuccess) {
        set_error("hipStreamEndCapture: %s", hipGetErrorString(err));
        return RKG_ERR_HIP;
    }
    err = hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (err != hipSuccess) {
        set_error("hipGraphInstantiate: %s", hipGetErrorString(err));
        return RKG_ERR_HIP;
    }
    e->graph_valid = true;
    return RKG_OK;
}

int ens_steps(rkg_ensemble* e, int n) {
    int rc = refresh_speeds(e);
    if (rc) return rc;
    int k = 0;
    if (e->use_graph && n >= 2) {
        if (!e->graph_valid && (rc = capture_graph(e))) return rc;
        if (e->parity == 1) {   // the graph starts at parity 0
            launch_stage(e, 1, 1);
            launch_stage(e, 2, 1);
            e->parity = 0;
            ++k;
        }
        for (; k + 2 <= n; k += 2) HIP_TRY(hipGraphLaunch(e->graph, e->stream));
    }
    for (; k < n; ++k) {
        launch_stage(e, 1, e->parity);
        launch_stage(e, 2, e->parity);
        e->parity ^= 1;
    }
    HIP_TRY(hipGetLastError());
    return RKG_OK;
}

double* ens_field(rkg_ensemble* e, int m, int field) {
    switch (field) {
    case RKG_FIELD_H: return ens_plane(e, m, 0);
    case RKG_FIELD_HU: return ens_plane(e, m, 1);
    case RKG_FIELD_HV: return ens_plane(e, m, 2);
    case RKG_FIELD_B: return ens_plane(e, m, 6);
    case RKG_FIELD_HS: return ens_plane(e, m, 3);
    case RKG_FIELD_HUS: return ens_plane(e, m, 4);
    case RKG_FIELD_HVS: return ens_plane(e, m, 5);
    default: return nullptr;
    }
}

int check_member(rkg_ensemble* e, int m) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    if (m < 0 || m >= e->n) { set_error("member %d out of range [0, %d)", m, e->n); return RKG_ERR_RANGE; }
    return RKG_OK;
}

}  // namespace

extern "C" {

int rkg_ensemble_create(const rkg_config* cfgs, int n, void* stream, rkg_ensemble** out) {
    if (!out) { set_error("null argument"); return RKG_ERR_RANGE; }
    *out = nullptr;
    if (!cfgs || n < 1) { set_error("an ensemble needs at least one member"); return RKG_ERR_RANGE; }
    for (int m = 0; m < n; ++m) {
        const rkg_config& c = cfgs[m];
        if (c.nx < 2 || c.ny < 2 || !(c.dx >= 1e-6 && c.dx <= 1e9) || !(c.dy >= 1e-6 && c.dy <= 1e9) ||
            c.cfl_interval < 1 || c.n_tidal_constituents < 0 ||
            c.n_tidal_constituents > RKG_MAX_TIDAL_CONSTITUENTS) {
            set_error("member %d: invalid configuration", m);
            return RKG_ERR_RANGE;
        }
        for (int s = 0; s < 4; ++s)
            if (c.bc[s] != RKG_BC_WALL && c.bc[s] != RKG_BC_OPEN && c.bc[s] != RKG_BC_TIDAL &&
                c.bc[s] != RKG_BC_INFLOW && c.bc[s] != RKG_BC_DISCHARGE && c.bc[s] != RKG_BC_CLAMPED &&
                c.bc[s] != RKG_BC_SPONGE) {
                set_error("member %d: boundary tag %d is not implemented", m, c.bc[s]);
                return RKG_ERR_UNSUPPORTED;
            }
        if ((c.use_hllc != 0) != (cfgs[0].use_hllc != 0)) {
            set_error("all members of an ensemble use the same Riemann solver (use_hllc)");
            return RKG_ERR_UNSUPPORTED;
        }
    }
    rkg_ensemble* e = new rkg_ensemble();
    e->magic = 0;
    e->n = n;
    e->cfg.assign(cfgs, cfgs + n);
    e->ty = 32;
    e->schedule = 0;
    e->use_hllc = cfgs[0].use_hllc != 0;
    e->speed_stale.assign(n, 1);
    e->buf = nullptr; e->ctrl = nullptr; e->ctrl_host = nullptr; e->descs = nullptr; e->tiles = nullptr;
    e->queue = nullptr; e->mass_buf = nullptr; e->mass_host = nullptr;
    e->graph_valid = false;
    e->use_graph = false;
    e->parity = 0;
    int cus = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&e->device) == hipSuccess && hipGetDeviceProperties(&prop, e->device) == hipSuccess)
        cus = prop.multiProcessorCount;
    e->cus = cus;
    e->persistent_blocks = cus * 4 * 2;
    size_t total = 0;
    for (int m = 0; m < n; ++m) {
        long long pitch = ((long long)cfgs[m].nx + 15) / 16 * 16;
        if (pitch % 1024 == 0) pitch += 32;
        e->pitch.push_back(pitch);
        e->plane.push_back((size_t)pitch * (size_t)cfgs[m].ny);
        e->offset.push_back(total);
        total += 7 * e->plane.back();
        total = (total + 31) / 32 * 32;   // 256-byte aligned members
    }
    hipError_t err = hipSuccess;
    if (stream) {
        e->stream = (hipStream_t)stream;
        e->own_stream = false;
    } else {
        err = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
        e->own_stream = err == hipSuccess;
    }
    if (err == hipSuccess) err = hipMalloc((void**)&e->buf, total * sizeof(double));
    if (err == hipSuccess) err = hipMalloc((void**)&e->ctrl, 2 * n * sizeof(CtrlSlot));
    if (err == hipSuccess) err = hipHostMalloc((void**)&e->ctrl_host, 2 * n * sizeof(CtrlSlot), hipHostMallocDefault);
    if (err == hipSuccess) err = hipMalloc((void**)&e->descs, n * sizeof(GridDesc));
    if (err == hipSuccess) err = hipMalloc((void**)&e->queue, 2 * sizeof(unsigned));
    if (err == hipSuccess) err = hipMalloc((void**)&e->mass_buf, (MASS_BLOCKS + 1) * sizeof(double));
    if (err == hipSuccess) err = hipHostMalloc((void**)&e->mass_host, sizeof(double), hipHostMallocDefault);
    if (err == hipSuccess) err = hipMemsetAsync(e->buf, 0, total * sizeof(double), e->stream);
    if (err == hipSuccess) err = hipMemsetAsync(e->ctrl, 0, 2 * n * sizeof(CtrlSlot), e->stream);
    if (err == hipSuccess) err = hipStreamSynchronize(e->stream);
    e->magic = ENS_MAGIC;
    if (err != hipSuccess) {
        set_error("ensemble allocation failed: %s", hipGetErrorString(err));
        rkg_ensemble_destroy(&e);
        return RKG_ERR_HIP;
    }
    int rc = build_tables(e);
    if (rc) {
        rkg_ensemble_destroy(&e);
        return rc;
    }
    *out = e;
    return RKG_OK;
}

void rkg_ensemble_destroy(rkg_ensemble** pe) {
    if (!pe || !*pe) return;
    rkg_ensemble* e = *pe;
    if (e->magic != ENS_MAGIC) { *pe = nullptr; return; }
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graph(e);
    if (e->buf) (void)hipFree(e->buf);
    if (e->ctrl) (void)hipFree(e->ctrl);
    if (e->ctrl_host) (void)hipHostFree(e->ctrl_host);
    if (e->descs) (void)hipFree(e->descs);
    if (e->tiles) (void)hipFree(e->tiles);
    if (e->queue) (void)hipFree(e->queue);
    if (e->mass_buf) (void)hipFree(e->mass_buf);
    if (e->mass_host) (void)hipHostFree(e->mass_host);
    if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
    e->magic = ENS_DEAD;
    delete e;
    *pe = nullptr;
}

int rkg_ensemble_size(rkg_ensemble* e, int* n, long long* cells, int* ntiles) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    long long c = 0;
    for (const rkg_config& k : e->cfg) c += (long long)k.nx * k.ny;
    if (n) *n = e->n;
    if (cells) *cells = c;
    if (ntiles) *ntiles = e->ntiles;
    return RKG_OK;
}

int rkg_ensemble_upload(rkg_ensemble* e, int m, int field, const double* src, int ld) {
    int rc = check_member(e, m);
    if (rc) return rc;
    if (field < RKG_FIELD_H || field > RKG_FIELD_B) { set_error("field %d is not writable", field); return RKG_ERR_RANGE; }
    const rkg_config& c = e->cfg[m];
    if (!src || ld < c.nx) { set_error("bad buffer or leading dimension"); return RKG_ERR_RANGE; }
    HIP_TRY(hipMemcpy2DAsync(ens_field(e, m, field), (size_t)e->pitch[m] * sizeof(double), src,
                             (size_t)ld * sizeof(double), (size_t)c.nx * sizeof(double), (size_t)c.ny,
                             hipMemcpyHostToDevice, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    e->speed_stale[m] = 1;
    return RKG_OK;
}

int rkg_ensemble_download(rkg_ensemble* e, int m, int field, double* dst, int ld) {
    int rc = check_member(e, m);
    if (rc) return rc;
    double* f = ens_field(e, m, field);
    const rkg_config& c = e->cfg[m];
    if (!f) { set_error("unknown field %d", field); return RKG_ERR_RANGE; }
    if (!dst || ld < c.nx) { set_error("bad buffer or leading dimension"); return RKG_ERR_RANGE; }
    HIP_TRY(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), f, (size_t)e->pitch[m] * sizeof(double),
                             (size_t)c.nx * sizeof(double), (size_t)c.ny, hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RKG_OK;
}

int rkg_ensemble_step(rkg_ensemble* e, int n) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    if (n < 0) return RKG_ERR_RANGE;
    return ens_steps(e, n);
}

int rkg_ensemble_sync(rkg_ensemble* e) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    HIP_TRY(hipStreamSynchronize(e->stream));
    return RKG_OK;
}

int rkg_ensemble_get_clock(rkg_ensemble* e, int m, double* t, double* dt, long long* steps) {
    int rc = check_member(e, m);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(e->ctrl_host + 2 * m, e->ctrl + 2 * m, 2 * sizeof(CtrlSlot), hipMemcpyDeviceToHost,
                           e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    const CtrlSlot c = e->ctrl_host[2 * m + e->parity];
    if (t) *t = c.t;
    if (dt) *dt = c.dt;
    if (steps) *steps = c.step_count;
    return RKG_OK;
}

int rkg_ensemble_total_mass(rkg_ensemble* e, int m, double* mass) {
    int rc = check_member(e, m);
    if (rc) return rc;
    if (!mass) return RKG_ERR_RANGE;
    const GridDesc d = ens_desc(e, m);
    hipLaunchKernelGGL(k_mass_partial, dim3(MASS_BLOCKS), dim3(256), 0, e->stream, d, e->mass_buf);
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, e->stream, e->mass_buf, MASS_BLOCKS,
                       e->mass_buf + MASS_BLOCKS);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(e->mass_host, e->mass_buf + MASS_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, e->stream));
    HIP_TRY(hipStreamSynchronize(e->stream));
    *mass = *e->mass_host;
    return RKG_OK;
}

int rkg_ensemble_set_schedule(rkg_ensemble* e, int schedule, int tile_rows, int persistent_blocks, int use_graph) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    if (schedule != 0 && schedule != 1) { set_error("schedule must be 0 (flat) or 1 (persistent)"); return RKG_ERR_RANGE; }
    if (tile_rows != 0 && (tile_rows < 4 || tile_rows > 64)) { set_error("tile_rows must lie in [4, 64]"); return RKG_ERR_RANGE; }
    if (persistent_blocks < 0 || persistent_blocks > (1 << 20)) { set_error("persistent_blocks out of range"); return RKG_ERR_RANGE; }
    HIP_TRY(hipStreamSynchronize(e->stream));
    const bool rebuild = tile_rows != 0 && tile_rows != e->ty;
    e->schedule = schedule;
    if (persistent_blocks) e->persistent_blocks = persistent_blocks;
    e->use_graph = use_graph != 0;
    if (rebuild) {
        e->ty = tile_rows;
        int rc = build_tables(e);
        if (rc) return rc;
    }
    drop_graph(e);
    return RKG_OK;
}

int rkg_ensemble_get_schedule(rkg_ensemble* e, int* schedule, int* tile_rows, int* persistent_blocks, int* use_graph) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    if (schedule) *schedule = e->schedule;
    if (tile_rows) *tile_rows = e->ty;
    if (persistent_blocks) *persistent_blocks = e->persistent_blocks;
    if (use_graph) *use_graph = e->use_graph ? 1 : 0;
    return RKG_OK;
}

// Times `steps` batched steps for each candidate schedule on the live
// ensemble, keeps the fastest and restores the state it started from
// (fields, control slots, parity).
int rkg_ensemble_autotune(rkg_ensemble* e, int steps, double* best_ms_per_step) {
    if (!valid(e)) return RKG_ERR_BAD_HANDLE;
    if (steps < 1 || steps > 1000) { set_error("steps must lie in [1, 1000]"); return RKG_ERR_RANGE; }
    int rc = refresh_speeds(e);
    if (rc) return rc;
    size_t total = 0;
    for (int m = 0; m < e->n; ++m) total = std::max(total, e->offset[m] + 7 * e->plane[m]);
    double* backup = nullptr;
    CtrlSlot* ctrl_backup = nullptr;
    HIP_TRY(hipMalloc((void**)&backup, total * sizeof(double)));
    if (hipMalloc((void**)&ctrl_backup, 2 * e->n * sizeof(CtrlSlot)) != hipSuccess) {
        (void)hipFree(backup);
        set_error("autotune: out of device memory");
        return RKG_ERR_HIP;
    }
    HIP_TRY(hipMemcpyAsync(backup, e->buf, total * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
    HIP_TRY(hipMemcpyAsync(ctrl_backup, e->ctrl, 2 * e->n * sizeof(CtrlSlot), hipMemcpyDeviceToDevice, e->stream));
    const int parity0 = e->parity;
    const int graph0 = e->use_graph ? 1 : 0;
    const int base_blocks = 4 * e->cus;   // one workgroup per SIMD
    struct Cand { int sched, ty, blocks; };
    std::vector<Cand> cands;
    for (int ty : {8, 16, 32, 64}) {
        cands.push_back({0, ty, 0});
        for (int mult : {1, 2}) cands.push_back({1, ty, base_blocks * mult});
    }
    hipEvent_t a, b;
    HIP_TRY(hipEventCreate(&a));
    HIP_TRY(hipEventCreate(&b));
    double best = 1e300;
    Cand win = {e->schedule, e->ty, e->persistent_blocks};
    for (const Cand& c : cands) {
        if ((rc = rkg_ensemble_set_schedule(e, c.sched, c.ty, c.blocks, 0))) break;
        ens_steps(e, 1);   // warm-up
        HIP_TRY(hipEventRecord(a, e->stream));
        ens_steps(e, steps);
        HIP_TRY(hipEventRecord(b, e->stream));
        HIP_TRY(hipEventSynchronize(b));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, a, b));
        if (ms / steps < best) {
            best = ms / steps;
            win = c;
        }
        // back to the starting state for the next candidate
        HIP_TRY(hipMemcpyAsync(e->buf, backup, total * sizeof(double), hipMemcpyDeviceToDevice, e->stream));
        HIP_TRY(hipMemcpyAsync(e->ctrl, ctrl_backup, 2 * e->n * sizeof(CtrlSlot), hipMemcpyDeviceToDevice, e->stream));
        e->parity = parity0;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    HIP_TRY(hipStreamSynchronize(e->stream));
    (void)hipFree(backup);
    (void)hipFree(ctrl_backup);
    if (rc) return rc;
    rc = rkg_ensemble_set_schedule(e, win.sched, win.ty, win.sched == 1 ? win.blocks : 0, graph0);
    if (best_ms_per_step) *best_ms_per_step = best;
    return rc;
}

}  // extern "C"
