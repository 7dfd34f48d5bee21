// rkg_noalloc.hip -- rkg_step_noalloc: SSP-RK2 steps on state buffers the
// caller owns, with every intermediate in a fixed device array, so that a
// call allocates no memory (for code under graph capture, serving loops and
// harnesses that forbid hipMalloc).
//
// The intermediates of rkg_step are U* (three planes, the stage-1 result)
// and the two control slots that carry t, dt, the step count and the CFL
// speed maximum between the launches.  Here they live in `scratch`, one
// array per GPU that the library runs on, held until the process ends:
//   [0, 256 B)                 control slots 0 and 1
//   [256 B, ...)               U*: h, hu, hv planes of pitch * ny doubles
// Its address on each device is taken once per device with
// hipGetSymbolAddress (the array's name in host code is not a device
// address) and kept per device.  Every call on a device shares that
// device's array: calls that overlap on one GPU (other streams, other host
// threads) race and must be serialised by the caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "../../include/rakali_amd.h"
#include "rkg_internal.hpp"
#include "stage_kernels.hpp"
#include "stage_launch.hpp"

using namespace rk;

namespace {

constexpr long long CTRL_DOUBLES = 32;   // 256 B: two 64-byte control slots, padded
constexpr long long SCRATCH_DOUBLES = CTRL_DOUBLES + 3LL * RKG_NOALLOC_MAX_CELLS;

static_assert(2 * sizeof(CtrlSlot) <= CTRL_DOUBLES * sizeof(double), "control slots exceed their region");

__device__ __align__(256) double scratch[SCRATCH_DOUBLES];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of this device's `scratch`, looked up on the first call
// on the device (with that device current) and kept.
int scratch_of(int dev, double** out) {
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("rkg_step_noalloc: device %d beyond the %d devices the library tracks", dev, MAX_DEVICES);
        return RKG_ERR_RANGE;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("rkg_step_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return RKG_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return RKG_OK;
}

}  // namespace

extern "C" int rkg_step_noalloc(const rkg_config* cfg, double* h, double* hu, double* hv, const double* b,
                                long long pitch, int n, double* t, double* dt, long long* step_count,
                                void* stream) {
    if (!cfg || !h || !hu || !hv || !b || !t || !dt || !step_count) {
        set_error("rkg_step_noalloc: null argument");
        return RKG_ERR_RANGE;
    }
    int rc = validate_config(cfg);
    if (rc) return rc;
    if (n < 0) { set_error("rkg_step_noalloc: n must be >= 0 (got %d)", n); return RKG_ERR_RANGE; }
    if (pitch < cfg->nx) {
        set_error("rkg_step_noalloc: pitch %lld is below nx %d", pitch, cfg->nx);
        return RKG_ERR_RANGE;
    }
    const long long cells = pitch * (long long)cfg->ny;
    if (cells > RKG_NOALLOC_MAX_CELLS) {
        set_error("rkg_step_noalloc: pitch x ny = %lld x %d = %lld cells exceeds the maximum of %lld "
                  "(RKG_NOALLOC_MAX_CELLS)", pitch, cfg->ny, cells, (long long)RKG_NOALLOC_MAX_CELLS);
        return RKG_ERR_RANGE;
    }
    if (n == 0) return RKG_OK;

    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    double* s = nullptr;
    rc = scratch_of(dev, &s);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;

    GridDesc d = desc_of_config(*cfg, pitch, default_tile_rows(cfg->nx, cfg->ny));
    d.ah = h; d.ahu = hu; d.ahv = hv;
    d.o2h = h; d.o2hu = hu; d.o2hv = hv;
    d.bed = b;
    d.bh = s + CTRL_DOUBLES;
    d.bhu = d.bh + cells;
    d.bhv = d.bhu + cells;
    d.ctrl = reinterpret_cast<CtrlSlot*>(s);
    bool chapman = false;
    for (int k = 0; k < 4; ++k) chapman = chapman || cfg->bc[k] == RKG_BC_CHAPMAN;
    const int ntiles = d.nstrips * d.nsegs;
    const bool hllc = cfg->use_hllc != 0;

    // Slot 0 starts from the caller's clock (a kernel writes it: no memset
    // node if the caller captures the call); the speed maximum of the state
    // is taken afresh, as rkg_step does after an upload.
    hipLaunchKernelGGL(k_ctrl_reset, dim3(1), dim3(64), 0, st, d.ctrl, 0, *t, *dt, *step_count, 0.0, 0.0);
    const int blocks = (int)std::min<long long>(((long long)cfg->nx * cfg->ny + 255) / 256, 4096);
    hipLaunchKernelGGL(k_init_speed, dim3(blocks), dim3(256), 0, st, d, 0);
    int parity = 0;
    for (int k = 0; k < n; ++k) {
        for (int stage = 1; stage <= 2; ++stage) {
            if (chapman)
                launch_fused_stage<FEAT_CHAPMAN>(stage, st, d, ntiles, parity, MODE_STEP, cfg->t_end, hllc, false);
            else
                launch_fused_stage<0>(stage, st, d, ntiles, parity, MODE_STEP, cfg->t_end, hllc, false);
        }
        parity ^= 1;
    }
    HIP_TRY(hipGetLastError());
    CtrlSlot c;
    HIP_TRY(hipMemcpyAsync(&c, d.ctrl + parity, sizeof c, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    *t = c.t;
    *dt = c.dt;
    *step_count = c.step_count;
    return RKG_OK;
}
