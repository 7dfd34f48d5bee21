// rkg_gauges.hip -- gauge stations: bilinear samples of eta, u and v at
// points given in metres (rki_gauges.F90), taken on the device, and a time
// series recorded after each step at the driver's cadence
// (rki_driver.F90:345-352,470-474) without a host round trip per step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/rakali_amd.h"
#include "gauge_eval.hpp"
#include "rkg_internal.hpp"

namespace rk {

// Station table on the device: lower-left interior cell (0-based) and the
// bilinear weights of each station inside the domain.
struct GaugeTable {
    const int* c0;
    const int* j0;
    const double* wx;
    const double* wy;
    int n;
};

struct GaugeState {
    int n_local = 0;
    std::vector<int> local_idx;     // station ids inside the domain, in input order
    int* d_cell = nullptr;          // [2 n]: c0 then j0
    double* d_w = nullptr;          // [2 n]: wx then wy
    double* d_now = nullptr;        // [3 n]: one sample (eta, u, v)
    // time series: records of (t, eta[n], u[n], v[n]) on the device,
    // drained to `host` when the device buffer may fill
    bool recording = false;
    double interval = 0.0;          // gauge_output_interval (<= 0: every step)
    int cap = 0;                    // records the device buffer holds
    double* d_series = nullptr;
    int* d_count = nullptr;         // [1] records in d_series
    double* d_next = nullptr;       // [1] next_gauge_time
    long long pending = 0;          // steps launched since the last drain (an upper bound of the records)
    std::vector<double> host;       // drained records, same layout
    long long host_records = 0;
};

namespace {

// Station g of the current state A (gauges_write, rki_gauges.F90:516-550).
__device__ __forceinline__ void gauge_value(const GridDesc& d, const GaugeTable& t, int g, double& eta, double& u,
                                            double& v) {
    const int c0 = t.c0[g], j0 = t.j0[g];
    const int c1 = min(c0 + 1, d.nx - 1), j1 = min(j0 + 1, d.ny - 1);
    const double wx = t.wx[g], wy = t.wy[g];
    const double hi = station_bilinear(d.ah, d.pitch, c0, c1, j0, j1, wx, wy);
    const double hui = station_bilinear(d.ahu, d.pitch, c0, c1, j0, j1, wx, wy);
    const double hvi = station_bilinear(d.ahv, d.pitch, c0, c1, j0, j1, wx, wy);
    const double bi = station_bilinear(d.bed, d.pitch, c0, c1, j0, j1, wx, wy);
    eta = hi + bi;
    const double hs = fmax(hi, 1.0e-6);   // MIN_DEPTH, :508
    u = hui / hs;
    v = hvi / hs;
}

__global__ __launch_bounds__(256) void k_gauges_now(GridDesc d, GaugeTable t, double* out) {
    for (int g = blockIdx.x * 256 + threadIdx.x; g < t.n; g += gridDim.x * 256) {
        double e, u, v;
        gauge_value(d, t, g, e, u, v);
        out[g] = e;
        out[t.n + g] = u;
        out[2 * t.n + g] = v;
    }
}

// After a step: slot p holds the state's clock.  A record is appended when
// the step ran and (interval <= 0 or t >= next), and next advances by the
// interval (write_gauges, rki_driver.F90:777-800).  One workgroup: thread 0
// decides, then all threads write the record.
__global__ __launch_bounds__(256) void k_gauges_record(GridDesc d, GaugeTable t, const CtrlSlot* ctrl, int p,
                                                       double interval, double* series, int* count, double* next,
                                                       int cap) {
    __shared__ int slot;
    if (threadIdx.x == 0) {
        slot = -1;
        const CtrlSlot c = ctrl[p];
        if (c.active && (interval <= 0.0 || c.t >= *next)) {
            const int k = *count;
            if (k < cap) {   // the host drains before the buffer can fill
                slot = k;
                *count = k + 1;
                series[(long long)k * (1 + 3 * t.n)] = c.t;
            }
            if (interval > 0.0) *next = *next + interval;
        }
    }
    __syncthreads();
    if (slot < 0) return;
    double* rec = series + (long long)slot * (1 + 3 * t.n) + 1;
    for (int g = threadIdx.x; g < t.n; g += 256) {
        double e, u, v;
        gauge_value(d, t, g, e, u, v);
        rec[g] = e;
        rec[t.n + g] = u;
        rec[2 * t.n + g] = v;
    }
}

__global__ void k_gauges_reset(int* count, double* next, const CtrlSlot* ctrl, int p, double interval) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        *count = 0;
        const double t = ctrl[p].t;
        *next = interval > 0.0 ? t + interval : t;   // rki_driver.F90:347-352
    }
}

GaugeTable table_of(const GaugeState* s) {
    return GaugeTable{s->d_cell, s->d_cell + s->n_local, s->d_w, s->d_w + s->n_local, s->n_local};
}

// floor(v / d) + 1, the reference's 1-based cell (compute_local_gauges,
// rki_gauges.F90:228-229), or -1 off the grid (finite for the cast).
int cell_of(double v, double d, int n) {
    const double f = std::floor(v / d);
    if (!(f >= -1.0) || f > (double)n) return -1;
    return (int)f + 1;
}

void free_tables(GaugeState* s) {
    for (void* ptr : {(void*)s->d_cell, (void*)s->d_w, (void*)s->d_now, (void*)s->d_series, (void*)s->d_count,
                      (void*)s->d_next})
        if (ptr) (void)hipFree(ptr);
    s->d_cell = nullptr; s->d_w = nullptr; s->d_now = nullptr;
    s->d_series = nullptr; s->d_count = nullptr; s->d_next = nullptr;
}

}  // namespace

bool station_weights(const rkg_config& c, double x, double y, int* c0, int* j0, double* wx, double* wy) {
    const int ig = cell_of(x, c.dx, c.nx), jg = cell_of(y, c.dy, c.ny);
    if (!(ig >= 1 && ig <= c.nx && jg >= 1 && jg <= c.ny)) return false;
    const double xc = ((double)ig - 0.5) * c.dx, yc = ((double)jg - 0.5) * c.dy;
    double u, v;
    if (x < xc) {
        if (ig > 1) { *c0 = ig - 2; u = (x - (xc - c.dx)) / c.dx; } else { *c0 = ig - 1; u = 0.0; }
    } else {
        *c0 = ig - 1;
        u = ig < c.nx ? (x - xc) / c.dx : 0.0;
    }
    if (y < yc) {
        if (jg > 1) { *j0 = jg - 2; v = (y - (yc - c.dy)) / c.dy; } else { *j0 = jg - 1; v = 0.0; }
    } else {
        *j0 = jg - 1;
        v = jg < c.ny ? (y - yc) / c.dy : 0.0;
    }
    *wx = std::max(0.0, std::min(1.0, u));
    *wy = std::max(0.0, std::min(1.0, v));
    return true;
}

long long gauges_room(const rkg_grid* g) {
    const GaugeState* s = g->gauges;
    if (!s || !s->recording || s->n_local == 0) return 1ll << 40;
    return std::max(0ll, (long long)s->cap - s->pending);
}

int gauges_drain(rkg_grid* g) {
    GaugeState* s = g->gauges;
    if (!s || !s->d_series) return RKG_OK;
    int count = 0;
    HIP_TRY(hipMemcpyAsync(&count, s->d_count, sizeof(int), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    const size_t rec = 1 + 3 * (size_t)s->n_local;
    if (count > 0) {
        const size_t old = s->host.size();
        s->host.resize(old + (size_t)count * rec);
        HIP_TRY(hipMemcpyAsync(s->host.data() + old, s->d_series, (size_t)count * rec * sizeof(double),
                               hipMemcpyDeviceToHost, g->stream));
        const int zero = 0;
        HIP_TRY(hipMemcpyAsync(s->d_count, &zero, sizeof(int), hipMemcpyHostToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        s->host_records += count;
    }
    s->pending = 0;
    return RKG_OK;
}

int gauges_after_step(rkg_grid* g, int p) {
    GaugeState* s = g->gauges;
    if (!s->recording || s->n_local == 0) return RKG_OK;
    const GridDesc d = make_desc(g);
    hipLaunchKernelGGL(k_gauges_record, dim3(1), dim3(256), 0, g->stream, d, table_of(s), g->ctrl, p, s->interval,
                       s->d_series, s->d_count, s->d_next, s->cap);
    s->pending += 1;
    return RKG_OK;
}

void gauges_free(rkg_grid* g) {
    if (!g->gauges) return;
    free_tables(g->gauges);
    delete g->gauges;
    g->gauges = nullptr;
}

}  // namespace rk

using namespace rk;

extern "C" {

int rkg_gauges_set(rkg_grid* g, int n, const double* x, const double* y, int* n_local) {
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    if (n < 0 || n > RKG_MAX_GAUGES || (n > 0 && (!x || !y))) {
        set_error("gauges: need 0 <= n <= %d and coordinate arrays", RKG_MAX_GAUGES);
        return RKG_ERR_RANGE;
    }
    HIP_TRY(hipStreamSynchronize(g->stream));
    gauges_free(g);
    if (n_local) *n_local = 0;
    if (n == 0) return RKG_OK;
    GaugeState* s = new GaugeState();
    g->gauges = s;
    const rkg_config& c = g->cfg;
    std::vector<int> cell;
    std::vector<double> w;
    for (int k = 0; k < n; ++k) {
        int c0, j0;
        double wx, wy;
        if (station_weights(c, x[k], y[k], &c0, &j0, &wx, &wy)) {
            s->local_idx.push_back(k);
            cell.push_back(c0);
            cell.push_back(j0);
            w.push_back(wx);
            w.push_back(wy);
        }
    }
    const int nl = (int)s->local_idx.size();
    s->n_local = nl;
    {   // (c0, j0) pairs and (wx, wy) pairs -> the table's planar layout
        std::vector<int> cp(2 * (size_t)nl);
        std::vector<double> wp(2 * (size_t)nl);
        for (int l = 0; l < nl; ++l) {
            cp[l] = cell[2 * l]; cp[nl + l] = cell[2 * l + 1];
            wp[l] = w[2 * l]; wp[nl + l] = w[2 * l + 1];
        }
        cell.swap(cp);
        w.swap(wp);
    }
    if (nl > 0) {
        HIP_TRY(hipMalloc((void**)&s->d_cell, 2 * (size_t)nl * sizeof(int)));
        HIP_TRY(hipMalloc((void**)&s->d_w, 2 * (size_t)nl * sizeof(double)));
        HIP_TRY(hipMalloc((void**)&s->d_now, 3 * (size_t)nl * sizeof(double)));
        HIP_TRY(hipMemcpy(s->d_cell, cell.data(), 2 * (size_t)nl * sizeof(int), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(s->d_w, w.data(), 2 * (size_t)nl * sizeof(double), hipMemcpyHostToDevice));
    }
    if (n_local) *n_local = nl;
    return RKG_OK;
}

int rkg_gauges_local(rkg_grid* g, int* local_idx) {
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!g->gauges) return RKG_OK;
    if (!local_idx) return RKG_ERR_RANGE;
    std::copy(g->gauges->local_idx.begin(), g->gauges->local_idx.end(), local_idx);
    return RKG_OK;
}

int rkg_gauges_sample(rkg_grid* g, double* eta, double* u, double* v) {
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    GaugeState* s = g->gauges;
    if (!s || s->n_local == 0) return RKG_OK;
    if (!eta || !u || !v) return RKG_ERR_RANGE;
    const GridDesc d = make_desc(g);
    const int n = s->n_local;
    hipLaunchKernelGGL(k_gauges_now, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, g->stream, d, table_of(s),
                       s->d_now);
    HIP_TRY(hipGetLastError());
    std::vector<double> h(3 * (size_t)n);
    HIP_TRY(hipMemcpyAsync(h.data(), s->d_now, h.size() * sizeof(double), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    std::copy(h.begin(), h.begin() + n, eta);
    std::copy(h.begin() + n, h.begin() + 2 * n, u);
    std::copy(h.begin() + 2 * n, h.end(), v);
    return RKG_OK;
}

int rkg_gauges_record(rkg_grid* g, double interval, int capacity) {
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    GaugeState* s = g->gauges;
    if (!s) { set_error("gauges: no stations set"); return RKG_ERR_RANGE; }
    if (capacity < 0 || !std::isfinite(interval)) { set_error("gauges: bad interval or capacity"); return RKG_ERR_RANGE; }
    HIP_TRY(hipStreamSynchronize(g->stream));
    if (s->d_series) { (void)hipFree(s->d_series); s->d_series = nullptr; }
    if (s->d_count) { (void)hipFree(s->d_count); s->d_count = nullptr; }
    if (s->d_next) { (void)hipFree(s->d_next); s->d_next = nullptr; }
    s->host.clear();
    s->host_records = 0;
    s->pending = 0;
    s->recording = capacity > 0 && s->n_local > 0;
    if (!s->recording) return RKG_OK;
    s->interval = interval;
    s->cap = capacity;
    HIP_TRY(hipMalloc((void**)&s->d_series, (size_t)capacity * (1 + 3 * (size_t)s->n_local) * sizeof(double)));
    HIP_TRY(hipMalloc((void**)&s->d_count, sizeof(int)));
    HIP_TRY(hipMalloc((void**)&s->d_next, sizeof(double)));
    hipLaunchKernelGGL(k_gauges_reset, dim3(1), dim3(64), 0, g->stream, s->d_count, s->d_next, g->ctrl, g->parity,
                       interval);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    return RKG_OK;
}

int rkg_gauges_series(rkg_grid* g, long long max_records, long long* n_records, double* t, double* eta, double* u,
                      double* v) {
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    if (!n_records) return RKG_ERR_RANGE;
    *n_records = 0;
    GaugeState* s = g->gauges;
    if (!s) return RKG_OK;
    int rc = gauges_drain(g);
    if (rc) return rc;
    const long long have = s->host_records;
    if (!t || max_records == 0) { *n_records = have; return RKG_OK; }
    if (!eta || !u || !v || max_records < 0) return RKG_ERR_RANGE;
    const long long k = std::min(have, max_records);
    const size_t n = (size_t)s->n_local, rec = 1 + 3 * n;
    for (long long r = 0; r < k; ++r) {
        const double* src = s->host.data() + (size_t)r * rec;
        t[r] = src[0];
        std::copy(src + 1, src + 1 + n, eta + (size_t)r * n);
        std::copy(src + 1 + n, src + 1 + 2 * n, u + (size_t)r * n);
        std::copy(src + 1 + 2 * n, src + rec, v + (size_t)r * n);
    }
    s->host.erase(s->host.begin(), s->host.begin() + (size_t)k * rec);
    s->host_records -= k;
    *n_records = k;
    return RKG_OK;
}

}  // extern "C"
