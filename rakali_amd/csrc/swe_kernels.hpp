// swe_kernels.hpp -- device-side data layout, ghost-cell synthesis and the
// strip-marching tile routine shared by the single-grid stage kernels.
//
// Layout in HBM: structure of arrays, one FP64 plane per field, row-major
// with x fastest (the reference's column-major (i, j) order), interior cells
// only, row pitch `pitch` elements (a multiple of 16, i.e. 128-byte rows).
// State buffers: A = U^n (h, hu, hv; overwritten in place by U^{n+1}),
// B = U* (stage-1 output), and the bed b.  Ghost cells are never stored:
// the tile loader synthesizes them from the boundary tags exactly as
// apply_boundary_conditions (rki_boundary.F90:40-470) would have filled
// them for the rows and columns the cross-shaped stencil reads.
#pragma once

#include "swe_core.hpp"

namespace rk {

constexpr int WAVE = 64;        // wavefront width on CDNA
constexpr int MAX_TC = 10;      // MAX_TIDAL_CONSTITUENTS

// Time-control slot.  Two slots ping-pong: a step reads slot p (state at
// the start of the step) and its stage-1 kernel writes slot 1-p.
struct alignas(64) CtrlSlot {
    double t;                     // time at the start of the step
    double dt;                    // dt of the step that produced this slot
    unsigned long long speed_bits;// max CFL speed of the state (as bits)
    long long step_count;         // state%step_count
    int active;                   // 1 if the step that wrote this slot ran
    int pad[5];
};

struct GridDesc {
    int nx, ny;
    long long pitch;              // elements per row
    double dx, dy;
    double* ah; double* ahu; double* ahv;   // A: U^n / U^{n+1}
    double* bh; double* bhu; double* bhv;   // B: U*
    const double* bed;
    double g, cfl, manning_n, coriolis_f;
    int use_hllc;
    int bc[4];                    // west, east, south, north
    double tidal_amplitude, tidal_period;
    int n_tc;
    double tc_amp[MAX_TC], tc_phase[MAX_TC], tc_omega[MAX_TC];
    double inflow_velocity, discharge_flux, clamped_elevation;
    double dt_fixed, t_end;
    int cfl_interval;
    int ty;                       // rows per tile segment (<= 64)
    int nstrips, nsegs;
    CtrlSlot* ctrl;               // [2]
};

// Boundary data resolved once per launch.
struct EdgeEta {
    double tidal;                 // tidal surface at this stage's time
};

// Tidal surface, rki_boundary.F90:92-101.
__device__ __forceinline__ double tidal_eta(const GridDesc& d, double time) {
    double e = 0.0;
    if (d.n_tc > 0) {
        for (int n = 0; n < d.n_tc; ++n) e = e + d.tc_amp[n] * crm::cos(d.tc_omega[n] * time - d.tc_phase[n]);
    } else if (d.tidal_period > 0.0) {
        e = d.tidal_amplitude * crm::sin(TWO_PI * time / d.tidal_period);
    }
    return e;
}

__device__ __forceinline__ Cell load_direct(const double* __restrict__ h, const double* __restrict__ hu,
                                            const double* __restrict__ hv, const double* __restrict__ b,
                                            long long pitch, int c, int j) {
    const long long k = (long long)j * pitch + c;
    Cell r;
    r.h = h[k]; r.hu = hu[k]; r.hv = hv[k]; r.b = b[k];
    return r;
}

// Ghost value for an edge, from the interior cells along its normal.
// `nrm_sign` is -1 on W/S and +1 on E/N; `normal_is_u` on W/E.
// kk = ghost layer (0 = adjacent), mir = mirror cell for wall edges,
// in0 = first interior cell.  rki_boundary.F90:220-466.
__device__ __forceinline__ Cell ghost_from(const GridDesc& d, int tag, double eta_tidal,
                                           const Cell& mir, const Cell& in0, double nrm_sign,
                                           bool normal_is_u) {
    Cell g;
    switch (tag) {
    case BC_WALL:
    case BC_SPONGE: {   // bc_wall_edge :281-289 (bed mirrored first)
        g.b = mir.b;
        g.h = fmax0(mir.h + mir.b - g.b);
        if (normal_is_u) { g.hu = -mir.hu; g.hv = mir.hv; }
        else { g.hv = -mir.hv; g.hu = mir.hu; }
        return g;
    }
    case BC_OPEN:
    case BC_TIDAL: {    // bc_flather_edge :331-343; bed ghost = first interior
        const bool prescribe = (tag == BC_TIDAL);
        const double eta_ext = prescribe ? eta_tidal : 0.0;
        g.b = in0.b;
        const double h_int = in0.h;
        const double eta_int = h_int + in0.b;
        const double hmax = dmax(h_int, 1.0e-6);
        const double speed = sqrt(d.g * hmax);
        const double h_ghost = prescribe ? fmax0(eta_ext - g.b) : h_int;
        g.h = h_ghost;
        const double nrm = normal_is_u ? in0.hu : in0.hv;
        const double rad = h_ghost * (nrm / hmax + nrm_sign * speed * (eta_int - eta_ext) / hmax);
        if (normal_is_u) { g.hu = rad; g.hv = in0.hv; }
        else { g.hv = rad; g.hu = in0.hu; }
        return g;
    }
    case BC_INFLOW:
    case BC_DISCHARGE: {  // bc_inflow_edge :377-387, bc_disch_edge :418-428
        const double in_sgn = -nrm_sign;
        g.b = in0.b;
        g.h = in0.h;
        const double nrm = (tag == BC_INFLOW) ? in_sgn * in0.h * d.inflow_velocity
                                              : in_sgn * d.discharge_flux;
        if (normal_is_u) { g.hu = nrm; g.hv = in0.hv; }
        else { g.hv = nrm; g.hu = in0.hu; }
        return g;
    }
    case BC_CLAMPED:
    default: {          // bc_clamped_edge :458-464
        g.b = in0.b;
        g.h = fmax0(d.clamped_elevation - g.b);
        g.hu = in0.hu;
        g.hv = in0.hv;
        return g;
    }
    }
}

// Cell (c, j) of the state (h, hu, hv) with bed d.bed, synthesizing the
// two ghost layers around the domain.  Corner cells and cells beyond the
// second ghost layer are never read by the stencil; they return zeros.
template <bool EDGE>
__device__ __forceinline__ Cell load_cell(const GridDesc& d, const double* __restrict__ h,
                                          const double* __restrict__ hu, const double* __restrict__ hv,
                                          int c, int j, double eta_tidal) {
    if (!EDGE) return load_direct(h, hu, hv, d.bed, d.pitch, c, j);
    const bool xin = (unsigned)c < (unsigned)d.nx;
    const bool yin = (unsigned)j < (unsigned)d.ny;
    if (xin && yin) return load_direct(h, hu, hv, d.bed, d.pitch, c, j);
    Cell z; z.h = 0.0; z.hu = 0.0; z.hv = 0.0; z.b = 0.0;
    if (!xin && !yin) return z;
    if (!xin) {
        const bool west = c < 0;
        const int kk = west ? (-c - 1) : (c - d.nx);
        if (kk > 1) return z;
        const int mc = west ? kk : (d.nx - 1 - kk);
        const int ic = west ? 0 : (d.nx - 1);
        const Cell mir = load_direct(h, hu, hv, d.bed, d.pitch, mc, j);
        const Cell in0 = load_direct(h, hu, hv, d.bed, d.pitch, ic, j);
        return ghost_from(d, d.bc[west ? 0 : 1], eta_tidal, mir, in0, west ? -1.0 : 1.0, true);
    } else {
        const bool south = j < 0;
        const int kk = south ? (-j - 1) : (j - d.ny);
        if (kk > 1) return z;
        const int mj = south ? kk : (d.ny - 1 - kk);
        const int ij = south ? 0 : (d.ny - 1);
        const Cell mir = load_direct(h, hu, hv, d.bed, d.pitch, c, mj);
        const Cell in0 = load_direct(h, hu, hv, d.bed, d.pitch, c, ij);
        return ghost_from(d, d.bc[south ? 2 : 3], eta_tidal, mir, in0, south ? -1.0 : 1.0, false);
    }
}

// Workgroup id -> tile id so that each XCD (workgroups are dispatched to
// the 8 XCDs round-robin) owns one contiguous band of tiles and the halo
// rows and columns it shares stay in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int x = b & 7, k = b >> 3;
    const int per = n >> 3, rem = n & 7;
    return x * per + (x < rem ? x : rem) + k;
}

__device__ __forceinline__ double shfl_down1(double v) { return __shfl_down(v, 1, WAVE); }

__device__ __forceinline__ Face shfl_down1(const Face& f) {
    Face r;
    r.fh = shfl_down1(f.fh); r.fhu = shfl_down1(f.fhu); r.fhv = shfl_down1(f.fhv);
    r.a = shfl_down1(f.a); r.c = shfl_down1(f.c); r.d = shfl_down1(f.d);
    return r;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off, WAVE));
    return v;
}

struct TileLds {
    double rh[WAVE + 3], rhu[WAVE + 3], rhv[WAVE + 3], rb[WAVE + 3];
    double bnd[6][WAVE];
};

struct StageCtx {
    const double* ih; const double* ihu; const double* ihv;   // stage input
    double* oh; double* ohu; double* ohv;                     // stage output
    double dt;
    double eta_tidal;
    SourceTerms src;
};

// One 64-column x d.ty-row tile of one RK stage, processed by one
// wavefront marching down the rows.  Lane l owns column c = c0 + l.
// Per row it evaluates the x-face on its left (columns c-2..c+1 read from
// an LDS copy of the row) and the y-face above it (rows j-1..j+2 held in a
// register window); the face below is carried from the previous row and
// the face right of lane 63 comes from a batch evaluated up front, so every
// face of the tile is evaluated exactly once.
// STAGE 1 writes U* to B; STAGE 2 reads U* and U^n, applies the RK2 average
// and writes U^{n+1} over A, returning the lane's max CFL speed.
template <int STAGE, bool HLLC, bool EDGE>
__device__ __forceinline__ double process_tile(const GridDesc& d, const StageCtx& s, TileLds& L,
                                               int strip, int seg) {
    const int lane = threadIdx.x;
    const int c0 = strip * WAVE;
    const int c = c0 + lane;
    const int j0 = seg * d.ty;
    const int j1 = min(j0 + d.ty, d.ny);
    const int nrows = j1 - j0;
    const double g = d.g;
    const long long pitch = d.pitch;
    double smax = 0.0;

    // (1) faces right of column c0+63, one row per lane.
    if (c0 + WAVE - 1 < d.nx && lane < nrows) {
        const int j = j0 + lane;
        const Cell m2 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c0 + 62, j, s.eta_tidal);
        const Cell m1 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c0 + 63, j, s.eta_tidal);
        const Cell p0 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c0 + 64, j, s.eta_tidal);
        const Cell p1 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c0 + 65, j, s.eta_tidal);
        const Face f = face<HLLC, false>(g, m2, m1, p0, p1);
        L.bnd[0][lane] = f.fh; L.bnd[1][lane] = f.fhu; L.bnd[2][lane] = f.fhv;
        L.bnd[3][lane] = f.a; L.bnd[4][lane] = f.c; L.bnd[5][lane] = f.d;
    }

    // (2) register window and the face below the first row.
    Cell w0 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j0 - 2, s.eta_tidal);
    Cell w1 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j0 - 1, s.eta_tidal);
    Cell w2 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j0, s.eta_tidal);
    Cell w3 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j0 + 1, s.eta_tidal);
    Face fyb = face<HLLC, true>(g, w0, w1, w2, w3);
    w0 = w1; w1 = w2; w2 = w3;
    w3 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j0 + 2, s.eta_tidal);

    // halo columns of the LDS row: lane 0 -> c0-2, 1 -> c0-1, 2 -> c0+64
    const int hcol = lane == 0 ? c0 - 2 : (lane == 1 ? c0 - 1 : c0 + WAVE);
    const int hidx = lane == 0 ? 0 : (lane == 1 ? 1 : WAVE + 2);
    Cell halo;
    if (lane < 3) halo = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, hcol, j0, s.eta_tidal);
    Cell un;  // U^n at (c, j) for the RK2 average
    const bool own = c < d.nx;
    if (STAGE == 2 && own) {
        const long long k = (long long)j0 * pitch + c;
        un.h = d.ah[k]; un.hu = d.ahu[k]; un.hv = d.ahv[k];
    }
    __syncthreads();

    for (int j = j0; j < j1; ++j) {
        // prefetch for row j+1
        const bool more = j + 1 < j1;
        Cell nw3, nhalo, nun;
        if (more) {
            nw3 = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j + 3, s.eta_tidal);
            if (lane < 3) nhalo = load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, hcol, j + 1, s.eta_tidal);
            if (STAGE == 2 && own) {
                const long long k = (long long)(j + 1) * pitch + c;
                nun.h = d.ah[k]; nun.hu = d.ahu[k]; nun.hv = d.ahv[k];
            }
        }

        // row j into LDS: own column at lane+2, halo columns from lanes 0..2
        L.rh[lane + 2] = w1.h; L.rhu[lane + 2] = w1.hu; L.rhv[lane + 2] = w1.hv; L.rb[lane + 2] = w1.b;
        if (lane < 3) {
            L.rh[hidx] = halo.h; L.rhu[hidx] = halo.hu; L.rhv[hidx] = halo.hv; L.rb[hidx] = halo.b;
        }
        __syncthreads();
        Cell xm2, xm1, xp1;
        xm2.h = L.rh[lane]; xm2.hu = L.rhu[lane]; xm2.hv = L.rhv[lane]; xm2.b = L.rb[lane];
        xm1.h = L.rh[lane + 1]; xm1.hu = L.rhu[lane + 1]; xm1.hv = L.rhv[lane + 1]; xm1.b = L.rb[lane + 1];
        xp1.h = L.rh[lane + 3]; xp1.hu = L.rhu[lane + 3]; xp1.hv = L.rhv[lane + 3]; xp1.b = L.rb[lane + 3];
        const int r = j - j0;
        Face bnd;
        bnd.fh = L.bnd[0][r]; bnd.fhu = L.bnd[1][r]; bnd.fhv = L.bnd[2][r];
        bnd.a = L.bnd[3][r]; bnd.c = L.bnd[4][r]; bnd.d = L.bnd[5][r];
        __syncthreads();

        const Face fxl = face<HLLC, false>(g, xm2, xm1, w1, xp1);
        const Face fyt = face<HLLC, true>(g, w0, w1, w2, w3);
        Face fxr = shfl_down1(fxl);
        if (lane == WAVE - 1) fxr = bnd;

        if (own) {
            double fh, fhu, fhv;
            divergence(g, d.dx, d.dy, fxl, fxr, fyb, fyt, fh, fhu, fhv);
            double h = w1.h, hu = w1.hu, hv = w1.hv;
            stage_body(s.src, s.dt, fh, fhu, fhv, h, hu, hv);
            if (STAGE == 2) {
                // rki_kernel_fused.F90:435-463
                h = 0.5 * (un.h + h);
                hu = 0.5 * (un.hu + hu);
                hv = 0.5 * (un.hv + hv);
                clamp_desing(g, h, hu, hv);
                smax = dmax(smax, cfl_speed(g, d.dx, d.dy, h, hu, hv));
            }
            const long long k = (long long)j * pitch + c;
            s.oh[k] = h; s.ohu[k] = hu; s.ohv[k] = hv;
        }

        fyb = fyt;
        w0 = w1; w1 = w2; w2 = w3;
        if (more) {
            w3 = nw3;
            if (lane < 3) halo = nhalo;
            if (STAGE == 2) un = nun;
        }
    }
    return smax;
}

}  // namespace rk
