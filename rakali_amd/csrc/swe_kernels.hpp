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
// Tiles a stage launch needs before the fast path (swe_core.hpp) pays: its
// wave-uniform tests and fallbacks cost latency, which a launch of a few
// waves per SIMD cannot hide (at 2048 tiles, 1024^2 cells, the stage
// kernels are 5 % slower with it; at 16384, 4096^2, 5 % faster).
constexpr long long FAST_MIN_TILES = 6144;

// Time-control slot.  Two slots ping-pong: a step reads slot p (state at
// the start of the step) and its stage-1 kernel writes slot 1-p.
struct alignas(64) CtrlSlot {
    double t;                     // time at the start of the step
    double dt;                    // dt of the step that produced this slot
    unsigned long long speed_bits;// max CFL speed of the state (as bits)
    long long step_count;         // state%step_count
    double mass_out;              // boundary mass budget (rki_solver.F90:507-537)
    int active;                   // 1 if the step that wrote this slot ran
    int pad;
    double mass_src;              // surface mass source (accumulate_mass_src, rki_solver.F90:539-606)
    double t_bc;                  // bc%time of the step's last boundary call
};

// Optional source terms of the stage kernels (template bit mask).
constexpr int FEAT_FORCING = 1;   // surface forcing, rki_kernel_fused.F90:268-289
constexpr int FEAT_SPONGE = 2;    // sponge relaxation, rki_kernel_sponge.F90:16-88
constexpr int FEAT_CHAPMAN = 4;   // Chapman radiation edges, rki_boundary.F90:471-552

struct GridDesc {
    int nx, ny;
    long long pitch;              // elements per row
    double dx, dy;
    double* ah; double* ahu; double* ahv;   // A: U^n / U^{n+1}
    double* bh; double* bhu; double* bhv;   // B: U*
    double* sh; double* shu; double* shv;   // optional copy of U^n (RK save), or null
    double* o2h; double* o2hu; double* o2hv; // stage-2 output (normally A itself)
    const double* bed;
    const double* nfield;         // per-cell Manning n, or null (scalar manning_n)
    double* rec;                  // 3 planes receiving each cell's flux divergence, or null
    double* mpart;                // per-tile sums of flux_h, [stage 1 tiles | stage 2 tiles], or null
    double* hist;                 // recording stage 1: (t, dt, dt recomputed) of the step, or null
    const double* fwx; const double* fwy;   // surface forcing planes (FEAT_FORCING kernels): wind
    const double* fdpx; const double* fdpy; // stress, the central pressure gradient (from the
    const double* frain;                    // pressure field, see k_pressure_gradient), precipitation
    double* spart;                // per-tile sums of the sponge's h change, [stage 1 | stage 2], or null
    int sponge_width;             // sponge relaxation band (FEAT_SPONGE kernels), bc_params_t
    double sponge_strength, sponge_h_ref;
    long long plane;              // elements per plane (pitch * ny)
    double ssp_coeff;             // stable-dt multiplier of the time scheme (rki_timestepper.F90)
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
    int fast;                     // the launch may take the fast path (FAST_MIN_TILES)
    int nstrips, nsegs;
    CtrlSlot* ctrl;               // [2]
};

// Boundary data resolved once per launch.
struct EdgeEta {
    double tidal;                 // tidal surface at this stage's time
};

// Tidal surface, rki_boundary.F90:92-101.
__host__ __device__ __forceinline__ double tidal_eta(const GridDesc& d, double time) {
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
__host__ __device__ __forceinline__ Cell ghost_from(const GridDesc& d, int tag, double eta_tidal,
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
__device__ __forceinline__ double shfl_up1(double v) { return __shfl_up(v, 1, WAVE); }

__device__ __forceinline__ Face shfl_down1(const Face& f) {
    Face r;
    r.fh = shfl_down1(f.fh); r.fhu = shfl_down1(f.fhu); r.fhv = shfl_down1(f.fhv);
    r.a = shfl_down1(f.a); r.c = shfl_down1(f.c); r.d = shfl_down1(f.d);
    return r;
}
__device__ __forceinline__ Cell shfl_up1(const Cell& x) {
    return Cell{shfl_up1(x.h), shfl_up1(x.hu), shfl_up1(x.hv), shfl_up1(x.b)};
}
__device__ __forceinline__ Cell shfl_down1(const Cell& x) {
    return Cell{shfl_down1(x.h), shfl_down1(x.hu), shfl_down1(x.hv), shfl_down1(x.b)};
}
__device__ __forceinline__ FState shfl_up1(const FState& x) {
    return FState{shfl_up1(x.h), shfl_up1(x.hu), shfl_up1(x.hv)};
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmax(v, __shfl_xor(v, off, WAVE));
    return v;
}

// Butterfly sum: every lane ends with the same bits (each pairwise add sees
// the same two operands), in an order fixed by the lane numbering.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_xor(v, off, WAVE);
    return v;
}

// Per-segment data of the two edge columns of a strip, one row per entry:
// raw cells c0-1 and c0+64, the high x-state of c0-1, and the x-face
// between c0+63 and c0+64.
struct TileLds {
    double hl[4][WAVE];
    double hr[4][WAVE];
    double ph[3][WAVE];
    double pf[6][WAVE];
};

struct StageCtx {
    const double* ih; const double* ihu; const double* ihv;   // stage input
    double* oh; double* ohu; double* ohv;                     // stage output
    double dt;
    double eta_tidal;
    Recip rdx, rdy;
    SourceTerms src;
    // Chapman edges (FEAT_CHAPMAN): whether this boundary call radiates
    // (its state differs from the previous call's), the previous call's
    // state prev_h (its eta_old = prev_h + b at the first interior cell),
    // the step's dt (bc%dt) and the tile's own rows
    bool radiate;
    const double* prev_h;
    double bc_dt;
    int row0, row1;
    // generic SSP stage only (STAGE 3)
    double alpha, a_prev;         // blend weights on U^n and on the Euler iterate
    bool blend, first, last;
};

// One stage of a generic explicit SSP scheme (rki_solver.F90:861-925):
// E = U^(i-1) + beta_i dt L(U^(i-1)), then U^(i) = alpha_i U^n +
// (1 - alpha_i) E when alpha_i != 0.  U^n stays in A until the last stage,
// which writes A.
struct GenStage {
    const double* ih; const double* ihu; const double* ihv;
    double* oh; double* ohu; double* ohv;
    const double* ph;             // h of the previous stage's input (Chapman eta_old), or null
    double alpha, beta, c;        // rki_timestepper.F90:53-58
    double c_last;                // c of the scheme's last stage (bc%time of its last call)
    int first, last;
};

// Ghost cell (c, j) of a Chapman edge (bc_chapman_edge, rki_boundary.F90:
// 471-552).  The reference calls the boundary before the dt computation
// (rki_solver.F90:718-720) and before each stage; each call sets eta_old
// to the interior eta of the state it sees (on the k = 1 layer, before the
// k = 2 layer of the same t reads it: t outer, k inner, the serial order of
// the reference's CPU build).  So
// * the first stage of a step sees eta_old = its own interior eta: the
//   ghost surface is the interior eta (the radiation term vanishes);
// * a later stage radiates on the adjacent layer against the previous
//   call's state, U^n for SSP-RK2 stage 2, the previous stage's input for
//   the generic schemes, with the step's dt; the outer layer again sees
//   eta_old = its interior eta.
// The bed ghost is the first interior bed (fill_bathymetry_ghosts_array).
// x-edge ghosts of rows outside the tile's own are never used (the y
// pipeline loads them for lanes past the grid); they skip the radiation,
// whose eta_old the owning tile may be overwriting in place.
__device__ __forceinline__ Cell chapman_ghost(const GridDesc& d, const StageCtx& s, int side, int c, int j) {
    const bool xedge = side < 2, lo = (side & 1) == 0;
    const int kk = xedge ? (lo ? -c - 1 : c - d.nx) : (lo ? -j - 1 : j - d.ny);
    Cell g;
    g.h = 0.0; g.hu = 0.0; g.hv = 0.0; g.b = 0.0;
    if (kk > 1) return g;
    int ci = c, ji = j, c2 = c, j2 = j;
    if (xedge) { ci = lo ? 0 : d.nx - 1; c2 = lo ? 1 : d.nx - 2; }
    else { ji = lo ? 0 : d.ny - 1; j2 = lo ? 1 : d.ny - 2; }
    const Cell in0 = load_direct(s.ih, s.ihu, s.ihv, d.bed, d.pitch, ci, ji);
    const double h_int = in0.h;
    const double eta_int = h_int + in0.b;
    const double hmax = dmax(h_int, 1.0e-6);
    const double speed = sqrt(d.g * hmax);
    double eta_g = eta_int;
    const bool own_row = !xedge || (j >= s.row0 && j < s.row1);
    if (kk == 0 && s.radiate && s.bc_dt > 0.0 && own_row) {
        const long long k0 = (long long)ji * d.pitch + ci;
        const long long k2 = (long long)j2 * d.pitch + c2;
        const double dd = xedge ? d.dx : d.dy;
        const double eta_old = s.prev_h[k0] + d.bed[k0];
        const double deta_dt = (eta_int - eta_old) / s.bc_dt;
        const double deta_dx = (s.ih[k2] + d.bed[k2] - eta_int) / dd;
        double cp = fabs(deta_dx) > 1.0e-10 ? -deta_dt / deta_dx : 0.0;
        cp = dmax(0.0, fmin(cp, speed));
        eta_g = eta_old - cp * s.bc_dt / dd * (eta_old - eta_int);
    }
    g.b = in0.b;
    g.h = fmax0(eta_g - g.b);
    const double sgn = lo ? -1.0 : 1.0;
    const double nrm = xedge ? in0.hu : in0.hv;
    const double rad = g.h * (nrm / hmax + sgn * speed * (eta_int - eta_g) / hmax);
    if (xedge) { g.hu = rad; g.hv = in0.hv; }
    else { g.hv = rad; g.hu = in0.hu; }
    return g;
}

template <bool EDGE, int FEAT = 0>
__device__ __forceinline__ Cell ld(const GridDesc& d, const StageCtx& s, int c, int j) {
    if constexpr (EDGE && (FEAT & FEAT_CHAPMAN) != 0) {
        const bool xin = (unsigned)c < (unsigned)d.nx, yin = (unsigned)j < (unsigned)d.ny;
        if (xin != yin) {
            const int side = !xin ? (c < 0 ? 0 : 1) : (j < 0 ? 2 : 3);
            if (d.bc[side] == BC_CHAPMAN) return chapman_ghost(d, s, side, c, j);
        }
    }
    return load_cell<EDGE>(d, s.ih, s.ihu, s.ihv, c, j, s.eta_tidal);
}

template <bool B>
struct ExactTag {
    static constexpr bool value = B;
};

// One 64-column x d.ty-row tile of one RK stage, processed by one
// wavefront marching down the rows; lane l owns column c = c0 + l.
// Row j: the y-slope of row j+1 is formed from the raw rows j, j+1, j+2
// held in registers, giving the face above row j together with the top
// state of row j carried from the previous row; the x-slope of row j takes
// the neighbouring columns by wave shuffles, the face left of each cell is
// evaluated by its lane and the face right of it is the next lane's.  The
// edge columns come from a per-segment batch (one row per lane) held in
// LDS.  Every slope and every face of the tile is evaluated once.
// A row whose stencil holds a non-zero value outside [2^-300, 2^40] in
// magnitude, or a non-finite one (see tiny()), anywhere in the wavefront is
// evaluated with plain divisions; every other row with the shared-reciprocal
// quotients, its faces and cell updates on the fast path of swe_core.hpp
// where the whole wavefront meets its assumptions.  All give the same bits
// where they apply.
// STAGE 1 writes U* to B; STAGE 2 reads U* and U^n, applies the RK2 average
// and writes U^{n+1} over A, returning the lane's max CFL speed.
// apply_sponge (rki_kernel_sponge.F90:16-88) on one interior cell (c, j),
// 0-based: the distance to the nearest BC_SPONGE edge closer than the
// width sets alpha = strength ((width - d) / width)^2.  dsum collects the
// change of h.
__device__ __forceinline__ void sponge_cell(const GridDesc& d, double dt, int c, int j, double& h,
                                            double& hu, double& hv, double& dsum) {
    const int sw = d.sponge_width;
    int dmin = sw + 1;
    const int dw = c, de = d.nx - 1 - c, ds = j, dn = d.ny - 1 - j;
    if (d.bc[0] == BC_SPONGE && dw < sw && dw < dmin) dmin = dw;
    if (d.bc[1] == BC_SPONGE && de < sw && de < dmin) dmin = de;
    if (d.bc[2] == BC_SPONGE && ds < sw && ds < dmin) dmin = ds;
    if (d.bc[3] == BC_SPONGE && dn < sw && dn < dmin) dmin = dn;
    if (dmin < sw) {
        const double ratio = (double)(sw - dmin) / (double)sw;
        const double alpha = d.sponge_strength * ratio * ratio;
        const double h_in = h;
        h = h - dt * alpha * (h - d.sponge_h_ref);
        hu = hu - dt * alpha * hu;
        hv = hv - dt * alpha * hv;
        dsum = dsum + (h - h_in);
    }
}

template <int STAGE, bool HLLC, bool EDGE, bool NF, bool REC, int FEAT = 0, bool FASTK = false>
__device__ __forceinline__ double process_tile(const GridDesc& d, const StageCtx& s, TileLds& L,
                                               int strip, int seg, double& fsum, double& ssum) {
    constexpr bool FRC = (FEAT & FEAT_FORCING) != 0;
    constexpr bool SPG = (FEAT & FEAT_SPONGE) != 0;
    const int lane = threadIdx.x;
    const int c0 = strip * WAVE;
    const int c = c0 + lane;
    const int j0 = seg * d.ty;
    const int j1 = min(j0 + d.ty, d.ny);
    const int nrows = j1 - j0;
    const double g = d.g;
    const long long pitch = d.pitch;
    // the fast path's ranges assume a gravity of ordinary size
    const bool try_fast = FASTK && g >= 1.0e-3 && g <= 1.0e6;
    // the value test of the rows: kernels with the fast path also keep its
    // range arguments (unbounded()), the others the division's (tiny())
    auto odd = [](const Cell& x) { return FASTK ? unbounded(x) : tiny(x); };
    double smax = 0.0;
    double facc = 0.0;  // sum of the lane's flux_h, rows in order
    double sacc = 0.0;  // sum of the lane's sponge change of h
    const bool read_un = STAGE == 2 || (STAGE == 3 && s.blend);

    // (1) edge columns of the segment, one row per lane
    bool edge_tiny = false;
    if (lane < nrows) {
        const int j = j0 + lane;
        const Cell a = ld<EDGE, FEAT>(d, s, c0 - 2, j);
        const Cell bl = ld<EDGE, FEAT>(d, s, c0 - 1, j);
        const Cell cc = ld<EDGE, FEAT>(d, s, c0, j);
        FState lo, hi;
        cell_states(a, bl, cc, lo, hi);
        L.hl[0][lane] = bl.h; L.hl[1][lane] = bl.hu; L.hl[2][lane] = bl.hv; L.hl[3][lane] = bl.b;
        L.ph[0][lane] = hi.h; L.ph[1][lane] = hi.hu; L.ph[2][lane] = hi.hv;
        const Cell m1 = ld<EDGE, FEAT>(d, s, c0 + 63, j);
        const Cell p0 = ld<EDGE, FEAT>(d, s, c0 + 64, j);
        L.hr[0][lane] = p0.h; L.hr[1][lane] = p0.hu; L.hr[2][lane] = p0.hv; L.hr[3][lane] = p0.b;
        edge_tiny = odd(a) || odd(bl) || odd(cc) || odd(m1) || odd(p0);
        if (c0 + WAVE - 1 < d.nx) {
            const Cell m2 = ld<EDGE, FEAT>(d, s, c0 + 62, j);
            const Cell p1 = ld<EDGE, FEAT>(d, s, c0 + 65, j);
            FState lo1, hi1, lo2, hi2;
            cell_states(m2, m1, p0, lo1, hi1);
            cell_states(m1, p0, p1, lo2, hi2);
            const Face f = face_flux<HLLC, false, true>(g, hi1, m1.b, lo2, p0.b);
            L.pf[0][lane] = f.fh; L.pf[1][lane] = f.fhu; L.pf[2][lane] = f.fhv;
            L.pf[3][lane] = f.a; L.pf[4][lane] = f.c; L.pf[5][lane] = f.d;
        }
    }
    // rows of the segment whose edge cells hold tiny values
    const unsigned long long edge_rows = __ballot(edge_tiny);

    // (2) y pipeline: top state of row j0-1, face below row j0
    Cell rc, rn, rnn;
    FState tc;
    Face fyb;
    bool t_prev, t_rc, t_rn, t_rnn;
    {
        const Cell a = ld<EDGE, FEAT>(d, s, c, j0 - 2);
        const Cell bm = ld<EDGE, FEAT>(d, s, c, j0 - 1);
        rc = ld<EDGE, FEAT>(d, s, c, j0);
        rn = ld<EDGE, FEAT>(d, s, c, j0 + 1);
        rnn = ld<EDGE, FEAT>(d, s, c, j0 + 2);
        FState lo, tm;
        cell_states(a, bm, rc, lo, tm);
        FState bc;
        cell_states(bm, rc, rn, bc, tc);
        fyb = face_flux<HLLC, true, true>(g, tm, bm.b, bc, rc.b);
        t_prev = odd(bm);
        t_rc = odd(rc);
        t_rn = odd(rn);
        t_rnn = odd(rnn);
    }
    const bool own = c < d.nx;
    Cell un;   // U^n at (c, j) for the RK2 average
    un.h = 0.0; un.hu = 0.0; un.hv = 0.0; un.b = 0.0;
    if (!FASTK && read_un && own) {
        const long long k = (long long)j0 * pitch + c;
        un.h = d.ah[k]; un.hu = d.ahu[k]; un.hv = d.ahv[k];
    }
    __syncthreads();

    for (int j = j0; j < j1; ++j) {
        const int r = j - j0;
        const bool more = j + 1 < j1;
        Cell nxt, nun;
        if (more) nxt = ld<EDGE, FEAT>(d, s, c, j + 3);
        // U^n at (c, j) for the RK2 average, needed at the end of the row.
        // Kernels with the fast path load it here (its load hides behind the
        // faces, and it saves the registers of a second row); the others
        // fetch it a row ahead, which a launch of few waves needs.
        if constexpr (FASTK) {
            if (read_un && own) {
                const long long k = (long long)j * pitch + c;
                un.h = d.ah[k]; un.hu = d.ahu[k]; un.hv = d.ahv[k];
            }
        } else if (more && read_un && own) {
            const long long k = (long long)(j + 1) * pitch + c;
            nun.h = d.ah[k]; nun.hu = d.ahu[k]; nun.hv = d.ahv[k];
        }
        FState tn;
        Face fyt;
        auto row = [&](auto exact) {
            constexpr bool EX = decltype(exact)::value;
            // faces and cells take the fast path (swe_core.hpp) where the
            // wavefront allows it; lanes up to the first column past the
            // grid feed an owned cell through their x-face
            const bool fast = FASTK && !EX && try_fast;
            const bool use_x = c <= d.nx;
            // y: slope of row j+1 and the face between rows j and j+1
            FState bn;
            cell_states(rc, rn, rnn, bn, tn);
            if (fast) fyt = face_flux_auto<HLLC, true>(g, tc, rc.b, bn, rn.b, own);
            else fyt = face_flux<HLLC, true, EX>(g, tc, rc.b, bn, rn.b);

            // x: slope of row j, face left of each cell, face right from the next lane
            Cell xm = shfl_up1(rc);
            Cell xp = shfl_down1(rc);
            if (lane == 0) xm = Cell{L.hl[0][r], L.hl[1][r], L.hl[2][r], L.hl[3][r]};
            if (lane == WAVE - 1) xp = Cell{L.hr[0][r], L.hr[1][r], L.hr[2][r], L.hr[3][r]};
            FState xlo, xhi;
            cell_states(xm, rc, xp, xlo, xhi);
            FState rprev = shfl_up1(xhi);
            if (lane == 0) rprev = FState{L.ph[0][r], L.ph[1][r], L.ph[2][r]};
            Face fxl;
            if (fast) fxl = face_flux_auto<HLLC, false>(g, rprev, xm.b, xlo, rc.b, use_x);
            else fxl = face_flux<HLLC, false, EX>(g, rprev, xm.b, xlo, rc.b);
            Face fxr = shfl_down1(fxl);
            if (lane == WAVE - 1)
                fxr = Face{L.pf[0][r], L.pf[1][r], L.pf[2][r], L.pf[3][r], L.pf[4][r], L.pf[5][r]};

            if (own) {
                double fh, fhu, fhv;
                divergence<EX>(g, s.rdx, s.rdy, fxl, fxr, fyb, fyt, fh, fhu, fhv);
                facc = facc + fh;
                const long long k = (long long)j * pitch + c;
                if (REC) {
                    d.rec[k] = fh; d.rec[d.plane + k] = fhu; d.rec[2 * d.plane + k] = fhv;
                }
                double h = rc.h, hu = rc.hu, hv = rc.hv;
                Frc fr{0.0, 0.0, 0.0, 0.0, 0.0};
                if (FRC) {
                    fr.wx = d.fwx[k];
                    fr.wy = d.fwy[k];
                    fr.dpdx = d.fdpx[k];
                    fr.dpdy = d.fdpy[k];
                    fr.rain = d.frain[k];
                }
                const double n_cell = NF ? d.nfield[k] : 0.0;
                bool body_fast = fast;
                if (fast) {
                    bool bad = false;
                    double h1 = h, hu1 = hu, hv1 = hv;
                    stage_body_fast<NF, FRC>(s.src, s.dt, fh, fhu, fhv, h1, hu1, hv1, n_cell, fr, bad);
                    if (__any(bad)) body_fast = false;
                    else { h = h1; hu = hu1; hv = hv1; }
                }
                if (!body_fast) stage_body<EX, NF, FRC>(s.src, s.dt, fh, fhu, fhv, h, hu, hv, n_cell, fr);
                // the sponge follows each fused stage (rki_solver.F90:787,814)
                if (STAGE == 1 && SPG) sponge_cell(d, s.dt, c, j, h, hu, hv, sacc);
                // U^n with a tiny value in any lane: the rest of the cell
                // update with plain divisions
                const bool ex_un = EX || (read_un && __any(odd(un)));
                if (STAGE == 2) {
                    // rki_kernel_fused.F90:435-463
                    h = 0.5 * (un.h + h);
                    hu = 0.5 * (un.hu + hu);
                    hv = 0.5 * (un.hv + hv);
                    if (ex_un) {
                        clamp_desing<true>(g, h, hu, hv);
                        if (SPG) sponge_cell(d, s.dt, c, j, h, hu, hv, sacc);
                        smax = dmax(smax, cfl_speed<true>(g, s.rdx, s.rdy, h, hu, hv));
                    } else {
                        bool cell_fast = fast;
                        if (fast) {
                            bool bad = false;
                            clamp_fast(g, h, hu, hv, bad);   // passes the values unchanged
                            cell_fast = !__any(bad);
                        }
                        if (!cell_fast) clamp_desing<false>(g, h, hu, hv);
                        if (SPG) sponge_cell(d, s.dt, c, j, h, hu, hv, sacc);
                        if (cell_fast && !SPG) smax = dmax(smax, cfl_speed_fast(g, s.rdx, s.rdy, h, hu, hv));
                        else smax = dmax(smax, cfl_speed<false>(g, s.rdx, s.rdy, h, hu, hv));
                    }
                }
                if (STAGE == 3) {
                    if (s.blend) {  // compute_ssp_combine, rki_kernel_update.F90:183-206
                        h = s.alpha * un.h + s.a_prev * h;
                        hu = s.alpha * un.hu + s.a_prev * hu;
                        hv = s.alpha * un.hv + s.a_prev * hv;
                        if (ex_un) clamp_desing<true>(g, h, hu, hv);
                        else clamp_desing<false>(g, h, hu, hv);
                    }
                    if (SPG) sponge_cell(d, s.dt, c, j, h, hu, hv, sacc);   // :919-920, at beta_i dt
                    if (s.last) {
                        if (ex_un) smax = dmax(smax, cfl_speed<true>(g, s.rdx, s.rdy, h, hu, hv));
                        else smax = dmax(smax, cfl_speed<false>(g, s.rdx, s.rdy, h, hu, hv));
                    }
                }
                s.oh[k] = h; s.ohu[k] = hu; s.ohv[k] = hv;
                if ((STAGE == 1 || (STAGE == 3 && s.first)) && d.sh != nullptr) {
                    // h0, hu0, hv0 (rki_kernel_fused.F90:224-226, rki_solver.F90:879-881)
                    d.sh[k] = rc.h; d.shu[k] = rc.hu; d.shv[k] = rc.hv;
                }
            }
        };
        const bool lane_tiny = t_prev || t_rc || t_rn || t_rnn;
        if (__any(lane_tiny) || ((edge_rows >> r) & 1ull))
            row(ExactTag<true>{});
        else
            row(ExactTag<false>{});

        fyb = fyt;
        tc = tn;
        rc = rn;
        rn = rnn;
        t_prev = t_rc;
        t_rc = t_rn;
        t_rn = t_rnn;
        if (more) {
            rnn = nxt;
            t_rnn = odd(nxt);
            if (!FASTK && read_un) un = nun;
        }
    }
    fsum = facc;
    ssum = sacc;
    return smax;
}

}  // namespace rk
