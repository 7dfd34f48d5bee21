// rkg_adjoint.hip -- discrete adjoint of the SSP-RK2 step: gradients of a
// gauge sea-surface misfit with respect to the bed, the per-cell Manning n
// and the initial state.
//
// Loss: L = sum over steps s and gauges k of (eta_k(t_s) - target_{s,k})^2,
// eta = h + b at the gauge cell after step s.  The forward pass runs the
// ordinary stage kernels (per-cell Manning n) and keeps the state every
// `checkpoint_interval` steps.  The reverse sweep takes segments last to
// first: it replays a segment from its checkpoint with the recording stage
// kernels, which keep U^n, U* and each stage's flux divergence for every
// step of the segment, then reverses the steps one by one:
//   k_adj_body  -- cell-local: RK2 average, clamp, KP07, Froude backstop,
//                  Manning and Coriolis transposed; emits the adjoint of the
//                  flux divergence of every cell and the Manning gradient;
//   k_adj_y     -- tile kernel, rows marched: every y face and y slope of
//                  the stage, with the south / north ghost rows;
//   k_adj_x     -- row-local: every x face and x slope, with the west /
//                  east ghost columns; both add onto the adjoint of the
//                  stage input and the bed.
//   k_dt_rev    -- one workgroup per step: the dependence of the loss on the
//                  step's dt (update, friction, Coriolis angle, the tide's
//                  clock t_{s+1} = t_s + dt_s) and, for a CFL dt, on the
//                  cell that set the max speed (k_adj_body<1> finds it).
// The gradient is that of the discrete scheme with its dt rule: fixed dt,
// the CFL dt every cfl_interval steps, and the t_end bound.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/rakali_amd.h"
#include "adjoint.hpp"
#include "gauge_eval.hpp"
#include "rkg_internal.hpp"
#include "stage_kernels.hpp"

using namespace rk;

namespace {

constexpr unsigned ADJ_MAGIC = 0x72616B41u;  // 'rakA'
constexpr unsigned ADJ_DEAD = 0xDEADBEEFu;
constexpr int AW = 60;                        // own columns per workgroup of k_adj_x

// Own range [lo, hi) of block `s` when [0, n) is cut into blocks of `w`,
// the last block holding at least two cells (a boundary ghost maps onto
// the two cells next to the edge, which must belong to one tile).
__host__ __device__ __forceinline__ void block_range(int n, int w, int s, int& lo, int& hi) {
    const int nb = (n + w - 1) / w;
    lo = s * w;
    hi = min(lo + w, n);
    if (nb >= 2 && n - (nb - 1) * w == 1) {
        if (s == nb - 1) lo = n - 2;
        if (s == nb - 2) hi = n - 2;
    }
}

struct AdjStage {
    GridDesc d;                 // geometry, physics, bed, n field; ah.. = stage input V
    const double* un;           // U^n (3 planes), stage 2
    const double* div;          // recorded flux divergence (3 planes)
    double* lam_in;             // adjoint of the stage input (3 planes)
    double* lam_out;            // adjoint of the stage output (3 planes); stage 2: the
                                // adjoint of U^{n+1} in, replaced by its U^n part
    double* lam_div;            // 3 planes
    double* grad_b;
    double* grad_n;
    const double* hist;         // (t, dt, dt recomputed) of the step
    double* part_dt;            // per body workgroup: sum of d L / d dt of its cells
    double* pval;               // stage 1 with the CFL dt: per body workgroup, max CFL speed of
    long long* pidx;            // the stage input and its first cell (x fastest), else null
    double* edge_lg;            // adjoint (h, normal momentum) of every tidal ghost cell, see edge_slot
    int edge_len;               // max(nx, ny)
    int ty;                     // rows per tile of k_adj_y
    int xrows;                  // rows per workgroup of k_adj_x
};

// Slot of ghost layer kk (0 = next to the edge) at position p along edge e
// (0 W, 1 E, 2 S, 3 N) in edge_lg (two doubles per slot).
__host__ __device__ __forceinline__ long long edge_slot(int e, int kk, int p, int len) {
    return ((long long)(e * 2 + kk) * len + p) * 2;
}

constexpr int HIST = 3;         // doubles per step in hist

__device__ __forceinline__ SourceTerms adj_src(const GridDesc& d, double dt) {
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

// Sum over a 256-thread workgroup in a fixed order; the result is valid in
// thread 0.
__device__ __forceinline__ double block_sum256(double v, double* red4) {
    v = wave_sum(v);
    if ((threadIdx.x & (WAVE - 1)) == 0) red4[threadIdx.x / WAVE] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) r = ((red4[0] + red4[1]) + red4[2]) + red4[3];
    __syncthreads();
    return r;
}

// Cell-local reverse of one stage.  STAGE 2: lam_out holds the adjoint of
// U^{n+1} and is replaced by the U^n part of the RK2 average; lam_in (the
// adjoint of U*) is written.  STAGE 1: lam_out is the adjoint of U* and
// lam_in (the adjoint of U^n) is added to.
template <int STAGE>
__device__ __forceinline__ void adj_body(const AdjStage& a, int blk, int nblk) {
    const GridDesc& d = a.d;
    const double dt = a.hist[1];
    const SourceTerms src = adj_src(d, dt);
    const long long P = d.plane;
    const long long n = (long long)d.nx * d.ny;
    double ldt = 0.0;
    double best = 0.0;   // stage 1: max CFL speed of U_s and its first cell
    long long bi = -1;
    for (long long idx = blk * 256ll + threadIdx.x; idx < n; idx += (long long)nblk * 256) {
        const int j = (int)(idx / d.nx), i = (int)(idx % d.nx);
        const long long k = (long long)j * d.pitch + i;
        const double v0 = d.ah[k], v1 = d.ahu[k], v2 = d.ahv[k];
        if (STAGE == 1 && a.pval) {
            const double sp = adj::cfl_speed_fwd(d.g, d.dx, d.dy, v0, v1, v2);
            if (sp > best) { best = sp; bi = idx; }
        }
        const adj::S3 div{a.div[k], a.div[P + k], a.div[2 * P + k]};
        const double nc = d.nfield ? d.nfield[k] : d.manning_n;
        const double nsq = d.nfield ? nc * nc : src.n_sq;
        const bool fr = d.nfield ? (nc > 0.0) : (src.has_friction != 0);
        adj::S3 lw;
        if (STAGE == 2) {
            double wh = v0, whu = v1, whv = v2;
            adj::body_fwd(src, dt, div, nsq, fr, wh, whu, whv);
            const double ah = 0.5 * (a.un[k] + wh), ahu = 0.5 * (a.un[P + k] + whu), ahv = 0.5 * (a.un[2 * P + k] + whv);
            const adj::S3 lo{a.lam_out[k], a.lam_out[P + k], a.lam_out[2 * P + k]};
            const adj::S3 la = adj::clamp_rev(d.g, ah, ahu, ahv, lo);
            lw = adj::S3{0.5 * la.h, 0.5 * la.hu, 0.5 * la.hv};
            a.lam_out[k] = lw.h; a.lam_out[P + k] = lw.hu; a.lam_out[2 * P + k] = lw.hv;
        } else {
            lw = adj::S3{a.lam_out[k], a.lam_out[P + k], a.lam_out[2 * P + k]};
        }
        const adj::BodyRev r = adj::body_rev(src, dt, div, nc, nsq, fr, v0, v1, v2, lw);
        if (STAGE == 2) {
            a.lam_in[k] = r.lv.h; a.lam_in[P + k] = r.lv.hu; a.lam_in[2 * P + k] = r.lv.hv;
        } else {
            a.lam_in[k] += r.lv.h; a.lam_in[P + k] += r.lv.hu; a.lam_in[2 * P + k] += r.lv.hv;
        }
        a.lam_div[k] = r.ldiv.h; a.lam_div[P + k] = r.ldiv.hu; a.lam_div[2 * P + k] = r.ldiv.hv;
        if (d.nfield && a.grad_n) a.grad_n[k] += r.ln;
        ldt = ldt + r.ldt;
    }
    __shared__ double red4[4];
    const double tot = block_sum256(ldt, red4);
    if (threadIdx.x == 0) a.part_dt[blk] = tot;
    if (STAGE == 1 && a.pval) {
        __shared__ double sv[256];
        __shared__ long long si[256];
        sv[threadIdx.x] = best;
        si[threadIdx.x] = bi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                const double v = sv[threadIdx.x + s];
                const long long x = si[threadIdx.x + s];
                if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && x >= 0 && (si[threadIdx.x] < 0 || x < si[threadIdx.x]))) {
                    sv[threadIdx.x] = v;
                    si[threadIdx.x] = x;
                }
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) { a.pval[blk] = sv[0]; a.pidx[blk] = si[0]; }
    }
}

template <int STAGE>
__global__ __launch_bounds__(256) void k_adj_body(AdjStage a) {
    adj_body<STAGE>(a, blockIdx.x, gridDim.x);
}

struct Acc {   // adjoint accumulator of one cell: (h, hu, hv, b)
    double h, hu, hv, b;
};

__device__ __forceinline__ void add(Acc& a, const adj::C4& c) { a.h += c.h; a.hu += c.hu; a.hv += c.hv; a.b += c.b; }
__device__ __forceinline__ adj::C4 c4(const Acc& a) { return adj::C4{a.h, a.hu, a.hv, a.b}; }
__device__ __forceinline__ double sh_dn(double v) { return __shfl_down(v, 1, WAVE); }
__device__ __forceinline__ double sh_up(double v) { return __shfl_up(v, 1, WAVE); }
__device__ __forceinline__ adj::C4 sh_dn(const adj::C4& c) { return adj::C4{sh_dn(c.h), sh_dn(c.hu), sh_dn(c.hv), sh_dn(c.b)}; }
__device__ __forceinline__ adj::C4 sh_up(const adj::C4& c) { return adj::C4{sh_up(c.h), sh_up(c.hu), sh_up(c.hv), sh_up(c.b)}; }
__device__ __forceinline__ adj::C4 sh_from(const adj::C4& c, int src) {
    return adj::C4{__shfl(c.h, src, WAVE), __shfl(c.hu, src, WAVE), __shfl(c.hv, src, WAVE), __shfl(c.b, src, WAVE)};
}

__device__ __forceinline__ adj::S3 load_ldiv(const AdjStage& a, int x, int y) {
    const GridDesc& d = a.d;
    if ((unsigned)x >= (unsigned)d.nx || (unsigned)y >= (unsigned)d.ny) return adj::S3{0.0, 0.0, 0.0};
    const long long k = (long long)y * d.pitch + x;
    return adj::S3{a.lam_div[k], a.lam_div[d.plane + k], a.lam_div[2 * d.plane + k]};
}

__device__ __forceinline__ Cell vload(const GridDesc& d, int x, int y, double eta) {
    return load_cell<true>(d, d.ah, d.ahu, d.ahv, x, y, eta);
}

// Adjoint of the row-y ghost at column x onto interior cells (x-direction
// ghosts: normal momentum hu).  Returns the contributions for the mirror
// and first-interior columns.
__device__ __forceinline__ void ghost_x_rev(const GridDesc& d, double eta, int x, int y, const adj::C4& lg, int& mcol,
                                            int& icol, adj::C4& lm, adj::C4& li) {
    const bool west = x < 0;
    const int kk = west ? (-x - 1) : (x - d.nx);
    mcol = west ? kk : (d.nx - 1 - kk);
    icol = west ? 0 : (d.nx - 1);
    const Cell mir = load_direct(d.ah, d.ahu, d.ahv, d.bed, d.pitch, mcol, y);
    const Cell in0 = load_direct(d.ah, d.ahu, d.ahv, d.bed, d.pitch, icol, y);
    lm = adj::C4{0, 0, 0, 0};
    li = adj::C4{0, 0, 0, 0};
    adj::ghost_rev(d.bc[west ? 0 : 1], d.g, eta, d.inflow_velocity, d.clamped_elevation, mir, in0, west ? -1.0 : 1.0,
                   true, lg, lm, li);
}

__device__ __forceinline__ void ghost_y_rev(const GridDesc& d, double eta, int x, int y, const adj::C4& lg, int& mrow,
                                            int& irow, adj::C4& lm, adj::C4& li) {
    const bool south = y < 0;
    const int kk = south ? (-y - 1) : (y - d.ny);
    mrow = south ? kk : (d.ny - 1 - kk);
    irow = south ? 0 : (d.ny - 1);
    const Cell mir = load_direct(d.ah, d.ahu, d.ahv, d.bed, d.pitch, x, mrow);
    const Cell in0 = load_direct(d.ah, d.ahu, d.ahv, d.bed, d.pitch, x, irow);
    lm = adj::C4{0, 0, 0, 0};
    li = adj::C4{0, 0, 0, 0};
    adj::ghost_rev(d.bc[south ? 2 : 3], d.g, eta, d.inflow_velocity, d.clamped_elevation, mir, in0,
                   south ? -1.0 : 1.0, false, lg, lm, li);
}

constexpr int NWIN = 6;   // accumulator rows r-3 .. r+2

// Reverse of the y faces and y slopes of one stage (and the south / north
// ghost rows) onto the stage input and the bed.  One wavefront per tile of
// 64 whole columns (lane l holds column c0 + l; this direction is
// column-local, so no halo columns) and rows [j0, j1).  The rows are
// marched from j0 - 2 to j1: iteration r reverses the y face between rows
// r and r+1 and the y slope of row r.  Row r-1 is complete after iteration
// r; rows are written two iterations later so that ghost rows can still be
// folded into their source rows.  TY: a south or north edge is tidal, and
// the adjoints of its ghost cells are kept in edge_lg for k_dt_rev.
template <bool HLLC, int STAGE, bool TY>
__device__ __forceinline__ void adj_y(const AdjStage& a, int blk) {
    const GridDesc& d = a.d;
    const double tau = STAGE == 1 ? a.hist[0] : a.hist[0] + a.hist[1];   // bc%time of the stage
    const int lane = threadIdx.x;
    const int nstr = (d.nx + WAVE - 1) / WAVE;
    const int sx = blk % nstr, sy = blk / nstr;
    const int c0 = sx * WAVE;
    int j0, j1;
    block_range(d.ny, a.ty, sy, j0, j1);
    const bool xin = c0 + lane < d.nx;
    const int x = xin ? c0 + lane : d.nx - 1;   // lanes past the edge repeat the last column, unwritten
    // rows j0 - 3 .. j1 + 2 are read; only ghost rows need the tide (whose
    // correctly rounded sine is a long dependent chain)
    const double eta = (j0 < 3 || j1 + 2 >= d.ny) ? tidal_eta(d, tau) : 0.0;
    const double g = d.g;
    const bool fast = g >= 1.0e-3 && g <= 1.0e6;   // face_rev_fast's ranges assume an ordinary gravity
    const double rdy = 1.0 / d.dy;
    const long long P = d.plane;

    Acc acc[NWIN];
#pragma unroll
    for (int q = 0; q < NWIN; ++q) acc[q] = Acc{0.0, 0.0, 0.0, 0.0};
    const int rs = j0 - 2, re = j1;
    Cell v0 = vload(d, x, rs - 1, eta), v1 = vload(d, x, rs, eta), v2 = vload(d, x, rs + 1, eta), v3;
    FState hi_r, lo_tmp;
    adj::states_fwd(v0, v1, v2, lo_tmp, hi_r);
    adj::S3 llo_r{0.0, 0.0, 0.0};        // adjoint of the y lo-state of row r (from face r - 1/2)
    adj::S3 ld_r = load_ldiv(a, x, rs);

    // Folds the completed row `rho` (window slot `q`): a ghost row onto its
    // source rows, an owned row into the adjoint arrays.
    auto flush = [&](int rho, int q, int r) {
        if (rho < -2 || rho > d.ny + 1) return;
        const adj::C4 lg = c4(acc[q]);
        if (rho < 0 || rho >= d.ny) {
            int mrow, irow;
            adj::C4 lm, li;
            ghost_y_rev(d, eta, x, rho, lg, mrow, irow, lm, li);
            // source rows are in the window: slot of row t is t - (r - 3)
            const int qm = mrow - (r - 3), qi = irow - (r - 3);
#pragma unroll
            for (int s = 0; s < NWIN; ++s) {
                if (s == qm) add(acc[s], lm);
                if (s == qi) add(acc[s], li);
            }
            // a tidal ghost row cell is kept once, by the tile holding the edge
            // row (a tile next to a short last tile also reaches the ghost
            // rows, with partial sums)
            const int e = rho < 0 ? 2 : 3;
            if (TY && xin && (rho < 0 ? j0 == 0 : j1 == d.ny) && d.bc[e] == BC_TIDAL) {
                const long long o = edge_slot(e, rho < 0 ? -rho - 1 : rho - d.ny, x, a.edge_len);
                a.edge_lg[o] = lg.h;
                a.edge_lg[o + 1] = lg.hv;
            }
            return;
        }
        if (xin && rho >= j0 && rho < j1) {
            const long long k = (long long)rho * d.pitch + x;
            a.lam_in[k] += lg.h;
            a.lam_in[P + k] += lg.hu;
            a.lam_in[2 * P + k] += lg.hv;
            a.grad_b[k] += lg.b;
        }
    };

    for (int r = rs; r <= re; ++r) {
        // window: acc[q] is row r - 3 + q; v0..v3 rows r-1 .. r+2
        v3 = vload(d, x, r + 2, eta);
        const adj::S3 ld_r1 = load_ldiv(a, x, r + 1);
        FState lo_r1, hi_r1;
        adj::states_fwd(v1, v2, v3, lo_r1, hi_r1);
        // y face r + 1/2 (faces -1/2 .. ny - 1/2)
        adj::S3 lhi_r{0.0, 0.0, 0.0}, llo_r1{0.0, 0.0, 0.0};
        if (r >= -1 && r <= d.ny - 1) {
            const Face lf = adj::face_seed<true>(g, rdy, ld_r, ld_r1);
            double lbL = 0.0, lbR = 0.0;
            if (!(fast && adj::face_rev_fast<HLLC, true>(g, hi_r, v1.b, lo_r1, v2.b, lf, lhi_r, lbL, llo_r1, lbR,
                                                          xin)))
                adj::face_rev<HLLC, true>(g, hi_r, v1.b, lo_r1, v2.b, lf, lhi_r, lbL, llo_r1, lbR);
            acc[3].b += lbL;
            acc[4].b += lbR;
        }
        // y slope of row r
        {
            adj::C4 lm{0, 0, 0, 0}, lc{0, 0, 0, 0}, lp{0, 0, 0, 0};
            adj::states_rev(v0, v1, v2, llo_r, lhi_r, lm, lc, lp);
            add(acc[2], lm);
            add(acc[3], lc);
            add(acc[4], lp);
        }
        // row r - 3 is complete and past the ghost folding window
        flush(r - 3, 0, r);
#pragma unroll
        for (int q = 0; q < NWIN - 1; ++q) acc[q] = acc[q + 1];
        acc[NWIN - 1] = Acc{0.0, 0.0, 0.0, 0.0};
        v0 = v1; v1 = v2; v2 = v3;
        hi_r = hi_r1;
        llo_r = llo_r1;
        ld_r = ld_r1;
    }
    // remaining rows re-2 .. re+2 in slots 0..4 (window base re + 1 - 3):
    // ghost rows first (their sources are among the owned rows), then the rest
    const int base = re + 1 - 3;
    for (int q = 0; q < NWIN; ++q) {
        const int rho = base + q;
        if (rho >= d.ny) flush(rho, q, re + 1);
    }
    for (int q = 0; q < NWIN; ++q) {
        const int rho = base + q;
        if (rho < d.ny) flush(rho, q, re + 1);
    }
}

template <bool HLLC, int STAGE, bool TY>
__global__ __launch_bounds__(WAVE) void k_adj_y(AdjStage a) {
    adj_y<HLLC, STAGE, TY>(a, blockIdx.x);
}

// Reverse of the x faces and x slopes of one stage (and the west / east
// ghost columns) onto the stage input and the bed.  This direction is
// row-local: every row is independent.  One wavefront per strip of at
// most 60 owned columns [c0, c1) and `xrows` rows; lane l holds column
// c0 - 2 + l, so the faces and slopes that touch an owned cell are all in
// the wavefront.  The face right of a cell is reversed by the cell's lane;
// neighbours' parts move by lane shuffles; ghost columns fold onto their
// source columns by lane permutes.  TX: a west or east edge is tidal (see
// k_adj_y's TY).
template <bool HLLC, int STAGE, bool TX>
__device__ __forceinline__ void adj_x(const AdjStage& a, int blk) {
    const GridDesc& d = a.d;
    const double tau = STAGE == 1 ? a.hist[0] : a.hist[0] + a.hist[1];
    const int lane = threadIdx.x;
    const int nstr = (d.nx + AW - 1) / AW;
    const int sx = blk % nstr, sy = blk / nstr;
    int c0, c1;
    block_range(d.nx, AW, sx, c0, c1);
    const int r0 = sy * a.xrows, r1 = min(r0 + a.xrows, d.ny);
    const int x = c0 - 2 + lane;
    const bool own_col = x >= c0 && x < c1;
    const bool edge_tile = c0 - 2 < 0 || c0 + WAVE - 3 >= d.nx;
    // columns c0 - 3 .. c0 + 62 are read (lanes 0 and 63 load one more);
    // only ghost columns need the tide
    const double eta = (c0 < 3 || c0 + 62 >= d.nx) ? tidal_eta(d, tau) : 0.0;
    const double g = d.g;
    const bool fast = g >= 1.0e-3 && g <= 1.0e6;
    const double rdx = 1.0 / d.dx;
    const long long P = d.plane;
    for (int r = r0; r < r1; ++r) {
        const Cell v1 = vload(d, x, r, eta);
        const adj::S3 ld_r = load_ldiv(a, x, r);
        Cell xm, xp;
        xm.h = sh_up(v1.h); xm.hu = sh_up(v1.hu); xm.hv = sh_up(v1.hv); xm.b = sh_up(v1.b);
        xp.h = sh_dn(v1.h); xp.hu = sh_dn(v1.hu); xp.hv = sh_dn(v1.hv); xp.b = sh_dn(v1.b);
        if (lane == 0) xm = vload(d, x - 1, r, eta);
        if (lane == WAVE - 1) xp = vload(d, x + 1, r, eta);
        FState lo_x, hi_x;
        adj::states_fwd(xm, v1, xp, lo_x, hi_x);
        FState R;
        R.h = sh_dn(lo_x.h); R.hu = sh_dn(lo_x.hu); R.hv = sh_dn(lo_x.hv);
        const adj::S3 ld_next{sh_dn(ld_r.h), sh_dn(ld_r.hu), sh_dn(ld_r.hv)};
        adj::S3 lhi_x{0.0, 0.0, 0.0}, lR{0.0, 0.0, 0.0};
        double lbL = 0.0, lbR = 0.0;
        if (lane < WAVE - 1 && x >= -1 && x <= d.nx - 1) {   // the face between x and x + 1
            const Face lf = adj::face_seed<false>(g, rdx, ld_r, ld_next);
            if (!(fast && adj::face_rev_fast<HLLC, false>(g, hi_x, v1.b, R, xp.b, lf, lhi_x, lbL, lR, lbR, true)))
                adj::face_rev<HLLC, false>(g, hi_x, v1.b, R, xp.b, lf, lhi_x, lbL, lR, lbR);
        }
        adj::S3 llo_x{sh_up(lR.h), sh_up(lR.hu), sh_up(lR.hv)};
        const double lb_from_left = sh_up(lbR);
        if (lane == 0) llo_x = adj::S3{0.0, 0.0, 0.0};
        adj::C4 mine{0.0, 0.0, 0.0, lbL + (lane == 0 ? 0.0 : lb_from_left)};
        adj::C4 lm{0, 0, 0, 0}, lc{0, 0, 0, 0}, lp{0, 0, 0, 0};
        adj::states_rev(xm, v1, xp, llo_x, lhi_x, lm, lc, lp);
        const adj::C4 from_right = sh_dn(lm);   // lane l+1's left-neighbour part
        const adj::C4 from_left = sh_up(lp);    // lane l-1's right-neighbour part
        mine.h += lc.h; mine.hu += lc.hu; mine.hv += lc.hv; mine.b += lc.b;
        if (lane < WAVE - 1) { mine.h += from_right.h; mine.hu += from_right.hu; mine.hv += from_right.hv; mine.b += from_right.b; }
        if (lane > 0) { mine.h += from_left.h; mine.hu += from_left.hu; mine.hv += from_left.hv; mine.b += from_left.b; }
        // ghost columns fold onto their source columns (tiles at an edge only)
        if (edge_tile) {
            adj::C4 gm{0, 0, 0, 0}, gi{0, 0, 0, 0};
            int mcol = -100, icol = -100;
            if (x == -1 || x == -2 || x == d.nx || x == d.nx + 1) {
                ghost_x_rev(d, eta, x, r, mine, mcol, icol, gm, gi);
                const int e = x < 0 ? 0 : 1;
                // kept once, by the tile holding the edge column
                if (TX && (x < 0 ? c0 == 0 : c1 == d.nx) && d.bc[e] == BC_TIDAL) {
                    const long long o = edge_slot(e, x < 0 ? -x - 1 : x - d.nx, r, a.edge_len);
                    a.edge_lg[o] = mine.h;
                    a.edge_lg[o + 1] = mine.hu;
                }
            }
            // lanes holding columns -2, -1, nx, nx + 1 (when inside the wavefront)
            const int srcs[4] = {-c0, 1 - c0, d.nx - c0 + 2, d.nx - c0 + 3};
#pragma unroll
            for (int s = 0; s < 4; ++s) {
                const int src = srcs[s];
                if (src < 0 || src >= WAVE) continue;
                const adj::C4 rm = sh_from(gm, src), ri = sh_from(gi, src);
                const int rmcol = __shfl(mcol, src, WAVE), ricol = __shfl(icol, src, WAVE);
                if (x == rmcol) { mine.h += rm.h; mine.hu += rm.hu; mine.hv += rm.hv; mine.b += rm.b; }
                if (x == ricol) { mine.h += ri.h; mine.hu += ri.hu; mine.hv += ri.hv; mine.b += ri.b; }
            }
        }
        if (own_col) {
            const long long k = (long long)r * d.pitch + x;
            a.lam_in[k] += mine.h;
            a.lam_in[P + k] += mine.hu;
            a.lam_in[2 * P + k] += mine.hv;
            a.grad_b[k] += mine.b;
        }
    }
}

template <bool HLLC, int STAGE, bool TX>
__global__ __launch_bounds__(WAVE) void k_adj_x(AdjStage a) {
    adj_x<HLLC, STAGE, TX>(a, blockIdx.x);
}

// d L / d eta_tidal from one tidal ghost cell (adjoints lg0 of h and lg1
// of the normal momentum; in0 its source cell of the stage input): the
// eta_ext terms of bc_flather_edge (rki_boundary.F90:331-343),
// h_g = max(eta_ext - b_g, 0) and the radiation term.
__device__ double eta_rev_one(const GridDesc& d, double eta, const Cell& in0, double lg0, double lg1, double sgn,
                              bool xe) {
    adj::C4 lg{lg0, 0.0, 0.0, 0.0};
    if (xe) lg.hu = lg1;
    else lg.hv = lg1;
    adj::C4 lm{0, 0, 0, 0}, li{0, 0, 0, 0};
    double le = 0.0;
    adj::ghost_rev(BC_TIDAL, d.g, eta, d.inflow_velocity, d.clamped_elevation, in0, in0, sgn, xe, lg, lm, li, &le);
    return le;
}

// This thread's share for both stages at once (stage 1: input d.ah at
// eta1, ghost adjoints lg1; stage 2: input in2 at eta2, lg2).  Items
// i = threadIdx.x + q blockDim.x go two at a time, their loads issued
// together; each stage's sum keeps the item order.
__device__ void eta_parts(const GridDesc& d, double eta1, double eta2, const double* in2, const double* lg1,
                          const double* lg2, int len, double* acc1, double* acc2) {
    double s1 = 0.0, s2 = 0.0;
    for (int e = 0; e < 4; ++e) {
        if (d.bc[e] != BC_TIDAL) continue;
        const bool xe = e < 2;
        const int n = xe ? d.ny : d.nx;
        const double sgn = (e == 0 || e == 2) ? -1.0 : 1.0;
        for (int i0 = threadIdx.x; i0 < 2 * n; i0 += 2 * blockDim.x) {
            Cell c1[2], c2[2];
            double g[2][4];
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int i = u == 0 ? i0 : min(i0 + (int)blockDim.x, 2 * n - 1);
                const int kk = i / n, p = i % n;
                const int ci = xe ? (e == 0 ? 0 : d.nx - 1) : p;
                const int cj = xe ? p : (e == 2 ? 0 : d.ny - 1);
                c1[u] = load_direct(d.ah, d.ahu, d.ahv, d.bed, d.pitch, ci, cj);
                c2[u] = load_direct(in2, in2 + d.plane, in2 + 2 * d.plane, d.bed, d.pitch, ci, cj);
                const long long o = edge_slot(e, kk, p, len);
                g[u][0] = lg1[o]; g[u][1] = lg1[o + 1];
                g[u][2] = lg2[o]; g[u][3] = lg2[o + 1];
            }
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (u == 1 && i0 + (int)blockDim.x >= 2 * n) break;
                s1 = s1 + eta_rev_one(d, eta1, c1[u], g[u][0], g[u][1], sgn, xe);
                s2 = s2 + eta_rev_one(d, eta2, c2[u], g[u][2], g[u][3], sgn, xe);
            }
        }
    }
    *acc1 = s1;
    *acc2 = s2;
}

// Gauges of a session: cells (eta = h + b of the cell) or bilinear
// stations (gauge_eval.hpp) with lower-left cell (gi, gj) and weights gw
// (wx then wy).
struct GaugeSpec {
    const int* gi;
    const int* gj;
    const double* gw;   // null: cell gauges
    int ng, nx, ny;
};

__device__ __forceinline__ double spec_eta(const GaugeSpec& G, const double* h, const double* bed, long long pitch,
                                           int k) {
    if (G.gw == nullptr) {
        const long long o = (long long)G.gj[k] * pitch + G.gi[k];
        return h[o] + bed[o];
    }
    return station_eta(h, bed, pitch, G.nx, G.ny, G.gi[k], G.gj[k], G.gw[k], G.gw[G.ng + k]);
}

// Gauge misfit of one state: loss += sum_k (eta_k - target_k)^2, eta
// recorded per gauge; a single workgroup, fixed summation order.
__global__ __launch_bounds__(256) void k_gauge_loss(const double* __restrict__ h, const double* __restrict__ bed,
                                                   long long pitch, GaugeSpec G, const double* __restrict__ tgt,
                                                   double* __restrict__ eta_out, double* __restrict__ loss) {
    __shared__ double red[256];
    double acc = 0.0;
    const int ng = G.ng;
    for (int k = threadIdx.x; k < ng; k += 256) {
        const double e = spec_eta(G, h, bed, pitch, k);
        if (eta_out) eta_out[k] = e;
        if (tgt) {
            const double m = e - tgt[k];
            acc = acc + m * m;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = *loss + red[0];
}

// Seeds d L / d U^{n+1} at the gauges.  One thread per distinct gauge cell
// (cell_pos[u]) adds its taps (cell_gauges[q], weight cell_w[q] for q in
// cell_start[u] .. cell_start[u+1]) in gauge order, as a single thread
// walking all gauges would.  A cell gauge has one tap of weight 1; a
// station the four of its bilinear weights ((1-wx)(1-wy), wx(1-wy),
// (1-wx)wy, wx wy), since d eta / d h = d eta / d b = weight there.
struct SeedArgs {
    const double* h;            // U^{n+1}
    const double* bed;
    long long pitch;
    GaugeSpec G;
    const double* tgt;          // the step's targets
    const int* cell_start;
    const int* cell_gauges;
    const double* cell_w;
    const long long* cell_pos;
    int ncells;
    double scale;
    double* lam_h;
    double* grad_b;
};

__device__ __forceinline__ void gauge_seed(const SeedArgs& A, int blk) {
    const double* __restrict__ h = A.h;
    const double* __restrict__ bed = A.bed;
    const long long pitch = A.pitch;
    const GaugeSpec& G = A.G;
    const double* __restrict__ tgt = A.tgt;
    const int* __restrict__ cell_start = A.cell_start;
    const int* __restrict__ cell_gauges = A.cell_gauges;
    const double* __restrict__ cell_w = A.cell_w;
    const double scale = A.scale;
    double* __restrict__ lam_h = A.lam_h;
    double* __restrict__ grad_b = A.grad_b;
    const int u = blk * 256 + threadIdx.x;
    if (u >= A.ncells) return;
    const long long* __restrict__ cell_pos = A.cell_pos;
    const long long o = cell_pos[u];
    double lh = lam_h[o], gb = grad_b[o];
    for (int q = cell_start[u]; q < cell_start[u + 1]; ++q) {
        const int k = cell_gauges[q];
        const double e = spec_eta(G, h, bed, pitch, k);
        double sk = 2.0 * scale * (e - tgt[k]);
        if (G.gw != nullptr) sk = sk * cell_w[q];
        lh += sk;
        gb += sk;
    }
    lam_h[o] = lh;
    grad_b[o] = gb;
}

__global__ __launch_bounds__(256) void k_gauge_seed(SeedArgs A) { gauge_seed(A, blockIdx.x); }



// Scalar part of the reverse of one step (one workgroup).  scal[0] is the
// adjoint of t_{s+1} carried from later steps, scal[1] that of a dt that
// the next step reused.  Adds d L / d dt * d dt / d U at the cell of
// maximum speed to lam (the adjoint of U_s) when dt was the CFL dt.
struct DtArgs {
    GridDesc d;                 // stage 1 (input U_s)
    const double* hist;
    const double* part_dt;      // 2 x nbody partial sums of d L / d dt
    int nbody;
    const double* in2;          // U* (the stage-2 input)
    const double* edge_lg;
    int edge_len;
    const double* pval;
    const long long* pidx;
    int nargmax;
    double* lam;                // the adjoint of U_s
    double* scal;
};

__device__ __forceinline__ void dt_rev(const DtArgs& A) {
    const GridDesc& d = A.d;
    const double* hist = A.hist;
    const double* part_dt = A.part_dt;
    const int nbody = A.nbody;
    const double* in2 = A.in2;
    const double* edge_lg = A.edge_lg;
    const int edge_len = A.edge_len;
    const double* pval = A.pval;
    const long long* pidx = A.pidx;
    const int nargmax = A.nargmax;
    double* lam = A.lam;
    double* scal = A.scal;
    __shared__ double red[3][1024];
    __shared__ double sv[1024];
    __shared__ long long si[1024];
    __shared__ double tide[4];   // eta(t), eta(t + dt), eta'(t), eta'(t + dt)
    const double t = hist[0], dt = hist[1];
    // The kernel is bound by dependent memory round trips: each loop below
    // issues four elements' loads before it waits, and adds or compares them
    // in the element order of a plain strided loop.
    double a = 0.0;
    for (int i0 = threadIdx.x; i0 < 2 * nbody; i0 += 4 * 1024) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = part_dt[min(i0 + u * 1024, 2 * nbody - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u * 1024 < 2 * nbody) a = a + v[u];
    }
    double best = 0.0;
    long long bi = -1;
    for (int i0 = threadIdx.x; i0 < nargmax; i0 += 4 * 1024) {
        double v[4];
        long long x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = min(i0 + u * 1024, nargmax - 1);
            v[u] = pval[i];
            x[u] = pidx[i];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (i0 + u * 1024 < nargmax && (v[u] > best || (v[u] == best && x[u] >= 0 && (bi < 0 || x[u] < bi)))) {
                best = v[u];
                bi = x[u];
            }
    }
    red[0][threadIdx.x] = a;
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    const int w = threadIdx.x / WAVE;
    if (w < 4) {
        const double tw = (w & 1) ? t + dt : t;
        const double v = w < 2 ? tidal_eta(d, tw) : adj::tidal_deta(d, tw);
        if (threadIdx.x % WAVE == 0) tide[w] = v;
    }
    __syncthreads();
    // d L / d eta_tidal of stage 1 (input A = d.ah, at t) and stage 2 (input B = in2, at t + dt)
    eta_parts(d, tide[0], tide[1], in2, edge_lg, edge_lg + edge_slot(4, 0, 0, edge_len), edge_len, &red[1][threadIdx.x],
              &red[2][threadIdx.x]);
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = red[q][threadIdx.x] + red[q][threadIdx.x + s];
            const double v = sv[threadIdx.x + s];
            const long long x = si[threadIdx.x + s];
            if (v > sv[threadIdx.x] || (v == sv[threadIdx.x] && x >= 0 && (si[threadIdx.x] < 0 || x < si[threadIdx.x]))) {
                sv[threadIdx.x] = v;
                si[threadIdx.x] = x;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const bool recomputed = hist[2] != 0.0;
    const double lam_t1 = scal[0] + red[2][0] * tide[3];   // stage 2 ran at t + dt
    const double gdt = red[0][0] + lam_t1 + scal[1];
    double lam_t = lam_t1 + red[1][0] * tide[2];
    double carry = 0.0;
    const double dt_max = d.t_end - t;
    if (d.dt_fixed > 0.0) {
        if (!(d.dt_fixed < dt_max)) lam_t -= gdt;           // dt = t_end - t
    } else if (recomputed) {
        const double ms = sv[0];
        double dc = ms > 0.0 ? d.cfl / ms : d.cfl * dmin(d.dx, d.dy) / sqrt(d.g * DRY_TOLERANCE);
        dc = dc * d.ssp_coeff;
        if (!(dc < dt_max)) {
            lam_t -= gdt;                                    // dt = t_end - t
        } else if (ms > 0.0 && si[0] >= 0) {                 // dt = cfl ssp / max speed
            const long long idx = si[0];
            const long long k = (idx / d.nx) * d.pitch + idx % d.nx;
            const double l0 = lam[k], l1 = lam[d.plane + k], l2 = lam[2 * d.plane + k];
            const double lms = gdt * (-(d.cfl * d.ssp_coeff) / (ms * ms));
            const adj::S3 l = adj::cfl_speed_rev(d.g, d.dx, d.dy, d.ah[k], d.ahu[k], d.ahv[k], lms);
            lam[k] = l0 + l.h;
            lam[d.plane + k] = l1 + l.hu;
            lam[2 * d.plane + k] = l2 + l.hv;
        }
    } else {
        carry = gdt;                                         // dt reused from the step before
    }
    scal[0] = lam_t;
    scal[1] = carry;
}

__global__ __launch_bounds__(1024) void k_dt_rev(DtArgs A) { dt_rev(A); }

// ---------------------------------------------------------------------
// One batched reverse sweep over several sessions: each kernel below is one
// launch over the blocks of every member, in the members' own block
// decomposition (a table entry names the member and its block), so that a
// member's arithmetic and sums are exactly those of its single-session
// sweep.

// A block of one member: the member, the block's index within the member's
// launch and that launch's block count.
struct BlockRef {
    int member, blk, nblk, pad;
};

// A replay tile of one member (strip, segment, leader of its launch).
struct RecTile {
    int member, strip, seg, leader;
};

struct RecTileOf {
    RecTile t;
    __device__ int3 operator()(const GridDesc&) const { return make_int3(t.strip, t.seg, t.leader); }
};

template <int STAGE, bool HLLC, bool FASTK>
__global__ __launch_bounds__(WAVE) __attribute__((amdgpu_waves_per_eu(RKG_WAVES_PER_EU, 8))) void k_replay_batch(
    const GridDesc* __restrict__ descs, const RecTile* __restrict__ tiles, const int* __restrict__ parity) {
    __shared__ TileLds L;
    const RecTile t = tiles[blockIdx.x];
    const GridDesc& d = descs[t.member];
    stage_tile<STAGE, HLLC, true, true, RecTileOf, 0, FASTK>(d, L, parity[t.member], MODE_STEP, d.t_end, RecTileOf{t});
}

__global__ __launch_bounds__(256) void k_gauge_seed_batch(const SeedArgs* __restrict__ sa,
                                                         const BlockRef* __restrict__ tab) {
    const BlockRef r = tab[blockIdx.x];
    gauge_seed(sa[r.member], r.blk);
}

template <int STAGE>
__global__ __launch_bounds__(256) void k_adj_body_batch(const AdjStage* __restrict__ as,
                                                       const BlockRef* __restrict__ tab) {
    const BlockRef r = tab[blockIdx.x];
    adj_body<STAGE>(as[r.member], r.blk, r.nblk);
}

// The tidal-edge variants (TY, TX) serve every member: they only add the
// bookkeeping of tidal ghost cells, which other edges do not produce.
template <bool HLLC, int STAGE>
__global__ __launch_bounds__(WAVE) void k_adj_y_batch(const AdjStage* __restrict__ as,
                                                     const BlockRef* __restrict__ tab) {
    const BlockRef r = tab[blockIdx.x];
    adj_y<HLLC, STAGE, true>(as[r.member], r.blk);
}

template <bool HLLC, int STAGE>
__global__ __launch_bounds__(WAVE) void k_adj_x_batch(const AdjStage* __restrict__ as,
                                                     const BlockRef* __restrict__ tab) {
    const BlockRef r = tab[blockIdx.x];
    adj_x<HLLC, STAGE, true>(as[r.member], r.blk);
}

__global__ __launch_bounds__(1024) void k_dt_rev_batch(const DtArgs* __restrict__ da) { dt_rev(da[blockIdx.x]); }

// Still-water initial state: h = max(eta0 - b, 0) and its bed adjoint.
__global__ __launch_bounds__(256) void k_fold_still_water(const double* __restrict__ bed, const double* __restrict__ lam_h,
                                                         double eta0, int nx, int ny, long long pitch,
                                                         double* __restrict__ grad_b) {
    const long long n = (long long)nx * ny;
    for (long long idx = blockIdx.x * 256ll + threadIdx.x; idx < n; idx += (long long)gridDim.x * 256) {
        const long long k = (idx / nx) * pitch + idx % nx;
        if (eta0 - bed[k] > 0.0) grad_b[k] -= lam_h[k];
    }
}

}  // namespace

struct rkg_adjoint {
    unsigned magic;
    rkg_grid* g;
    int nsteps, K, nseg, ng;
    int* gi;                    // device
    int* gj;
    double* gw;                 // device: station weights wx then wy, or null (cell gauges)
    int* cell_start;            // device CSR of the distinct gauge cells (ncells + 1)
    int* cell_gauges;           // device: gauge of each tap, per cell in gauge order
    double* cell_w;             // device: weight of each tap (stations)
    long long* cell_pos;        // device: offset of each distinct cell in a plane
    int ncells;
    double* tgt;                // device, nsteps x ng
    double* eta;                // device, nsteps x ng (gauge eta of the last forward run)
    double* ckpt;               // nseg x 3 planes
    CtrlSlot* ckpt_ctrl;        // nseg x 2 (device)
    std::vector<int> ckpt_parity;
    double* ws;                 // segment workspace: A (K+1) x 3, B K x 3, div1 K x 3, div2 K x 3
    double* lam;                // lamA (3), lamB (3), lamDiv (3), grad_b, grad_n
    double* hist;               // K x HIST
    double* part;               // part_dt (2 x body blocks), argmax values
    long long* part_idx;        // argmax cells
    double* scal;               // adjoint of the clock and of a reused dt
    double* edge_lg;            // tidal ghost adjoints of stage 1, then stage 2 (edge_slot)
    int edge_len;
    double* loss_dev;
    double* loss_host;
    CtrlSlot end_ctrl[2];
    int end_parity;
    bool have_forward;
    float fwd_ms, bwd_ms;
    hipEvent_t ev[2];
    // Overlapped replay (rkg_adjoint_backward): a second workspace and
    // history, so that segment s - 1 is replayed on rstream while segment s
    // is reversed on the grid's stream.  ws / hist and ws_alt / hist_alt
    // swap as the sweep moves between the sets; cur_set is the set in ws.
    double* ws_alt;
    double* hist_alt;
    int cur_set;
    bool overlap_off;           // the second set did not fit, or a single segment
    hipStream_t rstream;
    hipEvent_t seg_ev[4];       // replay into set 0 / 1 done, reverse of set 0 / 1 done
};

namespace {

// A live session on a live grid (the grid may have been destroyed by its
// caller while the session still holds its device state: then only
// rkg_adjoint_destroy is accepted).
bool valid(const rkg_adjoint* a) {
    if (a == nullptr || a->magic != ADJ_MAGIC) return false;
    if (!grid_valid(a->g)) {
        set_error("the grid of this adjoint session has been destroyed");
        return false;
    }
    return true;
}

size_t P(const rkg_adjoint* a) { return a->g->plane; }
double* wsA(rkg_adjoint* a, int k) { return a->ws + (size_t)k * 3 * P(a); }
double* wsB(rkg_adjoint* a, int k) { return a->ws + (size_t)(a->K + 1 + k) * 3 * P(a); }
double* wsD1(rkg_adjoint* a, int k) { return a->ws + (size_t)(2 * a->K + 1 + k) * 3 * P(a); }
double* wsD2(rkg_adjoint* a, int k) { return a->ws + (size_t)(3 * a->K + 1 + k) * 3 * P(a); }
double* lamA(rkg_adjoint* a) { return a->lam; }
double* lamB(rkg_adjoint* a) { return a->lam + 3 * P(a); }
double* lamDiv(rkg_adjoint* a) { return a->lam + 6 * P(a); }
double* gradB(rkg_adjoint* a) { return a->lam + 9 * P(a); }
double* gradN(rkg_adjoint* a) { return a->lam + 10 * P(a); }

GridDesc stage_desc(rkg_adjoint* a, double* in, double* in2, double* out2) {
    GridDesc d = make_desc(a->g);
    d.ah = in; d.ahu = in + P(a); d.ahv = in + 2 * P(a);
    d.bh = in2; d.bhu = in2 + P(a); d.bhv = in2 + 2 * P(a);
    d.o2h = out2; d.o2hu = out2 + P(a); d.o2hv = out2 + 2 * P(a);
    d.sh = d.shu = d.shv = nullptr;
    d.mpart = nullptr;
    return d;
}

template <int STAGE>
void launch_rec(rkg_grid* g, const GridDesc& d, int p, hipStream_t st) {
    const int ntiles = d.nstrips * d.nsegs;
    const double tlim = g->cfg.t_end;
    if (g->cfg.use_hllc && d.fast)
        hipLaunchKernelGGL((k_stage<STAGE, true, true, true, 0, true>), dim3(ntiles), dim3(WAVE), 0, st, d, p, MODE_STEP, tlim);
    else if (g->cfg.use_hllc)
        hipLaunchKernelGGL((k_stage<STAGE, true, true, true>), dim3(ntiles), dim3(WAVE), 0, st, d, p, MODE_STEP, tlim);
    else if (d.fast)
        hipLaunchKernelGGL((k_stage<STAGE, false, true, true, 0, true>), dim3(ntiles), dim3(WAVE), 0, st, d, p, MODE_STEP, tlim);
    else
        hipLaunchKernelGGL((k_stage<STAGE, false, true, true>), dim3(ntiles), dim3(WAVE), 0, st, d, p, MODE_STEP, tlim);
}

// The stage's face reverse: k_adj_y then k_adj_x (both add into lam_in and
// grad_b).
template <bool HLLC, int STAGE>
void launch_faces(rkg_grid* g, const AdjStage& st) {
    const int* bc = g->cfg.bc;
    const bool tx = bc[0] == BC_TIDAL || bc[1] == BC_TIDAL, ty = bc[2] == BC_TIDAL || bc[3] == BC_TIDAL;
    const int nx = g->cfg.nx, ny = g->cfg.ny;
    const int yblocks = ((nx + WAVE - 1) / WAVE) * ((ny + st.ty - 1) / st.ty);
    const int xblocks = ((nx + AW - 1) / AW) * ((ny + st.xrows - 1) / st.xrows);
    if (ty)
        hipLaunchKernelGGL((k_adj_y<HLLC, STAGE, true>), dim3(yblocks), dim3(WAVE), 0, g->stream, st);
    else
        hipLaunchKernelGGL((k_adj_y<HLLC, STAGE, false>), dim3(yblocks), dim3(WAVE), 0, g->stream, st);
    if (tx)
        hipLaunchKernelGGL((k_adj_x<HLLC, STAGE, true>), dim3(xblocks), dim3(WAVE), 0, g->stream, st);
    else
        hipLaunchKernelGGL((k_adj_x<HLLC, STAGE, false>), dim3(xblocks), dim3(WAVE), 0, g->stream, st);
}

int body_blocks(const rkg_grid* g) {
    const long long n = (long long)g->cfg.nx * g->cfg.ny;
    return (int)std::min<long long>((n + 255) / 256, 2048);
}


}  // namespace

namespace {

// A session over gauges given as lower-left cells gauge_i / gauge_j and, for
// stations, the weights wx / wy (null for cell gauges).
int create_session(rkg_grid* g, int nsteps, int K, int ng, const int* gauge_i, const int* gauge_j,
                   const double* wx, const double* wy, const double* targets, rkg_adjoint** out) {
    if (g->scheme != RKG_TIME_RK2) {
        set_error("the adjoint is that of the SSP-RK2 step; set time scheme rk2");
        return RKG_ERR_UNSUPPORTED;
    }
    if (g->forcing) {
        set_error("the adjoint does not cover the surface forcing; switch it off");
        return RKG_ERR_UNSUPPORTED;
    }
    if (sponge_on(g)) {
        set_error("the adjoint does not cover the sponge relaxation; set its width to 0");
        return RKG_ERR_UNSUPPORTED;
    }
    for (int e = 0; e < 4; ++e)
        if (g->cfg.bc[e] == RKG_BC_CHAPMAN) {
            set_error("the adjoint does not cover Chapman edges");
            return RKG_ERR_UNSUPPORTED;
        }
    int rc = rkg_use_manning_field(g, 1);
    if (rc) return rc;
    rkg_adjoint* a = new rkg_adjoint();
    memset((void*)a, 0, sizeof *a);
    a->ckpt_parity = std::vector<int>();
    a->g = g;
    grid_retain(g);   // released by rkg_adjoint_destroy
    a->nsteps = nsteps;
    a->K = std::min(K, nsteps);
    a->nseg = (nsteps + a->K - 1) / a->K;
    a->ng = ng;
    a->ckpt_parity.assign(a->nseg, 0);
    const size_t Pl = g->plane;
    hipError_t e = hipSuccess;
    if (ng > 0) {
        e = hipMalloc((void**)&a->gi, ng * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&a->gj, ng * sizeof(int));
        if (e == hipSuccess) e = hipMemcpy(a->gi, gauge_i, ng * sizeof(int), hipMemcpyHostToDevice);
        // taps: one per cell gauge (weight 1), four per station (its
        // bilinear weights, gauges_write :519-537); distinct cells in order
        // of first appearance, each with its taps in gauge order
        const int nx = g->cfg.nx, ny = g->cfg.ny;
        std::vector<long long> keys;
        std::vector<std::vector<std::pair<int, double>>> members;
        auto tap = [&](int k, int c, int j, double w) {
            const long long key = (long long)j * g->pitch + c;
            size_t u = 0;
            while (u < keys.size() && keys[u] != key) ++u;
            if (u == keys.size()) { keys.push_back(key); members.emplace_back(); }
            members[u].push_back({k, w});
        };
        for (int k = 0; k < ng; ++k) {
            if (!wx) { tap(k, gauge_i[k], gauge_j[k], 1.0); continue; }
            const int c0 = gauge_i[k], j0 = gauge_j[k];
            const int c1 = std::min(c0 + 1, nx - 1), j1 = std::min(j0 + 1, ny - 1);
            tap(k, c0, j0, (1.0 - wx[k]) * (1.0 - wy[k]));
            tap(k, c1, j0, wx[k] * (1.0 - wy[k]));
            tap(k, c0, j1, (1.0 - wx[k]) * wy[k]);
            tap(k, c1, j1, wx[k] * wy[k]);
        }
        std::vector<int> start(1, 0), list;
        std::vector<double> wlist;
        for (const auto& m : members) {
            for (const auto& t : m) { list.push_back(t.first); wlist.push_back(t.second); }
            start.push_back((int)list.size());
        }
        a->ncells = (int)keys.size();
        if (e == hipSuccess) e = hipMalloc((void**)&a->cell_start, start.size() * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&a->cell_gauges, list.size() * sizeof(int));
        if (e == hipSuccess) e = hipMalloc((void**)&a->cell_w, wlist.size() * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void**)&a->cell_pos, keys.size() * sizeof(long long));
        if (e == hipSuccess) e = hipMemcpy(a->cell_start, start.data(), start.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(a->cell_gauges, list.data(), list.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(a->cell_w, wlist.data(), wlist.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(a->cell_pos, keys.data(), keys.size() * sizeof(long long), hipMemcpyHostToDevice);
        if (wx) {
            std::vector<double> w(wx, wx + ng);
            w.insert(w.end(), wy, wy + ng);
            if (e == hipSuccess) e = hipMalloc((void**)&a->gw, 2 * (size_t)ng * sizeof(double));
            if (e == hipSuccess) e = hipMemcpy(a->gw, w.data(), 2 * (size_t)ng * sizeof(double), hipMemcpyHostToDevice);
        }
        if (e == hipSuccess) e = hipMemcpy(a->gj, gauge_j, ng * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMalloc((void**)&a->tgt, (size_t)nsteps * ng * sizeof(double));
        if (e == hipSuccess) e = hipMalloc((void**)&a->eta, (size_t)nsteps * ng * sizeof(double));
        if (e == hipSuccess) e = hipMemset(a->tgt, 0, (size_t)nsteps * ng * sizeof(double));
        if (e == hipSuccess && targets)
            e = hipMemcpy(a->tgt, targets, (size_t)nsteps * ng * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMalloc((void**)&a->ckpt, (size_t)a->nseg * 3 * Pl * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&a->ckpt_ctrl, (size_t)a->nseg * 2 * sizeof(CtrlSlot));
    if (e == hipSuccess) e = hipMalloc((void**)&a->ws, (size_t)(12 * a->K + 3) * Pl * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&a->lam, 11 * Pl * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&a->hist, (size_t)a->K * HIST * sizeof(double));
    a->edge_len = std::max(g->cfg.nx, g->cfg.ny);
    if (e == hipSuccess)
        e = hipMalloc((void**)&a->part, (size_t)(3 * body_blocks(g)) * sizeof(double));   // part_dt, argmax values
    if (e == hipSuccess) e = hipMalloc((void**)&a->edge_lg, 2 * (size_t)edge_slot(4, 0, 0, a->edge_len) * sizeof(double));
    if (e == hipSuccess) e = hipMemset(a->edge_lg, 0, 2 * (size_t)edge_slot(4, 0, 0, a->edge_len) * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&a->part_idx, (size_t)body_blocks(g) * sizeof(long long));
    if (e == hipSuccess) e = hipMalloc((void**)&a->scal, 2 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&a->loss_dev, sizeof(double));
    if (e == hipSuccess) e = hipHostMalloc((void**)&a->loss_host, sizeof(double), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMemset(a->ws, 0, (size_t)(12 * a->K + 3) * Pl * sizeof(double));
    if (e == hipSuccess) e = hipMemset(a->lam, 0, 11 * Pl * sizeof(double));
    if (e == hipSuccess) e = hipEventCreate(&a->ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&a->ev[1]);
    a->magic = ADJ_MAGIC;
    if (e != hipSuccess) {
        set_error("adjoint allocation failed: %s", hipGetErrorString(e));
        rkg_adjoint_destroy(&a);
        return RKG_ERR_HIP;
    }
    *out = a;
    return RKG_OK;
}

}  // namespace

extern "C" {

int rkg_adjoint_create(rkg_grid* g, int nsteps, int K, int ng, const int* gauge_i, const int* gauge_j,
                       const double* targets, rkg_adjoint** out) {
    if (!out) { set_error("null argument"); return RKG_ERR_RANGE; }
    *out = nullptr;
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    if (nsteps < 1 || K < 1 || ng < 0 || (ng > 0 && (!gauge_i || !gauge_j))) {
        set_error("need nsteps >= 1, checkpoint_interval >= 1 and gauge indices");
        return RKG_ERR_RANGE;
    }
    for (int k = 0; k < ng; ++k)
        if (gauge_i[k] < 0 || gauge_i[k] >= g->cfg.nx || gauge_j[k] < 0 || gauge_j[k] >= g->cfg.ny) {
            set_error("gauge %d at (%d, %d) lies outside the %dx%d grid", k, gauge_i[k], gauge_j[k], g->cfg.nx,
                      g->cfg.ny);
            return RKG_ERR_RANGE;
        }
    return create_session(g, nsteps, K, ng, gauge_i, gauge_j, nullptr, nullptr, targets, out);
}

int rkg_adjoint_create_stations(rkg_grid* g, int nsteps, int K, int n, const double* x, const double* y,
                                const double* targets, rkg_adjoint** out) {
    if (!out) { set_error("null argument"); return RKG_ERR_RANGE; }
    *out = nullptr;
    if (!grid_valid(g)) return RKG_ERR_BAD_HANDLE;
    if (nsteps < 1 || K < 1 || n < 0 || n > RKG_MAX_GAUGES || (n > 0 && (!x || !y))) {
        set_error("need nsteps >= 1, checkpoint_interval >= 1 and station coordinates");
        return RKG_ERR_RANGE;
    }
    std::vector<int> ci(n), cj(n);
    std::vector<double> wx(n), wy(n);
    for (int k = 0; k < n; ++k)
        if (!station_weights(g->cfg, x[k], y[k], &ci[k], &cj[k], &wx[k], &wy[k])) {
            set_error("station %d at (%g, %g) m lies outside the domain", k, x[k], y[k]);
            return RKG_ERR_RANGE;
        }
    return create_session(g, nsteps, K, n, ci.data(), cj.data(), wx.data(), wy.data(), targets, out);
}

void rkg_adjoint_destroy(rkg_adjoint** pa) {
    if (!pa || !*pa) return;
    rkg_adjoint* a = *pa;
    if (a->magic != ADJ_MAGIC) { *pa = nullptr; return; }
    if (a->g && a->g->stream) (void)hipStreamSynchronize(a->g->stream);
    if (a->rstream) (void)hipStreamSynchronize(a->rstream);
    for (void* p : {(void*)a->gi, (void*)a->gj, (void*)a->gw, (void*)a->cell_start, (void*)a->cell_gauges,
                    (void*)a->cell_w, (void*)a->cell_pos, (void*)a->tgt, (void*)a->eta, (void*)a->ckpt, (void*)a->ckpt_ctrl,
                    (void*)a->ws, (void*)a->lam, (void*)a->hist, (void*)a->loss_dev, (void*)a->part,
                    (void*)a->part_idx, (void*)a->scal, (void*)a->edge_lg, (void*)a->ws_alt, (void*)a->hist_alt})
        if (p) (void)hipFree(p);
    if (a->rstream) (void)hipStreamDestroy(a->rstream);
    for (hipEvent_t ev : a->seg_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (a->loss_host) (void)hipHostFree(a->loss_host);
    if (a->ev[0]) (void)hipEventDestroy(a->ev[0]);
    if (a->ev[1]) (void)hipEventDestroy(a->ev[1]);
    rkg_grid* g = a->g;
    a->magic = ADJ_DEAD;
    delete a;
    *pa = nullptr;
    if (g) grid_release(g);
}

int rkg_adjoint_set_targets(rkg_adjoint* a, const double* targets) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    if (a->ng == 0) return RKG_OK;
    if (!targets) return RKG_ERR_RANGE;
    HIP_TRY(hipMemcpy(a->tgt, targets, (size_t)a->nsteps * a->ng * sizeof(double), hipMemcpyHostToDevice));
    return RKG_OK;
}

int rkg_adjoint_gauge_eta(rkg_adjoint* a, double* eta) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    if (!a->have_forward) { set_error("no forward run yet"); return RKG_ERR_NOT_READY; }
    if (a->ng == 0) return RKG_OK;
    if (!eta) return RKG_ERR_RANGE;
    HIP_TRY(hipMemcpy(eta, a->eta, (size_t)a->nsteps * a->ng * sizeof(double), hipMemcpyDeviceToHost));
    return RKG_OK;
}

}  // extern "C"

namespace {

GaugeSpec gauge_spec(const rkg_adjoint* a) {
    return GaugeSpec{a->gi, a->gj, a->gw, a->ng, a->g->cfg.nx, a->g->cfg.ny};
}

// --- the forward run and the reverse sweep, as per-step pieces that queue
// work on the session's stream without waiting, so that several sessions on
// several streams can be interleaved step by step ---------------------------

int fwd_begin(rkg_adjoint* a) {
    rkg_grid* g = a->g;
    if (g->scheme != RKG_TIME_RK2) {
        set_error("the adjoint is that of the SSP-RK2 step; set time scheme rk2");
        return RKG_ERR_UNSUPPORTED;
    }
    if (g->forcing) {
        set_error("the adjoint does not cover the surface forcing; switch it off");
        return RKG_ERR_UNSUPPORTED;
    }
    if (sponge_on(g)) {
        set_error("the adjoint does not cover the sponge relaxation; set its width to 0");
        return RKG_ERR_UNSUPPORTED;
    }
    for (int e = 0; e < 4; ++e)
        if (g->cfg.bc[e] == RKG_BC_CHAPMAN) {
            set_error("the adjoint does not cover Chapman edges");
            return RKG_ERR_UNSUPPORTED;
        }
    HIP_TRY(hipMemsetAsync(a->loss_dev, 0, sizeof(double), g->stream));
    HIP_TRY(hipEventRecord(a->ev[0], g->stream));
    return RKG_OK;
}

// Step n of the forward run: the checkpoint at a segment start, the step,
// the gauge misfit.
int fwd_step(rkg_adjoint* a, int n) {
    rkg_grid* g = a->g;
    const size_t Pl = g->plane;
    if (n % a->K == 0) {
        const int s = n / a->K;
        if (g->speed_stale) {   // make the control slot current before saving it
            int rc = rkg_step(g, 0);
            if (rc) return rc;
        }
        HIP_TRY(hipMemcpyAsync(a->ckpt + (size_t)s * 3 * Pl, g->buf, 3 * Pl * sizeof(double), hipMemcpyDeviceToDevice,
                               g->stream));
        HIP_TRY(hipMemcpyAsync(a->ckpt_ctrl + 2 * s, g->ctrl, 2 * sizeof(CtrlSlot), hipMemcpyDeviceToDevice,
                               g->stream));
        a->ckpt_parity[s] = g->parity;
    }
    int rc = rkg_step(g, 1);
    if (rc) return rc;
    if (a->ng > 0)
        hipLaunchKernelGGL(k_gauge_loss, dim3(1), dim3(256), 0, g->stream, g->buf, g->buf + 6 * Pl,
                           (long long)g->pitch, gauge_spec(a), a->tgt + (size_t)n * a->ng,
                           a->eta + (size_t)n * a->ng, a->loss_dev);
    return RKG_OK;
}

int fwd_end(rkg_adjoint* a) {
    rkg_grid* g = a->g;
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(a->ev[1], g->stream));
    HIP_TRY(hipMemcpyAsync(a->end_ctrl, g->ctrl, 2 * sizeof(CtrlSlot), hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipMemcpyAsync(a->loss_host, a->loss_dev, sizeof(double), hipMemcpyDeviceToHost, g->stream));
    return RKG_OK;
}

// After the stream has drained.
int fwd_finish(rkg_adjoint* a, double* loss) {
    HIP_TRY(hipEventElapsedTime(&a->fwd_ms, a->ev[0], a->ev[1]));
    a->end_parity = a->g->parity;
    a->have_forward = true;
    if (loss) *loss = *a->loss_host;
    return RKG_OK;
}

// Per-sweep settings of one session.
struct BwdPlan {
    int tile_rows, xrows, nb, nargmax;
    bool cfl_dt;
    long long edge_size;
    double scale;
    int p;                      // parity of the replay
};

int bwd_begin(rkg_adjoint* a, double loss_scale, int tile_rows, BwdPlan& pl) {
    if (!a->have_forward) { set_error("rkg_adjoint_backward before rkg_adjoint_forward"); return RKG_ERR_NOT_READY; }
    rkg_grid* g = a->g;
    // rows per workgroup: about 2048 workgroups of k_adj_y and of k_adj_x
    // (two waves on each of the 1024 SIMDs); on the 1024^2 basin 8 rows of
    // k_adj_y (2048 workgroups) beat 9 (1821) by 1 %, and 9 rows of k_adj_x
    // (2048) beat 10 (1844) by about 1 %
    const long long ny = g->cfg.ny;
    const long long ystrips = (g->cfg.nx + WAVE - 1) / WAVE, xstrips = (g->cfg.nx + AW - 1) / AW;
    if (tile_rows <= 0) tile_rows = (int)std::min<long long>(32, std::max<long long>(4, (ny * ystrips + 2047) / 2048));
    if (tile_rows < 4 || tile_rows > 256) { set_error("tile_rows must lie in [4, 256]"); return RKG_ERR_RANGE; }
    pl.tile_rows = tile_rows;
    pl.xrows = (int)std::min<long long>(32, std::max<long long>(1, (ny * xstrips + 2047) / 2048));
    pl.nb = body_blocks(g);
    pl.cfl_dt = g->cfg.dt_fixed <= 0.0;
    pl.nargmax = pl.cfl_dt ? pl.nb : 0;   // argmax partials, one per body workgroup of stage 1
    pl.edge_size = edge_slot(4, 0, 0, a->edge_len);
    pl.scale = loss_scale;
    pl.p = 0;
    const size_t Pl = g->plane;
    HIP_TRY(hipEventRecord(a->ev[0], g->stream));
    HIP_TRY(hipMemsetAsync(a->lam, 0, 11 * Pl * sizeof(double), g->stream));
    HIP_TRY(hipMemsetAsync(a->scal, 0, 2 * sizeof(double), g->stream));
    return RKG_OK;
}

int seg_steps(const rkg_adjoint* a, int s) { return std::min(a->nsteps, (s + 1) * a->K) - s * a->K; }

// Start of the replay of segment s on stream st: its checkpoint into A_0
// and the clock.
int bwd_replay_begin(rkg_adjoint* a, int s, BwdPlan& pl, hipStream_t st) {
    rkg_grid* g = a->g;
    const size_t Pl = g->plane;
    HIP_TRY(hipMemcpyAsync(wsA(a, 0), a->ckpt + (size_t)s * 3 * Pl, 3 * Pl * sizeof(double), hipMemcpyDeviceToDevice,
                           st));
    HIP_TRY(hipMemcpyAsync(g->ctrl, a->ckpt_ctrl + 2 * s, 2 * sizeof(CtrlSlot), hipMemcpyDeviceToDevice, st));
    pl.p = a->ckpt_parity[s];
    return RKG_OK;
}

// Step k of the segment's replay, recording U^n, U*, both divergences and
// the clock.
void bwd_replay_step(rkg_adjoint* a, int k, BwdPlan& pl, hipStream_t st) {
    rkg_grid* g = a->g;
    GridDesc d1 = stage_desc(a, wsA(a, k), wsB(a, k), wsA(a, k + 1));
    d1.rec = wsD1(a, k);
    d1.hist = a->hist + HIST * k;
    launch_rec<1>(g, d1, pl.p, st);
    GridDesc d2 = d1;
    d2.rec = wsD2(a, k);
    d2.hist = nullptr;
    launch_rec<2>(g, d2, pl.p, st);
    pl.p ^= 1;
}

// Puts workspace set `set` (0 or 1) in ws / hist.
void use_set(rkg_adjoint* a, int set) {
    if (set == a->cur_set) return;
    std::swap(a->ws, a->ws_alt);
    std::swap(a->hist, a->hist_alt);
    a->cur_set = set;
}

size_t ws_bytes(const rkg_adjoint* a) { return (size_t)(12 * a->K + 3) * a->g->plane * sizeof(double); }

// The second workspace, stream and events of the overlapped replay,
// allocated at the first backward sweep that can use them: a session of
// two or more segments whose second workspace takes at most half of the
// device's free memory.  Otherwise, or when an allocation fails (what was
// allocated is released), the sweeps replay on the grid's stream.
void ensure_overlap(rkg_adjoint* a) {
    if (a->rstream || a->overlap_off) return;
    a->overlap_off = true;
    if (a->nseg < 2) return;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void)hipGetLastError(); return; }
    const size_t hist_b = (size_t)a->K * HIST * sizeof(double);
    if (ws_bytes(a) + hist_b > free_b / 2) return;
    hipError_t e = hipMalloc((void**)&a->ws_alt, ws_bytes(a));
    if (e == hipSuccess) e = hipMalloc((void**)&a->hist_alt, hist_b);
    if (e == hipSuccess) e = hipMemset(a->ws_alt, 0, ws_bytes(a));
    for (int q = 0; q < 4 && e == hipSuccess; ++q) e = hipEventCreateWithFlags(&a->seg_ev[q], hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&a->rstream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (a->ws_alt) (void)hipFree(a->ws_alt);
        if (a->hist_alt) (void)hipFree(a->hist_alt);
        for (hipEvent_t& ev : a->seg_ev)
            if (ev) { (void)hipEventDestroy(ev); ev = nullptr; }
        a->ws_alt = a->hist_alt = nullptr;
        a->rstream = nullptr;
        return;
    }
    a->overlap_off = false;
}

// Arguments of the kernels that reverse step k of the current segment
// (global step n): the gauge seed, the two stages (STAGE 2 first), and the
// dt/clock kernel.
SeedArgs seed_args(rkg_adjoint* a, int k, int n, const BwdPlan& pl) {
    rkg_grid* g = a->g;
    SeedArgs A;
    A.h = wsA(a, k + 1);
    A.bed = g->buf + 6 * g->plane;
    A.pitch = (long long)g->pitch;
    A.G = gauge_spec(a);
    A.tgt = a->tgt + (size_t)n * a->ng;
    A.cell_start = a->cell_start;
    A.cell_gauges = a->cell_gauges;
    A.cell_w = a->cell_w;
    A.cell_pos = a->cell_pos;
    A.ncells = a->ncells;
    A.scale = pl.scale;
    A.lam_h = lamA(a);
    A.grad_b = gradB(a);
    return A;
}

AdjStage stage_args(rkg_adjoint* a, int k, int stage, const BwdPlan& pl) {
    const int nb = pl.nb;
    AdjStage st;
    st.lam_div = lamDiv(a);
    st.grad_b = gradB(a);
    st.grad_n = gradN(a);
    st.hist = a->hist + HIST * k;
    st.edge_len = a->edge_len;
    st.ty = pl.tile_rows;
    st.xrows = pl.xrows;
    if (stage == 2) {   // input U* (B_k), U^n = A_k
        st.d = stage_desc(a, wsB(a, k), wsB(a, k), wsA(a, k + 1));
        st.un = wsA(a, k);
        st.div = wsD2(a, k);
        st.lam_in = lamB(a);
        st.lam_out = lamA(a);
        st.part_dt = a->part + nb;
        st.pval = nullptr;
        st.pidx = nullptr;
        st.edge_lg = a->edge_lg + pl.edge_size;
    } else {            // input U^n (A_k)
        st.d = stage_desc(a, wsA(a, k), wsB(a, k), wsA(a, k + 1));
        st.un = nullptr;
        st.div = wsD1(a, k);
        st.lam_in = lamA(a);
        st.lam_out = lamB(a);
        st.part_dt = a->part;
        st.pval = pl.cfl_dt ? a->part + 2 * nb : nullptr;
        st.pidx = pl.cfl_dt ? a->part_idx : nullptr;
        st.edge_lg = a->edge_lg;
    }
    return st;
}

DtArgs dt_args(rkg_adjoint* a, int k, const AdjStage& st1, const BwdPlan& pl) {
    DtArgs A;
    A.d = st1.d;
    A.hist = st1.hist;
    A.part_dt = a->part;
    A.nbody = pl.nb;
    A.in2 = wsB(a, k);
    A.edge_lg = a->edge_lg;
    A.edge_len = a->edge_len;
    A.pval = a->part + 2 * pl.nb;
    A.pidx = a->part_idx;
    A.nargmax = pl.nargmax;
    A.lam = lamA(a);
    A.scal = a->scal;
    return A;
}

// Reverse of step k of segment s (global step n0 + k).
void bwd_reverse_step(rkg_adjoint* a, int s, int k, const BwdPlan& pl) {
    rkg_grid* g = a->g;
    const int n = s * a->K + k;
    const int nb = pl.nb;
    if (a->ng > 0)
        hipLaunchKernelGGL(k_gauge_seed, dim3((a->ncells + 255) / 256), dim3(256), 0, g->stream,
                           seed_args(a, k, n, pl));
    const AdjStage st2 = stage_args(a, k, 2, pl);
    hipLaunchKernelGGL(k_adj_body<2>, dim3(nb), dim3(256), 0, g->stream, st2);
    if (g->cfg.use_hllc) launch_faces<true, 2>(g, st2);
    else launch_faces<false, 2>(g, st2);
    const AdjStage st1 = stage_args(a, k, 1, pl);
    hipLaunchKernelGGL(k_adj_body<1>, dim3(nb), dim3(256), 0, g->stream, st1);
    if (g->cfg.use_hllc) launch_faces<true, 1>(g, st1);
    else launch_faces<false, 1>(g, st1);
    // the step's dt and clock
    hipLaunchKernelGGL(k_dt_rev, dim3(1), dim3(1024), 0, g->stream, dt_args(a, k, st1, pl));
}

// The sweep with the replay of segment s - 1 (into workspace set (s-1)&1,
// on rstream) overlapping the reverse of segment s (set s&1, on the grid's
// stream).  Events order each set's replay before its reverse and a set's
// next replay after the reverse that last read it; the kernels and their
// inputs are those of the serial sweep, so the gradients are the same bits.
int bwd_overlapped(rkg_adjoint* a, BwdPlan& pl) {
    hipStream_t main = a->g->stream, rs = a->rstream;
    // the replays start after the work already queued on the grid's stream
    HIP_TRY(hipEventRecord(a->seg_ev[2], main));
    HIP_TRY(hipStreamWaitEvent(rs, a->seg_ev[2], 0));
    auto replay = [&](int s) -> int {
        use_set(a, s & 1);
        int rc = bwd_replay_begin(a, s, pl, rs);
        if (rc) return rc;
        for (int k = 0; k < seg_steps(a, s); ++k) bwd_replay_step(a, k, pl, rs);
        HIP_TRY(hipEventRecord(a->seg_ev[s & 1], rs));
        return RKG_OK;
    };
    int rc = replay(a->nseg - 1);
    for (int s = a->nseg - 1; s >= 0 && !rc; --s) {
        if (s > 0) {
            // set (s-1)&1 was last read by the reverse of segment s + 1
            if (s + 1 < a->nseg) HIP_TRY(hipStreamWaitEvent(rs, a->seg_ev[2 + ((s + 1) & 1)], 0));
            rc = replay(s - 1);
            if (rc) break;
        }
        HIP_TRY(hipStreamWaitEvent(main, a->seg_ev[s & 1], 0));
        use_set(a, s & 1);
        for (int k = seg_steps(a, s) - 1; k >= 0; --k) bwd_reverse_step(a, s, k, pl);
        HIP_TRY(hipEventRecord(a->seg_ev[2 + (s & 1)], main));
    }
    use_set(a, 0);
    return rc;
}

int bwd_end(rkg_adjoint* a) {
    rkg_grid* g = a->g;
    HIP_TRY(hipGetLastError());
    // the grid's control slots back to the end of the forward run
    HIP_TRY(hipMemcpyAsync(g->ctrl, a->end_ctrl, 2 * sizeof(CtrlSlot), hipMemcpyHostToDevice, g->stream));
    g->parity = a->end_parity;
    HIP_TRY(hipEventRecord(a->ev[1], g->stream));
    return RKG_OK;
}

int bwd_finish(rkg_adjoint* a) {
    HIP_TRY(hipEventElapsedTime(&a->bwd_ms, a->ev[0], a->ev[1]));
    return RKG_OK;
}

// The reverse sweeps of n sessions as one sweep on the first session's
// stream: per replay step two launches (both stages of every member) and
// per reverse step eight (gauge seeds, the two stages' bodies and faces, the
// dt/clock kernels), each over the blocks of every member.  Every kernel
// argument of the whole sweep is known in advance (the sweep uses the first
// workspace set of every session) and is uploaded once.  The members'
// kernels and their block decompositions are those of their single-session
// sweeps, so each member's gradients are the same bits.  All sessions must
// use the same Riemann solver (HLL or HLLC).
int bwd_batched(int n, rkg_adjoint* const* as, double loss_scale) {
    const bool hllc = as[0]->g->cfg.use_hllc != 0;
    const int nsteps = as[0]->nsteps, K = as[0]->K, nseg = as[0]->nseg;
    std::vector<BwdPlan> pl(n);
    int rc = RKG_OK;
    for (int m = 0; m < n && !rc; ++m) rc = bwd_begin(as[m], loss_scale, 0, pl[m]);
    if (rc) return rc;
    for (int m = 0; m < n; ++m) use_set(as[m], 0);
    hipStream_t S = as[0]->g->stream;

    // block tables
    std::vector<BlockRef> tb, ty, tx, tseed;
    std::vector<RecTile> trec;
    for (int m = 0; m < n; ++m) {
        rkg_grid* g = as[m]->g;
        const int nx = g->cfg.nx, ny = g->cfg.ny;
        const int nb = pl[m].nb;
        for (int b = 0; b < nb; ++b) tb.push_back(BlockRef{m, b, nb, 0});
        const int yb = ((nx + WAVE - 1) / WAVE) * ((ny + pl[m].tile_rows - 1) / pl[m].tile_rows);
        for (int b = 0; b < yb; ++b) ty.push_back(BlockRef{m, b, yb, 0});
        const int xb = ((nx + AW - 1) / AW) * ((ny + pl[m].xrows - 1) / pl[m].xrows);
        for (int b = 0; b < xb; ++b) tx.push_back(BlockRef{m, b, xb, 0});
        if (as[m]->ng > 0) {
            const int sb = (as[m]->ncells + 255) / 256;
            for (int b = 0; b < sb; ++b) tseed.push_back(BlockRef{m, b, sb, 0});
        }
        const GridDesc d0 = make_desc(g);
        const int ntiles = d0.nstrips * d0.nsegs;
        for (int t = 0; t < ntiles; ++t) {   // the single-grid kernel's tile order and leader
            RecTile r;
            r.member = m;
            r.strip = t % d0.nstrips;
            r.seg = t / d0.nstrips;
            r.leader = t == 0;
            trec.push_back(r);
        }
    }
    // per global step: replay descriptors (stage 1, 2), replay parities, and
    // the reverse arguments (stage 2, stage 1, seeds, dt), n of each
    const size_t szGD = sizeof(GridDesc), szAS = sizeof(AdjStage), szSA = sizeof(SeedArgs), szDA = sizeof(DtArgs);
    auto al = [](size_t x) { return (x + 255) / 256 * 256; };
    const size_t oRec1 = 0, oRec2 = oRec1 + al(n * szGD), oPar = oRec2 + al(n * szGD), oSt2 = oPar + al(n * sizeof(int)),
                 oSt1 = oSt2 + al(n * szAS), oSeed = oSt1 + al(n * szAS), oDt = oSeed + al(n * szSA),
                 step_bytes = oDt + al(n * szDA);
    const size_t tab_bytes = al(tb.size() * sizeof(BlockRef)) + al(ty.size() * sizeof(BlockRef)) +
                             al(tx.size() * sizeof(BlockRef)) + al(std::max<size_t>(1, tseed.size()) * sizeof(BlockRef)) +
                             al(trec.size() * sizeof(RecTile));
    const size_t total = tab_bytes + step_bytes * (size_t)nsteps;
    char* host = nullptr;
    char* dev = nullptr;
    hipEvent_t join = nullptr;
    auto cleanup = [&]() {
        if (host) (void)hipHostFree(host);
        if (dev) (void)hipFree(dev);
        if (join) (void)hipEventDestroy(join);
    };
    hipError_t e = hipHostMalloc((void**)&host, total, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&dev, total);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&join, hipEventDisableTiming);
    if (e != hipSuccess) {
        cleanup();
        set_error("batched adjoint sweep: cannot allocate its %zu-byte argument tables: %s", total, hipGetErrorString(e));
        return RKG_ERR_HIP;
    }
    size_t off = 0;
    auto put = [&](const void* src, size_t bytes) {
        const size_t o = off;
        if (bytes) std::memcpy(host + o, src, bytes);
        off += al(bytes);
        return dev + o;
    };
    const BlockRef* dTb = (const BlockRef*)put(tb.data(), tb.size() * sizeof(BlockRef));
    const BlockRef* dTy = (const BlockRef*)put(ty.data(), ty.size() * sizeof(BlockRef));
    const BlockRef* dTx = (const BlockRef*)put(tx.data(), tx.size() * sizeof(BlockRef));
    const BlockRef* dTs = (const BlockRef*)put(tseed.data(), tseed.size() * sizeof(BlockRef));
    if (tseed.empty()) off += 256;
    const RecTile* dTr = (const RecTile*)put(trec.data(), trec.size() * sizeof(RecTile));
    off = tab_bytes;
    for (int s = 0; s < nseg; ++s) {
        for (int k = 0; k < seg_steps(as[0], s); ++k) {
            char* h = host + tab_bytes + step_bytes * (size_t)(s * K + k);
            for (int m = 0; m < n; ++m) {
                rkg_adjoint* a = as[m];
                GridDesc d1 = stage_desc(a, wsA(a, k), wsB(a, k), wsA(a, k + 1));
                d1.rec = wsD1(a, k);
                d1.hist = a->hist + HIST * k;
                GridDesc d2 = d1;
                d2.rec = wsD2(a, k);
                d2.hist = nullptr;
                std::memcpy(h + oRec1 + m * szGD, &d1, szGD);
                std::memcpy(h + oRec2 + m * szGD, &d2, szGD);
                const int par = a->ckpt_parity[s] ^ (k & 1);
                std::memcpy(h + oPar + m * sizeof(int), &par, sizeof(int));
                const AdjStage st2 = stage_args(a, k, 2, pl[m]);
                const AdjStage st1 = stage_args(a, k, 1, pl[m]);
                const SeedArgs sa = seed_args(a, k, s * K + k, pl[m]);
                const DtArgs da = dt_args(a, k, st1, pl[m]);
                std::memcpy(h + oSt2 + m * szAS, &st2, szAS);
                std::memcpy(h + oSt1 + m * szAS, &st1, szAS);
                std::memcpy(h + oSeed + m * szSA, &sa, szSA);
                std::memcpy(h + oDt + m * szDA, &da, szDA);
            }
        }
    }
    // the sweep starts after every member's queued work (bwd_begin's clears)
    for (int m = 1; m < n && e == hipSuccess; ++m) {
        e = hipEventRecord(join, as[m]->g->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(S, join, 0);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(dev, host, total, hipMemcpyHostToDevice, S);
    if (e != hipSuccess) {
        cleanup();
        set_error("batched adjoint sweep: %s", hipGetErrorString(e));
        return RKG_ERR_HIP;
    }
    const int nrec = (int)trec.size();
    for (int s = nseg - 1; s >= 0 && !rc; --s) {
        for (int m = 0; m < n && !rc; ++m) rc = bwd_replay_begin(as[m], s, pl[m], S);
        if (rc) break;
        const int Ks = seg_steps(as[0], s);
        for (int k = 0; k < Ks; ++k) {
            const char* d = dev + tab_bytes + step_bytes * (size_t)(s * K + k);
            const GridDesc* r1 = (const GridDesc*)(d + oRec1);
            const GridDesc* r2 = (const GridDesc*)(d + oRec2);
            const int* par = (const int*)(d + oPar);
            // one launch over every member: the fast path by the total tile count
            const bool fk = (long long)nrec >= FAST_MIN_TILES;
            if (hllc && fk) {
                hipLaunchKernelGGL((k_replay_batch<1, true, true>), dim3(nrec), dim3(WAVE), 0, S, r1, dTr, par);
                hipLaunchKernelGGL((k_replay_batch<2, true, true>), dim3(nrec), dim3(WAVE), 0, S, r2, dTr, par);
            } else if (hllc) {
                hipLaunchKernelGGL((k_replay_batch<1, true, false>), dim3(nrec), dim3(WAVE), 0, S, r1, dTr, par);
                hipLaunchKernelGGL((k_replay_batch<2, true, false>), dim3(nrec), dim3(WAVE), 0, S, r2, dTr, par);
            } else if (fk) {
                hipLaunchKernelGGL((k_replay_batch<1, false, true>), dim3(nrec), dim3(WAVE), 0, S, r1, dTr, par);
                hipLaunchKernelGGL((k_replay_batch<2, false, true>), dim3(nrec), dim3(WAVE), 0, S, r2, dTr, par);
            } else {
                hipLaunchKernelGGL((k_replay_batch<1, false, false>), dim3(nrec), dim3(WAVE), 0, S, r1, dTr, par);
                hipLaunchKernelGGL((k_replay_batch<2, false, false>), dim3(nrec), dim3(WAVE), 0, S, r2, dTr, par);
            }
        }
        for (int k = Ks - 1; k >= 0; --k) {
            const char* d = dev + tab_bytes + step_bytes * (size_t)(s * K + k);
            const AdjStage* st2 = (const AdjStage*)(d + oSt2);
            const AdjStage* st1 = (const AdjStage*)(d + oSt1);
            if (!tseed.empty())
                hipLaunchKernelGGL(k_gauge_seed_batch, dim3((int)tseed.size()), dim3(256), 0, S,
                                   (const SeedArgs*)(d + oSeed), dTs);
            hipLaunchKernelGGL(k_adj_body_batch<2>, dim3((int)tb.size()), dim3(256), 0, S, st2, dTb);
            if (hllc) {
                hipLaunchKernelGGL((k_adj_y_batch<true, 2>), dim3((int)ty.size()), dim3(WAVE), 0, S, st2, dTy);
                hipLaunchKernelGGL((k_adj_x_batch<true, 2>), dim3((int)tx.size()), dim3(WAVE), 0, S, st2, dTx);
            } else {
                hipLaunchKernelGGL((k_adj_y_batch<false, 2>), dim3((int)ty.size()), dim3(WAVE), 0, S, st2, dTy);
                hipLaunchKernelGGL((k_adj_x_batch<false, 2>), dim3((int)tx.size()), dim3(WAVE), 0, S, st2, dTx);
            }
            hipLaunchKernelGGL(k_adj_body_batch<1>, dim3((int)tb.size()), dim3(256), 0, S, st1, dTb);
            if (hllc) {
                hipLaunchKernelGGL((k_adj_y_batch<true, 1>), dim3((int)ty.size()), dim3(WAVE), 0, S, st1, dTy);
                hipLaunchKernelGGL((k_adj_x_batch<true, 1>), dim3((int)tx.size()), dim3(WAVE), 0, S, st1, dTx);
            } else {
                hipLaunchKernelGGL((k_adj_y_batch<false, 1>), dim3((int)ty.size()), dim3(WAVE), 0, S, st1, dTy);
                hipLaunchKernelGGL((k_adj_x_batch<false, 1>), dim3((int)tx.size()), dim3(WAVE), 0, S, st1, dTx);
            }
            hipLaunchKernelGGL(k_dt_rev_batch, dim3(n), dim3(1024), 0, S, (const DtArgs*)(d + oDt));
        }
    }
    if (!rc) e = hipGetLastError();
    // every member's stream continues after the sweep
    if (!rc && e == hipSuccess) e = hipEventRecord(join, S);
    for (int m = 1; m < n && !rc && e == hipSuccess; ++m) e = hipStreamWaitEvent(as[m]->g->stream, join, 0);
    if (!rc && e != hipSuccess) { set_error("batched adjoint sweep: %s", hipGetErrorString(e)); rc = RKG_ERR_HIP; }
    for (int m = 0; m < n && !rc; ++m) rc = bwd_end(as[m]);
    // the tables are freed once every launch that reads them has finished
    for (int m = 0; m < n; ++m) (void)hipStreamSynchronize(as[m]->g->stream);
    cleanup();
    if (rc) return rc;
    for (int m = 0; m < n && !rc; ++m) rc = bwd_finish(as[m]);
    return rc;
}

// Sessions of one batch: valid, distinct grids on distinct streams, with
// one step count and checkpoint interval.
int check_batch(int n, rkg_adjoint* const* as) {
    if (n < 1 || !as) { set_error("need at least one session"); return RKG_ERR_RANGE; }
    for (int m = 0; m < n; ++m) {
        if (!valid(as[m])) return RKG_ERR_BAD_HANDLE;
        if (as[m]->nsteps != as[0]->nsteps || as[m]->K != as[0]->K) {
            set_error("sessions of a batch need the same step count and checkpoint interval");
            return RKG_ERR_RANGE;
        }
        for (int q = 0; q < m; ++q)
            if (as[q]->g == as[m]->g || as[q]->g->stream == as[m]->g->stream) {
                set_error("sessions %d and %d share a grid or a stream", q, m);
                return RKG_ERR_RANGE;
            }
    }
    return RKG_OK;
}

}  // namespace

extern "C" {

// nsteps steps from the grid's current state, recording checkpoints and the
// gauge misfit; the grid is left at the final state.
int rkg_adjoint_forward(rkg_adjoint* a, double* loss) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    int rc = fwd_begin(a);
    for (int n = 0; !rc && n < a->nsteps; ++n) rc = fwd_step(a, n);
    if (!rc) rc = fwd_end(a);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(a->g->stream));
    return fwd_finish(a, loss);
}

// Reverse sweep: gradients of loss_scale * L with respect to the bed, the
// per-cell Manning n and the initial state (h, hu, hv).  The grid's state
// is left as the forward run left it.
int rkg_adjoint_backward(rkg_adjoint* a, double loss_scale, int tile_rows) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    if (!a->have_forward) { set_error("rkg_adjoint_backward before rkg_adjoint_forward"); return RKG_ERR_NOT_READY; }
    ensure_overlap(a);   // before the sweep's start event: the first call allocates
    BwdPlan pl;
    int rc = bwd_begin(a, loss_scale, tile_rows, pl);
    if (rc) return rc;
    if (a->rstream) {
        rc = bwd_overlapped(a, pl);
    } else {
        for (int s = a->nseg - 1; s >= 0 && !rc; --s) {
            rc = bwd_replay_begin(a, s, pl, a->g->stream);
            if (rc) break;
            for (int k = 0; k < seg_steps(a, s); ++k) bwd_replay_step(a, k, pl, a->g->stream);
            for (int k = seg_steps(a, s) - 1; k >= 0; --k) bwd_reverse_step(a, s, k, pl);
        }
    }
    if (rc) return rc;
    rc = bwd_end(a);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(a->g->stream));
    return bwd_finish(a);
}

// The forward runs of n sessions (one grid and stream each), queued step
// by step across the sessions so that their kernels overlap.
int rkg_adjoint_forward_many(int n, rkg_adjoint* const* as, double* losses) {
    int rc = check_batch(n, as);
    if (rc) return rc;
    for (int m = 0; m < n && !rc; ++m) rc = fwd_begin(as[m]);
    for (int k = 0; k < as[0]->nsteps && !rc; ++k)
        for (int m = 0; m < n && !rc; ++m) rc = fwd_step(as[m], k);
    for (int m = 0; m < n && !rc; ++m) rc = fwd_end(as[m]);
    if (rc) return rc;
    for (int m = 0; m < n; ++m) HIP_TRY(hipStreamSynchronize(as[m]->g->stream));
    for (int m = 0; m < n && !rc; ++m) rc = fwd_finish(as[m], losses ? losses + m : nullptr);
    return rc;
}

// The reverse sweeps of n sessions, queued step by step across the
// sessions (replays, then reverse steps, segment by segment).  Each
// session's gradients are those of rkg_adjoint_backward.
int rkg_adjoint_backward_many(int n, rkg_adjoint* const* as, double loss_scale) {
    int rc = check_batch(n, as);
    if (rc) return rc;
    bool same_solver = true;
    for (int m = 1; m < n; ++m) same_solver = same_solver && as[m]->g->cfg.use_hllc == as[0]->g->cfg.use_hllc;
    if (same_solver) return bwd_batched(n, as, loss_scale);
    // sessions mixing HLL and HLLC: their own kernels, interleaved on their streams
    std::vector<BwdPlan> pl(n);
    for (int m = 0; m < n && !rc; ++m) rc = bwd_begin(as[m], loss_scale, 0, pl[m]);
    for (int s = as[0]->nseg - 1; s >= 0 && !rc; --s) {
        for (int m = 0; m < n && !rc; ++m) rc = bwd_replay_begin(as[m], s, pl[m], as[m]->g->stream);
        if (rc) break;
        const int Ks = seg_steps(as[0], s);
        for (int k = 0; k < Ks; ++k)
            for (int m = 0; m < n; ++m) bwd_replay_step(as[m], k, pl[m], as[m]->g->stream);
        for (int k = Ks - 1; k >= 0; --k)
            for (int m = 0; m < n; ++m) bwd_reverse_step(as[m], s, k, pl[m]);
    }
    for (int m = 0; m < n && !rc; ++m) rc = bwd_end(as[m]);
    if (rc) return rc;
    for (int m = 0; m < n; ++m) HIP_TRY(hipStreamSynchronize(as[m]->g->stream));
    for (int m = 0; m < n && !rc; ++m) rc = bwd_finish(as[m]);
    return rc;
}

// Adds the adjoint of a still-water initial state (h = max(eta0 - b, 0),
// hu = hv = 0) to the bed gradient.
int rkg_adjoint_fold_still_water(rkg_adjoint* a, double eta0) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    rkg_grid* g = a->g;
    hipLaunchKernelGGL(k_fold_still_water, dim3(body_blocks(g)), dim3(256), 0, g->stream, g->buf + 6 * g->plane,
                       lamA(a), eta0, g->cfg.nx, g->cfg.ny, (long long)g->pitch, gradB(a));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(g->stream));
    return RKG_OK;
}

// field: RKG_FIELD_B (bed), RKG_FIELD_N (Manning n), RKG_FIELD_H/HU/HV
// (initial state).  on_device != 0: dst is device memory.
int rkg_adjoint_gradient(rkg_adjoint* a, int field, double* dst, int ld, int on_device) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    rkg_grid* g = a->g;
    const double* src;
    switch (field) {
    case RKG_FIELD_H: src = lamA(a); break;
    case RKG_FIELD_HU: src = lamA(a) + g->plane; break;
    case RKG_FIELD_HV: src = lamA(a) + 2 * g->plane; break;
    case RKG_FIELD_B: src = gradB(a); break;
    case RKG_FIELD_N: src = gradN(a); break;
    default: set_error("no gradient for field %d", field); return RKG_ERR_RANGE;
    }
    if (!dst || ld < g->cfg.nx) { set_error("bad buffer or leading dimension"); return RKG_ERR_RANGE; }
    HIP_TRY(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), src, (size_t)g->pitch * sizeof(double),
                             (size_t)g->cfg.nx * sizeof(double), (size_t)g->cfg.ny,
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, g->stream));
    HIP_TRY(hipStreamSynchronize(g->stream));
    return RKG_OK;
}

int rkg_adjoint_timing(rkg_adjoint* a, double* forward_ms, double* backward_ms) {
    if (!valid(a)) return RKG_ERR_BAD_HANDLE;
    if (forward_ms) *forward_ms = a->fwd_ms;
    if (backward_ms) *backward_ms = a->bwd_ms;
    return RKG_OK;
}

}  // extern "C"
