// swe_core.hpp -- per-face and per-cell arithmetic of the structured 2-D
// shallow-water step, as gfx950 device functions.
//
// The step is Rakali's SSP-RK2 barotropic step (rki_solver.F90:761-815):
// Audusse hydrostatic reconstruction with minmod MUSCL slopes, HLL/HLLC
// Riemann fluxes with Ritter dry-front speeds, the well-balanced pressure
// correction, and the fused Euler update with clamp, KP07
// desingularization, Froude backstop, semi-implicit Manning and Coriolis.
//
// Arithmetic is written expression-for-expression in the reference's
// evaluation order and built with -ffp-contract=off, so each face and cell
// value is the IEEE result of the same operation sequence.  The reference
// evaluates every interior face twice (once from each cell,
// rki_kernel_flux.F90:371-372) and both evaluations are bit-identical,
// because a face depends only on the four cells across it; here each
// cell's slopes are evaluated once (cell_states) and each face once
// (face_flux), and the face is handed to both of its cells.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crmath.hpp"

namespace rk {

// rki_constants.F90:53,59,68-73,120
constexpr double DRY_TOLERANCE = 1.0e-6;
constexpr double THIN_LAYER_THRESHOLD = 1.0e-3;
constexpr double DESING_EPS = 1.5e-4;
constexpr double FROUDE_BACKSTOP = 50.0;
constexpr double RHO_WATER = 1035.0;          // rki_constants.F90:40
constexpr double TWO_PI = 6.283185307179586;  // 8*atan(1), rki_constants.F90:114
constexpr double SQRT2 = 1.4142135623730951;  // sqrt(2.0_wp) folded

// Boundary tags (rki_boundary_types.F90:18-32).
enum : int { BC_WALL = 1, BC_OPEN = 2, BC_TIDAL = 3, BC_NESTED = 4, BC_INFLOW = 5,
             BC_DISCHARGE = 6, BC_CLAMPED = 7, BC_SPONGE = 8, BC_CHAPMAN = 9 };

struct Cell {
    double h, hu, hv, b;
};

// One face's contribution to its two cells.  a = h_hyd_r^2 - h_R^2 is the
// right cell's "left" pressure term; c = h_hyd_l^2 and d = h_L^2 are the
// left cell's "right" terms (rki_kernel_flux.F90:361-369 evaluates
// ((hyd_left^2 - pre_left^2) - hyd_right^2) + pre_right^2).
struct Face {
    double fh, fhu, fhv, a, c, d;
};

__host__ __device__ __forceinline__ double fmax0(double x) { return x > 0.0 ? x : 0.0; }
__host__ __device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
// x**4 as LLVM flang lowers it for a run-time base: ((x*x)*x)*x.
__device__ __forceinline__ double pow4(double x) { return ((x * x) * x) * x; }

// IEEE division by a denominator shared between several numerators.
// The compiler's f64 division is v_div_scale (identity for operands away
// from the exponent extremes), v_rcp, two Newton refinements of the
// reciprocal, a product, a residual fma and the correction fma
// (v_div_fmas), then v_div_fixup (identity for finite, non-zero operands).
// Recip keeps the refined reciprocal so each further quotient by the same
// denominator costs a product and two fmas and returns the same bits as
// `n / d`.  The residual is formed as -fma(d, q0, -n) so that zero
// numerators keep their sign.  The unscaled sequence equals the IEEE
// quotient while |n| >= 2^-969 (or n = 0) and the denominator lies in
// [1e-8, 1e12]; callers guarantee the numerator bound by running a row
// with EXACT = true (plain division) whenever a cell of its stencil holds
// a non-zero h, hu, hv or b below 2^-300 in magnitude (see tiny()).
struct Recip {
    double d, y;
};

__device__ __forceinline__ Recip recip(double d) {
    const double y0 = __builtin_amdgcn_rcp(d);
    const double y1 = fma(y0, fma(-d, y0, 1.0), y0);
    const double y2 = fma(y1, fma(-d, y1, 1.0), y1);
    return Recip{d, y2};
}

template <bool EXACT>
__device__ __forceinline__ double rdiv(double n, const Recip& r) {
    if (EXACT) return n / r.d;
    const double q0 = n * r.y;
    const double t = fma(r.d, q0, -n);
    return fma(-t, r.y, q0);
}

// Non-zero and below 2^-300 in magnitude: a value for which some quotient
// of the step could leave the range where rdiv<false> is exact.
__device__ __forceinline__ bool tiny(double x) { return fabs(x) < 0x1p-300 && x != 0.0; }

// Non-zero and outside [2^-300, 2^40] in magnitude, or not finite: besides
// tiny values, the values for which the fast path's range arguments (below)
// do not hold; the kernels that carry the fast path test their rows with it.
__device__ __forceinline__ bool unbounded(double x) {
    const double a = fabs(x);
    return x != 0.0 && !(a >= 0x1p-300 && a <= 0x1p40);
}

// Square root of a finite x >= 2^-767: the correctly rounded sequence the
// compiler emits for sqrt (reciprocal square root, Goldschmidt refinement,
// two fma residual corrections) without its scaling of inputs below
// 2^-767 and its fixup of 0 and infinity, which are identities on that
// range.  Same bits as sqrt(x) there.
__device__ __forceinline__ double sqrt_pos(double x) {
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r;
    double h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
__device__ __forceinline__ bool tiny(const Cell& c) {
    return tiny(c.h) || tiny(c.hu) || tiny(c.hv) || tiny(c.b);
}
__device__ __forceinline__ bool unbounded(const Cell& c) {
    return unbounded(c.h) || unbounded(c.hu) || unbounded(c.hv) || unbounded(c.b);
}

// minmod, rki_kernel_flux.F90:380-393
__device__ __forceinline__ double minmod(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}

// HLL / HLLC normal flux in x: rki_kernel_flux.F90:410-495 and 584-684.
template <bool HLLC, bool EXACT>
__device__ __forceinline__ void riemann_x(double g, double h_l, double hu_l, double hv_l,
                                          double h_r, double hu_r, double hv_r,
                                          double& f_h, double& f_hu, double& f_hv) {
    double u_l = 0.0, v_l = 0.0, u_r = 0.0, v_r = 0.0;
    if (h_l > DRY_TOLERANCE) { const Recip r = recip(h_l); u_l = rdiv<EXACT>(hu_l, r); v_l = rdiv<EXACT>(hv_l, r); }
    if (h_r > DRY_TOLERANCE) { const Recip r = recip(h_r); u_r = rdiv<EXACT>(hu_r, r); v_r = rdiv<EXACT>(hv_r, r); }
    const double a_l = sqrt(g * fmax0(h_l));
    const double a_r = sqrt(g * fmax0(h_r));
    double s_l, s_r;
    if (h_r <= DRY_TOLERANCE) {          // dry right: Ritter front speed
        s_l = u_l - a_l;
        s_r = u_l + 2.0 * a_l;
    } else if (h_l <= DRY_TOLERANCE) {   // dry left
        s_l = u_r - 2.0 * a_r;
        s_r = u_r + a_r;
    } else {                             // Davis
        s_l = dmin(u_l - a_l, u_r - a_r);
        s_r = dmax(u_l + a_l, u_r + a_r);
    }
    const double fl_h = hu_l, fl_hu = hu_l * u_l + 0.5 * g * h_l * h_l, fl_hv = hu_l * v_l;
    const double fr_h = hu_r, fr_hu = hu_r * u_r + 0.5 * g * h_r * h_r, fr_hv = hu_r * v_r;
    if (s_l >= 0.0) {
        f_h = fl_h; f_hu = fl_hu; f_hv = fl_hv;
    } else if (s_r <= 0.0) {
        f_h = fr_h; f_hu = fr_hu; f_hv = fr_hv;
    } else {
        const double denom = 1.0 / (s_r - s_l);
        f_h = (s_r * fl_h - s_l * fr_h + s_l * s_r * (h_r - h_l)) * denom;
        f_hu = (s_r * fl_hu - s_l * fr_hu + s_l * s_r * (hu_r - hu_l)) * denom;
        if (!HLLC) {
            f_hv = (s_r * fl_hv - s_l * fr_hv + s_l * s_r * (hv_r - hv_l)) * denom;
        } else {
            // contact speed, rki_kernel_flux.F90:670-681
            const double denom_star = h_r * (u_r - s_r) - h_l * (u_l - s_l);
            double s_star;
            if (fabs(denom_star) > DRY_TOLERANCE)
                s_star = (s_l * h_r * (u_r - s_r) - s_r * h_l * (u_l - s_l)) / denom_star;
            else
                s_star = 0.5 * (u_l + u_r);
            f_hv = (s_star >= 0.0) ? f_h * v_l : f_h * v_r;
        }
    }
}

// HLL / HLLC normal flux in y: rki_kernel_flux.F90:497-582 and 686-783.
// Arguments are (h, hu, hv) of the bottom then the top state, as there.
template <bool HLLC, bool EXACT>
__device__ __forceinline__ void riemann_y(double g, double h_b, double hu_b, double hv_b,
                                          double h_t, double hu_t, double hv_t,
                                          double& g_h, double& g_hu, double& g_hv) {
    double u_b = 0.0, v_b = 0.0, u_t = 0.0, v_t = 0.0;
    if (h_b > DRY_TOLERANCE) { const Recip r = recip(h_b); u_b = rdiv<EXACT>(hu_b, r); v_b = rdiv<EXACT>(hv_b, r); }
    if (h_t > DRY_TOLERANCE) { const Recip r = recip(h_t); u_t = rdiv<EXACT>(hu_t, r); v_t = rdiv<EXACT>(hv_t, r); }
    const double a_b = sqrt(g * fmax0(h_b));
    const double a_t = sqrt(g * fmax0(h_t));
    double s_l, s_r;
    if (h_t <= DRY_TOLERANCE) {
        s_l = v_b - a_b;
        s_r = v_b + 2.0 * a_b;
    } else if (h_b <= DRY_TOLERANCE) {
        s_l = v_t - 2.0 * a_t;
        s_r = v_t + a_t;
    } else {
        s_l = dmin(v_b - a_b, v_t - a_t);
        s_r = dmax(v_b + a_b, v_t + a_t);
    }
    const double gl_h = hv_b, gl_hu = hu_b * v_b, gl_hv = hv_b * v_b + 0.5 * g * h_b * h_b;
    const double gr_h = hv_t, gr_hu = hu_t * v_t, gr_hv = hv_t * v_t + 0.5 * g * h_t * h_t;
    if (s_l >= 0.0) {
        g_h = gl_h; g_hu = gl_hu; g_hv = gl_hv;
    } else if (s_r <= 0.0) {
        g_h = gr_h; g_hu = gr_hu; g_hv = gr_hv;
    } else {
        const double denom = 1.0 / (s_r - s_l);
        g_h = (s_r * gl_h - s_l * gr_h + s_l * s_r * (h_t - h_b)) * denom;
        g_hv = (s_r * gl_hv - s_l * gr_hv + s_l * s_r * (hv_t - hv_b)) * denom;
        if (!HLLC) {
            g_hu = (s_r * gl_hu - s_l * gr_hu + s_l * s_r * (hu_t - hu_b)) * denom;
        } else {
            const double denom_star = h_t * (v_t - s_r) - h_b * (v_b - s_l);
            double s_star;
            if (fabs(denom_star) > DRY_TOLERANCE)
                s_star = (s_l * h_t * (v_t - s_r) - s_r * h_b * (v_b - s_l)) / denom_star;
            else
                s_star = 0.5 * (v_b + v_t);
            g_hu = (s_star >= 0.0) ? g_h * u_b : g_h * u_t;
        }
    }
}

// Reconstructed state of a cell at one of its faces along a direction:
// h = max(0, eta_face - b) and the reconstructed momenta (the bed is the
// cell's own).
struct FState {
    double h, hu, hv;
};

// Minmod-limited MUSCL states of cell c at its low (lo) and high (hi) face
// along one direction, from its neighbours m (below) and p (above):
// rki_kernel_flux.F90:158-181 for the slope (zeroed near wet/dry fronts),
// :186-189 / :229-232 for the face states.  Each face of the reference
// re-evaluates this slope for both of its cells; a cell's slope is the
// same from both of its faces, so it is evaluated once here.
__device__ __forceinline__ void cell_states(const Cell& m, const Cell& c, const Cell& p, FState& lo,
                                            FState& hi) {
    const double e_m = m.h + m.b;
    const double e_c = c.h + c.b;
    const double e_p = p.h + p.b;
    double s_eta, s_hu, s_hv;
    if (c.h < THIN_LAYER_THRESHOLD || m.h < DRY_TOLERANCE || p.h < DRY_TOLERANCE) {
        s_eta = 0.0; s_hu = 0.0; s_hv = 0.0;
    } else {
        s_eta = minmod(e_c - e_m, e_p - e_c);
        s_hu = minmod(c.hu - m.hu, p.hu - c.hu);
        s_hv = minmod(c.hv - m.hv, p.hv - c.hv);
    }
    lo.h = fmax0((e_c - 0.5 * s_eta) - c.b);
    lo.hu = c.hu - 0.5 * s_hu;
    lo.hv = c.hv - 0.5 * s_hv;
    hi.h = fmax0((e_c + 0.5 * s_eta) - c.b);
    hi.hu = c.hu + 0.5 * s_hu;
    hi.hv = c.hv + 0.5 * s_hv;
}

// One face between the high state L of its left cell (bed bL) and the low
// state R of its right cell (bed bR): Audusse hydrostatic reconstruction
// (rki_kernel_flux.F90:208-210), velocities (extract_velocity, :395-408)
// and the HLL/HLLC flux of the h_hyd-weighted momenta (:216-224).
template <bool HLLC, bool YDIR, bool EXACT>
__device__ __forceinline__ Face face_flux(double g, const FState& L, double bL, const FState& R,
                                          double bR) {
    const double b_max = dmax(bL, bR);
    const double h_hyd_l = fmax0(L.h + bL - b_max);
    const double h_hyd_r = fmax0(R.h + bR - b_max);
    double u_l = 0.0, v_l = 0.0, u_r = 0.0, v_r = 0.0;
    if (L.h > DRY_TOLERANCE) { const Recip r = recip(L.h); u_l = rdiv<EXACT>(L.hu, r); v_l = rdiv<EXACT>(L.hv, r); }
    if (R.h > DRY_TOLERANCE) { const Recip r = recip(R.h); u_r = rdiv<EXACT>(R.hu, r); v_r = rdiv<EXACT>(R.hv, r); }
    Face f;
    if (YDIR)
        riemann_y<HLLC, EXACT>(g, h_hyd_l, h_hyd_l * u_l, h_hyd_l * v_l, h_hyd_r, h_hyd_r * u_r,
                        h_hyd_r * v_r, f.fh, f.fhu, f.fhv);
    else
        riemann_x<HLLC, EXACT>(g, h_hyd_l, h_hyd_l * u_l, h_hyd_l * v_l, h_hyd_r, h_hyd_r * u_r,
                        h_hyd_r * v_r, f.fh, f.fhu, f.fhv);
    f.a = h_hyd_r * h_hyd_r - R.h * R.h;
    f.c = h_hyd_l * h_hyd_l;
    f.d = L.h * L.h;
    return f;
}

// ---------------------------------------------------------------------
// The fast path.  Away from wet/dry fronts and shocks a wavefront's faces
// and cells satisfy the assumptions below; each *_fast function performs
// exactly the IEEE operations of its general counterpart for the branch
// those assumptions select, and raises `bad` on a lane where one fails.
// The callers test the assumptions they can before the work (wave-uniform
// branches) and redo a piece on the general path when any lane that
// matters raised `bad`, so results never depend on the path taken.  The
// assumptions and what they remove:
//  * every face state and hydrostatic depth exceeds DRY_TOLERANCE: the
//    velocity quotients need no dry branch, the wave-speed arguments g h
//    lie above 2^-767 and are finite (the row's inputs are bounded by
//    unbounded()), so sqrt_pos gives sqrt's bits, and the speeds are Davis's;
//  * s_l < 0 < s_r (the HLL star region) with s_r - s_l in [1e-8, 1e12]:
//    1/(s_r - s_l) is the shared-reciprocal quotient of 1.0, the same bits
//    as the division;
//  * HLLC: |denom_star| > DRY_TOLERANCE and num = 0 or |num| >= 2^-900.
//    Only the sign of s_star = num / denom_star is used; |denom_star| stays
//    below 2^83 (depths below 2^41 for inputs bounded by unbounded(), and
//    |u - s| <= 2 (s_r - s_l) <= 2e12), so the quotient does not underflow
//    to zero and
//    s_star >= 0 exactly when num is zero or num and denom_star have the
//    same sign: no division;
//  * updated cells: h >= DRY_TOLERANCE with h^4 >= DESING_EPS^4 (no clamp,
//    no KP07) and speed^2 below the Froude backstop by a margin checked
//    without the division; friction's x^(1/3) below 1e12 and its
//    numerator g n^2 at least 2^-969, so its quotient is a shared-reciprocal
//    one.

// riemann_x restricted to the HLL star region, rki_kernel_flux.F90:410-495
// and 584-684 (Davis speeds, :463-466).
template <bool HLLC>
__device__ __forceinline__ void riemann_x_fast(double g, double h_l, double hu_l, double hv_l, double h_r,
                                               double hu_r, double hv_r, double& f_h, double& f_hu,
                                               double& f_hv, bool& bad) {
    const Recip rl = recip(h_l), rr = recip(h_r);
    const double u_l = rdiv<false>(hu_l, rl), v_l = rdiv<false>(hv_l, rl);
    const double u_r = rdiv<false>(hu_r, rr), v_r = rdiv<false>(hv_r, rr);
    const double a_l = sqrt_pos(g * h_l);
    const double a_r = sqrt_pos(g * h_r);
    const double s_l = dmin(u_l - a_l, u_r - a_r);
    const double s_r = dmax(u_l + a_l, u_r + a_r);
    const double fl_h = hu_l, fl_hu = hu_l * u_l + 0.5 * g * h_l * h_l, fl_hv = hu_l * v_l;
    const double fr_h = hu_r, fr_hu = hu_r * u_r + 0.5 * g * h_r * h_r, fr_hv = hu_r * v_r;
    const double span = s_r - s_l;
    bad = bad || !(s_l < 0.0 && s_r > 0.0 && span >= 1.0e-8 && span <= 1.0e12);
    const double denom = rdiv<false>(1.0, recip(span));
    f_h = (s_r * fl_h - s_l * fr_h + s_l * s_r * (h_r - h_l)) * denom;
    f_hu = (s_r * fl_hu - s_l * fr_hu + s_l * s_r * (hu_r - hu_l)) * denom;
    if (!HLLC) {
        f_hv = (s_r * fl_hv - s_l * fr_hv + s_l * s_r * (hv_r - hv_l)) * denom;
    } else {
        const double denom_star = h_r * (u_r - s_r) - h_l * (u_l - s_l);
        const double num = s_l * h_r * (u_r - s_r) - s_r * h_l * (u_l - s_l);
        bad = bad || !(fabs(denom_star) > DRY_TOLERANCE && (num == 0.0 || fabs(num) >= 0x1p-900));
        const bool star_nonneg = num == 0.0 || ((num > 0.0) == (denom_star > 0.0));
        f_hv = star_nonneg ? f_h * v_l : f_h * v_r;
    }
}

// riemann_y restricted to the HLL star region, rki_kernel_flux.F90:497-582
// and 686-783.
template <bool HLLC>
__device__ __forceinline__ void riemann_y_fast(double g, double h_b, double hu_b, double hv_b, double h_t,
                                               double hu_t, double hv_t, double& g_h, double& g_hu,
                                               double& g_hv, bool& bad) {
    const Recip rb = recip(h_b), rt = recip(h_t);
    const double u_b = rdiv<false>(hu_b, rb), v_b = rdiv<false>(hv_b, rb);
    const double u_t = rdiv<false>(hu_t, rt), v_t = rdiv<false>(hv_t, rt);
    const double a_b = sqrt_pos(g * h_b);
    const double a_t = sqrt_pos(g * h_t);
    const double s_l = dmin(v_b - a_b, v_t - a_t);
    const double s_r = dmax(v_b + a_b, v_t + a_t);
    const double gl_h = hv_b, gl_hu = hu_b * v_b, gl_hv = hv_b * v_b + 0.5 * g * h_b * h_b;
    const double gr_h = hv_t, gr_hu = hu_t * v_t, gr_hv = hv_t * v_t + 0.5 * g * h_t * h_t;
    const double span = s_r - s_l;
    bad = bad || !(s_l < 0.0 && s_r > 0.0 && span >= 1.0e-8 && span <= 1.0e12);
    const double denom = rdiv<false>(1.0, recip(span));
    g_h = (s_r * gl_h - s_l * gr_h + s_l * s_r * (h_t - h_b)) * denom;
    g_hv = (s_r * gl_hv - s_l * gr_hv + s_l * s_r * (hv_t - hv_b)) * denom;
    if (!HLLC) {
        g_hu = (s_r * gl_hu - s_l * gr_hu + s_l * s_r * (hu_t - hu_b)) * denom;
    } else {
        const double denom_star = h_t * (v_t - s_r) - h_b * (v_b - s_l);
        const double num = s_l * h_t * (v_t - s_r) - s_r * h_b * (v_b - s_l);
        bad = bad || !(fabs(denom_star) > DRY_TOLERANCE && (num == 0.0 || fabs(num) >= 0x1p-900));
        const bool star_nonneg = num == 0.0 || ((num > 0.0) == (denom_star > 0.0));
        g_hu = star_nonneg ? g_h * u_b : g_h * u_t;
    }
}

// face_flux<HLLC, YDIR, false> on the fast path where the wavefront allows
// it: when every lane that matters (`use`) has both face states and both
// hydrostatic depths wet, the velocities need no dry branch and the Riemann
// solver is riemann_*_fast; lanes outside its star region then take the
// general solver on the same arguments.  Otherwise face_flux.
template <bool HLLC, bool YDIR>
__device__ __forceinline__ Face face_flux_auto(double g, const FState& L, double bL, const FState& R, double bR,
                                               bool use) {
    const double b_max = dmax(bL, bR);
    const double h_hyd_l = fmax0(L.h + bL - b_max);
    const double h_hyd_r = fmax0(R.h + bR - b_max);
    const bool wet = L.h > DRY_TOLERANCE && R.h > DRY_TOLERANCE && h_hyd_l > DRY_TOLERANCE &&
                     h_hyd_r > DRY_TOLERANCE;
    if (__any(use && !wet)) return face_flux<HLLC, YDIR, false>(g, L, bL, R, bR);
    const Recip rl = recip(L.h), rr = recip(R.h);
    const double u_l = rdiv<false>(L.hu, rl), v_l = rdiv<false>(L.hv, rl);
    const double u_r = rdiv<false>(R.hu, rr), v_r = rdiv<false>(R.hv, rr);
    const double qu_l = h_hyd_l * u_l, qv_l = h_hyd_l * v_l, qu_r = h_hyd_r * u_r, qv_r = h_hyd_r * v_r;
    Face f;
    bool bad = false;
    if (YDIR) riemann_y_fast<HLLC>(g, h_hyd_l, qu_l, qv_l, h_hyd_r, qu_r, qv_r, f.fh, f.fhu, f.fhv, bad);
    else riemann_x_fast<HLLC>(g, h_hyd_l, qu_l, qv_l, h_hyd_r, qu_r, qv_r, f.fh, f.fhu, f.fhv, bad);
    if (__any(use && bad)) {
        if (YDIR) riemann_y<HLLC, false>(g, h_hyd_l, qu_l, qv_l, h_hyd_r, qu_r, qv_r, f.fh, f.fhu, f.fhv);
        else riemann_x<HLLC, false>(g, h_hyd_l, qu_l, qv_l, h_hyd_r, qu_r, qv_r, f.fh, f.fhu, f.fhv);
    }
    f.a = h_hyd_r * h_hyd_r - R.h * R.h;
    f.c = h_hyd_l * h_hyd_l;
    f.d = L.h * L.h;
    return f;
}

// Net flux divergence of one cell from its four faces,
// rki_kernel_flux.F90:358-369 (rdx, rdy: reciprocals of dx, dy).
template <bool EXACT>
__device__ __forceinline__ void divergence(double g, const Recip& rdx, const Recip& rdy,
                                           const Face& xl, const Face& xr, const Face& yb,
                                           const Face& yt, double& fh, double& fhu, double& fhv) {
    const double half_g = 0.5 * g;
    fh = rdiv<EXACT>(-(xr.fh - xl.fh), rdx) - rdiv<EXACT>(yt.fh - yb.fh, rdy);
    fhu = rdiv<EXACT>(-(xr.fhu - xl.fhu), rdx) - rdiv<EXACT>(yt.fhu - yb.fhu, rdy) -
          rdiv<EXACT>(half_g * (xl.a - xr.c + xr.d), rdx);
    fhv = rdiv<EXACT>(-(xr.fhv - xl.fhv), rdx) - rdiv<EXACT>(yt.fhv - yb.fhv, rdy) -
          rdiv<EXACT>(half_g * (yb.a - yt.c + yt.d), rdy);
}

// Clamp + KP07 desingularization + Froude backstop,
// rki_kernel_fused.F90:234-260 and :440-463.
template <bool EXACT>
__device__ __forceinline__ void clamp_desing(double g, double& h, double& hu, double& hv) {
    const double eps4 = 5.062499999999998e-16;  // DESING_EPS**4 folded (0x3CC23D5504170727)
    if (h < 0.0) {
        h = 0.0; hu = 0.0; hv = 0.0;
    } else if (h < DRY_TOLERANCE) {
        hu = 0.0; hv = 0.0;
    } else {
        const double h4 = pow4(h);
        if (h4 < eps4) {
            const Recip r = recip(sqrt(h4 + eps4));
            const double u = rdiv<EXACT>(SQRT2 * h * hu, r);
            const double v = rdiv<EXACT>(SQRT2 * h * hv, r);
            hu = h * u;
            hv = h * v;
        }
        const double speed_sq = (hu * hu + hv * hv) / (h * h);
        const double fr_max_sq = (FROUDE_BACKSTOP * FROUDE_BACKSTOP * g) * h;
        if (speed_sq > fr_max_sq) {
            const double fr_scale = sqrt(fr_max_sq / speed_sq);
            hu = hu * fr_scale;
            hv = hv * fr_scale;
        }
    }
}

struct SourceTerms {
    double g;
    double n_sq;        // manning_n**2
    int has_friction;   // manning_n > 0
    int has_coriolis;   // |f| > 1e-20
    double cosf, sinf;  // cos(f dt), sin(f dt)
    double coriolis_f;  // f (used by the reverse sweep's d/d dt)
};

// Surface forcing at one cell (rki_surface_forcing.F90): wind stress (Pa),
// the central pressure gradient (Pa/m) and the precipitation rate (m/s).
struct Frc {
    double wx, wy, dpdx, dpdy, rain;
};

// Euler update + clamp + Manning + forcing + Coriolis + precipitation,
// rki_kernel_fused.F90:228-290.  With NF the Manning coefficient is the
// cell's own n_cell (friction where n_cell > 0) instead of the scalar one;
// FRC applies the surface forcing f (forcing_active).
template <bool EXACT, bool NF, bool FRC = false>
__device__ __forceinline__ void stage_body(const SourceTerms& s, double dt, double fh, double fhu,
                                           double fhv, double& h, double& hu, double& hv, double n_cell,
                                           const Frc& f = Frc{}) {
    h = h + dt * fh;
    hu = hu + dt * fhu;
    hv = hv + dt * fhv;
    clamp_desing<EXACT>(s.g, h, hu, hv);
    const bool friction = NF ? (n_cell > 0.0) : (s.has_friction != 0);
    const double n_sq = NF ? n_cell * n_cell : s.n_sq;
    if (h > DRY_TOLERANCE) {
        if (friction) {
            const Recip rh = recip(h);
            const double u = rdiv<EXACT>(hu, rh);
            const double v = rdiv<EXACT>(hv, rh);
            const double speed = sqrt(u * u + v * v);
            const double cf = s.g * n_sq / crm::pow13(h);
            const Recip rd = recip(1.0 + rdiv<EXACT>(dt * cf * speed, rh));
            hu = rdiv<EXACT>(hu, rd);
            hv = rdiv<EXACT>(hv, rd);
        }
        if (FRC) {   // rki_kernel_fused.F90:273-280
            hu = hu + dt * f.wx / RHO_WATER;
            hv = hv + dt * f.wy / RHO_WATER;
            hu = hu - dt * h * f.dpdx / RHO_WATER;
            hv = hv - dt * h * f.dpdy / RHO_WATER;
        }
        if (s.has_coriolis) {
            const double hu_old = hu, hv_old = hv;
            hu = hu_old * s.cosf + hv_old * s.sinf;
            hv = -hu_old * s.sinf + hv_old * s.cosf;
        }
    }
    if (FRC) h = h + dt * f.rain;   // every cell, :286-289
}

// clamp_desing of the fast row path: the cell is wet (h >= DRY_TOLERANCE),
// outside the KP07 band (h^4 >= DESING_EPS^4) and below the Froude
// backstop, so the values pass unchanged.  The backstop test
// (hu^2 + hv^2) / h^2 > fr_max_sq is decided without the quotient:
// num < h^2 (fr_max_sq (1 - 2^-50)), rounded, implies num / h^2 < fr_max_sq
// exactly, hence the rounded quotient is at most fr_max_sq (a double).
__device__ __forceinline__ void clamp_fast(double g, double h, double hu, double hv, bool& bad) {
    const double eps4 = 5.062499999999998e-16;
    const double h4 = pow4(h);
    const double num = hu * hu + hv * hv;
    const double fr_max_sq = (FROUDE_BACKSTOP * FROUDE_BACKSTOP * g) * h;
    const double lim = (h * h) * (fr_max_sq * (1.0 - 0x1p-50));
    bad = bad || !(h >= DRY_TOLERANCE && h4 >= eps4 && num < lim);
}

// stage_body of the fast row path (clamp_fast above; friction's quotient
// by x^(1/3) as a shared-reciprocal one while x^(1/3) <= 1e12 and
// g n^2 >= 2^-969).
template <bool NF, bool FRC = false>
__device__ __forceinline__ void stage_body_fast(const SourceTerms& s, double dt, double fh, double fhu, double fhv,
                                                double& h, double& hu, double& hv, double n_cell, const Frc& f,
                                                bool& bad) {
    h = h + dt * fh;
    hu = hu + dt * fhu;
    hv = hv + dt * fhv;
    clamp_fast(s.g, h, hu, hv, bad);
    const bool friction = NF ? (n_cell > 0.0) : (s.has_friction != 0);
    const double n_sq = NF ? n_cell * n_cell : s.n_sq;
    if (friction) {
        const Recip rh = recip(h);
        const double u = rdiv<false>(hu, rh);
        const double v = rdiv<false>(hv, rh);
        const double speed = sqrt(u * u + v * v);
        const double p13 = crm::pow13(h);
        const double gn2 = s.g * n_sq;
        bad = bad || !(p13 <= 1.0e12 && gn2 >= 0x1p-969);
        const double cf = rdiv<false>(gn2, recip(p13));
        const Recip rd = recip(1.0 + rdiv<false>(dt * cf * speed, rh));
        hu = rdiv<false>(hu, rd);
        hv = rdiv<false>(hv, rd);
    }
    if (FRC) {
        hu = hu + dt * f.wx / RHO_WATER;
        hv = hv + dt * f.wy / RHO_WATER;
        hu = hu - dt * h * f.dpdx / RHO_WATER;
        hv = hv - dt * h * f.dpdy / RHO_WATER;
    }
    if (s.has_coriolis) {
        const double hu_old = hu, hv_old = hv;
        hu = hu_old * s.cosf + hv_old * s.sinf;
        hv = -hu_old * s.sinf + hv_old * s.cosf;
    }
    if (FRC) h = h + dt * f.rain;
}

// cfl_speed for a cell that passed clamp_fast (h > DRY_TOLERANCE; g h and
// h^4 + eps^4 are positive and finite, so sqrt_pos gives sqrt's bits).
__device__ __forceinline__ double cfl_speed_fast(double g, const Recip& rdx, const Recip& rdy, double h, double hu,
                                                 double hv) {
    const double eps4 = 5.062499999999998e-16;
    const double wave_speed = sqrt_pos(g * h);
    const double h4 = pow4(h);
    const Recip rd = recip(sqrt_pos(h4 + eps4));
    const double u_abs = rdiv<false>(fabs(hu) * SQRT2 * h, rd);
    const double v_abs = rdiv<false>(fabs(hv) * SQRT2 * h, rd);
    const double sx = rdiv<false>(u_abs + wave_speed, rdx);
    const double sy = rdiv<false>(v_abs + wave_speed, rdy);
    return dmax(sx, sy);
}

// Per-cell CFL speed term, rki_kernel_timestep.F90:56-69; 0 for dry cells.
template <bool EXACT>
__device__ __forceinline__ double cfl_speed(double g, const Recip& rdx, const Recip& rdy, double h,
                                            double hu, double hv) {
    const double eps4 = 5.062499999999998e-16;
    if (!(h > DRY_TOLERANCE)) return 0.0;
    const double wave_speed = sqrt(g * h);
    const double h4 = pow4(h);
    const Recip rd = recip(sqrt(h4 + eps4));
    const double u_abs = rdiv<EXACT>(fabs(hu) * SQRT2 * h, rd);
    const double v_abs = rdiv<EXACT>(fabs(hv) * SQRT2 * h, rd);
    const double sx = rdiv<EXACT>(u_abs + wave_speed, rdx);
    const double sy = rdiv<EXACT>(v_abs + wave_speed, rdy);
    return dmax(sx, sy);
}

}  // namespace rk
