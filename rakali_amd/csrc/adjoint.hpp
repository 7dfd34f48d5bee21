// adjoint.hpp -- reverse-mode derivatives of the structured step's face,
// slope, cell-update and boundary arithmetic (host and device).
//
// Each *_rev function takes the inputs of its forward counterpart in
// swe_core.hpp / swe_kernels.hpp and the adjoints of its outputs, recomputes
// the forward values it needs with IEEE division (the same quotients as
// the forward kernels, so every branch is taken as in the forward run) and
// ADDS the adjoints of its inputs to the accumulators it is given.
// Branches (wet/dry tests, minmod, min/max wave speeds, HLL vs upwind
// states, the HLLC contact side, clamps and the Froude backstop) are
// piecewise: the derivative of the branch taken in the forward run.  At a
// kink the derivative of the taken branch is used (x > 0 ? 1 : 0 for
// max(x, 0)); sqrt(u^2 + v^2) at rest contributes no derivative.
#pragma once

#include <hip/hip_runtime.h>

#include "crmath.hpp"
#include "swe_kernels.hpp"

namespace rk {
namespace adj {

struct C4 {   // adjoint of a Cell (h, hu, hv, b)
    double h, hu, hv, b;
};
struct S3 {   // adjoint of an FState / a flux triple
    double h, hu, hv;
};

__host__ __device__ __forceinline__ double max0(double x) { return x > 0.0 ? x : 0.0; }
__host__ __device__ __forceinline__ double amax(double a, double b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ double amin(double a, double b) { return a < b ? a : b; }

__host__ __device__ __forceinline__ double minmod_f(double a, double b) {
    if (a * b <= 0.0) return 0.0;
    return (fabs(a) < fabs(b)) ? a : b;
}
__host__ __device__ __forceinline__ void minmod_rev(double a, double b, double ls, double& la, double& lb) {
    if (a * b <= 0.0) return;
    if (fabs(a) < fabs(b)) la += ls; else lb += ls;
}

// ---------------------------------------------------------------- slopes
// Forward of cell_states (swe_core.hpp) on host and device.
__host__ __device__ __forceinline__ void states_fwd(const Cell& m, const Cell& c, const Cell& p, FState& lo,
                                                    FState& hi) {
    const double e_m = m.h + m.b, e_c = c.h + c.b, e_p = p.h + p.b;
    double s_eta = 0.0, s_hu = 0.0, s_hv = 0.0;
    if (!(c.h < THIN_LAYER_THRESHOLD || m.h < DRY_TOLERANCE || p.h < DRY_TOLERANCE)) {
        s_eta = minmod_f(e_c - e_m, e_p - e_c);
        s_hu = minmod_f(c.hu - m.hu, p.hu - c.hu);
        s_hv = minmod_f(c.hv - m.hv, p.hv - c.hv);
    }
    lo.h = max0((e_c - 0.5 * s_eta) - c.b);
    lo.hu = c.hu - 0.5 * s_hu;
    lo.hv = c.hv - 0.5 * s_hv;
    hi.h = max0((e_c + 0.5 * s_eta) - c.b);
    hi.hu = c.hu + 0.5 * s_hu;
    hi.hv = c.hv + 0.5 * s_hv;
}

__host__ __device__ __forceinline__ void states_rev(const Cell& m, const Cell& c, const Cell& p, const S3& llo,
                                                    const S3& lhi, C4& lm, C4& lc, C4& lp) {
    const double e_m = m.h + m.b, e_c = c.h + c.b, e_p = p.h + p.b;
    const bool flat = c.h < THIN_LAYER_THRESHOLD || m.h < DRY_TOLERANCE || p.h < DRY_TOLERANCE;
    double s_eta = 0.0;
    if (!flat) s_eta = minmod_f(e_c - e_m, e_p - e_c);
    double le_c = 0.0, ls_eta = 0.0, ls_hu = 0.0, ls_hv = 0.0;
    if ((e_c - 0.5 * s_eta) - c.b > 0.0) { le_c += llo.h; ls_eta -= 0.5 * llo.h; lc.b -= llo.h; }
    if ((e_c + 0.5 * s_eta) - c.b > 0.0) { le_c += lhi.h; ls_eta += 0.5 * lhi.h; lc.b -= lhi.h; }
    lc.hu += llo.hu + lhi.hu;
    ls_hu += 0.5 * (lhi.hu - llo.hu);
    lc.hv += llo.hv + lhi.hv;
    ls_hv += 0.5 * (lhi.hv - llo.hv);
    if (!flat) {
        double la = 0.0, lb = 0.0;
        minmod_rev(e_c - e_m, e_p - e_c, ls_eta, la, lb);
        le_c += la - lb;
        const double le_m = -la, le_p = lb;
        lm.h += le_m; lm.b += le_m;
        lp.h += le_p; lp.b += le_p;
        la = 0.0; lb = 0.0;
        minmod_rev(c.hu - m.hu, p.hu - c.hu, ls_hu, la, lb);
        lc.hu += la - lb; lm.hu -= la; lp.hu += lb;
        la = 0.0; lb = 0.0;
        minmod_rev(c.hv - m.hv, p.hv - c.hv, ls_hv, la, lb);
        lc.hv += la - lb; lm.hv -= la; lp.hv += lb;
    }
    lc.h += le_c;
    lc.b += le_c;
}

// --------------------------------------------------------------- Riemann
// Riemann flux in normal/tangential form: states (h, qn, qt) left and
// right; x faces have qn = hu, qt = hv, y faces qn = hv, qt = hu.  Returns
// (mass, normal momentum, tangential momentum) fluxes; this is
// riemann_x / riemann_y of swe_core.hpp with the momenta renamed.
template <bool HLLC>
__host__ __device__ __forceinline__ void riemann_fwd(double g, double hl, double nl, double tl, double hr, double nr,
                                                     double tr, double& fm, double& fn, double& ft) {
    double ul = 0.0, vl = 0.0, ur = 0.0, vr = 0.0;   // normal, tangential velocities
    if (hl > DRY_TOLERANCE) { ul = nl / hl; vl = tl / hl; }
    if (hr > DRY_TOLERANCE) { ur = nr / hr; vr = tr / hr; }
    const double al = sqrt(g * max0(hl)), ar = sqrt(g * max0(hr));
    double sl, sr;
    if (hr <= DRY_TOLERANCE) { sl = ul - al; sr = ul + 2.0 * al; }
    else if (hl <= DRY_TOLERANCE) { sl = ur - 2.0 * ar; sr = ur + ar; }
    else { sl = amin(ul - al, ur - ar); sr = amax(ul + al, ur + ar); }
    const double flm = nl, fln = nl * ul + 0.5 * g * hl * hl, flt = nl * vl;
    const double frm = nr, frn = nr * ur + 0.5 * g * hr * hr, frt = nr * vr;
    if (sl >= 0.0) { fm = flm; fn = fln; ft = flt; return; }
    if (sr <= 0.0) { fm = frm; fn = frn; ft = frt; return; }
    const double den = 1.0 / (sr - sl);
    fm = (sr * flm - sl * frm + sl * sr * (hr - hl)) * den;
    fn = (sr * fln - sl * frn + sl * sr * (nr - nl)) * den;
    if (!HLLC) {
        ft = (sr * flt - sl * frt + sl * sr * (tr - tl)) * den;
    } else {
        const double dstar = hr * (ur - sr) - hl * (ul - sl);
        double sstar;
        if (fabs(dstar) > DRY_TOLERANCE) sstar = (sl * hr * (ur - sr) - sr * hl * (ul - sl)) / dstar;
        else sstar = 0.5 * (ul + ur);
        ft = (sstar >= 0.0) ? fm * vl : fm * vr;
    }
}

template <bool HLLC>
__host__ __device__ __forceinline__ void riemann_rev(double g, double hl, double nl, double tl, double hr, double nr,
                                                     double tr, double lfm, double lfn, double lft, double& lhl,
                                                     double& lnl, double& ltl, double& lhr, double& lnr,
                                                     double& ltr) {
    double ul = 0.0, vl = 0.0, ur = 0.0, vr = 0.0;
    const bool wl = hl > DRY_TOLERANCE, wr = hr > DRY_TOLERANCE;
    if (wl) { ul = nl / hl; vl = tl / hl; }
    if (wr) { ur = nr / hr; vr = tr / hr; }
    const double al = sqrt(g * max0(hl)), ar = sqrt(g * max0(hr));
    int kase;
    double sl, sr;
    if (!wr) { kase = 0; sl = ul - al; sr = ul + 2.0 * al; }
    else if (!wl) { kase = 1; sl = ur - 2.0 * ar; sr = ur + ar; }
    else { kase = 2; sl = amin(ul - al, ur - ar); sr = amax(ul + al, ur + ar); }
    const double flm = nl, fln = nl * ul + 0.5 * g * hl * hl, flt = nl * vl;
    const double frm = nr, frn = nr * ur + 0.5 * g * hr * hr, frt = nr * vr;
    double lflm = 0.0, lfln = 0.0, lflt = 0.0, lfrm = 0.0, lfrn = 0.0, lfrt = 0.0;
    double lsl = 0.0, lsr = 0.0, lul = 0.0, lvl = 0.0, lur = 0.0, lvr = 0.0;
    if (sl >= 0.0) {
        lflm = lfm; lfln = lfn; lflt = lft;
    } else if (sr <= 0.0) {
        lfrm = lfm; lfrn = lfn; lfrt = lft;
    } else {
        const double den = 1.0 / (sr - sl);
        const double nm = sr * flm - sl * frm + sl * sr * (hr - hl);
        const double fm = nm * den;
        double lfm_tot = lfm, lden = 0.0;
        if (HLLC) {
            const double dstar = hr * (ur - sr) - hl * (ul - sl);
            double sstar;
            if (fabs(dstar) > DRY_TOLERANCE) sstar = (sl * hr * (ur - sr) - sr * hl * (ul - sl)) / dstar;
            else sstar = 0.5 * (ul + ur);
            if (sstar >= 0.0) { lfm_tot += lft * vl; lvl += lft * fm; }
            else { lfm_tot += lft * vr; lvr += lft * fm; }
        }
        // HLL combination of q = (h, qn[, qt]): f = (sr fl - sl fr + sl sr (qr - ql)) den
        auto hll = [&](double lf, double fl, double fr, double ql, double qr, double& lfl, double& lfr,
                       double& lql, double& lqr) {
            const double n = sr * fl - sl * fr + sl * sr * (qr - ql);
            const double ln = lf * den;
            lden += lf * n;
            lsr += ln * (fl + sl * (qr - ql));
            lsl += ln * (-fr + sr * (qr - ql));
            lfl += ln * sr;
            lfr -= ln * sl;
            lqr += ln * sl * sr;
            lql -= ln * sl * sr;
        };
        hll(lfm_tot, flm, frm, hl, hr, lflm, lfrm, lhl, lhr);
        hll(lfn, fln, frn, nl, nr, lfln, lfrn, lnl, lnr);
        if (!HLLC) hll(lft, flt, frt, tl, tr, lflt, lfrt, ltl, ltr);
        // den = 1 / (sr - sl)
        const double ld = -lden * den * den;
        lsr += ld;
        lsl -= ld;
    }
    // physical fluxes
    lnl += lflm + lfln * ul + lflt * vl;
    lul += lfln * nl;
    lvl += lflt * nl;
    lhl += lfln * g * hl;
    lnr += lfrm + lfrn * ur + lfrt * vr;
    lur += lfrn * nr;
    lvr += lfrt * nr;
    lhr += lfrn * g * hr;
    // wave speeds
    double lal = 0.0, lar = 0.0;
    if (kase == 0) { lul += lsl + lsr; lal += -lsl + 2.0 * lsr; }
    else if (kase == 1) { lur += lsl + lsr; lar += -2.0 * lsl + lsr; }
    else {
        if ((ul - al) < (ur - ar)) { lul += lsl; lal -= lsl; } else { lur += lsl; lar -= lsl; }
        if ((ul + al) > (ur + ar)) { lul += lsr; lal += lsr; } else { lur += lsr; lar += lsr; }
    }
    // a = sqrt(g h): d a / d h = g / (2 a) = a / (2 h); u = n / h.  These
    // adjoint terms share one reciprocal of h per side (no branch of the
    // forward run depends on them); a wet side (h > DRY_TOLERANCE) has a > 0.
    if (hl > 0.0 && al > 0.0) {
        const double r = 1.0 / hl;
        lhl += lal * (0.5 * al) * r;
        if (wl) { lnl += lul * r; ltl += lvl * r; lhl -= (lul * ul + lvl * vl) * r; }
    }
    if (hr > 0.0 && ar > 0.0) {
        const double r = 1.0 / hr;
        lhr += lar * (0.5 * ar) * r;
        if (wr) { lnr += lur * r; ltr += lvr * r; lhr -= (lur * ur + lvr * vr) * r; }
    }
}

// ------------------------------------------------------------------ faces
// face_flux (swe_core.hpp) with host-side arithmetic.
template <bool HLLC, bool YDIR>
__host__ __device__ __forceinline__ Face face_fwd(double g, const FState& L, double bL, const FState& R, double bR) {
    const double b_max = amax(bL, bR);
    const double hl = max0(L.h + bL - b_max), hr = max0(R.h + bR - b_max);
    double u_l = 0.0, v_l = 0.0, u_r = 0.0, v_r = 0.0;
    if (L.h > DRY_TOLERANCE) { u_l = L.hu / L.h; v_l = L.hv / L.h; }
    if (R.h > DRY_TOLERANCE) { u_r = R.hu / R.h; v_r = R.hv / R.h; }
    Face f;
    double fm, fn, ft;
    if (YDIR) {
        riemann_fwd<HLLC>(g, hl, hl * v_l, hl * u_l, hr, hr * v_r, hr * u_r, fm, fn, ft);
        f.fh = fm; f.fhv = fn; f.fhu = ft;
    } else {
        riemann_fwd<HLLC>(g, hl, hl * u_l, hl * v_l, hr, hr * u_r, hr * v_r, fm, fn, ft);
        f.fh = fm; f.fhu = fn; f.fhv = ft;
    }
    f.a = hr * hr - R.h * R.h;
    f.c = hl * hl;
    f.d = L.h * L.h;
    return f;
}

// lf = adjoint of the face's (fh, fhu, fhv, a, c, d).
template <bool HLLC, bool YDIR>
__host__ __device__ __forceinline__ void face_rev(double g, const FState& L, double bL, const FState& R, double bR,
                                                  const Face& lf, S3& lL, double& lbL, S3& lR, double& lbR) {
    const double b_max = amax(bL, bR);
    const double pl = L.h + bL - b_max, pr = R.h + bR - b_max;
    const double hl = max0(pl), hr = max0(pr);
    double u_l = 0.0, v_l = 0.0, u_r = 0.0, v_r = 0.0;
    const bool wl = L.h > DRY_TOLERANCE, wr = R.h > DRY_TOLERANCE;
    if (wl) { u_l = L.hu / L.h; v_l = L.hv / L.h; }
    if (wr) { u_r = R.hu / R.h; v_r = R.hv / R.h; }
    double lhl = 2.0 * hl * lf.c, lhr = 2.0 * hr * lf.a;
    lR.h -= 2.0 * R.h * lf.a;
    lL.h += 2.0 * L.h * lf.d;
    // Riemann inputs: (hl, hl*un_l, hl*ut_l, hr, ...)
    const double unl = YDIR ? v_l : u_l, utl = YDIR ? u_l : v_l;
    const double unr = YDIR ? v_r : u_r, utr = YDIR ? u_r : v_r;
    const double lfm = lf.fh, lfn = YDIR ? lf.fhv : lf.fhu, lft = YDIR ? lf.fhu : lf.fhv;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    riemann_rev<HLLC>(g, hl, hl * unl, hl * utl, hr, hr * unr, hr * utr, lfm, lfn, lft, a0, a1, a2, b0, b1, b2);
    lhl += a0 + a1 * unl + a2 * utl;
    lhr += b0 + b1 * unr + b2 * utr;
    const double lunl = a1 * hl, lutl = a2 * hl, lunr = b1 * hr, lutr = b2 * hr;
    const double lul = YDIR ? lutl : lunl, lvl = YDIR ? lunl : lutl;
    const double lur = YDIR ? lutr : lunr, lvr = YDIR ? lunr : lutr;
    if (wl) { const double r = 1.0 / L.h; lL.hu += lul * r; lL.hv += lvl * r; lL.h -= (lul * u_l + lvl * v_l) * r; }
    if (wr) { const double r = 1.0 / R.h; lR.hu += lur * r; lR.hv += lvr * r; lR.h -= (lur * u_r + lvr * v_r) * r; }
    double lbmax = 0.0;
    if (pl > 0.0) { lL.h += lhl; lbL += lhl; lbmax -= lhl; }
    if (pr > 0.0) { lR.h += lhr; lbR += lhr; lbmax -= lhr; }
    if (bL > bR) lbL += lbmax; else lbR += lbmax;
}

// face_rev on the device for a face whose states and hydrostatic depths
// are all wet and whose Riemann fan straddles the face: the branch the
// forward run took there, with the forward quotients as shared-reciprocal
// ones and the square roots as sqrt_pos (swe_core.hpp: the same bits as the
// IEEE operations on the ranges checked here), the HLLC contact side from
// the sign of its quotient, and the adjoint-only reciprocals 1/h from the
// same refined reciprocals (the same bits as 1.0 / h).  Every operation
// and accumulation is otherwise face_rev's, in its order, so the results
// are face_rev's bits.  Returns false, having touched nothing, when a lane
// that matters (`use`) breaks an assumption or holds a face value outside
// the ranges of unbounded(); the caller then runs face_rev.
template <bool HLLC, bool YDIR>
__device__ __forceinline__ bool face_rev_fast(double g, const FState& L, double bL, const FState& R, double bR,
                                              const Face& lf, S3& lL, double& lbL, S3& lR, double& lbR, bool use) {
    const double b_max = amax(bL, bR);
    const double pl = L.h + bL - b_max, pr = R.h + bR - b_max;
    const double hl = max0(pl), hr = max0(pr);
    bool bad = !(L.h > DRY_TOLERANCE && R.h > DRY_TOLERANCE && hl > DRY_TOLERANCE && hr > DRY_TOLERANCE &&
                 L.h <= 1.0e12 && R.h <= 1.0e12);
    bad = bad || unbounded(L.h) || unbounded(L.hu) || unbounded(L.hv) || unbounded(bL) || unbounded(R.h) ||
          unbounded(R.hu) || unbounded(R.hv) || unbounded(bR);
    if (__any(use && bad)) return false;
    const Recip rL = recip(L.h), rR = recip(R.h);
    const double u_l = rdiv<false>(L.hu, rL), v_l = rdiv<false>(L.hv, rL);
    const double u_r = rdiv<false>(R.hu, rR), v_r = rdiv<false>(R.hv, rR);
    const double unl = YDIR ? v_l : u_l, utl = YDIR ? u_l : v_l;
    const double unr = YDIR ? v_r : u_r, utr = YDIR ? u_r : v_r;
    // riemann_rev's forward part (states (hl, nl, tl), (hr, nr, tr))
    const double nl = hl * unl, tl = hl * utl, nr = hr * unr, tr = hr * utr;
    const Recip rhl = recip(hl), rhr = recip(hr);
    const double ul = rdiv<false>(nl, rhl), vl = rdiv<false>(tl, rhl);
    const double ur = rdiv<false>(nr, rhr), vr = rdiv<false>(tr, rhr);
    const double al = sqrt_pos(g * hl), ar = sqrt_pos(g * hr);
    const double sl = amin(ul - al, ur - ar), sr = amax(ul + al, ur + ar);
    const double span = sr - sl;
    bad = !(sl < 0.0 && sr > 0.0 && span >= 1.0e-8 && span <= 1.0e12);
    bool star_nonneg = true;
    if (HLLC) {
        const double dstar = hr * (ur - sr) - hl * (ul - sl);
        const double num = sl * hr * (ur - sr) - sr * hl * (ul - sl);
        bad = bad || !(fabs(dstar) > DRY_TOLERANCE && (num == 0.0 || fabs(num) >= 0x1p-900));
        star_nonneg = num == 0.0 || ((num > 0.0) == (dstar > 0.0));
    }
    if (__any(use && bad)) return false;

    // face_rev, before the Riemann solver
    double lhl = 2.0 * hl * lf.c, lhr = 2.0 * hr * lf.a;
    lR.h -= 2.0 * R.h * lf.a;
    lL.h += 2.0 * L.h * lf.d;
    const double lfm = lf.fh, lfn = YDIR ? lf.fhv : lf.fhu, lft = YDIR ? lf.fhu : lf.fhv;
    // riemann_rev, Davis speeds and the HLL branch
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, b0 = 0.0, b1 = 0.0, b2 = 0.0;
    {
        const double flm = nl, fln = nl * ul + 0.5 * g * hl * hl, flt = nl * vl;
        const double frm = nr, frn = nr * ur + 0.5 * g * hr * hr, frt = nr * vr;
        double lflm = 0.0, lfln = 0.0, lflt = 0.0, lfrm = 0.0, lfrn = 0.0, lfrt = 0.0;
        double lsl = 0.0, lsr = 0.0, lul = 0.0, lvl = 0.0, lur = 0.0, lvr = 0.0;
        const double den = rdiv<false>(1.0, recip(span));
        const double nm = sr * flm - sl * frm + sl * sr * (hr - hl);
        const double fm = nm * den;
        double lfm_tot = lfm, lden = 0.0;
        if (HLLC) {
            if (star_nonneg) { lfm_tot += lft * vl; lvl += lft * fm; }
            else { lfm_tot += lft * vr; lvr += lft * fm; }
        }
        auto hll = [&](double lfq, double fl, double fr, double ql, double qr, double& lfl, double& lfr, double& lql,
                       double& lqr) {
            const double n = sr * fl - sl * fr + sl * sr * (qr - ql);
            const double ln = lfq * den;
            lden += lfq * n;
            lsr += ln * (fl + sl * (qr - ql));
            lsl += ln * (-fr + sr * (qr - ql));
            lfl += ln * sr;
            lfr -= ln * sl;
            lqr += ln * sl * sr;
            lql -= ln * sl * sr;
        };
        hll(lfm_tot, flm, frm, hl, hr, lflm, lfrm, a0, b0);
        hll(lfn, fln, frn, nl, nr, lfln, lfrn, a1, b1);
        if (!HLLC) hll(lft, flt, frt, tl, tr, lflt, lfrt, a2, b2);
        const double ld = -lden * den * den;
        lsr += ld;
        lsl -= ld;
        a1 += lflm + lfln * ul + lflt * vl;
        lul += lfln * nl;
        lvl += lflt * nl;
        a0 += lfln * g * hl;
        b1 += lfrm + lfrn * ur + lfrt * vr;
        lur += lfrn * nr;
        lvr += lfrt * nr;
        b0 += lfrn * g * hr;
        double lal = 0.0, lar = 0.0;
        if ((ul - al) < (ur - ar)) { lul += lsl; lal -= lsl; } else { lur += lsl; lar -= lsl; }
        if ((ul + al) > (ur + ar)) { lul += lsr; lal += lsr; } else { lur += lsr; lar += lsr; }
        {
            const double r = rdiv<false>(1.0, rhl);
            a0 += lal * (0.5 * al) * r;
            a1 += lul * r; a2 += lvl * r; a0 -= (lul * ul + lvl * vl) * r;
        }
        {
            const double r = rdiv<false>(1.0, rhr);
            b0 += lar * (0.5 * ar) * r;
            b1 += lur * r; b2 += lvr * r; b0 -= (lur * ur + lvr * vr) * r;
        }
    }
    // face_rev, after the Riemann solver
    lhl += a0 + a1 * unl + a2 * utl;
    lhr += b0 + b1 * unr + b2 * utr;
    const double lunl = a1 * hl, lutl = a2 * hl, lunr = b1 * hr, lutr = b2 * hr;
    const double lul = YDIR ? lutl : lunl, lvl = YDIR ? lunl : lutl;
    const double lur = YDIR ? lutr : lunr, lvr = YDIR ? lunr : lutr;
    { const double r = rdiv<false>(1.0, rL); lL.hu += lul * r; lL.hv += lvl * r; lL.h -= (lul * u_l + lvl * v_l) * r; }
    { const double r = rdiv<false>(1.0, rR); lR.hu += lur * r; lR.hv += lvr * r; lR.h -= (lur * u_r + lvr * v_r) * r; }
    double lbmax = 0.0;
    if (pl > 0.0) { lL.h += lhl; lbL += lhl; lbmax -= lhl; }
    if (pr > 0.0) { lR.h += lhr; lbR += lhr; lbmax -= lhr; }
    if (bL > bR) lbL += lbmax; else lbR += lbmax;
    return true;
}

// Adjoint of a face from the flux-divergence adjoints of the cells on its
// low (ldl) and high (ldh) side: divergence (swe_core.hpp) transposed.
// `inv` = 1/dx (x faces) or 1/dy (y faces).
template <bool YDIR>
__host__ __device__ __forceinline__ Face face_seed(double g, double inv, const S3& ldl, const S3& ldh) {
    const double hg = 0.5 * g;
    Face f;
    f.fh = (ldh.h - ldl.h) * inv;
    f.fhu = (ldh.hu - ldl.hu) * inv;
    f.fhv = (ldh.hv - ldl.hv) * inv;
    const double nh = YDIR ? ldh.hv : ldh.hu, nl = YDIR ? ldl.hv : ldl.hu;
    f.a = -hg * nh * inv;
    f.c = hg * nl * inv;
    f.d = -hg * nl * inv;
    return f;
}

// ------------------------------------------------------------- cell body
__host__ __device__ __forceinline__ void clamp_fwd(double g, double& h, double& hu, double& hv) {
    const double eps4 = 5.062499999999998e-16;
    if (h < 0.0) { h = 0.0; hu = 0.0; hv = 0.0; return; }
    if (h < DRY_TOLERANCE) { hu = 0.0; hv = 0.0; return; }
    const double h4 = ((h * h) * h) * h;
    if (h4 < eps4) {
        const double r = sqrt(h4 + eps4);
        const double u = (SQRT2 * h * hu) / r, v = (SQRT2 * h * hv) / r;
        hu = h * u;
        hv = h * v;
    }
    const double ss = (hu * hu + hv * hv) / (h * h);
    const double fm = (FROUDE_BACKSTOP * FROUDE_BACKSTOP * g) * h;
    if (ss > fm) {
        const double sc = sqrt(fm / ss);
        hu = hu * sc;
        hv = hv * sc;
    }
}

// (h, hu, hv) in, lo = adjoint of the output; returns the adjoint of the input.
__host__ __device__ __forceinline__ S3 clamp_rev(double g, double h, double hu, double hv, const S3& lo) {
    const double eps4 = 5.062499999999998e-16;
    if (h < 0.0) return S3{0.0, 0.0, 0.0};
    if (h < DRY_TOLERANCE) return S3{lo.h, 0.0, 0.0};
    const double h4 = ((h * h) * h) * h;
    const bool des = h4 < eps4;
    double r = 0.0, u = 0.0, v = 0.0, hu1 = hu, hv1 = hv;
    if (des) {
        r = sqrt(h4 + eps4);
        u = (SQRT2 * h * hu) / r;
        v = (SQRT2 * h * hv) / r;
        hu1 = h * u;
        hv1 = h * v;
    }
    const double ss = (hu1 * hu1 + hv1 * hv1) / (h * h);
    const double K = FROUDE_BACKSTOP * FROUDE_BACKSTOP * g;
    const double fm = K * h;
    double lh = lo.h, lhu1 = lo.hu, lhv1 = lo.hv;
    if (ss > fm) {
        const double q = fm / ss, sc = sqrt(q);
        lhu1 = lo.hu * sc;
        lhv1 = lo.hv * sc;
        const double lsc = lo.hu * hu1 + lo.hv * hv1;
        const double lq = lsc / (2.0 * sc);
        const double lfm = lq / ss, lss = -lq * q / ss;
        lh += lfm * K;
        lhu1 += lss * 2.0 * hu1 / (h * h);
        lhv1 += lss * 2.0 * hv1 / (h * h);
        lh += lss * (-2.0 * ss / h);
    }
    S3 li{lh, 0.0, 0.0};
    if (des) {
        lh = 0.0;
        double lr = 0.0;
        li.h += lhu1 * u + lhv1 * v;
        const double lu = lhu1 * h, lv = lhv1 * h;
        const double lU = lu / r, lV = lv / r;
        lr -= (lu * u + lv * v) / r;
        li.h += (lU * hu + lV * hv) * SQRT2;
        li.hu += lU * SQRT2 * h;
        li.hv += lV * SQRT2 * h;
        const double lh4 = lr / (2.0 * r);
        li.h += lh4 * 4.0 * ((h * h) * h);
    } else {
        li.hu += lhu1;
        li.hv += lhv1;
    }
    return li;
}

// stage_body (swe_core.hpp): W = coriolis(friction(clamp(V + dt div))).
struct BodyRev {
    S3 lv;      // adjoint of the stage input V at the cell
    S3 ldiv;    // adjoint of the flux divergence at the cell
    double ln;  // adjoint of the cell's Manning n
    double ldt; // adjoint of the stage's dt (this cell's share)
};

__host__ __device__ __forceinline__ void body_fwd(const SourceTerms& s, double dt, const S3& div, double n_sq,
                                                  bool friction, double& h, double& hu, double& hv) {
    h = h + dt * div.h;
    hu = hu + dt * div.hu;
    hv = hv + dt * div.hv;
    clamp_fwd(s.g, h, hu, hv);
    if (h > DRY_TOLERANCE) {
        if (friction) {
            const double u = hu / h, v = hv / h;
            const double speed = sqrt(u * u + v * v);
            const double cf = s.g * n_sq / crm::pow13(h);
            const double D = 1.0 + (dt * cf * speed) / h;
            hu = hu / D;
            hv = hv / D;
        }
        if (s.has_coriolis) {
            const double a = hu, b = hv;
            hu = a * s.cosf + b * s.sinf;
            hv = -a * s.sinf + b * s.cosf;
        }
    }
}

// n = the cell's Manning n (n*n is n_sq; ignored without friction).
__host__ __device__ __forceinline__ BodyRev body_rev(const SourceTerms& s, double dt, const S3& div, double n,
                                                     double n_sq, bool friction, double h0, double hu0, double hv0,
                                                     const S3& lw) {
    const double ph = h0 + dt * div.h, phu = hu0 + dt * div.hu, phv = hv0 + dt * div.hv;
    double h = ph, hu = phu, hv = phv;
    clamp_fwd(s.g, h, hu, hv);
    S3 l = lw;
    double ln = 0.0, ldt = 0.0;
    if (h > DRY_TOLERANCE) {
        double fhu = hu, fhv = hv;   // after friction
        double D = 1.0, u = 0.0, v = 0.0, speed = 0.0, cf = 0.0, h13 = 1.0;
        if (friction) {
            u = hu / h;
            v = hv / h;
            speed = sqrt(u * u + v * v);
            h13 = crm::pow13(h);
            cf = s.g * n_sq / h13;
            D = 1.0 + (dt * cf * speed) / h;
            fhu = hu / D;
            fhv = hv / D;
        }
        if (s.has_coriolis) {
            // rotation by the angle f dt: d(out)/d dt = f (hv_out, -hu_out)
            const double ohu = fhu * s.cosf + fhv * s.sinf, ohv = -fhu * s.sinf + fhv * s.cosf;
            ldt += s.coriolis_f * (l.hu * ohv - l.hv * ohu);
            const double a = l.hu, b = l.hv;
            l.hu = a * s.cosf - b * s.sinf;
            l.hv = a * s.sinf + b * s.cosf;
        }
        if (friction) {
            const double lD = -(l.hu * fhu + l.hv * fhv) / D;
            ldt += lD * (cf * speed) / h;
            l.hu = l.hu / D;
            l.hv = l.hv / D;
            const double X = (dt * cf * speed) / h;
            const double lcf = lD * dt * speed / h, lsp = lD * dt * cf / h;
            l.h += -lD * X / h;
            const double lnsq = lcf * s.g / h13;
            l.h += lcf * cf * (-1.0 / (3.0 * h));
            ln += lnsq * 2.0 * n;
            double lu = 0.0, lv = 0.0;
            if (speed > 0.0) { lu = lsp * u / speed; lv = lsp * v / speed; }
            l.hu += lu / h;
            l.hv += lv / h;
            l.h -= (lu * u + lv * v) / h;
        }
    }
    const S3 lp = clamp_rev(s.g, ph, phu, phv, l);
    ldt += lp.h * div.h + lp.hu * div.hu + lp.hv * div.hv;
    BodyRev r;
    r.lv = lp;
    r.ldiv = S3{dt * lp.h, dt * lp.hu, dt * lp.hv};
    r.ln = ln;
    r.ldt = ldt;
    return r;
}

// d eta_tidal / d t of tidal_eta (rki_boundary.F90:92-101).
__host__ __device__ __forceinline__ double tidal_deta(const GridDesc& d, double time) {
    double e = 0.0;
    if (d.n_tc > 0) {
        for (int n = 0; n < d.n_tc; ++n)
            e = e - d.tc_amp[n] * d.tc_omega[n] * crm::sin(d.tc_omega[n] * time - d.tc_phase[n]);
    } else if (d.tidal_period > 0.0) {
        const double w = TWO_PI / d.tidal_period;
        e = d.tidal_amplitude * w * crm::cos(TWO_PI * time / d.tidal_period);
    }
    return e;
}

// CFL speed of one cell (compute_timestep's term, rki_kernel_timestep.F90
// :56-69) with IEEE division: the same value as the stage kernels' max.
__host__ __device__ __forceinline__ double cfl_speed_fwd(double g, double dx, double dy, double h, double hu,
                                                         double hv) {
    const double eps4 = 5.062499999999998e-16;
    if (!(h > DRY_TOLERANCE)) return 0.0;
    const double ws = sqrt(g * h);
    const double S = sqrt(((h * h) * h) * h + eps4);
    const double ua = (fabs(hu) * SQRT2 * h) / S;
    const double va = (fabs(hv) * SQRT2 * h) / S;
    const double sx = (ua + ws) / dx;
    const double sy = (va + ws) / dy;
    return sx > sy ? sx : sy;
}

// Adjoint of cfl_speed_fwd: ls is the adjoint of the speed; the branch of
// the max taken in the forward run.
__host__ __device__ __forceinline__ S3 cfl_speed_rev(double g, double dx, double dy, double h, double hu, double hv,
                                                     double ls) {
    const double eps4 = 5.062499999999998e-16;
    S3 l{0.0, 0.0, 0.0};
    if (!(h > DRY_TOLERANCE)) return l;
    const double ws = sqrt(g * h);
    const double h4 = ((h * h) * h) * h;
    const double S = sqrt(h4 + eps4);
    const double ua = (fabs(hu) * SQRT2 * h) / S;
    const double va = (fabs(hv) * SQRT2 * h) / S;
    const double sx = (ua + ws) / dx;
    const double sy = (va + ws) / dy;
    const bool xs = sx > sy;
    const double lsum = xs ? ls / dx : ls / dy;        // adjoint of (u_abs + ws)
    const double m = xs ? hu : hv;
    // u_abs = |m| sqrt2 h / S, S = sqrt(h^4 + eps^4): d/dh = |m| sqrt2 (S^2 - 2 h^4) / S^3
    const double dm = (m > 0.0 ? 1.0 : (m < 0.0 ? -1.0 : 0.0)) * SQRT2 * h / S;
    const double dh = fabs(m) * SQRT2 * ((h4 + eps4) - 2.0 * h4) / (S * S * S);
    l.h = lsum * (dh + g / (2.0 * ws));
    if (xs) l.hu = lsum * dm;
    else l.hv = lsum * dm;
    return l;
}

// ------------------------------------------------------------- boundaries
// Adjoint of one ghost cell (ghost_from, swe_kernels.hpp) onto its mirror
// cell (wall) or first interior cell (others).
__host__ __device__ __forceinline__ void ghost_rev(int tag, double g, double eta_tidal, double inflow_velocity,
                                                   double clamped_elevation, const Cell& mir, const Cell& in0,
                                                   double nrm_sign, bool normal_is_u, const C4& lg, C4& lmir,
                                                   C4& lin0, double* leta = nullptr) {
    const double lgn = normal_is_u ? lg.hu : lg.hv;
    const double lgt = normal_is_u ? lg.hv : lg.hu;
    double* lin0n = normal_is_u ? &lin0.hu : &lin0.hv;
    double* lin0t = normal_is_u ? &lin0.hv : &lin0.hu;
    switch (tag) {
    case BC_WALL:
    case BC_SPONGE: {
        if ((mir.h + mir.b) - mir.b > 0.0) lmir.h += lg.h;
        lmir.b += lg.b;
        if (normal_is_u) { lmir.hu -= lg.hu; lmir.hv += lg.hv; }
        else { lmir.hv -= lg.hv; lmir.hu += lg.hu; }
        return;
    }
    case BC_OPEN:
    case BC_TIDAL: {
        const bool prescribe = tag == BC_TIDAL;
        const double eta_ext = prescribe ? eta_tidal : 0.0;
        const double gb = in0.b, h_int = in0.h, eta_int = h_int + in0.b;
        const double hmax = amax(h_int, 1.0e-6);
        const double speed = sqrt(g * hmax);
        const double pre = eta_ext - gb;
        const double hg = prescribe ? max0(pre) : h_int;
        const double nrm = normal_is_u ? in0.hu : in0.hv;
        const double A = nrm / hmax, B = nrm_sign * speed * (eta_int - eta_ext) / hmax;
        *lin0t += lgt;
        const double lhg = lg.h + lgn * (A + B);
        const double lA = lgn * hg, lB = lgn * hg;
        *lin0n += lA / hmax;
        double lhmax = -lA * A / hmax - lB * B / hmax;
        const double lspeed = lB * nrm_sign * (eta_int - eta_ext) / hmax;
        const double leta_int = lB * nrm_sign * speed / hmax;
        double lh_int = leta_int;
        double lgb = lg.b;
        lin0.b += leta_int;
        if (speed > 0.0) lhmax += lspeed * g / (2.0 * speed);
        if (h_int > 1.0e-6) lh_int += lhmax;
        if (prescribe) { if (pre > 0.0) lgb -= lhg; }
        else lh_int += lhg;
        if (prescribe && leta) {   // d/d eta_ext of h_ghost and of the Flather term
            double le = -lB * nrm_sign * speed / hmax;
            if (pre > 0.0) le += lhg;
            *leta += le;
        }
        lin0.b += lgb;
        lin0.h += lh_int;
        return;
    }
    case BC_INFLOW:
    case BC_DISCHARGE: {
        const double in_sgn = -nrm_sign;
        lin0.b += lg.b;
        lin0.h += lg.h;
        if (tag == BC_INFLOW) lin0.h += lgn * in_sgn * inflow_velocity;
        *lin0t += lgt;
        return;
    }
    case BC_CLAMPED:
    default: {
        lin0.b += lg.b;
        if (clamped_elevation - in0.b > 0.0) lin0.b -= lg.h;
        lin0.hu += lg.hu;
        lin0.hv += lg.hv;
        return;
    }
    }
}

}  // namespace adj
}  // namespace rk
