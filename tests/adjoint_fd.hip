// Finite-difference check of the reverse-mode derivatives in
// rakali_amd/csrc/adjoint.hpp, on the host.  For each routine: random
// inputs in a regime chosen for it, a random output adjoint lambda, and the
// comparison of the reverse-mode gradient with central differences of
// lambda . F.  Prints one line per routine: name, cases, worst relative
// error; exits non-zero if any exceeds its tolerance.
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <vector>

#include "adjoint.hpp"
#include "swe_kernels.hpp"

using namespace rk;
using namespace rk::adj;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

using Fwd = std::function<std::vector<double>(const std::vector<double>&)>;
using Rev = std::function<std::vector<double>(const std::vector<double>&, const std::vector<double>&)>;

// max over inputs of |rev - fd| / (|rev| + |fd| + scale)
static double check(const Fwd& f, const Rev& r, const std::vector<double>& x, const std::vector<double>& lam) {
    const std::vector<double> g = r(x, lam);
    double worst = 0.0, scale = 0.0;
    std::vector<double> fd(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        const double d = 1e-6 * std::max(std::fabs(x[i]), 1e-3);
        std::vector<double> xp = x, xm = x;
        xp[i] += d;
        xm[i] -= d;
        const std::vector<double> fp = f(xp), fm = f(xm);
        double s = 0.0;
        for (size_t k = 0; k < lam.size(); ++k) s += lam[k] * (fp[k] - fm[k]);
        fd[i] = s / (2.0 * d);
        scale = std::max(scale, std::max(std::fabs(fd[i]), std::fabs(g[i])));
    }
    for (size_t i = 0; i < x.size(); ++i)
        worst = std::max(worst, std::fabs(g[i] - fd[i]) / (scale + 1e-30));
    return worst;
}

static std::vector<double> rand_lam(int n) {
    std::vector<double> l(n);
    for (double& v : l) v = U(-1.0, 1.0);
    return l;
}

static int report(const char* name, int cases, double worst, double tol) {
    std::printf("%-28s %5d %.3e %s\n", name, cases, worst, worst <= tol ? "ok" : "FAIL");
    return worst <= tol ? 0 : 1;
}

int main() {
    const double g = 9.80665;
    int fails = 0;
    const int N = 400;

    {   // cell states
        double worst = 0.0;
        for (int t = 0; t < N; ++t) {
            std::vector<double> x(12);
            for (int k = 0; k < 3; ++k) {
                x[4 * k + 0] = U(0.01, 3.0);
                x[4 * k + 1] = U(-2.0, 2.0);
                x[4 * k + 2] = U(-2.0, 2.0);
                x[4 * k + 3] = U(-2.0, 1.0);
            }
            auto cells = [](const std::vector<double>& v, Cell& m, Cell& c, Cell& p) {
                m = Cell{v[0], v[1], v[2], v[3]};
                c = Cell{v[4], v[5], v[6], v[7]};
                p = Cell{v[8], v[9], v[10], v[11]};
            };
            Fwd f = [&](const std::vector<double>& v) {
                Cell m, c, p;
                cells(v, m, c, p);
                FState lo, hi;
                states_fwd(m, c, p, lo, hi);
                return std::vector<double>{lo.h, lo.hu, lo.hv, hi.h, hi.hu, hi.hv};
            };
            Rev r = [&](const std::vector<double>& v, const std::vector<double>& l) {
                Cell m, c, p;
                cells(v, m, c, p);
                C4 lm{0, 0, 0, 0}, lc{0, 0, 0, 0}, lp{0, 0, 0, 0};
                states_rev(m, c, p, S3{l[0], l[1], l[2]}, S3{l[3], l[4], l[5]}, lm, lc, lp);
                return std::vector<double>{lm.h, lm.hu, lm.hv, lm.b, lc.h, lc.hu, lc.hv, lc.b,
                                           lp.h, lp.hu, lp.hv, lp.b};
            };
            worst = std::max(worst, check(f, r, x, rand_lam(6)));
        }
        fails += report("cell_states", N, worst, 1e-6);
    }

    for (int hllc = 0; hllc < 2; ++hllc)
        for (int regime = 0; regime < 3; ++regime) {   // wet-wet, dry right, dry left
            double worst = 0.0;
            for (int t = 0; t < N; ++t) {
                std::vector<double> x = {U(0.05, 3.0), U(-3.0, 3.0), U(-3.0, 3.0),
                                         U(0.05, 3.0), U(-3.0, 3.0), U(-3.0, 3.0)};
                if (regime == 1) { x[3] = U(0.0, 5e-7); x[4] = 0.0; x[5] = 0.0; }
                if (regime == 2) { x[0] = U(0.0, 5e-7); x[1] = 0.0; x[2] = 0.0; }
                Fwd f = [&](const std::vector<double>& v) {
                    double a, b, c;
                    if (hllc) riemann_fwd<true>(g, v[0], v[1], v[2], v[3], v[4], v[5], a, b, c);
                    else riemann_fwd<false>(g, v[0], v[1], v[2], v[3], v[4], v[5], a, b, c);
                    return std::vector<double>{a, b, c};
                };
                Rev r = [&](const std::vector<double>& v, const std::vector<double>& l) {
                    std::vector<double> o(6, 0.0);
                    if (hllc) riemann_rev<true>(g, v[0], v[1], v[2], v[3], v[4], v[5], l[0], l[1], l[2],
                                                o[0], o[1], o[2], o[3], o[4], o[5]);
                    else riemann_rev<false>(g, v[0], v[1], v[2], v[3], v[4], v[5], l[0], l[1], l[2],
                                            o[0], o[1], o[2], o[3], o[4], o[5]);
                    if (regime == 1) { o[3] = o[4] = o[5] = 0.0; }   // FD across h = 1e-6 not meaningful
                    if (regime == 2) { o[0] = o[1] = o[2] = 0.0; }
                    return o;
                };
                Fwd fz = [&](const std::vector<double>& v) {   // freeze the dry side for the FD
                    std::vector<double> w = v;
                    if (regime == 1) { w[3] = x[3]; w[4] = x[4]; w[5] = x[5]; }
                    if (regime == 2) { w[0] = x[0]; w[1] = x[1]; w[2] = x[2]; }
                    return f(w);
                };
                worst = std::max(worst, check(fz, r, x, rand_lam(3)));
            }
            char name[64];
            std::snprintf(name, sizeof name, "riemann %s %s", hllc ? "hllc" : "hll",
                          regime == 0 ? "wet" : (regime == 1 ? "dry-right" : "dry-left"));
            fails += report(name, N, worst, 1e-6);
        }

    for (int ydir = 0; ydir < 2; ++ydir)
        for (int hllc = 0; hllc < 2; ++hllc) {
            double worst = 0.0;
            for (int t = 0; t < N; ++t) {
                std::vector<double> x = {U(0.05, 3.0), U(-2.0, 2.0), U(-2.0, 2.0), U(-1.0, 1.0),
                                         U(0.05, 3.0), U(-2.0, 2.0), U(-2.0, 2.0), U(-1.0, 1.0)};
                Fwd f = [&](const std::vector<double>& v) {
                    const FState L{v[0], v[1], v[2]}, R{v[4], v[5], v[6]};
                    Face o;
                    if (ydir) o = hllc ? face_fwd<true, true>(g, L, v[3], R, v[7]) : face_fwd<false, true>(g, L, v[3], R, v[7]);
                    else o = hllc ? face_fwd<true, false>(g, L, v[3], R, v[7]) : face_fwd<false, false>(g, L, v[3], R, v[7]);
                    return std::vector<double>{o.fh, o.fhu, o.fhv, o.a, o.c, o.d};
                };
                Rev r = [&](const std::vector<double>& v, const std::vector<double>& l) {
                    const FState L{v[0], v[1], v[2]}, R{v[4], v[5], v[6]};
                    const Face lf{l[0], l[1], l[2], l[3], l[4], l[5]};
                    S3 lL{0, 0, 0}, lR{0, 0, 0};
                    double lbL = 0, lbR = 0;
                    if (ydir) {
                        if (hllc) face_rev<true, true>(g, L, v[3], R, v[7], lf, lL, lbL, lR, lbR);
                        else face_rev<false, true>(g, L, v[3], R, v[7], lf, lL, lbL, lR, lbR);
                    } else {
                        if (hllc) face_rev<true, false>(g, L, v[3], R, v[7], lf, lL, lbL, lR, lbR);
                        else face_rev<false, false>(g, L, v[3], R, v[7], lf, lL, lbL, lR, lbR);
                    }
                    return std::vector<double>{lL.h, lL.hu, lL.hv, lbL, lR.h, lR.hu, lR.hv, lbR};
                };
                worst = std::max(worst, check(f, r, x, rand_lam(6)));
            }
            char name[64];
            std::snprintf(name, sizeof name, "face %s %s", ydir ? "y" : "x", hllc ? "hllc" : "hll");
            fails += report(name, N, worst, 1e-6);
        }

    for (int regime = 0; regime < 3; ++regime) {   // plain, KP07 thin film, Froude backstop
        double worst = 0.0;
        for (int t = 0; t < N; ++t) {
            std::vector<double> x = {U(0.01, 3.0), U(-2.0, 2.0), U(-2.0, 2.0)};
            if (regime == 1) { x[0] = U(2e-6, 1.2e-4); x[1] = U(-1e-4, 1e-4); x[2] = U(-1e-4, 1e-4); }
            if (regime == 2) { x[0] = U(0.01, 0.1); x[1] = U(60.0, 90.0) * x[0]; x[2] = U(-60.0, 60.0) * x[0]; }
            Fwd f = [&](const std::vector<double>& v) {
                double h = v[0], hu = v[1], hv = v[2];
                clamp_fwd(g, h, hu, hv);
                return std::vector<double>{h, hu, hv};
            };
            Rev r = [&](const std::vector<double>& v, const std::vector<double>& l) {
                const S3 o = clamp_rev(g, v[0], v[1], v[2], S3{l[0], l[1], l[2]});
                return std::vector<double>{o.h, o.hu, o.hv};
            };
            worst = std::max(worst, check(f, r, x, rand_lam(3)));
        }
        const char* names[] = {"clamp plain", "clamp kp07", "clamp backstop"};
        fails += report(names[regime], N, worst, 2e-6);
    }

    for (int cor = 0; cor < 2; ++cor) {   // body: update + clamp + Manning (+ Coriolis)
        double worst = 0.0;
        SourceTerms s;
        s.g = g;
        s.n_sq = 0.0;
        s.has_friction = 1;
        s.has_coriolis = cor;
        s.cosf = std::cos(1e-4 * 0.3);
        s.sinf = std::sin(1e-4 * 0.3);
        for (int t = 0; t < N; ++t) {
            const double dt = 0.3;
            std::vector<double> x = {U(0.05, 3.0), U(-2.0, 2.0), U(-2.0, 2.0),
                                     U(-0.1, 0.1), U(-0.5, 0.5), U(-0.5, 0.5), U(0.01, 0.05)};
            Fwd f = [&](const std::vector<double>& v) {
                double h = v[0], hu = v[1], hv = v[2];
                body_fwd(s, dt, S3{v[3], v[4], v[5]}, v[6] * v[6], v[6] > 0.0, h, hu, hv);
                return std::vector<double>{h, hu, hv};
            };
            Rev r = [&](const std::vector<double>& v, const std::vector<double>& l) {
                const BodyRev b = body_rev(s, dt, S3{v[3], v[4], v[5]}, v[6], v[6] * v[6], v[6] > 0.0, v[0], v[1],
                                           v[2], S3{l[0], l[1], l[2]});
                return std::vector<double>{b.lv.h, b.lv.hu, b.lv.hv, b.ldiv.h, b.ldiv.hu, b.ldiv.hv, b.ln};
            };
            worst = std::max(worst, check(f, r, x, rand_lam(3)));
        }
        fails += report(cor ? "body friction+coriolis" : "body friction", N, worst, 2e-6);
    }

    const int tags[] = {BC_WALL, BC_OPEN, BC_TIDAL, BC_INFLOW, BC_DISCHARGE, BC_CLAMPED};
    const char* tag_names[] = {"wall", "open", "tidal", "inflow", "discharge", "clamped"};
    for (int ti = 0; ti < 6; ++ti)
        for (int dir = 0; dir < 2; ++dir) {
            double worst = 0.0;
            GridDesc d{};
            d.g = g;
            d.inflow_velocity = 0.4;
            d.discharge_flux = 0.7;
            d.clamped_elevation = 0.3;
            for (int t = 0; t < N; ++t) {
                std::vector<double> x = {U(0.05, 3.0), U(-2.0, 2.0), U(-2.0, 2.0), U(-1.0, 0.2),
                                         U(0.05, 3.0), U(-2.0, 2.0), U(-2.0, 2.0), U(-1.0, 0.2)};
                const double eta = 0.6, sgn = dir ? 1.0 : -1.0;
                const bool nu = (t % 2) == 0;
                Fwd f = [&](const std::vector<double>& v) {
                    const Cell mir{v[0], v[1], v[2], v[3]}, in0{v[4], v[5], v[6], v[7]};
                    const Cell o = ghost_from(d, tags[ti], eta, mir, in0, sgn, nu);
                    return std::vector<double>{o.h, o.hu, o.hv, o.b};
                };
                Rev r = [&](const std::vector<double>& v, const std::vector<double>& l) {
                    const Cell mir{v[0], v[1], v[2], v[3]}, in0{v[4], v[5], v[6], v[7]};
                    C4 lm{0, 0, 0, 0}, li{0, 0, 0, 0};
                    ghost_rev(tags[ti], g, eta, d.inflow_velocity, d.clamped_elevation, mir, in0, sgn, nu,
                              C4{l[0], l[1], l[2], l[3]}, lm, li);
                    return std::vector<double>{lm.h, lm.hu, lm.hv, lm.b, li.h, li.hu, li.hv, li.b};
                };
                worst = std::max(worst, check(f, r, x, rand_lam(4)));
            }
            char name[64];
            std::snprintf(name, sizeof name, "ghost %s %s", tag_names[ti], dir ? "high" : "low");
            fails += report(name, N, worst, 1e-6);
        }

    std::printf("%s\n", fails ? "FAILED" : "ALL OK");
    return fails ? 1 : 0;
}
