/*
 * rk_oracle.c -- CPU restatement (plain C, OpenMP over rows) of the
 * reference's structured 2D barotropic step.  TEST INFRASTRUCTURE: the
 * checker for the HIP path, never the product path.
 *
 * Operation order follows the reference expression by expression (Fortran
 * evaluates left to right; `x**2` is x*x; `x**4` is ((x*x)*x)*x, which is
 * what LLVM flang -O3 emits for a run-time base; constant expressions such
 * as DESING_EPS**4 and FROUDE_BACKSTOP**2*GRAVITY are folded to the correctly
 * rounded value).  Build with -ffp-contract=off so no multiply-add is fused,
 * matching an x86-64 build of the reference (no FMA in the baseline ISA).
 *
 * Every function cites the reference routine it restates.
 */
#include "rk_oracle.h"
#include "rk_crmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* rki_constants.F90:53,59,68,73,120 */
#define DRY_TOLERANCE 1.0e-6
#define THIN_LAYER_THRESHOLD 1.0e-3
#define DESING_EPS 1.5e-4
#define FROUDE_BACKSTOP 50.0
/* rki_constants.F90:114, TWO_PI = 8*atan(1) folded */
#define TWO_PI 6.283185307179586
#define RHO_WATER 1035.0   /* rki_constants.F90 */

static int g_threads = 0;
static int g_libm = 0; /* 0: correctly rounded pow/sin/cos (rk_crmath.h), 1: platform libm */

void orc_set_libm(int use_platform_libm) { g_libm = use_platform_libm ? 1 : 0; }
int orc_get_libm(void) { return g_libm; }

/* h**(1.0/3.0), sin, cos as the reference calls them (see rk_crmath.h) */
static inline double orc_pow13(double x) { return g_libm ? pow(x, 1.0 / 3.0) : crm_pow13(x); }
static inline double orc_sin(double x) { return g_libm ? sin(x) : crm_sin(x); }
static inline double orc_cos(double x) { return g_libm ? cos(x) : crm_cos(x); }

double orc_eval(int fn, double x)
{
    switch (fn) {
    case 0: return orc_pow13(x);
    case 1: return orc_sin(x);
    case 2: return orc_cos(x);
    default: return NAN;
    }
}

void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#define NT(p) ((p)->nx + 2 * (p)->ng)
#define MT(p) ((p)->ny + 2 * (p)->ng)
/* 1-based Fortran (i, j) -> linear offset, column-major. */
#define IX(i, j) ((size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)nxt)

static inline double fmax0(double x) { return x > 0.0 ? x : 0.0; }
static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline double pow4(double x) { return ((x * x) * x) * x; }

void orc_default_params(orc_params *p)
{
    memset(p, 0, sizeof(*p));
    p->nx = 200; p->ny = 1; p->ng = 3;   /* rki_config.F90 config_t */
    p->dx = 1.0; p->dy = 1.0;
    p->g = 9.80665;                      /* rki_constants.F90:38 */
    p->cfl = 0.45;
    p->manning_n = 0.0;
    p->coriolis_f = 0.0;
    p->dt_fixed = 0.0;
    p->t_end = 1.0;
    p->cfl_interval = 1;
    p->use_hllc = 0;                     /* rki_config.F90:191 */
    p->bc[0] = p->bc[1] = p->bc[2] = p->bc[3] = ORC_BC_WALL;
    p->tidal_amplitude = 0.0;
    p->tidal_period = 44712.0;           /* rki_boundary_types.F90 */
    p->time_scheme = ORC_TIME_RK2;       /* rki_timestepper.F90:47 */
}

/* ---------------------------------------------------------------------- */
/* Flux (rki_kernel_flux.F90)                                             */
/* ---------------------------------------------------------------------- */

/* minmod, rki_kernel_flux.F90:380-393 */
static inline double minmod(double a, double b)
{
    if (a * b <= 0.0) return 0.0;
    if (fabs(a) < fabs(b)) return a;
    return b;
}

/* extract_velocity, rki_kernel_flux.F90:395-408 */
static inline void extract_velocity(double h, double hu, double hv,
                                    double *u, double *v)
{
    if (h > DRY_TOLERANCE) { *u = hu / h; *v = hv / h; }
    else { *u = 0.0; *v = 0.0; }
}

/* hll_flux_x / hllc_flux_x, rki_kernel_flux.F90:410-495, 584-684 */
static void riemann_x(double g, int hllc, double h_l, double hu_l, double hv_l,
                      double h_r, double hu_r, double hv_r,
                      double *f_h, double *f_hu, double *f_hv)
{
    double u_l, v_l, u_r, v_r, a_l, a_r, s_l, s_r;
    if (h_l > DRY_TOLERANCE) { u_l = hu_l / h_l; v_l = hv_l / h_l; }
    else { u_l = 0.0; v_l = 0.0; }
    if (h_r > DRY_TOLERANCE) { u_r = hu_r / h_r; v_r = hv_r / h_r; }
    else { u_r = 0.0; v_r = 0.0; }
    a_l = sqrt(g * fmax0(h_l));
    a_r = sqrt(g * fmax0(h_r));
    if (h_r <= DRY_TOLERANCE) {
        s_l = u_l - a_l;
        s_r = u_l + 2.0 * a_l;
    } else if (h_l <= DRY_TOLERANCE) {
        s_l = u_r - 2.0 * a_r;
        s_r = u_r + a_r;
    } else {
        s_l = dmin(u_l - a_l, u_r - a_r);
        s_r = dmax(u_l + a_l, u_r + a_r);
    }
    double fl_h = hu_l, fl_hu = hu_l * u_l + 0.5 * g * h_l * h_l, fl_hv = hu_l * v_l;
    double fr_h = hu_r, fr_hu = hu_r * u_r + 0.5 * g * h_r * h_r, fr_hv = hu_r * v_r;
    if (s_l >= 0.0) {
        *f_h = fl_h; *f_hu = fl_hu; *f_hv = fl_hv;
    } else if (s_r <= 0.0) {
        *f_h = fr_h; *f_hu = fr_hu; *f_hv = fr_hv;
    } else {
        double denom = 1.0 / (s_r - s_l);
        *f_h = (s_r * fl_h - s_l * fr_h + s_l * s_r * (h_r - h_l)) * denom;
        *f_hu = (s_r * fl_hu - s_l * fr_hu + s_l * s_r * (hu_r - hu_l)) * denom;
        if (!hllc) {
            *f_hv = (s_r * fl_hv - s_l * fr_hv + s_l * s_r * (hv_r - hv_l)) * denom;
        } else {
            double s_star;
            double denom_star = h_r * (u_r - s_r) - h_l * (u_l - s_l);
            if (fabs(denom_star) > DRY_TOLERANCE)
                s_star = (s_l * h_r * (u_r - s_r) - s_r * h_l * (u_l - s_l)) / denom_star;
            else
                s_star = 0.5 * (u_l + u_r);
            *f_hv = (s_star >= 0.0) ? *f_h * v_l : *f_h * v_r;
        }
    }
}

/* hll_flux_y / hllc_flux_y, rki_kernel_flux.F90:497-582, 686-783.
 * Arguments keep the reference order (h, hu, hv) of bottom then top. */
static void riemann_y(double g, int hllc, double h_b, double hu_b, double hv_b,
                      double h_t, double hu_t, double hv_t,
                      double *g_h, double *g_hu, double *g_hv)
{
    double u_b, v_b, u_t, v_t, a_b, a_t, s_l, s_r;
    if (h_b > DRY_TOLERANCE) { u_b = hu_b / h_b; v_b = hv_b / h_b; }
    else { u_b = 0.0; v_b = 0.0; }
    if (h_t > DRY_TOLERANCE) { u_t = hu_t / h_t; v_t = hv_t / h_t; }
    else { u_t = 0.0; v_t = 0.0; }
    a_b = sqrt(g * fmax0(h_b));
    a_t = sqrt(g * fmax0(h_t));
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
    double gl_h = hv_b, gl_hu = hu_b * v_b, gl_hv = hv_b * v_b + 0.5 * g * h_b * h_b;
    double gr_h = hv_t, gr_hu = hu_t * v_t, gr_hv = hv_t * v_t + 0.5 * g * h_t * h_t;
    if (s_l >= 0.0) {
        *g_h = gl_h; *g_hu = gl_hu; *g_hv = gl_hv;
    } else if (s_r <= 0.0) {
        *g_h = gr_h; *g_hu = gr_hu; *g_hv = gr_hv;
    } else {
        double denom = 1.0 / (s_r - s_l);
        *g_h = (s_r * gl_h - s_l * gr_h + s_l * s_r * (h_t - h_b)) * denom;
        if (!hllc) {
            *g_hu = (s_r * gl_hu - s_l * gr_hu + s_l * s_r * (hu_t - hu_b)) * denom;
            *g_hv = (s_r * gl_hv - s_l * gr_hv + s_l * s_r * (hv_t - hv_b)) * denom;
        } else {
            double s_star;
            *g_hv = (s_r * gl_hv - s_l * gr_hv + s_l * s_r * (hv_t - hv_b)) * denom;
            double denom_star = h_t * (v_t - s_r) - h_b * (v_b - s_l);
            if (fabs(denom_star) > DRY_TOLERANCE)
                s_star = (s_l * h_t * (v_t - s_r) - s_r * h_b * (v_b - s_l)) / denom_star;
            else
                s_star = 0.5 * (v_b + v_t);
            *g_hu = (s_star >= 0.0) ? *g_h * u_b : *g_h * u_t;
        }
    }
}

/* flux_cell, rki_kernel_flux.F90:104-378 (the cell-centric form: every
 * interior face is evaluated from both of its cells). */
static void flux_cell(const orc_params *p, int nxt, const double *h,
                      const double *hu, const double *hv, const double *b,
                      int i, int j, int hllc,
                      double *out_fh, double *out_fhu, double *out_fhv)
{
    const double g = p->g, dx = p->dx, dy = p->dy;
    const double half_g = 0.5 * g;
    double eta_c = h[IX(i, j)] + b[IX(i, j)];
    double eta_im1 = h[IX(i - 1, j)] + b[IX(i - 1, j)];
    double eta_ip1 = h[IX(i + 1, j)] + b[IX(i + 1, j)];
    double eta_im2 = h[IX(i - 2, j)] + b[IX(i - 2, j)];
    double eta_ip2 = h[IX(i + 2, j)] + b[IX(i + 2, j)];
    double eta_jm1 = h[IX(i, j - 1)] + b[IX(i, j - 1)];
    double eta_jp1 = h[IX(i, j + 1)] + b[IX(i, j + 1)];
    double eta_jm2 = h[IX(i, j - 2)] + b[IX(i, j - 2)];
    double eta_jp2 = h[IX(i, j + 2)] + b[IX(i, j + 2)];
    double seta_x, shu_x, shv_x, seta_y, shu_y, shv_y;
    double snb_eta, snb_hu, snb_hv;
    double eta_L, eta_R, hu_L, hv_L, hu_R, hv_R, h_L, h_R, b_max;
    double h_hyd_l, h_hyd_r, u_l, v_l, u_r, v_r;
    double fx_h_l, fx_hu_l, fx_hv_l, fx_h_r, fx_hu_r, fx_hv_r;
    double fy_h_b, fy_hu_b, fy_hv_b, fy_h_t, fy_hu_t, fy_hv_t;
    double h_hyd_left, h_pre_left, h_hyd_right, h_pre_right;
    double h_hyd_bottom, h_pre_bottom, h_hyd_top, h_pre_top;

    /* :158-181 first-order fallback near wet/dry fronts */
    if (h[IX(i, j)] < THIN_LAYER_THRESHOLD || h[IX(i - 1, j)] < DRY_TOLERANCE ||
        h[IX(i + 1, j)] < DRY_TOLERANCE) {
        seta_x = shu_x = shv_x = 0.0;
    } else {
        seta_x = minmod(eta_c - eta_im1, eta_ip1 - eta_c);
        shu_x = minmod(hu[IX(i, j)] - hu[IX(i - 1, j)], hu[IX(i + 1, j)] - hu[IX(i, j)]);
        shv_x = minmod(hv[IX(i, j)] - hv[IX(i - 1, j)], hv[IX(i + 1, j)] - hv[IX(i, j)]);
    }
    if (h[IX(i, j)] < THIN_LAYER_THRESHOLD || h[IX(i, j - 1)] < DRY_TOLERANCE ||
        h[IX(i, j + 1)] < DRY_TOLERANCE) {
        seta_y = shu_y = shv_y = 0.0;
    } else {
        seta_y = minmod(eta_c - eta_jm1, eta_jp1 - eta_c);
        shu_y = minmod(hu[IX(i, j)] - hu[IX(i, j - 1)], hu[IX(i, j + 1)] - hu[IX(i, j)]);
        shv_y = minmod(hv[IX(i, j)] - hv[IX(i, j - 1)], hv[IX(i, j + 1)] - hv[IX(i, j)]);
    }

    /* :183-224 x LEFT interface (i-1/2) */
    eta_R = eta_c - 0.5 * seta_x;
    hu_R = hu[IX(i, j)] - 0.5 * shu_x;
    hv_R = hv[IX(i, j)] - 0.5 * shv_x;
    h_R = fmax0(eta_R - b[IX(i, j)]);
    h_pre_left = h_R;
    if (h[IX(i - 1, j)] < THIN_LAYER_THRESHOLD || h[IX(i - 2, j)] < DRY_TOLERANCE ||
        h[IX(i, j)] < DRY_TOLERANCE) {
        snb_eta = snb_hu = snb_hv = 0.0;
    } else {
        snb_eta = minmod(eta_im1 - eta_im2, eta_c - eta_im1);
        snb_hu = minmod(hu[IX(i - 1, j)] - hu[IX(i - 2, j)], hu[IX(i, j)] - hu[IX(i - 1, j)]);
        snb_hv = minmod(hv[IX(i - 1, j)] - hv[IX(i - 2, j)], hv[IX(i, j)] - hv[IX(i - 1, j)]);
    }
    eta_L = eta_im1 + 0.5 * snb_eta;
    hu_L = hu[IX(i - 1, j)] + 0.5 * snb_hu;
    hv_L = hv[IX(i - 1, j)] + 0.5 * snb_hv;
    h_L = fmax0(eta_L - b[IX(i - 1, j)]);
    b_max = dmax(b[IX(i - 1, j)], b[IX(i, j)]);
    h_hyd_l = fmax0(h_L + b[IX(i - 1, j)] - b_max);
    h_hyd_r = fmax0(h_R + b[IX(i, j)] - b_max);
    h_hyd_left = h_hyd_r;
    extract_velocity(h_L, hu_L, hv_L, &u_l, &v_l);
    extract_velocity(h_R, hu_R, hv_R, &u_r, &v_r);
    riemann_x(g, hllc, h_hyd_l, h_hyd_l * u_l, h_hyd_l * v_l,
              h_hyd_r, h_hyd_r * u_r, h_hyd_r * v_r, &fx_h_l, &fx_hu_l, &fx_hv_l);

    /* :226-267 x RIGHT interface (i+1/2) */
    eta_L = eta_c + 0.5 * seta_x;
    hu_L = hu[IX(i, j)] + 0.5 * shu_x;
    hv_L = hv[IX(i, j)] + 0.5 * shv_x;
    h_L = fmax0(eta_L - b[IX(i, j)]);
    h_pre_right = h_L;
    if (h[IX(i + 1, j)] < THIN_LAYER_THRESHOLD || h[IX(i, j)] < DRY_TOLERANCE ||
        h[IX(i + 2, j)] < DRY_TOLERANCE) {
        snb_eta = snb_hu = snb_hv = 0.0;
    } else {
        snb_eta = minmod(eta_ip1 - eta_c, eta_ip2 - eta_ip1);
        snb_hu = minmod(hu[IX(i + 1, j)] - hu[IX(i, j)], hu[IX(i + 2, j)] - hu[IX(i + 1, j)]);
        snb_hv = minmod(hv[IX(i + 1, j)] - hv[IX(i, j)], hv[IX(i + 2, j)] - hv[IX(i + 1, j)]);
    }
    eta_R = eta_ip1 - 0.5 * snb_eta;
    hu_R = hu[IX(i + 1, j)] - 0.5 * snb_hu;
    hv_R = hv[IX(i + 1, j)] - 0.5 * snb_hv;
    h_R = fmax0(eta_R - b[IX(i + 1, j)]);
    b_max = dmax(b[IX(i, j)], b[IX(i + 1, j)]);
    h_hyd_l = fmax0(h_L + b[IX(i, j)] - b_max);
    h_hyd_r = fmax0(h_R + b[IX(i + 1, j)] - b_max);
    h_hyd_right = h_hyd_l;
    extract_velocity(h_L, hu_L, hv_L, &u_l, &v_l);
    extract_velocity(h_R, hu_R, hv_R, &u_r, &v_r);
    riemann_x(g, hllc, h_hyd_l, h_hyd_l * u_l, h_hyd_l * v_l,
              h_hyd_r, h_hyd_r * u_r, h_hyd_r * v_r, &fx_h_r, &fx_hu_r, &fx_hv_r);

    /* :269-310 y BOTTOM interface (j-1/2) */
    eta_R = eta_c - 0.5 * seta_y;
    hu_R = hu[IX(i, j)] - 0.5 * shu_y;
    hv_R = hv[IX(i, j)] - 0.5 * shv_y;
    h_R = fmax0(eta_R - b[IX(i, j)]);
    h_pre_bottom = h_R;
    if (h[IX(i, j - 1)] < THIN_LAYER_THRESHOLD || h[IX(i, j - 2)] < DRY_TOLERANCE ||
        h[IX(i, j)] < DRY_TOLERANCE) {
        snb_eta = snb_hu = snb_hv = 0.0;
    } else {
        snb_eta = minmod(eta_jm1 - eta_jm2, eta_c - eta_jm1);
        snb_hu = minmod(hu[IX(i, j - 1)] - hu[IX(i, j - 2)], hu[IX(i, j)] - hu[IX(i, j - 1)]);
        snb_hv = minmod(hv[IX(i, j - 1)] - hv[IX(i, j - 2)], hv[IX(i, j)] - hv[IX(i, j - 1)]);
    }
    eta_L = eta_jm1 + 0.5 * snb_eta;
    hu_L = hu[IX(i, j - 1)] + 0.5 * snb_hu;
    hv_L = hv[IX(i, j - 1)] + 0.5 * snb_hv;
    h_L = fmax0(eta_L - b[IX(i, j - 1)]);
    b_max = dmax(b[IX(i, j - 1)], b[IX(i, j)]);
    h_hyd_l = fmax0(h_L + b[IX(i, j - 1)] - b_max);
    h_hyd_r = fmax0(h_R + b[IX(i, j)] - b_max);
    h_hyd_bottom = h_hyd_r;
    extract_velocity(h_L, hu_L, hv_L, &u_l, &v_l);
    extract_velocity(h_R, hu_R, hv_R, &u_r, &v_r);
    riemann_y(g, hllc, h_hyd_l, h_hyd_l * u_l, h_hyd_l * v_l,
              h_hyd_r, h_hyd_r * u_r, h_hyd_r * v_r, &fy_h_b, &fy_hu_b, &fy_hv_b);

    /* :312-353 y TOP interface (j+1/2) */
    eta_L = eta_c + 0.5 * seta_y;
    hu_L = hu[IX(i, j)] + 0.5 * shu_y;
    hv_L = hv[IX(i, j)] + 0.5 * shv_y;
    h_L = fmax0(eta_L - b[IX(i, j)]);
    h_pre_top = h_L;
    if (h[IX(i, j + 1)] < THIN_LAYER_THRESHOLD || h[IX(i, j)] < DRY_TOLERANCE ||
        h[IX(i, j + 2)] < DRY_TOLERANCE) {
        snb_eta = snb_hu = snb_hv = 0.0;
    } else {
        snb_eta = minmod(eta_jp1 - eta_c, eta_jp2 - eta_jp1);
        snb_hu = minmod(hu[IX(i, j + 1)] - hu[IX(i, j)], hu[IX(i, j + 2)] - hu[IX(i, j + 1)]);
        snb_hv = minmod(hv[IX(i, j + 1)] - hv[IX(i, j)], hv[IX(i, j + 2)] - hv[IX(i, j + 1)]);
    }
    eta_R = eta_jp1 - 0.5 * snb_eta;
    hu_R = hu[IX(i, j + 1)] - 0.5 * snb_hu;
    hv_R = hv[IX(i, j + 1)] - 0.5 * snb_hv;
    h_R = fmax0(eta_R - b[IX(i, j + 1)]);
    b_max = dmax(b[IX(i, j)], b[IX(i, j + 1)]);
    h_hyd_l = fmax0(h_L + b[IX(i, j)] - b_max);
    h_hyd_r = fmax0(h_R + b[IX(i, j + 1)] - b_max);
    h_hyd_top = h_hyd_l;
    extract_velocity(h_L, hu_L, hv_L, &u_l, &v_l);
    extract_velocity(h_R, hu_R, hv_R, &u_r, &v_r);
    riemann_y(g, hllc, h_hyd_l, h_hyd_l * u_l, h_hyd_l * v_l,
              h_hyd_r, h_hyd_r * u_r, h_hyd_r * v_r, &fy_h_t, &fy_hu_t, &fy_hv_t);

    /* :355-369 net divergence with the well-balanced correction */
    *out_fh = -(fx_h_r - fx_h_l) / dx - (fy_h_t - fy_h_b) / dy;
    *out_fhu = -(fx_hu_r - fx_hu_l) / dx - (fy_hu_t - fy_hu_b) / dy -
               half_g * (h_hyd_left * h_hyd_left - h_pre_left * h_pre_left -
                         h_hyd_right * h_hyd_right + h_pre_right * h_pre_right) / dx;
    *out_fhv = -(fx_hv_r - fx_hv_l) / dx - (fy_hv_t - fy_hv_b) / dy -
               half_g * (h_hyd_bottom * h_hyd_bottom - h_pre_bottom * h_pre_bottom -
                         h_hyd_top * h_hyd_top + h_pre_top * h_pre_top) / dy;
}

/* compute_flux_hll / compute_flux_hllc, rki_kernel_flux.F90:17-102 */
void orc_compute_flux(const orc_params *p, const double *h, const double *hu,
                      const double *hv, const double *b, double *flux_h,
                      double *flux_hu, double *flux_hv, int use_hllc)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    int j;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (j = ng + 1; j <= nyt - ng; ++j) {
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            double fh, fhu, fhv;
            flux_cell(p, nxt, h, hu, hv, b, i, j, use_hllc, &fh, &fhu, &fhv);
            flux_h[IX(i, j)] = fh;
            flux_hu[IX(i, j)] = fhu;
            flux_hv[IX(i, j)] = fhv;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Fused updates (rki_kernel_fused.F90)                                   */
/* ---------------------------------------------------------------------- */

/* Clamp + KP07 desingularization + Froude backstop, rki_kernel_fused.F90
 * :234-260 (stage bodies) and :441-463 (post-average). */
static inline void clamp_desing(double g, double eps4, double *h_val,
                                double *hu_val, double *hv_val)
{
    double h = *h_val, hu = *hu_val, hv = *hv_val;
    if (h < 0.0) {
        h = 0.0; hu = 0.0; hv = 0.0;
    } else if (h < DRY_TOLERANCE) {
        hu = 0.0; hv = 0.0;
    } else {
        double h4 = pow4(h);
        if (h4 < eps4) {
            double u = sqrt(2.0) * h * hu / sqrt(h4 + eps4);
            double v = sqrt(2.0) * h * hv / sqrt(h4 + eps4);
            hu = h * u;
            hv = h * v;
        }
        double speed_sq = (hu * hu + hv * hv) / (h * h);
        double fr_max_sq = (FROUDE_BACKSTOP * FROUDE_BACKSTOP * g) * h;
        if (speed_sq > fr_max_sq) {
            double fr_scale = sqrt(fr_max_sq / speed_sq);
            hu = hu * fr_scale;
            hv = hv * fr_scale;
        }
    }
    *h_val = h; *hu_val = hu; *hv_val = hv;
}

/* Forcing at one cell: wind stress, the central pressure gradient and the
 * precipitation rate (rki_kernel_fused.F90:268-275 and :286-289).  Off-grid
 * pressure is the ghost value, which no setter writes (0). */
typedef struct { int on; double wx, wy, dpdx, dpdy, rain; } orc_frc;

static inline double frc_at(const double *a, size_t k) { return a ? a[k] : 0.0; }

static inline orc_frc forcing_at(const orc_params *p, int nxt, int i, int j)
{
    orc_frc f = {0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (!p->forcing_active) return f;
    f.on = 1;
    f.wx = frc_at(p->wind_x, IX(i, j));
    f.wy = frc_at(p->wind_y, IX(i, j));
    f.dpdx = (frc_at(p->atm_pressure, IX(i + 1, j)) - frc_at(p->atm_pressure, IX(i - 1, j))) / (2.0 * p->dx);
    f.dpdy = (frc_at(p->atm_pressure, IX(i, j + 1)) - frc_at(p->atm_pressure, IX(i, j - 1))) / (2.0 * p->dy);
    f.rain = frc_at(p->precipitation, IX(i, j));
    return f;
}

/* Euler update + clamp + Manning + forcing + Coriolis + precipitation for
 * one cell, rki_kernel_fused.F90:228-290. */
static inline void stage_body(const orc_params *p, double dt, double eps4,
                              double n_sq, int has_friction, int has_coriolis,
                              double cosf, double sinf, double fh, double fhu,
                              double fhv, double *h_val, double *hu_val,
                              double *hv_val, orc_frc frc)
{
    const double g = p->g;
    double h = *h_val + dt * fh;
    double hu = *hu_val + dt * fhu;
    double hv = *hv_val + dt * fhv;
    clamp_desing(g, eps4, &h, &hu, &hv);
    if (h > DRY_TOLERANCE) {
        if (has_friction) {
            double u = hu / h;
            double v = hv / h;
            double speed = sqrt(u * u + v * v);
            double cf = g * n_sq / orc_pow13(h);
            double denom = 1.0 + dt * cf * speed / h;
            hu = hu / denom;
            hv = hv / denom;
        }
        if (frc.on) {   /* :268-275 */
            hu = hu + dt * frc.wx / RHO_WATER;
            hv = hv + dt * frc.wy / RHO_WATER;
            hu = hu - dt * h * frc.dpdx / RHO_WATER;
            hv = hv - dt * h * frc.dpdy / RHO_WATER;
        }
        if (has_coriolis) {
            double hu_old = hu, hv_old = hv;
            hu = hu_old * cosf + hv_old * sinf;
            hv = -hu_old * sinf + hv_old * cosf;
        }
    }
    if (frc.on) h = h + dt * frc.rain;   /* :286-289, every cell */
    *h_val = h; *hu_val = hu; *hv_val = hv;
}

#define STAGE_SETUP                                                        \
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;                        \
    const double eps4 = pow4(DESING_EPS);                                  \
    const double n_sq = p->manning_n * p->manning_n;                       \
    const int has_friction = p->manning_n > 0.0;                           \
    const int has_coriolis = fabs(p->coriolis_f) > 1.0e-20;                \
    double cosf = 0.0, sinf = 0.0;                                         \
    if (has_coriolis) { cosf = orc_cos(p->coriolis_f * dt); sinf = orc_sin(p->coriolis_f * dt); }

/* fused_update_stage1, rki_kernel_fused.F90:164-303 */
void orc_fused_stage1(const orc_params *p, double *h, double *hu, double *hv,
                      double *h0, double *hu0, double *hv0,
                      const double *fh, const double *fhu, const double *fhv,
                      double dt)
{
    STAGE_SETUP
    int j;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (j = ng + 1; j <= nyt - ng; ++j) {
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            size_t k = IX(i, j);
            double hh = h[k], hhu = hu[k], hhv = hv[k];
            h0[k] = hh; hu0[k] = hhu; hv0[k] = hhv;
            stage_body(p, dt, eps4, n_sq, has_friction, has_coriolis, cosf, sinf,
                       fh[k], fhu[k], fhv[k], &hh, &hhu, &hhv, forcing_at(p, nxt, i, j));
            h[k] = hh; hu[k] = hhu; hv[k] = hhv;
        }
    }
}

/* fused_update_stage2, rki_kernel_fused.F90:305-471 */
void orc_fused_stage2(const orc_params *p, double *h, double *hu, double *hv,
                      const double *h0, const double *hu0, const double *hv0,
                      const double *fh, const double *fhu, const double *fhv,
                      double dt)
{
    STAGE_SETUP
    int j;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (j = ng + 1; j <= nyt - ng; ++j) {
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            size_t k = IX(i, j);
            double hh = h[k], hhu = hu[k], hhv = hv[k];
            stage_body(p, dt, eps4, n_sq, has_friction, has_coriolis, cosf, sinf,
                       fh[k], fhu[k], fhv[k], &hh, &hhu, &hhv, forcing_at(p, nxt, i, j));
            /* :435-438 RK2 average, then :440-463 final clamp */
            hh = 0.5 * (h0[k] + hh);
            hhu = 0.5 * (hu0[k] + hhu);
            hhv = 0.5 * (hv0[k] + hhv);
            clamp_desing(p->g, eps4, &hh, &hhu, &hhv);
            h[k] = hh; hu[k] = hhu; hv[k] = hhv;
        }
    }
}

/* fused_update_stage_generic, rki_kernel_fused.F90:22-162: the stage-1
 * body without the U^n save, at the stage's dt (beta_i * dt). */
void orc_fused_stage_generic(const orc_params *p, double *h, double *hu, double *hv,
                             const double *fh, const double *fhu, const double *fhv,
                             double dt)
{
    STAGE_SETUP
    int j;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (j = ng + 1; j <= nyt - ng; ++j) {
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            size_t k = IX(i, j);
            double hh = h[k], hhu = hu[k], hhv = hv[k];
            stage_body(p, dt, eps4, n_sq, has_friction, has_coriolis, cosf, sinf,
                       fh[k], fhu[k], fhv[k], &hh, &hhu, &hhv, forcing_at(p, nxt, i, j));
            h[k] = hh; hu[k] = hhu; hv[k] = hhv;
        }
    }
}

/* compute_ssp_combine, rki_kernel_update.F90:152-208:
 * U = a_n U^n + (1 - a_n) U, then the clamp / KP07 / backstop. */
void orc_ssp_combine(const orc_params *p, double *h, double *hu, double *hv,
                     const double *h0, const double *hu0, const double *hv0,
                     double a_n)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const double eps4 = pow4(DESING_EPS);
    const double a_prev = 1.0 - a_n;
    int j;
#pragma omp parallel for schedule(static) num_threads(orc_get_threads())
    for (j = ng + 1; j <= nyt - ng; ++j) {
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            size_t k = IX(i, j);
            double hh = a_n * h0[k] + a_prev * h[k];
            double hhu = a_n * hu0[k] + a_prev * hu[k];
            double hhv = a_n * hv0[k] + a_prev * hv[k];
            clamp_desing(p->g, eps4, &hh, &hhu, &hhv);
            h[k] = hh; hu[k] = hhu; hv[k] = hhv;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* Gauges (rki_gauges.F90)                                                */
/* ---------------------------------------------------------------------- */

/* floor(v) + 1 as the reference's integer cell index, or a value outside
 * [1, n] when v lies off the grid (kept finite for the cast). */
static int gauge_cell(double v, double d, int n)
{
    const double f = floor(v / d);
    if (!(f >= -1.0) || f > (double)n) return -1;
    return (int)f + 1;
}

int orc_gauges_setup(const orc_params *p, int n, const double *x, const double *y,
                     int *local_idx, int *ix0, int *iy0, double *wx, double *wy)
{
    const double dx = p->dx, dy = p->dy;
    const int ng = p->ng, nx = p->nx, ny = p->ny;
    int nl = 0;
    for (int g = 0; g < n; ++g) {   /* compute_local_gauges :208-243 */
        const int ig = gauge_cell(x[g], dx, nx), jg = gauge_cell(y[g], dy, ny);
        if (ig >= 1 && ig <= nx && jg >= 1 && jg <= ny) local_idx[nl++] = g;
    }
    for (int l = 0; l < nl; ++l) {   /* compute_interp_weights :245-336 */
        const double gx = x[local_idx[l]], gy = y[local_idx[l]];
        const int ig = gauge_cell(gx, dx, nx), jg = gauge_cell(gy, dy, ny);
        const int ia = ig + ng, ja = jg + ng;
        const double xc = ((double)ig - 0.5) * dx, yc = ((double)jg - 0.5) * dy;
        double w;
        if (gx < xc) {
            if (ig > 1) { ix0[l] = ia - 1; w = (gx - (xc - dx)) / dx; }
            else { ix0[l] = ia; w = 0.0; }
        } else {
            if (ig < nx) { ix0[l] = ia; w = (gx - xc) / dx; }
            else { ix0[l] = ia; w = 0.0; }
        }
        wx[l] = fmax(0.0, fmin(1.0, w));
        if (gy < yc) {
            if (jg > 1) { iy0[l] = ja - 1; w = (gy - (yc - dy)) / dy; }
            else { iy0[l] = ja; w = 0.0; }
        } else {
            if (jg < ny) { iy0[l] = ja; w = (gy - yc) / dy; }
            else { iy0[l] = ja; w = 0.0; }
        }
        wy[l] = fmax(0.0, fmin(1.0, w));
    }
    return nl;
}

static double bilinear(const orc_params *p, const double *f, int i0, int j0, double wx, double wy)
{
    const int nxt = NT(p);
    return (1.0 - wx) * (1.0 - wy) * f[IX(i0, j0)] + wx * (1.0 - wy) * f[IX(i0 + 1, j0)]
           + (1.0 - wx) * wy * f[IX(i0, j0 + 1)] + wx * wy * f[IX(i0 + 1, j0 + 1)];
}

void orc_gauges_sample(const orc_params *p, const double *h, const double *hu,
                       const double *hv, const double *b, int n_local, const int *ix0,
                       const int *iy0, const double *wx, const double *wy,
                       double *eta, double *u, double *v)
{
    for (int g = 0; g < n_local; ++g) {
        const double hi = bilinear(p, h, ix0[g], iy0[g], wx[g], wy[g]);
        const double hui = bilinear(p, hu, ix0[g], iy0[g], wx[g], wy[g]);
        const double hvi = bilinear(p, hv, ix0[g], iy0[g], wx[g], wy[g]);
        const double bi = bilinear(p, b, ix0[g], iy0[g], wx[g], wy[g]);
        eta[g] = hi + bi;
        const double hs = fmax(hi, 1.0e-6);
        u[g] = hui / hs;
        v[g] = hvi / hs;
    }
}

/* apply_sponge, rki_kernel_sponge.F90:16-88: interior cells closer than
 * sponge_width cells to a BC_SPONGE edge relax toward (h_ref, 0, 0) with
 * alpha = strength ((width - d) / width)^2, d the distance in cells from
 * the nearest such edge (0 for the first interior cell). */
void orc_apply_sponge(const orc_params *p, double *h, double *hu, double *hv, double dt)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const int sw = p->sponge_width;
    if (sw <= 0) return;
    const double strength = p->sponge_strength, h_ref = p->sponge_h_ref;
    if (strength <= 0.0) return;
    const int i_first = ng + 1, i_last = nxt - ng, j_first = ng + 1, j_last = nyt - ng;
    const int has_w = p->bc[0] == ORC_BC_SPONGE, has_e = p->bc[1] == ORC_BC_SPONGE;
    const int has_s = p->bc[2] == ORC_BC_SPONGE, has_n = p->bc[3] == ORC_BC_SPONGE;
    for (int j = j_first; j <= j_last; ++j) {
        for (int i = i_first; i <= i_last; ++i) {
            int d_min = sw + 1;
            if (has_w && i - i_first < sw && i - i_first < d_min) d_min = i - i_first;
            if (has_e && i_last - i < sw && i_last - i < d_min) d_min = i_last - i;
            if (has_s && j - j_first < sw && j - j_first < d_min) d_min = j - j_first;
            if (has_n && j_last - j < sw && j_last - j < d_min) d_min = j_last - j;
            if (d_min < sw) {
                const double ratio = (double)(sw - d_min) / (double)sw;
                const double alpha = strength * ratio * ratio;
                const size_t k = IX(i, j);
                h[k] = h[k] - dt * alpha * (h[k] - h_ref);
                hu[k] = hu[k] - dt * alpha * hu[k];
                hv[k] = hv[k] - dt * alpha * hv[k];
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* CFL timestep (rki_kernel_timestep.F90:16-78)                           */
/* ---------------------------------------------------------------------- */
double orc_compute_timestep(const orc_params *p, const double *h,
                            const double *hu, const double *hv)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const double g = p->g, dx = p->dx, dy = p->dy;
    const double eps4 = pow4(DESING_EPS);
    double max_speed = 0.0;
    int j;
#pragma omp parallel for schedule(static) reduction(max : max_speed) num_threads(orc_get_threads())
    for (j = ng + 1; j <= nyt - ng; ++j) {
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            double hh = h[IX(i, j)];
            if (hh > DRY_TOLERANCE) {
                double wave_speed = sqrt(g * hh);
                double h4 = pow4(hh);
                double denom = sqrt(h4 + eps4);
                double u_abs = fabs(hu[IX(i, j)]) * sqrt(2.0) * hh / denom;
                double v_abs = fabs(hv[IX(i, j)]) * sqrt(2.0) * hh / denom;
                double sx = (u_abs + wave_speed) / dx;
                double sy = (v_abs + wave_speed) / dy;
                max_speed = dmax(max_speed, dmax(sx, sy));
            }
        }
    }
    if (max_speed > 0.0) return p->cfl / max_speed;
    return p->cfl * dmin(dx, dy) / sqrt(g * DRY_TOLERANCE);
}

/* ---------------------------------------------------------------------- */
/* Boundary conditions (rki_boundary.F90:40-470)                          */
/* ---------------------------------------------------------------------- */

/* Tidal surface, rki_boundary.F90:92-101 */
double orc_tidal_eta(const orc_params *p, double time)
{
    double tidal_eta = 0.0;
    if (p->n_tc > 0) {
        for (int n = 0; n < p->n_tc; ++n)
            tidal_eta = tidal_eta + p->tc_amp[n] * orc_cos(p->tc_omega[n] * time - p->tc_phase[n]);
    } else if (p->tidal_period > 0.0) {
        tidal_eta = p->tidal_amplitude * orc_sin(TWO_PI * time / p->tidal_period);
    }
    return tidal_eta;
}

typedef struct {
    int n_tan, ig0, jg0, din, djn, dit, djt;
    double nrm_sign;
    int normal_is_u;
} edge_t;

/* bc_wall_edge, rki_boundary.F90:220-291 */
static void bc_wall_edge(const orc_params *p, double *h, double *hu, double *hv,
                         double *b, const edge_t *e)
{
    const int nxt = NT(p), ng = p->ng;
    for (int t = 1; t <= e->n_tan; ++t) {
        for (int k = 1; k <= ng; ++k) {
            int ig = e->ig0 + (k - 1) * e->din + (t - 1) * e->dit;
            int jg = e->jg0 + (k - 1) * e->djn + (t - 1) * e->djt;
            int im = ig - (2 * k - 1) * e->din;
            int jm = jg - (2 * k - 1) * e->djn;
            b[IX(ig, jg)] = b[IX(im, jm)];
            h[IX(ig, jg)] = fmax0(h[IX(im, jm)] + b[IX(im, jm)] - b[IX(ig, jg)]);
            if (e->normal_is_u) {
                hu[IX(ig, jg)] = -hu[IX(im, jm)];
                hv[IX(ig, jg)] = hv[IX(im, jm)];
            } else {
                hv[IX(ig, jg)] = -hv[IX(im, jm)];
                hu[IX(ig, jg)] = hu[IX(im, jm)];
            }
        }
    }
}

/* bc_flather_edge, rki_boundary.F90:293-345 */
static void bc_flather_edge(const orc_params *p, double *h, double *hu, double *hv,
                            const double *b, const edge_t *e, double eta_ext,
                            int prescribe_h)
{
    const int nxt = NT(p), ng = p->ng;
    const double g = p->g, sgn = e->nrm_sign;
    for (int t = 1; t <= e->n_tan; ++t) {
        for (int k = 1; k <= ng; ++k) {
            int ig = e->ig0 + (k - 1) * e->din + (t - 1) * e->dit;
            int jg = e->jg0 + (k - 1) * e->djn + (t - 1) * e->djt;
            int ii = e->ig0 + (t - 1) * e->dit - e->din;
            int ji = e->jg0 + (t - 1) * e->djt - e->djn;
            double h_int = h[IX(ii, ji)];
            double eta_int = h_int + b[IX(ii, ji)];
            double hmax = dmax(h_int, 1.0e-6);
            double speed = sqrt(g * hmax);
            double h_ghost = prescribe_h ? fmax0(eta_ext - b[IX(ig, jg)]) : h_int;
            h[IX(ig, jg)] = h_ghost;
            if (e->normal_is_u) {
                hu[IX(ig, jg)] = h_ghost * (hu[IX(ii, ji)] / hmax + sgn * speed * (eta_int - eta_ext) / hmax);
                hv[IX(ig, jg)] = hv[IX(ii, ji)];
            } else {
                hv[IX(ig, jg)] = h_ghost * (hv[IX(ii, ji)] / hmax + sgn * speed * (eta_int - eta_ext) / hmax);
                hu[IX(ig, jg)] = hu[IX(ii, ji)];
            }
        }
    }
}

/* bc_inflow_edge (:347-390), bc_disch_edge (:392-430), bc_clamped_edge
 * (:432-466): zero-gradient / prescribed ghost fills. */
static void bc_simple_edge(const orc_params *p, double *h, double *hu, double *hv,
                           const double *b, const edge_t *e, int tag)
{
    const int nxt = NT(p), ng = p->ng;
    const double in_sgn = -e->nrm_sign;
    for (int t = 1; t <= e->n_tan; ++t) {
        for (int k = 1; k <= ng; ++k) {
            int ig = e->ig0 + (k - 1) * e->din + (t - 1) * e->dit;
            int jg = e->jg0 + (k - 1) * e->djn + (t - 1) * e->djt;
            int ii = e->ig0 + (t - 1) * e->dit - e->din;
            int ji = e->jg0 + (t - 1) * e->djt - e->djn;
            if (tag == ORC_BC_CLAMPED) {
                h[IX(ig, jg)] = fmax0(p->clamped_elevation - b[IX(ig, jg)]);
                hu[IX(ig, jg)] = hu[IX(ii, ji)];
                hv[IX(ig, jg)] = hv[IX(ii, ji)];
                continue;
            }
            double h_int = h[IX(ii, ji)];
            double nrm = (tag == ORC_BC_INFLOW) ? in_sgn * h_int * p->inflow_velocity
                                                : in_sgn * p->discharge_flux;
            h[IX(ig, jg)] = h_int;
            if (e->normal_is_u) {
                hu[IX(ig, jg)] = nrm;
                hv[IX(ig, jg)] = hv[IX(ii, ji)];
            } else {
                hv[IX(ig, jg)] = nrm;
                hu[IX(ig, jg)] = hu[IX(ii, ji)];
            }
        }
    }
}

/* bc_chapman_edge, rki_boundary.F90:471-552: Chapman / Orlanski radiation.
 * The ghost surface advects at the phase speed c = -(d eta/dt)/(d eta/dn)
 * from the interior tendency against eta_old and the along-normal
 * gradient, clamped to [0, sqrt(g hmax)]; the normal momentum is the
 * Flather radiation about that surface.  The loops run t outer, k inner
 * (the serial order of the reference's CPU build), so eta_old(t), set to
 * the interior eta on the k = 1 layer, is already current for k = 2. */
static void bc_chapman_edge(const orc_params *p, double *h, double *hu, double *hv,
                            const double *b, const edge_t *e, double *eta_old, double dt,
                            double dd)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const double g = p->g, sgn = e->nrm_sign;
    for (int t = 1; t <= e->n_tan; ++t) {
        for (int k = 1; k <= ng; ++k) {
            const int ig = e->ig0 + (k - 1) * e->din + (t - 1) * e->dit;
            const int jg = e->jg0 + (k - 1) * e->djn + (t - 1) * e->djt;
            const int ii = e->ig0 + (t - 1) * e->dit - e->din;
            const int ji = e->jg0 + (t - 1) * e->djt - e->djn;
            const int i2 = ii - e->din, j2 = ji - e->djn;
            const double h_int = h[IX(ii, ji)];
            const double eta_int = h_int + b[IX(ii, ji)];
            const double eta_old_val = eta_old[t - 1];
            const double hmax = dmax(h_int, 1.0e-6);
            const double speed = sqrt(g * hmax);
            double eta_ghost;
            if (dt > 0.0) {
                const double deta_dt = (eta_int - eta_old_val) / dt;
                double deta_dx = 0.0;
                if (i2 >= 1 && i2 <= nxt && j2 >= 1 && j2 <= nyt)
                    deta_dx = (h[IX(i2, j2)] + b[IX(i2, j2)] - eta_int) / dd;
                double c_phase = fabs(deta_dx) > 1.0e-10 ? -deta_dt / deta_dx : 0.0;
                c_phase = dmax(0.0, fmin(c_phase, speed));
                eta_ghost = eta_old_val - c_phase * dt / dd * (eta_old_val - eta_int);
            } else {
                eta_ghost = eta_int;
            }
            if (k == 1) eta_old[t - 1] = eta_int;
            const double h_ghost = fmax0(eta_ghost - b[IX(ig, jg)]);
            h[IX(ig, jg)] = h_ghost;
            if (e->normal_is_u) {
                hu[IX(ig, jg)] = h_ghost * (hu[IX(ii, ji)] / hmax + sgn * speed * (eta_int - eta_ghost) / hmax);
                hv[IX(ig, jg)] = hv[IX(ii, ji)];
            } else {
                hv[IX(ig, jg)] = h_ghost * (hv[IX(ii, ji)] / hmax + sgn * speed * (eta_int - eta_ghost) / hmax);
                hu[IX(ig, jg)] = hu[IX(ii, ji)];
            }
        }
    }
}

void orc_chapman_seed(const orc_params *p, const double *h, const double *b)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const int any_we = p->bc[0] == ORC_BC_CHAPMAN || p->bc[1] == ORC_BC_CHAPMAN;
    const int any_sn = p->bc[2] == ORC_BC_CHAPMAN || p->bc[3] == ORC_BC_CHAPMAN;
    for (int j = 1; j <= nyt && any_we; ++j) {
        if (p->chap_eta_old[0]) p->chap_eta_old[0][j - 1] = h[IX(ng + 1, j)] + b[IX(ng + 1, j)];
        if (p->chap_eta_old[1]) p->chap_eta_old[1][j - 1] = h[IX(nxt - ng, j)] + b[IX(nxt - ng, j)];
    }
    for (int i = 1; i <= nxt && any_sn; ++i) {
        if (p->chap_eta_old[2]) p->chap_eta_old[2][i - 1] = h[IX(i, ng + 1)] + b[IX(i, ng + 1)];
        if (p->chap_eta_old[3]) p->chap_eta_old[3][i - 1] = h[IX(i, nyt - ng)] + b[IX(i, nyt - ng)];
    }
}

/* apply_boundary_conditions, rki_boundary.F90:40-218: edges W, E, S, N in
 * that order; W/E span the full ny_total and S/N the full nx_total. */
void orc_apply_bc(const orc_params *p, double time, double *h, double *hu,
                  double *hv, double *b)
{
    orc_apply_bc_dt(p, time, 0.0, h, hu, hv, b);
}

void orc_apply_bc_dt(const orc_params *p, double time, double bc_dt, double *h,
                     double *hu, double *hv, double *b)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const int i_first = ng + 1, i_last = nxt - ng, j_first = ng + 1, j_last = nyt - ng;
    edge_t e[4] = {
        {nyt, i_first - 1, 1, -1, 0, 0, 1, -1.0, 1},
        {nyt, i_last + 1, 1, 1, 0, 0, 1, 1.0, 1},
        {nxt, 1, j_first - 1, 0, -1, 1, 0, -1.0, 0},
        {nxt, 1, j_last + 1, 0, 1, 1, 0, 1.0, 0},
    };
    double tidal_eta = orc_tidal_eta(p, time);
    for (int s = 0; s < 4; ++s) {
        switch (p->bc[s]) {
        case ORC_BC_WALL:
        case ORC_BC_SPONGE: bc_wall_edge(p, h, hu, hv, b, &e[s]); break;
        case ORC_BC_OPEN: bc_flather_edge(p, h, hu, hv, b, &e[s], 0.0, 0); break;
        case ORC_BC_TIDAL: bc_flather_edge(p, h, hu, hv, b, &e[s], tidal_eta, 1); break;
        case ORC_BC_INFLOW:
        case ORC_BC_DISCHARGE:
        case ORC_BC_CLAMPED: bc_simple_edge(p, h, hu, hv, b, &e[s], p->bc[s]); break;
        case ORC_BC_CHAPMAN:
            if (p->chap_eta_old[s])
                bc_chapman_edge(p, h, hu, hv, b, &e[s], p->chap_eta_old[s], bc_dt, s < 2 ? p->dx : p->dy);
            break;
        default: break; /* nested: not restated */
        }
    }
}

/* fill_bathymetry_ghosts_array, rki_bathymetry.F90:165-196 */
void orc_fill_bathymetry_ghosts(const orc_params *p, double *b)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng, nx = p->nx, ny = p->ny;
    for (int j = 1; j <= nyt; ++j)
        for (int i = 1; i <= ng; ++i) {
            b[IX(i, j)] = b[IX(ng + 1, j)];
            b[IX(ng + nx + i, j)] = b[IX(ng + nx, j)];
        }
    for (int j = 1; j <= ng; ++j)
        for (int i = 1; i <= nxt; ++i) {
            b[IX(i, j)] = b[IX(i, ng + 1)];
            b[IX(i, ng + ny + j)] = b[IX(i, ng + ny)];
        }
}

/* init_still_water, rki_init.F90:176-194 */
void orc_init_still_water(const orc_params *p, double eta0, const double *b,
                          double *h, double *hu, double *hv)
{
    size_t n = (size_t)NT(p) * (size_t)MT(p);
    for (size_t k = 0; k < n; ++k) {
        h[k] = dmax(eta0 - b[k], 0.0);
        hu[k] = 0.0;
        hv[k] = 0.0;
    }
}

/* compute_total_mass, rki_init.F90:287-317 (serial sum, interior only). */
double orc_total_mass(const orc_params *p, const double *h)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    const double area = p->dx * p->dy;
    double acc = 0.0;
    for (int j = ng + 1; j <= nyt - ng; ++j)
        for (int i = ng + 1; i <= nxt - ng; ++i) {
            double hh = h[IX(i, j)];
            if (hh > 0.0 && hh < 1.0e6) acc = acc + hh * area;
        }
    return acc;
}

static double interior_sum(const orc_params *p, const double *f)
{
    const int nxt = NT(p), nyt = MT(p), ng = p->ng;
    double acc = 0.0;
    for (int j = ng + 1; j <= nyt - ng; ++j)
        for (int i = ng + 1; i <= nxt - ng; ++i) acc = acc + f[IX(i, j)];
    return acc;
}

/* accumulate_sponge_mass_src, rki_solver.F90:608-657: the sponge, and the
 * change of the interior sum of h it made, weighted 0.5 per RK2 stage. */
static void sponge_with_budget(const orc_params *p, orc_clock *clk, double *h, double *hu,
                               double *hv, double dt, double area)
{
    if (p->sponge_width <= 0) return;   /* apply_sponge is a no-op then */
    const double before = interior_sum(p, h);
    orc_apply_sponge(p, h, hu, hv, dt);
    const double after = interior_sum(p, h);
    clk->mass_src = clk->mass_src + 0.5 * (after - before) * area;
}

/* ---------------------------------------------------------------------- */
/* One SSP-RK2 step: rki_solver.F90:659-815 (2-D barotropic fused path)   */
/* driven by rki_step / rki_advance (rki_api.F90:358-410).                */
/* ---------------------------------------------------------------------- */
long long orc_run(const orc_params *p, double *h, double *hu, double *hv,
                  double *b, orc_clock *clk, int mode, long long nsteps,
                  double target)
{
    if (p->ng < 2 || p->nx < 1 || p->ny < 1) return -2;
    size_t n = (size_t)NT(p) * (size_t)MT(p);
    double *work = (double *)calloc(6 * n, sizeof(double));
    if (!work) return -1;
    double *fh = work, *fhu = work + n, *fhv = work + 2 * n;
    double *h0 = work + 3 * n, *hu0 = work + 4 * n, *hv0 = work + 5 * n;
    const double area = p->dx * p->dy;
    /* Shu-Osher table, make_timestepper (rki_timestepper.F90:60-100) */
    const int generic = p->time_scheme != ORC_TIME_RK2;
    int n_stages = 2;
    double ssp_coeff = 1.0, alpha[4] = {0.0, 0.5, 0.0, 0.0}, beta[4] = {1.0, 1.0, 0.0, 0.0},
           cabs[4] = {0.0, 1.0, 0.0, 0.0};
    if (p->time_scheme == ORC_TIME_EULER) {
        n_stages = 1; alpha[0] = 0.0; beta[0] = 1.0; cabs[0] = 0.0;
    } else if (p->time_scheme == ORC_TIME_RK3) {
        n_stages = 3;
        alpha[0] = 0.0; alpha[1] = 0.75; alpha[2] = 1.0 / 3.0;
        beta[0] = beta[1] = beta[2] = 1.0;
        cabs[0] = 0.0; cabs[1] = 1.0; cabs[2] = 0.5;
    } else if (p->time_scheme == ORC_TIME_RK43) {
        n_stages = 4; ssp_coeff = 2.0;
        alpha[0] = 0.0; alpha[1] = 0.0; alpha[2] = 2.0 / 3.0; alpha[3] = 0.0;
        beta[0] = beta[1] = beta[2] = beta[3] = 0.5;
        cabs[0] = 0.0; cabs[1] = 0.5; cabs[2] = 1.0; cabs[3] = 0.5;
    } else if (p->time_scheme != ORC_TIME_RK2 && p->time_scheme != ORC_TIME_RK2_GENERIC) {
        free(work);
        return -2;
    }
    long long taken = 0;
    for (;;) {
        double dt_max;
        if (mode == 0) {
            if (taken >= nsteps) break;
            dt_max = p->t_end - clk->t;
        } else {
            if (!(clk->t < target - 1.0e-10)) break;
            dt_max = target - clk->t;
        }
        double time = clk->t;
        /* solver_step_impl :718-750; bc%dt still holds the previous step's
         * dt (set at :748), 0 before the first step */
        orc_apply_bc_dt(p, time, clk->dt, h, hu, hv, b);
        if (p->dt_fixed > 0.0) {
            clk->dt = dmin(p->dt_fixed, dt_max);
        } else if (clk->step_count % p->cfl_interval == 0 || clk->dt <= 0.0) {
            clk->dt = orc_compute_timestep(p, h, hu, hv);
            clk->dt = clk->dt * ssp_coeff; /* :736-741 */
            clk->dt = dmin(clk->dt, dt_max);
        }
        clk->step_count += 1;
        const double dt = clk->dt;
        if (generic) {
            /* solver_advance_2d_generic, rki_solver.F90:861-925 */
            memcpy(h0, h, n * sizeof(double));
            memcpy(hu0, hu, n * sizeof(double));
            memcpy(hv0, hv, n * sizeof(double));
            for (int st = 0; st < n_stages; ++st) {
                const double dt_stage = beta[st] * dt;
                orc_apply_bc_dt(p, time + cabs[st] * dt, dt, h, hu, hv, b);
                orc_compute_flux(p, h, hu, hv, b, fh, fhu, fhv, p->use_hllc);
                orc_fused_stage_generic(p, h, hu, hv, fh, fhu, fhv, dt_stage);
                if (alpha[st] != 0.0) orc_ssp_combine(p, h, hu, hv, h0, hu0, hv0, alpha[st]);
                orc_apply_sponge(p, h, hu, hv, dt_stage);   /* :919-920 */
            }
            clk->t = clk->t + dt;
            ++taken;
            continue;
        }
        /* stage 1 :763-787, bc%dt = state%dt (:748) */
        orc_apply_bc_dt(p, time, dt, h, hu, hv, b);
        orc_compute_flux(p, h, hu, hv, b, fh, fhu, fhv, p->use_hllc);
        clk->mass_out = clk->mass_out - 0.5 * dt * interior_sum(p, fh) * area;
        orc_fused_stage1(p, h, hu, hv, h0, hu0, hv0, fh, fhu, fhv, dt);
        if (p->forcing_active && p->precipitation)   /* accumulate_mass_src :569-601 */
            clk->mass_src = clk->mass_src + 0.5 * dt * interior_sum(p, p->precipitation) * area;
        sponge_with_budget(p, clk, h, hu, hv, dt, area);   /* :787 */
        /* stage 2 :789-814, bc%time = time + dt */
        double time2 = time + dt;
        orc_apply_bc_dt(p, time2, dt, h, hu, hv, b);
        orc_compute_flux(p, h, hu, hv, b, fh, fhu, fhv, p->use_hllc);
        clk->mass_out = clk->mass_out - 0.5 * dt * interior_sum(p, fh) * area;
        orc_fused_stage2(p, h, hu, hv, h0, hu0, hv0, fh, fhu, fhv, dt);
        if (p->forcing_active && p->precipitation)
            clk->mass_src = clk->mass_src + 0.5 * dt * interior_sum(p, p->precipitation) * area;
        sponge_with_budget(p, clk, h, hu, hv, dt, area);   /* :814 */
        /* rki_api.F90:378 */
        clk->t = clk->t + dt;
        ++taken;
    }
    free(work);
    return taken;
}
