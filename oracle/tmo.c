/*
 * tmo.c -- CPU oracle for the Tramonto nonlocal-density fill.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates the reference algorithm in
 * plain C so that the GPU product can be checked against it; it is linked only
 * into oracle/_build/libtmo.so, which only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load.  The product never calls it.
 *
 * Parity status of each function is recorded in its comment: "pinned by" names
 * the independent check in tests/test_oracle_*.py.
 */
#include "tmo.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <setjmp.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI 3.141592653589793238462643383279502884197169399375
#define NC TMO_NC

typedef struct {
    int L;
    int *off;     /* [L][3] */
    double *w;    /* [L] */
    double *hw;   /* [L][8]: the part of w from elements where the point is local node ln (HW_Weight) */
} sten_t;

struct tmo_ctx {
    tmo_params p;
    int N, nrb, nrbs, nunk, iu_rho, iu_rb, iu_mf, nmf;
    double *vext;            /* [N][C] or NULL */
    unsigned char *zero;     /* [N][C] */
    unsigned char *hwmask;   /* [C][N] or NULL: fluid elements of boundary nodes (0: not a boundary node) */
    double inv_4pi, inv_rad[NC], inv_4pir[NC], inv_4pirsq[NC];
    double rhobar_b[10], dphi_b[10], betamu_hs_ex[NC], betamu_att[NC], avdw[NC][NC], rmin[NC][NC];
    sten_t sten[3][NC * NC];
};

/* errors inside tmo_create unwind to it (it returns NULL); anywhere else they
 * are fatal, since the oracle has no partial results to return */
static jmp_buf g_create_env;
static int g_in_create = 0;
static char g_errmsg[256];
static tmo_ctx *volatile g_pending = NULL;

static void die(const char *msg)
{
    snprintf(g_errmsg, sizeof g_errmsg, "%s", msg);
    if (g_in_create) longjmp(g_create_env, 1);
    fprintf(stderr, "tmo oracle: %s\n", msg);
    exit(2);
}

const char *tmo_last_error(void) { return g_errmsg; }

int tmo_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------- */
/* Gauss-Legendre rules on [0,1], in the point order of set_gauss_quad
 * (dft_quadrature.c:40-159): positive roots of P_n in ascending order, then
 * their mirror images.  The 1/3/6/12-point rules use the reference's
 * truncated decimals (they differ from the exact rule by up to 4e-11); the
 * 20/40-point rules are computed here to full precision.                     */
static void gauss_exact(int n, double *xpos, double *wpos)
{
    /* Newton iteration on P_n from the Chebyshev-like initial guesses */
    int h = n / 2;
    for (int k = 1; k <= h; k++) {
        double x = cos(PI * (k - 0.25) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = x;
            for (int j = 2; j <= n; j++) {
                double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (fabs(dx) < 1e-17) break;
        }
        {
            double p0 = 1.0, p1 = x;
            for (int j = 2; j <= n; j++) {
                double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
        }
        /* k=1 is the largest root: store ascending */
        xpos[h - k] = x;
        wpos[h - k] = 2.0 / ((1.0 - x * x) * dp * dp);
    }
}

static void set_gauss(int n, double *gp, double *gw)
{
    static const double x12[6] = {0.125233408511469, 0.367831498998180, 0.587317954286617,
                                  0.769902674194305, 0.904117256370475, 0.981560634246719};
    static const double w12[6] = {0.249147045813403, 0.233492536538355, 0.203167426723066,
                                  0.160078328543346, 0.106939325995318, 0.047175336386512};
    static const double x6[3] = {0.238619186083197, 0.661209386466265, 0.932469514203152};
    static const double w6[3] = {0.467913934572691, 0.360761573048139, 0.171324492379170};
    double xp[20], wp[20];
    if (n == 1) {
        gp[0] = (0.0 + 1.0) / 2.0;
        gw[0] = 2.0 / 2.0;
    } else if (n == 3) {
        gp[0] = (0.7745966692 + 1.0) / 2.0;
        gw[0] = 0.555555555555556 / 2.0;
        gp[1] = (0.0 + 1.0) / 2.0;
        gw[1] = 0.888888888888888 / 2.0;
        gp[2] = (-0.7745966692 + 1.0) / 2.0;
        gw[2] = 0.555555555555556 / 2.0;
    } else if (n == 6 || n == 12) {
        const double *xs = (n == 6) ? x6 : x12, *ws = (n == 6) ? w6 : w12;
        int h = n / 2;
        for (int i = 0; i < h; i++) {
            gp[i] = (xs[i] + 1.0) / 2.0;
            gw[i] = ws[i] / 2.0;
            gp[h + i] = (-xs[i] + 1.0) / 2.0;
            gw[h + i] = ws[i] / 2.0;
        }
    } else if (n == 20 || n == 40) {
        /* computed once per rule (the tables are pure functions of n) */
        static double cx[2][20], cw[2][20];
        static int ready[2] = {0, 0};
        int slot = (n == 40), h = n / 2;
        if (!ready[slot]) {
#pragma omp critical(tmo_gauss)
            {
                if (!ready[slot]) {
                    gauss_exact(n, cx[slot], cw[slot]);
                    ready[slot] = 1;
                }
            }
        }
        for (int i = 0; i < h; i++) {
            xp[i] = cx[slot][i];
            wp[i] = cw[slot][i];
        }
        for (int i = 0; i < h; i++) {
            gp[i] = (xp[i] + 1.0) / 2.0;
            gw[i] = wp[i] / 2.0;
            gp[h + i] = (-xp[i] + 1.0) / 2.0;
            gw[h + i] = gw[i];
        }
    } else {
        die("unsupported Gauss rule");
    }
}

/* ------------------------------------------------------------------------- */
/* LJ 12-6 attraction, cut and shifted, with the WCA-style constant core
 * (dft_pairPot_LJ12_6.c:38-48 uLJ12_6_CS, :128-170 uLJ12_6_ATT_CS,
 * :174-205 uLJ12_6_ATT_noCS, :210-228 uLJ12_6_Integral; Type_attr 0 ->
 * ATTCORE_UMIN / CORECONST_UCONST, dft_defs_stencil.c:84-87).
 * Pinned by: Avdw vs the closed-form WCA volume and the reference BULK examples. */
static double lj_cs(double r, double sigma, double eps, double rcut)
{
    double u;
    if (r <= rcut)
        u = (pow(sigma / r, 12) - pow(sigma / rcut, 12)) - (pow(sigma / r, 6) - pow(sigma / rcut, 6));
    else
        u = 0.0;
    return 4.0 * eps * u;
}

static double u_att_cs(const tmo_ctx *c, double r, int i, int j)
{
    double sigma2 = c->p.sigma_ff[i][j] * c->p.sigma_ff[i][j];
    double sigma6 = sigma2 * sigma2 * sigma2;
    double rc_inv = 1.0 / c->p.cut_ff[i][j];
    double rc2_inv = rc_inv * rc_inv;
    double rc6_inv = rc2_inv * rc2_inv * rc2_inv;
    double rc12_inv = rc6_inv * rc6_inv;
    double r_min = c->rmin[i][j];
    double uatt;
    if (r <= c->p.cut_ff[i][j]) {
        if (r < r_min) r = r_min;
        double r_inv = 1.0 / r;
        double r2_inv = r_inv * r_inv;
        double r6_inv = r2_inv * r2_inv * r2_inv;
        double r12_inv = r6_inv * r6_inv;
        uatt = 4.0 * c->p.eps_ff[i][j] * sigma6 * (sigma6 * (r12_inv - rc12_inv) - (r6_inv - rc6_inv));
    } else {
        uatt = 0.0;
    }
    return uatt;
}

static double u_att_nocs(const tmo_ctx *c, double r, int i, int j)
{
    double sigma2 = c->p.sigma_ff[i][j] * c->p.sigma_ff[i][j];
    double sigma6 = sigma2 * sigma2 * sigma2;
    double r_min = c->rmin[i][j];
    if (r < r_min) r = r_min;
    double r_inv = 1.0 / r;
    double r2_inv = r_inv * r_inv;
    double r6_inv = r2_inv * r2_inv * r2_inv;
    double r12_inv = r6_inv * r6_inv;
    return 4.0 * c->p.eps_ff[i][j] * sigma6 * (sigma6 * r12_inv - r6_inv);
}

static double u_integral(const tmo_ctx *c, double r, int i, int j)
{
    double sigma2 = c->p.sigma_ff[i][j] * c->p.sigma_ff[i][j];
    double sigma6 = sigma2 * sigma2 * sigma2;
    double r_inv = 1.0 / r;
    double r3_inv = r_inv * r_inv * r_inv;
    double r9_inv = r3_inv * r3_inv * r3_inv;
    return 16 * PI * c->p.eps_ff[i][j] * sigma6 * (-sigma6 * r9_inv / 9.0 + r3_inv / 3.0);
}

/* grid search for the potential minimum, dft_pairPot.c:98-148 (pairPot_find_rmin).
 * Pinned by: |r_min - 2^(1/6) sigma| < 1e-3 sigma and u(r_min) <= u on a fine grid. */
static double find_rmin(double sigma, double eps, double rcut, double dx0)
{
    int updated = 0;
    double r, rmin = sigma, umin, umin_old, rstart, delr, tol = 1.e-8, error = 1.0, u, uleft, delr_old,
              rlast = 0.0;
    r = 2.0 * dx0;
    delr = dx0;
    umin = lj_cs(r, sigma, eps, rcut);
    while (r < rcut) {
        r += delr;
        u = lj_cs(r, sigma, eps, rcut);
        if (u < umin) {
            umin = u;
            rmin = r;
            updated = 1;
        }
        rlast = r;
    }
    if (!updated || fabs(rmin - rlast) < 1.e-6) return sigma;
    while (error > tol) {
        umin_old = umin;
        delr_old = delr;
        uleft = lj_cs(rmin - delr_old, sigma, eps, rcut);
        rstart = rmin - delr_old;
        r = rstart;
        umin = uleft;
        delr = delr_old / 100;
        while (r < rstart + 2.0 * delr_old) {
            r += delr;
            u = lj_cs(r, sigma, eps, rcut);
            if (u < umin) {
                umin = u;
                rmin = r;
            }
        }
        error = 100 * fabs(umin - umin_old) / fabs(umin);
    }
    return rmin;
}

/* analytic volume of the attraction stencil, dft_stencil_theta_uattr.c:41-63
 * (WCA-like branch: one constant in the core) */
static double uattr_volume(const tmo_ctx *c, int i, int j)
{
    double r_cut = c->p.cut_ff[i][j];
    double r_min = c->rmin[i][j];
    return (4.0 / 3.0) * PI * pow(r_min, 3.0) * u_att_nocs(c, r_min, i, j) -
           (4.0 / 3.0) * PI * pow(r_cut, 3.0) * u_att_nocs(c, r_cut, i, j) + u_integral(c, r_cut, i, j) -
           u_integral(c, r_min, i, j);
}

/* ------------------------------------------------------------------------- */
/* stencil construction, dft_stencil.c:40-287 (calc_stencils) for one zone.
 * Pinned by: sum of weights equals the analytic volume (pi d^2, pi d^3/6,
 * attraction volume), w(k) = w(-k), the 1D projections equal the closed forms
 * to quadrature accuracy, and the product library's independent builder.     */
static double sten_radius(const tmo_ctx *c, int isten, int i, int j)
{
    if (isten == TMO_UATT) return c->p.cut_ff[i][j];
    return c->p.hs_diam[i] / 2.0;
}

static double sten_volume(const tmo_ctx *c, int isten, int i, int j)
{
    if (isten == TMO_DELTA) return PI * pow(c->p.hs_diam[i], 2);
    if (isten == TMO_THETA) return PI * pow(c->p.hs_diam[i], 3) / 6.0;
    return uattr_volume(c, i, j);
}

/* the n-point rule on [0, 1] in set_gauss_quad's order (test access) */
void tmo_gauss_rule(int n, double *gp, double *gw) { set_gauss(n, gp, gw); }

/* Rule overrides for the stencil-pin tests only (tmo_set_stencil_rule): a changed interior Gauss
 * order, straddling-element sub-point count or 3D delta shell band must move the weights away from
 * their exact integrals.  Zero keeps the reference's value. */
static struct {
    int gauss, npt;
    double band;
} g_rule = {0, 0, 0.0};

void tmo_set_stencil_rule(int gauss_order, int npt_boundary, double band_factor)
{
    g_rule.gauss = gauss_order;
    g_rule.npt = npt_boundary;
    g_rule.band = band_factor;
}

static int quad_boundary(const tmo_ctx *c, int isten)
{
    int nd = c->p.ndim;
    if (g_rule.npt > 0 && isten != TMO_UATT) return g_rule.npt;
    if (isten == TMO_DELTA) return nd == 3 ? 40 : 150;   /* dft_stencil_delta_R.c:50-59 */
    if (isten == TMO_THETA) return nd == 3 ? 20 : 40;    /* dft_stencil_theta_R.c:46-55 */
    return nd == 3 ? 1 : 20;                             /* dft_stencil_theta_uattr.c:99-108 */
}

static int quad_gauss(int isten, double r)
{
    if (isten == TMO_UATT) return 1;
    if (g_rule.gauss > 0) return g_rule.gauss;
    return (r <= 4.000000001) ? 6 : 3;
}

static int quad_integrand(int isten, double r)
{
    if (isten != TMO_UATT) return 0;
    if (r <= 4.000000001) return 40;
    if (r <= 16.00000001) return 20;
    return 12;
}

/* dft_stencil_delta_R.c:70-78, theta_R.c:66-74, theta_uattr.c:121-157 */
static double get_weight(const tmo_ctx *c, int isten, int i, int j, double rsq, double R, int ngpu,
                         const double *gpu, const double *gwu)
{
    int nd = c->p.ndim;
    if (isten == TMO_DELTA) {
        if (nd == 1) return 2.0 * PI * R;
        if (nd == 2) return 2.0 / sqrt(1.0 - rsq);
        return 1.0;
    }
    if (isten == TMO_THETA) {
        if (nd == 1) return (1.0 - rsq) * PI * R * R;
        if (nd == 2) return sqrt(1.0 - rsq) * 2.0 * R;
        return 1.0;
    }
    double rc = c->p.cut_ff[i][j];
    double weight = 0.0, zmax = sqrt(1 - rsq), z, rho;
    if (nd == 1) {
        for (int q = 0; q < ngpu; q++) {
            z = zmax * gpu[q];
            rho = sqrt(rsq + z * z) * rc;
            weight += gwu[q] * z * u_att_cs(c, rho, i, j);
        }
        return 2.0 * PI * weight * rc * rc * zmax;
    }
    if (nd == 2) {
        for (int q = 0; q < ngpu; q++) {
            z = zmax * gpu[q];
            rho = sqrt(rsq + z * z) * rc;
            weight += gwu[q] * u_att_cs(c, rho, i, j);
        }
        return 2.0 * weight * rc * zmax;
    }
    rho = sqrt(rsq) * rc;
    return u_att_cs(c, rho, i, j);
}

/* dft_stencil.c:630-690 distribute_weight_to_nodes */
static void distribute(int nd, double weight, const double *qp, double *elw, int ig, int jg, int kg)
{
    if (nd == 3) {
        elw[0] += weight * (1.0 - qp[ig]) * (1.0 - qp[jg]) * (1.0 - qp[kg]);
        elw[1] += weight * qp[ig] * (1.0 - qp[jg]) * (1.0 - qp[kg]);
        elw[2] += weight * (1.0 - qp[ig]) * qp[jg] * (1.0 - qp[kg]);
        elw[3] += weight * qp[ig] * qp[jg] * (1.0 - qp[kg]);
        elw[4] += weight * (1.0 - qp[ig]) * (1.0 - qp[jg]) * qp[kg];
        elw[5] += weight * qp[ig] * (1.0 - qp[jg]) * qp[kg];
        elw[6] += weight * (1.0 - qp[ig]) * qp[jg] * qp[kg];
        elw[7] += weight * qp[ig] * qp[jg] * qp[kg];
    } else if (nd == 2) {
        elw[0] += weight * (1.0 - qp[ig]) * (1.0 - qp[jg]);
        elw[1] += weight * qp[ig] * (1.0 - qp[jg]);
        elw[2] += weight * (1.0 - qp[ig]) * qp[jg];
        elw[3] += weight * qp[ig] * qp[jg];
    } else {
        elw[0] += weight * (1.0 - qp[ig]);
        elw[1] += weight * qp[ig];
    }
}

/* dft_stencil.c:329-397 calc_sten_weight_el_in_radius */
static void el_in_radius(const tmo_ctx *c, int isten, double R, int i, int j, double rc2, const double *xl,
                         const double *es, double *elw)
{
    int nd = c->p.ndim;
    double point[3], weight, inv_r2 = 1.0 / (R * R), el_vol = 1.0, radius_sq;
    double gp[12], gw[12], gpu[40], gwu[40];
    for (int d = 0; d < nd; d++) el_vol *= es[d];
    int ngp = quad_gauss(isten, rc2);
    int ngpu = quad_integrand(isten, rc2);
    set_gauss(ngp, gp, gw);
    if (ngpu > 0) set_gauss(ngpu, gpu, gwu);
    for (int ig = 0; ig < ngp; ig++) {
        point[0] = xl[0] + gp[ig] * es[0];
        if (nd == 3) {
            for (int jg = 0; jg < ngp; jg++) {
                point[1] = xl[1] + gp[jg] * es[1];
                for (int kg = 0; kg < ngp; kg++) {
                    point[2] = xl[2] + gp[kg] * es[2];
                    radius_sq = (point[0] * point[0] + point[1] * point[1] + point[2] * point[2]) * inv_r2;
                    weight = get_weight(c, isten, i, j, radius_sq, R, ngpu, gpu, gwu);
                    weight *= gw[ig] * gw[jg] * gw[kg] * el_vol;
                    distribute(nd, weight, gp, elw, ig, jg, kg);
                }
            }
        } else if (nd == 2) {
            for (int jg = 0; jg < ngp; jg++) {
                point[1] = xl[1] + gp[jg] * es[1];
                radius_sq = (point[0] * point[0] + point[1] * point[1]) * inv_r2;
                weight = get_weight(c, isten, i, j, radius_sq, R, ngpu, gpu, gwu);
                weight *= gw[ig] * gw[jg] * el_vol;
                distribute(nd, weight, gp, elw, ig, jg, 0);
            }
        } else {
            radius_sq = point[0] * point[0] * inv_r2;
            weight = get_weight(c, isten, i, j, radius_sq, R, ngpu, gpu, gwu);
            weight *= gw[ig] * el_vol;
            distribute(nd, weight, gp, elw, ig, 0, 0);
        }
    }
}

/* dft_stencil.c:399-496 calc_sten_weight_el_on_boundary */
static void el_on_boundary(const tmo_ctx *c, int isten, double R, int i, int j, double rc2, const double *xl,
                           const double *es, double *elw)
{
    int nd = c->p.ndim;
    int delta = (isten == TMO_DELTA);
    double inv_r2 = 1.0 / (R * R), el_vol = 1.0, ok_distance = 0.0, point[3], radius_sq, tmp_sum = 0.0;
    double gpu[40], gwu[40], qp[3], weight;
    for (int d = 0; d < nd; d++) el_vol *= es[d];
    int npt = quad_boundary(c, isten);
    int ngpu = quad_integrand(isten, rc2);
    if (ngpu > 0) set_gauss(ngpu, gpu, gwu);
    double inv_npt = 1.0 / (double)npt;
    if (nd == 3 && delta) {
        for (int d = 0; d < nd; d++) tmp_sum += es[d] * es[d];
        ok_distance = 0.5 * inv_npt * sqrt(tmp_sum);
    }
    for (int ig = 0; ig < npt; ig++) {
        qp[0] = ((double)ig + 0.5) * inv_npt;
        point[0] = xl[0] + qp[0] * es[0];
        if (nd == 1) {
            radius_sq = point[0] * point[0] * inv_r2;
            if (radius_sq <= 1.000001) {
                weight = get_weight(c, isten, i, j, radius_sq, R, ngpu, gpu, gwu);
                if (fabs(radius_sq - 1.0) <= 1.e-8) weight *= 0.5;
                weight *= inv_npt * el_vol;
                distribute(nd, weight, qp, elw, 0, 1, 2);
            }
        } else if (nd == 2) {
            for (int jg = 0; jg < npt; jg++) {
                qp[1] = ((double)jg + 0.5) * inv_npt;
                point[1] = xl[1] + qp[1] * es[1];
                radius_sq = (point[0] * point[0] + point[1] * point[1]) * inv_r2;
                if (radius_sq <= 1.0000001) {
                    weight = get_weight(c, isten, i, j, radius_sq, R, ngpu, gpu, gwu);
                    if (fabs(radius_sq - 1.0) <= 1.e-8) weight *= 0.5;
                    weight *= inv_npt * inv_npt * el_vol;
                    distribute(nd, weight, qp, elw, 0, 1, 2);
                }
            }
        } else {
            for (int jg = 0; jg < npt; jg++) {
                qp[1] = ((double)jg + 0.5) * inv_npt;
                point[1] = xl[1] + qp[1] * es[1];
                for (int kg = 0; kg < npt; kg++) {
                    qp[2] = ((double)kg + 0.5) * inv_npt;
                    point[2] = xl[2] + qp[2] * es[2];
                    radius_sq = (point[0] * point[0] + point[1] * point[1] + point[2] * point[2]) * inv_r2;
                    const double band = g_rule.band > 0.0 ? g_rule.band : 2.0;
                    if (((radius_sq <= 1.00000001) && !delta) ||
                        (delta && (fabs(radius_sq - 1.0) < band * ok_distance))) {
                        weight = get_weight(c, isten, i, j, radius_sq, R, ngpu, gpu, gwu);
                        if (fabs(radius_sq - 1.0) <= 1.e-8 && !delta) weight *= 0.5;
                        weight *= (inv_npt * inv_npt * inv_npt) * el_vol;
                        distribute(nd, weight, qp, elw, 0, 1, 2);
                    }
                }
            }
        }
    }
}

/* dft_stencil.c:753-812 calc_in_out_on: 1 outside, -1 inside, 0 straddling */
static int in_out_on(int nd, const double *xl, const double *xr, double R)
{
    int in = 0, out = 0;
    int nc = 1 << nd;
    for (int corner = 0; corner < nc; corner++) {
        double s = 0.0;
        for (int d = 0; d < nd; d++) {
            double v = ((corner >> d) & 1) ? xr[d] : xl[d];
            s += v * v;
        }
        double r = (nd == 1) ? fabs(xl[0] * (corner ? 0 : 1) + xr[0] * (corner ? 1 : 0)) : sqrt(s);
        if (r > R) out = 1; else in = 1;
    }
    if (!in && out) return 1;
    if (in && !out) return -1;
    return 0;
}

static void build_stencil(tmo_ctx *c, int isten, int icomp, int jcomp)
{
    int nd = c->p.ndim;
    sten_t *s = &c->sten[isten][icomp + c->p.ncomp * jcomp];
    double vol = sten_volume(c, isten, icomp, jcomp);
    double R = sten_radius(c, isten, icomp, jcomp);
    int e[3] = {0, 0, 0}, nxs[3] = {1, 1, 1}, maxn = 1;
    for (int d = 0; d < nd; d++) {
        e[d] = (int)((float)R / (float)c->p.dx[d] + 0.999);
        nxs[d] = 2 * e[d] + 1;
        maxn *= nxs[d];
        /* ijk_to_isten_index (dft_stencil.c:622-651) folds offsets only when the
           stencil is wider than a periodic domain; that case is not restated */
        if (c->p.bc[d][0] == TMO_PERIODIC && nxs[d] > c->p.nx[d]) die("stencil wider than periodic domain");
    }
    s->L = 0;
    s->off = (int *)calloc((size_t)maxn * 3, sizeof(int));
    s->w = (double *)calloc((size_t)maxn, sizeof(double));
    s->hw = (double *)calloc((size_t)maxn * 8, sizeof(double));
    int *index = (int *)malloc((size_t)maxn * sizeof(int));
    for (int q = 0; q < maxn; q++) index[q] = -1;
    int nnode_el = 1 << nd;
    int i0 = -e[0], i1 = e[0];
    int j0 = (nd >= 2) ? -e[1] : 0, j1 = (nd >= 2) ? e[1] : 1;
    int k0 = (nd == 3) ? -e[2] : 0, k1 = (nd == 3) ? e[2] : 1;
    for (int i = i0; i < i1; i++)
        for (int j = j0; j < j1; j++)
            for (int k = k0; k < k1; k++) {
                int elo[8][3];
                /* setup_stencil_offsets, dft_stencil.c:289-327: local node bits (x, y, z) */
                for (int ln = 0; ln < nnode_el; ln++) {
                    elo[ln][0] = i + (ln & 1);
                    elo[ln][1] = j + ((ln >> 1) & 1);
                    elo[ln][2] = k + ((ln >> 2) & 1);
                }
                double xl[3], xr[3], xm[3], rc2 = 0.0, elw[8];
                int ijk[3] = {i, j, k};
                for (int d = 0; d < nd; d++) {
                    xl[d] = ijk[d] * c->p.dx[d];
                    xr[d] = (ijk[d] + 1) * c->p.dx[d];
                }
                for (int d = 0; d < nd; d++) xm[d] = xl[d] + 0.5 * c->p.dx[d];
                for (int d = 0; d < nd; d++) rc2 += xm[d] * xm[d];
                for (int q = 0; q < nnode_el; q++) elw[q] = 0.0;
                int flag = in_out_on(nd, xl, xr, R);
                if (flag == 1 || (flag == -1 && nd == 3 && isten == TMO_DELTA)) continue;
                if (flag == -1)
                    el_in_radius(c, isten, R, icomp, jcomp, rc2, xl, c->p.dx, elw);
                else
                    el_on_boundary(c, isten, R, icomp, jcomp, rc2, xl, c->p.dx, elw);
                /* merge_load_stencil, dft_stencil.c:570-620 */
                for (int ln = 0; ln < nnode_el; ln++) {
                    if (elw[ln] == 0.0) continue;
                    int id = 0, mul = 1;
                    for (int d = 0; d < nd; d++) {
                        id += (elo[ln][d] + e[d]) * mul;
                        mul *= nxs[d];
                    }
                    if (index[id] == -1) {
                        s->w[s->L] = elw[ln];
                        s->hw[8 * s->L + ln] = elw[ln];
                        for (int d = 0; d < 3; d++) s->off[3 * s->L + d] = (d < nd) ? elo[ln][d] : 0;
                        index[id] = s->L;
                        s->L++;
                    } else {
                        s->w[index[id]] += elw[ln];
                        s->hw[8 * index[id] + ln] += elw[ln];
                    }
                }
            }
    free(index);
    /* renormalize_stencil, dft_stencil.c:653-683 */
    double sum = 0.0;
    for (int q = 0; q < s->L; q++) sum += s->w[q];
    for (int q = 0; q < s->L; q++)
        if (s->w[q] != s->w[q]) die("NaN stencil weight");
    if (sum != vol) {
        double ratio = vol / sum;
        for (int q = 0; q < s->L; q++) s->w[q] *= ratio;
        for (int q = 0; q < 8 * s->L; q++) s->hw[q] *= ratio;   /* dft_stencil.c:700-706 */
    }
}

/* HW_boundary_weight (dft_utils.c:41-62): the weight of stencil point k at a wall-fluid boundary
 * node is the sum of its parts hw[ln] over the local nodes ln whose element is fluid.  The element
 * is node_box_to_elem_box_reflect's (dft_mesh_lib.c:896-937): where offset_to_node_box reflected the
 * point in dimension d (refl[d]), the element on the mirror side, i.e. local node ln with bit d
 * flipped in the node's own element numbering. */
static double hw_weight(const tmo_ctx *c, const sten_t *s, int k, int jcomp, int jnode, const int *refl, double w)
{
    if (!c->hwmask || jnode < 0) return w;
    const unsigned m = c->hwmask[(size_t)jcomp * c->N + jnode];
    if (!m) return w;
    unsigned flip = 0;
    for (int d = 0; d < c->p.ndim; d++)
        if (refl[d]) flip |= 1u << d;
    double v = 0.0;
    for (int ln = 0; ln < (1 << c->p.ndim); ln++)
        if (m & (1u << (ln ^ flip))) v += s->hw[8 * k + ln];
    return v;
}

void tmo_set_hw_masks(tmo_ctx *c, const unsigned char *mask)
{
    if (!c->p.lhard_surf) die("hard-surface masks need lhard_surf");
    free(c->hwmask);
    c->hwmask = NULL;
    if (!mask) return;
    c->hwmask = (unsigned char *)malloc((size_t)c->p.ncomp * c->N);
    memcpy(c->hwmask, mask, (size_t)c->p.ncomp * c->N);
}

/* ------------------------------------------------------------------------- */
/* FMT physics in n order (n0 n1 n2 n3 nv1[D] nv2[D]).
 * FMT1: dft_physics_FMT1.c:39-75 (Phi, first derivatives), :79-168 (second).
 * FMT3: dft_physics_FMT3.c:39-96, :101-266.
 * FMT2: dft_physics_FMT2.c:40-250.  FMT4: dft_physics_FMT4.c:38-276.
 * Pinned by: finite differences of Phi (first derivatives) and of the first
 * derivatives (second), bulk pressure = Percus-Yevick (FMT1) and
 * Carnahan-Starling (FMT3), exact hard-sphere limits.                        */
double tmo_fmt_phi(const tmo_ctx *c, const double *n)
{
    int nd = c->p.ndim;
    double d12 = 0.0, d22 = 0.0;
    for (int d = 0; d < nd; d++) {
        d12 += n[4 + d] * n[4 + nd + d];
        d22 += n[4 + nd + d] * n[4 + nd + d];
    }
    double p1 = -n[0] * log(1.0 - n[3]);
    double p2 = (n[1] * n[2] - d12) / (1.0 - n[3]);
    double p3;
    if (c->p.type_func == TMO_FMT1) {
        p3 = (n[2] * n[2] * n[2] - 3.0 * n[2] * d22) / (24.0 * PI * (1.0 - n[3]) * (1.0 - n[3]));
    } else if (c->p.type_func == TMO_FMT2) {
        /* FMT2_energy_density, dft_physics_FMT2.c:40-58; (n2 - nv2.nv2/n2)^3 read as n2^3 where
           n2 <= 1e-15, as its derivatives do (0/0 in the reference) */
        const double al = n[2] > 1.e-15 ? n[2] - d22 / n[2] : n[2];
        p3 = al * al * al / (24.0 * PI * (1.0 - n[3]) * (1.0 - n[3]));
    } else if (c->p.type_func == TMO_FMT4) {
        /* FMT4_energy_density, dft_physics_FMT4.c:38-58 (Oleksy and Hansen, Mol. Phys. 104, 2871) */
        const double f = n[2] > 1.e-15 ? 1.0 - d22 / (n[2] * n[2]) : 1.0;
        p3 = n[2] * n[2] * n[2] * f * f * f * (n[3] + (1.0 - n[3]) * (1.0 - n[3]) * log(1.0 - n[3])) /
             (36.0 * PI * n[3] * n[3] * (1.0 - n[3]) * (1.0 - n[3]));
    } else {
        p3 = (n[2] * n[2] * n[2] - 3.0 * n[2] * d22) * (n[3] + (1.0 - n[3]) * (1.0 - n[3]) * log(1.0 - n[3])) /
             (36.0 * PI * n[3] * n[3] * (1.0 - n[3]) * (1.0 - n[3]));
    }
    return p1 + p2 + p3;
}

static void dphi_loc(const tmo_ctx *c, const double *n, double d12, double d22, const double *inv, double *out)
{
    int nd = c->p.ndim;
    if (c->p.type_func == TMO_FMT1) {
        out[0] = log(inv[1]);
        out[1] = n[2] * inv[1];
        out[2] = n[1] * inv[1] + (n[2] * n[2] - d22) * inv[2] / (8.0 * PI);
        out[3] = n[0] * inv[1] + (n[1] * n[2] - d12) * inv[2] + (n[2] * n[2] * n[2] - 3. * n[2] * d22) * inv[3] / (12.0 * PI);
        for (int d = 0; d < nd; d++) {
            out[4 + d] = -n[4 + nd + d] * inv[1];
            out[4 + nd + d] = -n[4 + d] * inv[1] - n[2] * n[4 + nd + d] * inv[2] / (4.0 * PI);
        }
        return;
    }
    if (c->p.type_func == TMO_FMT2) {
        /* FMT2_1stderiv, dft_physics_FMT2.c:60-100 */
        out[0] = log(inv[1]);
        out[1] = n[2] * inv[1];
        out[2] = n[1] * inv[1];
        out[3] = n[0] * inv[1] + (n[1] * n[2] - d12) * inv[2];
        for (int d = 0; d < nd; d++) {
            out[4 + d] = -n[4 + nd + d] * inv[1];
            out[4 + nd + d] = -n[4 + d] * inv[1];
        }
        double alpha, beta, gamma[3];
        if (n[2] > 1.e-15) {
            alpha = n[2] - d22 / n[2];
            beta = 1.0 + d22 / (n[2] * n[2]);
            for (int d = 0; d < nd; d++) gamma[d] = n[4 + nd + d] / n[2];
        } else {
            alpha = n[2];
            beta = 1.0;
            for (int d = 0; d < nd; d++) gamma[d] = 0.0;
        }
        const double asq = alpha * alpha, acb = asq * alpha;
        out[2] += asq * beta * inv[2] / (8.0 * PI);
        out[3] += acb * inv[3] / (12.0 * PI);
        for (int d = 0; d < nd; d++) out[4 + nd + d] -= inv[2] * asq * gamma[d] / (4.0 * PI);
        return;
    }
    if (c->p.type_func == TMO_FMT4) {
        /* FMT4_1stderiv, dft_physics_FMT4.c:60-96; the n2 <= 1e-15 guard of fac is this oracle's (the
           reference divides by n2^2 there) */
        out[0] = log(inv[1]);
        out[1] = n[2] * inv[1];
        double fac = 0.0, fac2 = 0.0;
        if (n[3] > 1.e-10) {
            fac = n[2] > 1.e-15 ? 1 - d22 / (n[2] * n[2]) : 1.0;
            const double fac1 = n[3] - 2. * (1. - n[3]) * log(1. - n[3]);
            fac2 = (n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3]));
            out[2] = n[1] * inv[1] + (n[2] * n[2] * fac * fac * fac + 2 * fac * fac * d22) * inv[2] * fac2 /
                                         (12.0 * PI * n[3] * n[3]);
            out[3] = n[0] * inv[1] + (n[1] * n[2] - d12) * inv[2] +
                     fac1 * n[2] * n[2] * n[2] * fac * fac * fac * inv[2] / (36.0 * PI * n[3] * n[3]) +
                     fac2 * n[2] * n[2] * n[2] * fac * fac * fac *
                         (-inv[2] / (n[3] * n[3] * n[3]) + inv[3] / (n[3] * n[3])) / (18.0 * PI);
        } else {
            out[2] = 0.0;
            out[3] = 0.0;
        }
        for (int d = 0; d < nd; d++) {
            const int i1 = 4 + d, i2 = 4 + nd + d;
            out[i1] = -n[i2] * inv[1];
            if (n[3] > 1.e-10)
                out[i2] = -n[i1] * inv[1] - n[2] * n[i2] * fac2 * inv[2] * fac * fac / (6.0 * PI * n[3] * n[3]);
            else
                out[i2] = 0.0;
        }
        return;
    }
    /* FMT3 (White Bear) */
    out[0] = log(inv[1]);
    out[1] = n[2] * inv[1];
    if (n[3] > 1.e-10) {
        out[2] = n[1] * inv[1] + (n[2] * n[2] - d22) * inv[2] * (n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3])) /
                                     (12.0 * PI * n[3] * n[3]);
        out[3] = n[0] * inv[1] + (n[1] * n[2] - d12) * inv[2] +
                 (((n[2] * n[2] * n[2] - 3.0 * n[2] * d22) * inv[2]) / (36. * PI * n[3] * n[3])) *
                     (+2. * (n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3])) * inv[1] +
                      (-2. * (1. - n[3]) * log(1. - n[3]) + n[3]) -
                      2. * ((n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3])) / n[3]));
    } else {
        out[2] = 0.0;
        out[3] = 0.0;
    }
    for (int d = 0; d < nd; d++) {
        int i1 = 4 + d, i2 = 4 + nd + d;
        out[i1] = -n[i2] * inv[1];
        if (n[3] > 1.e-10)
            out[i2] = -n[i1] * inv[1] -
                      n[2] * n[i2] * inv[2] * (n[3] + (1.0 - n[3]) * (1.0 - n[3]) * log(1.0 - n[3])) /
                          (6.0 * PI * n[3] * n[3]);
        else
            out[i2] = 0.0;
    }
}

/* calc_FMT_derivatives, dft_fill_FMT.c:349-385 (takes n order) */
void tmo_fmt_dphi(const tmo_ctx *c, const double *n, double *dphi)
{
    int nd = c->p.ndim;
    double inv[5];
    inv[0] = (1.0 - n[3]);
    inv[1] = 1.0 / inv[0];
    inv[2] = inv[1] * inv[1];
    inv[3] = inv[2] * inv[1];
    inv[4] = inv[3] * inv[1];
    double d22 = 0.0, d12 = 0.0;
    for (int d = 0; d < nd; d++) {
        d22 += n[4 + nd + d] * n[4 + nd + d];
        d12 += n[4 + d] * n[4 + nd + d];
    }
    dphi_loc(c, n, d12, d22, inv, dphi);
}

/* RB_Struct in n order: out[0..3]=S0..S3, out[4+d]=V1[d], out[4+D+d]=V2[d] */
void tmo_fmt_d2_delta(const tmo_ctx *c, const double *n, const int *offset, const double *sign, int icomp,
                      double *out)
{
    int nd = c->p.ndim;
    double vec[3];
    for (int d = 0; d < nd; d++) vec[d] = offset[d] * c->p.dx[d] * c->inv_rad[icomp];
    const double i4pi = c->inv_4pi, i4pir = c->inv_4pir[icomp], i4pirsq = c->inv_4pirsq[icomp];
    if (c->p.type_func == TMO_FMT1) {
        double n1v[3], n2v[3];
        for (int d = 0; d < nd; d++) {
            n2v[d] = sign[d] * n[4 + nd + d];
            n1v[d] = sign[d] * n[4 + d];
        }
        double a = 1.0 / (1.0 - n[3]), a2 = a * a, a3 = a2 * a;
        double S2 = (a * i4pir + n[2] * i4pi * a2);
        for (int d = 0; d < nd; d++) S2 += vec[d] * n2v[d] * i4pi * a2;
        double S3 = (i4pirsq * a + n[2] * i4pir * a2 + n[1] * a2 + n[2] * n[2] * i4pi * a3);
        for (int d = 0; d < nd; d++)
            S3 += a2 * (-n2v[d] * n2v[d] * i4pi * a + vec[d] * (n2v[d] * i4pir + n1v[d] + n[2] * n2v[d] * a / (2 * PI)));
        out[0] = 0.0;
        out[1] = a;
        out[2] = S2;
        out[3] = S3;
        for (int d = 0; d < nd; d++) out[4 + d] = sign[d] * a * vec[d];
        for (int d = 0; d < nd; d++)
            out[4 + nd + d] = sign[d] * (-n2v[d] * i4pi * a2 - (-i4pir * a - n[2] * i4pi * a2) * vec[d]);
        return;
    }
    if (c->p.type_func == TMO_FMT2) {
        /* d2phi_drb2_delta_rb_FMT2, dft_physics_FMT2.c:106-184 */
        double n1v[3], n2v[3], gam[3], d22 = 0.0, dg = 0.0, alpha, beta, eps;
        for (int d = 0; d < nd; d++) {
            n2v[d] = sign[d] * n[4 + nd + d];
            n1v[d] = sign[d] * n[4 + d];
            d22 += n2v[d] * n2v[d];
        }
        if (n[2] > 1.e-15) {
            alpha = n[2] - d22 / n[2];
            beta = 1.0 + d22 / (n[2] * n[2]);
            for (int d = 0; d < nd; d++) {
                gam[d] = n2v[d] / n[2];
                dg += gam[d] * gam[d];
            }
            eps = alpha / n[2];
        } else {
            alpha = n[2];
            beta = 1.0;
            for (int d = 0; d < nd; d++) gam[d] = 0.0;
            eps = 1.0;
        }
        const double asq = alpha * alpha;
        const double a = 1.0 / (1.0 - n[3]), a2 = a * a, a3 = a2 * a;
        double S2 = a * i4pir + i4pi * a2 * alpha * (beta * beta - eps * dg);
        for (int d = 0; d < nd; d++) S2 -= i4pi * a2 * alpha * (eps - 2 * beta) * gam[d] * vec[d];
        double S3 = (i4pirsq * a + n[2] * i4pir * a2 + n[1] * a2);
        for (int d = 0; d < nd; d++) S3 += a2 * vec[d] * (n2v[d] * i4pir + n1v[d]);
        S3 += i4pi * a3 * asq * beta;
        for (int d = 0; d < nd; d++) S3 += 2.0 * i4pi * a3 * asq * gam[d] * vec[d];
        out[0] = 0.0;
        out[1] = a;
        out[2] = S2;
        out[3] = S3;
        for (int d = 0; d < nd; d++) {
            out[4 + d] = sign[d] * a * vec[d];
            out[4 + nd + d] = sign[d] * i4pir * a * vec[d];
        }
        for (int d = 0; d < nd; d++) {
            out[4 + nd + d] += sign[d] * i4pi * a2 * alpha * (eps - 2.0 * beta) * gam[d];
            out[4 + nd + d] -= sign[d] * i4pi * a2 * alpha * (4.0 * dg - eps) * vec[d];
        }
        return;
    }
    if (c->p.type_func == TMO_FMT4) {
        /* d2phi_drb2_delta_rb_FMT4, dft_physics_FMT4.c:101-191 (the vector densities enter unsigned
           except through sign[] on V1 and V2, as there) */
        double d22 = 0.0, d12 = 0.0;
        for (int d = 0; d < nd; d++) {
            d22 += n[4 + nd + d] * n[4 + nd + d];
            d12 += n[4 + d] * n[4 + nd + d];
        }
        double inv[5];
        inv[0] = (1. - n[3]);
        inv[1] = 1.0 / (1.0 - n[3]);
        inv[2] = inv[1] * inv[1];
        inv[3] = inv[2] * inv[1];
        inv[4] = inv[3] * inv[1];
        const double n3sq = n[3] * n[3], n3cb = n3sq * n[3], n2sq = n[2] * n[2], n2cb = n2sq * n[2];
        if (n[3] > 1.e-10) {
            const double fac1 = n[3] - 2 * (1. - n[3]) * log(1. - n[3]);
            const double fac2 = n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3]);
            const double fac3 = -(inv[2] / n3cb) + (inv[3] / n3sq);
            const double fac = n[2] > 1.e-15 ? 1 - d22 / (n[2] * n[2]) : 1.0;
            const double i2 = n[2] > 1.e-15 ? 1.0 / n[2] : 0.0;   /* 1/n2, 1/n2^2, 1/n2^3 of the terms in d22 */
            double S2 = (inv[1] * i4pir +
                         fac2 * inv[2] * (n[2] * fac * fac * fac + 3 * d22 * fac * fac * i2 + 4 * d22 * d22 * fac * i2 * i2 * i2) /
                             (6. * PI * n3sq));
            for (int d = 0; d < nd; d++)
                S2 += vec[d] * n[4 + nd + d] * (fac2 * inv[2] / (6. * PI * n3sq)) * (fac * fac + 4 * d22 * fac * i2 * i2);
            double S3 = (i4pirsq * inv[1] + i4pir * n[2] * inv[2] + n[1] * inv[2] +
                         (n2sq * fac * fac * fac + 2 * d22 * fac * fac) * ((2. * fac2 * fac3) + (fac1 * inv[2] / n3sq)) / (12. * PI));
            for (int d = 0; d < nd; d++)
                S3 += vec[d] * (n[4 + nd + d] * i4pir * inv[2] + n[4 + d] * inv[2] +
                                n[2] * n[4 + nd + d] * fac * fac * ((2. * fac2 * fac3) + (fac1 * inv[2] / n3sq)) / (6. * PI));
            out[0] = 0.0;
            out[1] = inv[1];
            out[2] = S2;
            out[3] = S3;
            for (int d = 0; d < nd; d++) out[4 + d] = sign[d] * (inv[1]) * vec[d];
            double dnw = 0.0;
            for (int d = 0; d < nd; d++) dnw += n[4 + nd + d] * vec[d];
            for (int d = 0; d < nd; d++)
                out[4 + nd + d] = sign[d] * (-(n[4 + nd + d] * inv[2] * fac2 / (6. * PI * n3sq)) *
                                                 (fac * fac - 4 * d22 * fac * i2 + 4 * dnw * fac * i2) -
                                             (-i4pir * inv[1] - n[2] * fac * fac * fac2 * inv[2] / (6. * PI * n3sq)) * vec[d]);
        } else {
            for (int q = 0; q < 4 + 2 * nd; q++) out[q] = 0.0;
        }
        (void)n2cb;
        (void)d12;
        return;
    }
    /* FMT3 */
    double d22 = 0.0, d12 = 0.0;
    for (int d = 0; d < nd; d++) {
        d22 += n[4 + nd + d] * n[4 + nd + d];
        d12 += n[4 + d] * n[4 + nd + d];
    }
    double inv[5];
    inv[0] = (1. - n[3]);
    inv[1] = 1.0 / (1.0 - n[3]);
    inv[2] = inv[1] * inv[1];
    inv[3] = inv[2] * inv[1];
    inv[4] = inv[3] * inv[1];
    double n3sq = n[3] * n[3], n3cb = n3sq * n[3], n2sq = n[2] * n[2];
    if (n[3] > 1.e-10) {
        double fac1 = n[3] - 2 * (1. - n[3]) * log(1. - n[3]);
        double fac2 = n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3]);
        double fac3 = -(inv[2] / n3cb) + (inv[3] / n3sq);
        double S2 = (inv[1] * i4pir + n[2] * fac2 * inv[2] / (6. * PI * n3sq));
        for (int d = 0; d < nd; d++) S2 += vec[d] * n[4 + nd + d] * fac2 * inv[2] / (6. * PI * n3sq);
        double S3 = (i4pirsq * inv[1] + i4pir * n[2] * inv[2] + n[1] * inv[2] +
                     (n2sq - d22) * ((2. * fac2 * fac3) + (fac1 * inv[2] / n3sq)) / (12. * PI));
        for (int d = 0; d < nd; d++)
            S3 += vec[d] * (n[4 + nd + d] * i4pir * inv[2] + n[4 + d] * inv[2] +
                            n[2] * n[4 + nd + d] * ((2. * fac2 * fac3) + (fac1 * inv[2] / n3sq)) / (6. * PI));
        out[0] = 0.0;
        out[1] = inv[1];
        out[2] = S2;
        out[3] = S3;
        for (int d = 0; d < nd; d++) out[4 + d] = sign[d] * (inv[1]) * vec[d];
        for (int d = 0; d < nd; d++)
            out[4 + nd + d] = sign[d] * (-n[4 + nd + d] * inv[2] * fac2 / (6. * PI * n3sq) -
                                         (-i4pir * inv[1] - n[2] * fac2 * inv[2] / (6. * PI * n3sq)) * vec[d]);
    } else {
        for (int q = 0; q < 4 + 2 * nd; q++) out[q] = 0.0;
    }
}

void tmo_fmt_d2_theta(const tmo_ctx *c, const double *n, double *out)
{
    int nd = c->p.ndim;
    const double i4pi = c->inv_4pi;
    if (c->p.type_func == TMO_FMT1) {
        double a = 1.0 / (1.0 - n[3]), a2 = a * a, a3 = a2 * a, a4 = a3 * a;
        double S2 = (n[1] * a2 + n[2] * n[2] * i4pi * a3);
        for (int d = 0; d < nd; d++) S2 += -n[4 + nd + d] * n[4 + nd + d] * i4pi * a3;
        double S3 = n[0] * a2 + 2 * n[1] * n[2] * a3 + n[2] * n[2] * n[2] * i4pi * a4;
        for (int d = 0; d < nd; d++)
            S3 += -(2 * n[4 + d] * n[4 + nd + d] * a3 + 3 * n[2] * n[4 + nd + d] * n[4 + nd + d] * i4pi * a4);
        out[0] = a;
        out[1] = n[2] * a2;
        out[2] = S2;
        out[3] = S3;
        for (int d = 0; d < nd; d++) out[4 + d] = -n[4 + nd + d] * a2;
        for (int d = 0; d < nd; d++) out[4 + nd + d] = -(n[4 + d] * a2 + n[2] * n[4 + nd + d] * a3 / (2.0 * PI));
        return;
    }
    if (c->p.type_func == TMO_FMT2) {
        /* d2phi_drb2_theta_rb_FMT2, dft_physics_FMT2.c:189-250 */
        double d22 = 0.0, d12 = 0.0, alpha, beta, gam[3];
        for (int d = 0; d < nd; d++) {
            d22 += n[4 + nd + d] * n[4 + nd + d];
            d12 += n[4 + d] * n[4 + nd + d];
        }
        if (n[2] > 1.e-15) {
            alpha = n[2] - d22 / n[2];
            beta = 1.0 + d22 / (n[2] * n[2]);
            for (int d = 0; d < nd; d++) gam[d] = n[4 + nd + d] / n[2];
        } else {
            alpha = n[2];
            beta = 1.0;
            for (int d = 0; d < nd; d++) gam[d] = 0.0;
        }
        const double a = 1.0 / (1.0 - n[3]), a2 = a * a, a3 = a2 * a, a4 = a3 * a;
        const double asq = alpha * alpha, acb = asq * alpha;
        out[2] = n[1] * a2 + i4pi * a3 * asq * beta;
        out[3] = (n[0] * a2 + 2.0 * (n[1] * n[2] - d12) * a3) + i4pi * a4 * acb;
        out[0] = a;
        out[1] = n[2] * a2;
        for (int d = 0; d < nd; d++) {
            out[4 + d] = -n[4 + nd + d] * a2;
            out[4 + nd + d] = -n[4 + d] * a2 - 2.0 * i4pi * a3 * asq * gam[d];
        }
        return;
    }
    if (c->p.type_func == TMO_FMT4) {
        /* d2phi_drb2_theta_rb_FMT4, dft_physics_FMT4.c:196-276 */
        double d22 = 0.0, d12 = 0.0;
        for (int d = 0; d < nd; d++) {
            d22 += n[4 + nd + d] * n[4 + nd + d];
            d12 += n[4 + d] * n[4 + nd + d];
        }
        double inv[5];
        inv[0] = (1. - n[3]);
        inv[1] = 1.0 / (1.0 - n[3]);
        inv[2] = inv[1] * inv[1];
        inv[3] = inv[2] * inv[1];
        inv[4] = inv[3] * inv[1];
        const double n3sq = n[3] * n[3], n3cb = n3sq * n[3], n3_4 = n3cb * n[3], n2sq = n[2] * n[2], n2cb = n2sq * n[2];
        if (n[3] > 1.e-10) {
            const double fac1 = n[3] - 2 * (1. - n[3]) * log(1. - n[3]);
            const double fac2 = n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3]);
            const double fac3 = -(inv[2] / n3cb) + (inv[3] / n3sq);
            const double fac = n[2] > 1.e-15 ? 1 - d22 / (n[2] * n[2]) : 1.0;
            out[2] = (n[1] * inv[2] + (n2sq * fac * fac * fac + 2 * fac * fac * d22) * (fac1 * inv[2] / n3sq + 2. * fac2 * fac3) / (12 * PI));
            out[3] = (n[0] * inv[2] + 2 * (n[1] * n[2] - d12) * inv[3] +
                      ((n2cb * fac * fac * fac) / (36. * PI)) *
                          (+(((2. * log(1. - n[3]) + 3) * inv[2]) / n3sq) + (4. * fac1 * fac3) +
                           (fac2 * (6. * inv[2] / n3_4 + 6. * inv[4] / n3sq - 8. * inv[3] / n3cb))));
            out[0] = inv[1];
            out[1] = n[2] * inv[2];
            for (int d = 0; d < nd; d++) out[4 + d] = -n[4 + nd + d] * inv[2];
            for (int d = 0; d < nd; d++)
                out[4 + nd + d] = -(n[4 + d] * inv[2] + n[2] * n[4 + nd + d] * fac * fac * ((fac1 * inv[2] / n3sq) + (2. * fac2 * fac3)) / (6. * PI));
        } else {
            for (int q = 0; q < 4 + 2 * nd; q++) out[q] = 0.0;
        }
        return;
    }
    double d22 = 0.0, d12 = 0.0;
    for (int d = 0; d < nd; d++) {
        d22 += n[4 + nd + d] * n[4 + nd + d];
        d12 += n[4 + d] * n[4 + nd + d];
    }
    double inv[5];
    inv[0] = (1. - n[3]);
    inv[1] = 1.0 / (1.0 - n[3]);
    inv[2] = inv[1] * inv[1];
    inv[3] = inv[2] * inv[1];
    inv[4] = inv[3] * inv[1];
    double n3sq = n[3] * n[3], n3cb = n3sq * n[3], n3_4 = n3cb * n[3], n2sq = n[2] * n[2], n2cb = n2sq * n[2];
    if (n[3] > 1.e-10) {
        double fac1 = n[3] - 2 * (1. - n[3]) * log(1. - n[3]);
        double fac2 = n[3] + (1. - n[3]) * (1. - n[3]) * log(1. - n[3]);
        double fac3 = -(inv[2] / n3cb) + (inv[3] / n3sq);
        out[2] = (n[1] * inv[2] + (n[2] * n[2] - d22) * (fac1 * inv[2] / n3sq + 2. * fac2 * fac3) / (12 * PI));
        out[3] = (n[0] * inv[2] + 2 * (n[1] * n[2] - d12) * inv[3] +
                  ((n2cb - 3. * n[2] * d22) / (36. * PI)) *
                      (+(((2. * log(1. - n[3]) + 3) * inv[2]) / n3sq) + (4. * fac1 * fac3) +
                       (fac2 * (6. * inv[2] / n3_4 + 6. * inv[4] / n3sq - 8. * inv[3] / n3cb))));
        out[0] = inv[1];
        out[1] = n[2] * inv[2];
        for (int d = 0; d < nd; d++) out[4 + d] = -n[4 + nd + d] * inv[2];
        for (int d = 0; d < nd; d++)
            out[4 + nd + d] = -(n[4 + d] * inv[2] + n[2] * n[4 + nd + d] * ((fac1 * inv[2] / n3sq) + (2. * fac2 * fac3)) / (6. * PI));
    } else {
        for (int q = 0; q < 4 + 2 * nd; q++) out[q] = 0.0;
    }
}

/* solutionVec_to_nOrdering, dft_utils.c:205-221 */
static void sv_to_n(int nd, const double *rb, double *n)
{
    n[0] = rb[3];
    n[1] = rb[2];
    n[2] = rb[1];
    n[3] = rb[0];
    for (int d = 0; d < nd; d++) {
        n[4 + d] = rb[4 + nd + d];
        n[4 + nd + d] = rb[4 + d];
    }
}

/* ------------------------------------------------------------------------- */
/* bulk constants, dft_thermo_hs.c:58-75 (calc_InvR_params), :226-334
 * (compute_bulk_FMT_properties, rhobar_icomp, dphi_drb_bulk), :136-154
 * (chempot_FMT_hs); dft_thermo_att.c:44-61 (calc_Avdw_att), :86-97 (chempot_att).
 * Pinned by: FMT1 bulk pressure/chemical potential = Percus-Yevick closed form,
 * FMT3 = Carnahan-Starling/BMCSL, uniform periodic state has F == 0, and the
 * reference BULK examples' betamu.                                          */
static void bulk_constants(tmo_ctx *c)
{
    int nc = c->p.ncomp, nd = c->p.ndim;
    c->inv_4pi = 1.0 / (4.0 * PI);
    for (int i = 0; i < nc; i++) {
        c->inv_rad[i] = 2.0 / c->p.hs_diam[i];
        c->inv_4pir[i] = c->inv_4pi * c->inv_rad[i];
        c->inv_4pirsq[i] = c->inv_4pir[i] * c->inv_rad[i];
    }
    for (int q = 0; q < 10; q++) c->rhobar_b[q] = c->dphi_b[q] = 0.0;
    for (int i = 0; i < nc; i++) {
        double hs = c->p.hs_diam[i];
        double vol = PI * hs * hs * hs / 6.0, area = PI * hs * hs, rho = c->p.rho_b[i];
        c->rhobar_b[0] += 1.0 * vol * rho;
        c->rhobar_b[1] += 1.0 * area * rho;
        c->rhobar_b[2] += 1.0 * area * rho * c->inv_4pir[i];
        c->rhobar_b[3] += 1.0 * area * rho * c->inv_4pirsq[i];
    }
    double n[10] = {0}, inv[5];
    n[3] = c->rhobar_b[0];
    n[2] = c->rhobar_b[1];
    n[1] = c->rhobar_b[2];
    n[0] = c->rhobar_b[3];
    inv[0] = (1.0 - n[3]);
    inv[1] = 1.0 / inv[0];
    inv[2] = inv[1] * inv[1];
    inv[3] = inv[2] * inv[1];
    inv[4] = inv[3] * inv[1];
    dphi_loc(c, n, 0.0, 0.0, inv, c->dphi_b);
    for (int i = 0; i < nc; i++) {
        double ss[4];
        ss[0] = 1.;
        ss[1] = c->p.hs_diam[i] / 2.;
        ss[2] = (4. * PI) * pow(c->p.hs_diam[i] / 2., 2);
        ss[3] = (4. * PI / 3.) * pow(c->p.hs_diam[i] / 2., 3);
        c->betamu_hs_ex[i] = 0.0;
        for (int q = 0; q < 4; q++) c->betamu_hs_ex[i] += c->dphi_b[q] * ss[q] * 1.0;
    }
    (void)nd;
}

/* ------------------------------------------------------------------------- */
static tmo_ctx *create_impl(const tmo_params *p, const double *vext, const unsigned char *zero_tf);

tmo_ctx *tmo_create(const tmo_params *p, const double *vext, const unsigned char *zero_tf)
{
    g_pending = NULL;
    g_errmsg[0] = 0;
    g_in_create = 1;
    if (setjmp(g_create_env)) {
        g_in_create = 0;
        if (g_pending) tmo_destroy(g_pending);
        return NULL;
    }
    tmo_ctx *c = create_impl(p, vext, zero_tf);
    g_in_create = 0;
    return c;
}

static tmo_ctx *create_impl(const tmo_params *p, const double *vext, const unsigned char *zero_tf)
{

    if (p->type_func < TMO_FMT1 || p->type_func > TMO_FMT4) die("type_func");
    if (p->ncomp < 1 || p->ncomp > NC || p->ndim < 1 || p->ndim > 3) die("dims");
    tmo_ctx *c = (tmo_ctx *)calloc(1, sizeof(tmo_ctx));
    g_pending = c;
    c->p = *p;
    c->N = 1;
    for (int d = 0; d < p->ndim; d++) c->N *= p->nx[d];
    c->nrbs = 4;
    c->nrb = 4 + 2 * p->ndim;
    c->nmf = (p->type_attr >= 0) ? p->ncomp : 0;
    c->iu_rho = 0;
    c->iu_rb = p->ncomp;
    c->iu_mf = p->ncomp + c->nrb;
    c->nunk = p->ncomp + c->nrb + c->nmf;
    size_t nc = (size_t)c->N * p->ncomp;
    c->vext = (double *)calloc(nc, sizeof(double));
    c->zero = (unsigned char *)calloc(nc, 1);
    if (vext) memcpy(c->vext, vext, nc * sizeof(double));
    if (zero_tf) memcpy(c->zero, zero_tf, nc);
    bulk_constants(c);
    if (p->type_attr >= 0)
        for (int i = 0; i < p->ncomp; i++)
            for (int j = 0; j < p->ncomp; j++)
                c->rmin[i][j] = find_rmin(p->sigma_ff[i][j], p->eps_ff[i][j], p->cut_ff[i][j], p->dx[0]);
    for (int i = 0; i < p->ncomp; i++) {
        build_stencil(c, TMO_DELTA, i, 0);
        build_stencil(c, TMO_THETA, i, 0);
    }
    if (p->type_attr >= 0) {
        for (int i = 0; i < p->ncomp; i++)
            for (int j = 0; j < p->ncomp; j++) build_stencil(c, TMO_UATT, i, j);
        for (int i = 0; i < p->ncomp; i++)
            for (int j = 0; j < p->ncomp; j++) {
                const sten_t *s = &c->sten[TMO_UATT][i + p->ncomp * j];
                double sum = 0.0;
                for (int q = 0; q < s->L; q++) sum += s->w[q];
                c->avdw[i][j] = sum;
            }
        for (int i = 0; i < p->ncomp; i++) {
            c->betamu_att[i] = 0.0;
            for (int j = 0; j < p->ncomp; j++) c->betamu_att[i] += p->rho_b[j] * c->avdw[i][j];
        }
    }
    return c;
}

void tmo_destroy(tmo_ctx *c)
{
    if (!c) return;
    for (int t = 0; t < 3; t++)
        for (int q = 0; q < NC * NC; q++) {
            free(c->sten[t][q].off);
            free(c->sten[t][q].w);
            free(c->sten[t][q].hw);
        }
    free(c->hwmask);
    free(c->vext);
    free(c->zero);
    free(c);
}

int tmo_nunk(const tmo_ctx *c) { return c->nunk; }
int tmo_nnodes(const tmo_ctx *c) { return c->N; }

void tmo_unk_layout(const tmo_ctx *c, int *iu_rho, int *iu_rb, int *n_rb, int *iu_mf, int *n_mf)
{
    *iu_rho = c->iu_rho;
    *iu_rb = c->iu_rb;
    *n_rb = c->nrb;
    *iu_mf = c->iu_mf;
    *n_mf = c->nmf;
}

int tmo_stencil_length(const tmo_ctx *c, int isten, int icomp, int jcomp)
{
    return c->sten[isten][icomp + c->p.ncomp * jcomp].L;
}

int tmo_stencil_get(const tmo_ctx *c, int isten, int icomp, int jcomp, int *off, double *w)
{
    const sten_t *s = &c->sten[isten][icomp + c->p.ncomp * jcomp];
    memcpy(off, s->off, sizeof(int) * 3 * (size_t)s->L);
    memcpy(w, s->w, sizeof(double) * (size_t)s->L);
    return s->L;
}

void tmo_bulk(const tmo_ctx *c, double *out)
{
    memset(out, 0, 60 * sizeof(double));
    for (int q = 0; q < 10; q++) out[q] = c->rhobar_b[q];
    for (int q = 0; q < 10; q++) out[10 + q] = c->dphi_b[q];
    for (int i = 0; i < c->p.ncomp; i++) {
        out[20 + i] = c->betamu_hs_ex[i];
        out[24 + i] = c->betamu_att[i];
        for (int j = 0; j < c->p.ncomp; j++) {
            out[28 + 4 * i + j] = c->avdw[i][j];
            out[44 + 4 * i + j] = c->rmin[i][j];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* offset_to_node_box, dft_mesh_lib.c:203-297, for a single box equal to the
 * domain (uniform interface: bulk codes are -1, walls -2).
 * Pinned by: wrap/reflect/bulk cases against hand-computed indices.          */
static int o2n(const tmo_ctx *c, const int *ijk, const int *off, int *refl)
{
    int s[3] = {0, 0, 0};
    int nd = c->p.ndim;
    for (int d = 0; d < nd; d++) {
        int nx = c->p.nx[d];
        s[d] = ijk[d] + off[d];
        refl[d] = 0;
        if (s[d] < 0) {
            switch (c->p.bc[d][0]) {
            case TMO_IN_BULK: return -1;
            case TMO_IN_WALL: return -2;
            case TMO_REFLECT:
                refl[d] = 1;
                s[d] = -s[d];
                if (s[d] >= nx) {
                    if (c->p.bc[d][1] == TMO_IN_BULK) return -1;
                    if (c->p.bc[d][1] == TMO_IN_WALL) return -2;
                    die("multiple reflections");
                }
                break;
            case TMO_PERIODIC: s[d] += nx; break;
            }
        } else if (s[d] >= nx) {
            switch (c->p.bc[d][1]) {
            case TMO_IN_BULK: return -1;
            case TMO_IN_WALL: return -2;
            case TMO_REFLECT:
                refl[d] = 1;
                s[d] = 2 * (nx - 1) - s[d];
                if (s[d] < 0) {
                    if (c->p.bc[d][0] == TMO_IN_BULK) return -1;
                    if (c->p.bc[d][0] == TMO_IN_WALL) return -2;
                    die("multiple reflections");
                }
                break;
            case TMO_PERIODIC: s[d] -= nx; break;
            }
        }
    }
    if (nd == 1) return s[0];
    if (nd == 2) return s[0] + s[1] * c->p.nx[0];
    return s[0] + s[1] * c->p.nx[0] + s[2] * c->p.nx[0] * c->p.nx[1];
}

static void node_ijk(const tmo_ctx *c, int node, int *ijk)
{
    ijk[0] = node % c->p.nx[0];
    ijk[1] = (c->p.ndim >= 2) ? (node / c->p.nx[0]) % c->p.nx[1] : 0;
    ijk[2] = (c->p.ndim == 3) ? node / (c->p.nx[0] * c->p.nx[1]) : 0;
}

#define X(u, node) x[(size_t)(u) * (size_t)N + (size_t)(node)]

/* resid_rho_bar, dft_fill_FMT.c:257-279 with constant_boundary (dft_utils.c:90) */
static double rho_val(const tmo_ctx *c, const double *x, int jcomp, int jnode)
{
    int N = c->N;
    if (jnode >= 0 && !c->zero[(size_t)jnode * c->p.ncomp + jcomp]) return X(c->iu_rho + jcomp, jnode);
    if (jnode == -1) return c->p.rho_b[jcomp];
    return 0.0;
}

/* ------------------------------------------------------------------------- */
/* coordinate-form Jacobian sink (NULL rows => count only) */
typedef struct {
    int *rows, *cols;
    double *vals;
    long n, cap;
} coo_t;

static void coo_add(coo_t *J, int row, int col, double v)
{
    if (J->rows && J->n < J->cap) {
        J->rows[J->n] = row;
        J->cols[J->n] = col;
        J->vals[J->n] = v;
    }
    J->n++;
}

/* one node: residual rows (F != NULL) and/or Jacobian entries (J != NULL),
 * following fill_resid_and_matrix / load_standard_node (dft_fill_main.c:38-145,
 * :451-560) with dphi precomputed at every node (FMT1stDeriv_switch,
 * dft_switch_FMT.c:68-96). */
static int one_size(const tmo_ctx *c);
static void fill_node(const tmo_ctx *c, const double *x, const double *dphi, int inode, double *F, coo_t *J)
{
    const int N = c->N, nd = c->p.ndim, ncomp = c->p.ncomp, nrb = c->nrb;
    int ijk[3], refl[3];
    node_ijk(c, inode, ijk);
    for (int iunk = 0; iunk < c->nunk; iunk++) {
        double r = 0.0;
        const int row = iunk * N + inode;
        if (iunk < c->iu_rb) {
            /* ---- Euler-Lagrange row, load_euler_lagrange dft_fill_EL.c:38-197 ---- */
            int icomp = iunk - c->iu_rho;
            double vx = c->vext[(size_t)inode * ncomp + icomp];
            if (c->zero[(size_t)inode * ncomp + icomp] || (c->p.nwall > 0 && vx == c->p.vext_max)) {
                /* fill_zero_value :224-240 */
                r = X(iunk, inode);
                if (J) coo_add(J, row, row, 1.0);
                if (F) F[row] = r;
                continue;
            }
            /* fill_EL_ideal_gas :351-386 */
            r += log(X(iunk, inode));
            if (J) coo_add(J, row, row, 1.0 / X(iunk, inode));
            /* fill_EL_chem_pot :388-433 (uniform interface) */
            {
                double mu = 0.0;
                mu -= log(c->p.rho_b[icomp]);
                mu -= c->betamu_hs_ex[icomp];
                if (c->p.type_attr >= 0) mu -= c->betamu_att[icomp];
                r += mu;
            }
            /* fill_EL_ext_field :435-444 */
            if (c->p.nwall > 0) r += vx;
            /* load_nonlocal_hs_rosen_rb, dft_fill_FMT.c:41-189 */
            for (int st = TMO_DELTA; st <= TMO_THETA; st++) {
                const sten_t *s = &c->sten[st][icomp];
                for (int k = 0; k < s->L; k++) {
                    const int *off = &s->off[3 * k];
                    double w = s->w[k], t = 0.0;
                    int jn = o2n(c, ijk, off, refl);
                    if (jn >= 0) {
                        const double *dp = &dphi[(size_t)jn * nrb];
                        if (st == TMO_DELTA) {
                            t = 1.0 * w * (dp[0] * c->inv_4pirsq[icomp] + dp[1] * c->inv_4pir[icomp] + dp[2]);
                            for (int d = 0; d < nd; d++) {
                                double sg = refl[d] ? -1.0 : 1.0;
                                t -= 1.0 * sg * w * (dp[4 + d] * c->inv_4pir[icomp] + dp[4 + nd + d]) *
                                     (off[d] * c->p.dx[d] * c->inv_rad[icomp]);
                            }
                        } else {
                            t = 1.0 * w * dp[3];
                        }
                    } else if (jn == -1) {
                        if (st == TMO_DELTA)
                            t = 1.0 * w * (c->dphi_b[0] * c->inv_4pirsq[icomp] + c->dphi_b[1] * c->inv_4pir[icomp] + c->dphi_b[2]);
                        else
                            t = 1.0 * w * c->dphi_b[3];
                    }
                    r += t;
                    if (J && jn >= 0) {
                        double sg[3] = {1.0, 1.0, 1.0}, rb[10], n[10], tmp[10];
                        for (int d = 0; d < nd; d++) sg[d] = refl[d] ? -1.0 : 1.0;
                        for (int q = 0; q < nrb; q++) rb[q] = X(c->iu_rb + q, jn);
                        sv_to_n(nd, rb, n);
                        if (st == TMO_DELTA)
                            tmo_fmt_d2_delta(c, n, off, sg, icomp, tmp);
                        else
                            tmo_fmt_d2_theta(c, n, tmp);
                        int junk = c->iu_rb;
                        coo_add(J, row, (junk + 0) * N + jn, 1.0 * w * tmp[3]);
                        coo_add(J, row, (junk + 1) * N + jn, 1.0 * w * tmp[2]);
                        coo_add(J, row, (junk + 2) * N + jn, 1.0 * w * tmp[1]);
                        coo_add(J, row, (junk + 3) * N + jn, 1.0 * w * tmp[0]);
                        for (int d = 0; d < nd; d++) {
                            coo_add(J, row, (junk + 4 + d) * N + jn, 1.0 * w * tmp[4 + nd + d]);
                            coo_add(J, row, (junk + 4 + nd + d) * N + jn, 1.0 * w * tmp[4 + d]);
                        }
                    }
                }
            }
            /* mean-field unknown, dft_fill_EL.c:143-156 (ATTInA22Block = FALSE) */
            if (c->p.type_attr >= 0) {
                int iu_att = c->iu_mf + icomp;
                r += X(iu_att, inode);
                if (J) coo_add(J, row, iu_att * N + inode, 1.0);
            }
            if (F) F[row] = r;
        } else if (iunk < c->iu_rb + nrb) {
            /* ---- rho-bar definition rows, load_rho_bar_s/_v dft_fill_FMT.c:195-347 ---- */
            int irb = iunk - c->iu_rb;
            int st = (irb == 0) ? TMO_THETA : TMO_DELTA;
            r += -X(iunk, inode);
            if (J) coo_add(J, row, row, -1.0);
            if (one_size(c) && (irb == 2 || irb == 3 || irb >= 4 + nd)) {
                /* dft_fill_FMT.c:215-236 (n1, n0), :305-320 (nv1): multiples of n2 and nv2 */
                const int ind = (irb < 4) ? 1 : irb - nd;
                const double f = (irb == 3) ? c->inv_4pirsq[0] : c->inv_4pir[0];
                r += X(c->iu_rb + ind, inode) * f;
                if (J) coo_add(J, row, (c->iu_rb + ind) * N + inode, f);
                if (F) F[row] = r;
                continue;
            }
            /* resid_and_Jac_sten_fill_sum_Ncomp, dft_fill_CLSmf.c:38-171 */
            for (int jc = 0; jc < ncomp; jc++) {
                const sten_t *s = &c->sten[st][jc];
                for (int k = 0; k < s->L; k++) {
                    const int *off = &s->off[3 * k];
                    double w = s->w[k], fac;
                    if (irb < 4) { /* prefactor_rho_bar_s :242-252 */
                        if (irb <= 1) fac = 1.0;
                        else if (irb == 2) fac = c->inv_4pir[jc];
                        else fac = c->inv_4pirsq[jc];
                        fac *= 1.0;
                    } else { /* prefactor_rho_bar_v :330-347 */
                        int idim;
                        if (irb < 4 + nd) {
                            fac = 1.0;
                            idim = irb - 4;
                        } else {
                            fac = c->inv_4pir[jc];
                            idim = irb - 4 - nd;
                        }
                        double vec = (double)off[idim] * c->p.dx[idim] * c->inv_rad[jc];
                        fac *= 1.0 * vec;
                    }
                    int jn = o2n(c, ijk, off, refl);
                    if (c->p.lhard_surf) w = hw_weight(c, s, k, jc, jn, refl, w);   /* dft_fill_CLSmf.c:140-160 */
                    r += fac * (w * rho_val(c, x, jc, jn) - 0.0);
                    if (J && jn >= 0 && !c->zero[(size_t)jn * ncomp + jc])
                        coo_add(J, row, (c->iu_rho + jc) * N + jn, w * fac * 1.0);
                }
            }
            if (F) F[row] = r;
        } else {
            /* ---- mean-field rows, load_mean_field dft_fill_mf.c:38-63 ---- */
            int icomp = iunk - c->iu_mf;
            r = -X(iunk, inode);
            if (J) coo_add(J, row, row, -1.0);
            for (int jc = 0; jc < ncomp; jc++) {
                const sten_t *s = &c->sten[TMO_UATT][icomp + ncomp * jc];
                for (int k = 0; k < s->L; k++) {
                    const int *off = &s->off[3 * k];
                    double w = s->w[k];
                    int jn = o2n(c, ijk, off, refl);
                    if (c->p.lhard_surf) w = hw_weight(c, s, k, jc, jn, refl, w);   /* dft_fill_CLSmf.c:140-160 */
                    r += 1.0 * (w * rho_val(c, x, jc, jn) - 0.0);
                    if (J && jn >= 0 && !c->zero[(size_t)jn * ncomp + jc])
                        coo_add(J, row, (c->iu_rho + jc) * N + jn, w * 1.0 * 1.0);
                }
            }
            if (F) F[row] = r;
        }
    }
}

static double *compute_dphi(const tmo_ctx *c, const double *x)
{
    const int N = c->N, nrb = c->nrb, nd = c->p.ndim;
    double *dphi = (double *)malloc(sizeof(double) * (size_t)N * nrb);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        double rb[10], n[10], out[10];
        for (int q = 0; q < nrb; q++) rb[q] = X(c->iu_rb + q, i);
        sv_to_n(nd, rb, n);
        tmo_fmt_dphi(c, n, out);
        for (int q = 0; q < nrb; q++) dphi[(size_t)i * nrb + q] = out[q];
    }
    return dphi;
}

double tmo_residual(const tmo_ctx *c, const double *x, double *F)
{
    double *dphi = compute_dphi(c, x);
    const int N = c->N;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; i++) fill_node(c, x, dphi, i, F, NULL);
    free(dphi);
    double s = 0.0;
    for (long q = 0; q < (long)c->nunk * N; q++) s += F[q] * F[q];
    return s;
}

/* The residual rows of the listed nodes only, F[u][k] for node nodes[k] (compact [nunk][count]):
 * the same per-node fill after the pointwise derivatives of every node, so a bounded sample of a
 * large grid costs count/N of the stencil sums.  Returns the sum of their F^2. */
double tmo_residual_nodes(const tmo_ctx *c, const double *x, const int *nodes, int count, double *F)
{
    double *dphi = compute_dphi(c, x);
    const int N = c->N, nunk = c->nunk;
    double s = 0.0;
    /* fill_node writes F[u * N + inode]: give it a full-size scratch and copy the listed columns */
    double *Fs = (double *)malloc(sizeof(double) * (size_t)nunk * N);
    if (!Fs) die("out of memory");
#pragma omp parallel for schedule(dynamic, 16)
    for (int k = 0; k < count; k++) {
        const int i = nodes[k];
        if (i >= 0 && i < N) fill_node(c, x, dphi, i, Fs, NULL);
    }
    for (int k = 0; k < count; k++) {
        const int i = nodes[k];
        for (int u = 0; u < nunk; u++) {
            const double f = (i >= 0 && i < N) ? Fs[(size_t)u * N + i] : 0.0;
            F[(size_t)u * count + k] = f;
            s += f * f;
        }
    }
    free(Fs);
    free(dphi);
    return s;
}

long tmo_jacobian_coo(const tmo_ctx *c, const double *x, int *rows, int *cols, double *vals, long cap)
{
    double *dphi = compute_dphi(c, x);
    coo_t J = {rows, cols, vals, 0, cap};
    for (int i = 0; i < c->N; i++) fill_node(c, x, dphi, i, NULL, &J);
    free(dphi);
    return J.n;
}

/* calc_init_rho_bar / int_stencil_HSFMT (dft_guess_HSFMT.c:70-83,
 * dft_energy_CLSintsten.c:195-255) and calc_init_mf_attract / int_stencil
 * (dft_guess_MFATT.c:66-79, dft_energy_CLSintsten.c:73-129) */
void tmo_init_fields(const tmo_ctx *c, double *x)
{
    const int N = c->N, nd = c->p.ndim, ncomp = c->p.ncomp;
#pragma omp parallel for schedule(dynamic, 64)
    for (int inode = 0; inode < N; inode++) {
        int ijk[3], refl[3];
        node_ijk(c, inode, ijk);
        for (int irb = 0; irb < c->nrb; irb++) {
            int st = (irb == 0) ? TMO_THETA : TMO_DELTA;
            double sum = 0.0;
            for (int jc = 0; jc < ncomp; jc++) {
                const sten_t *s = &c->sten[st][jc];
                for (int k = 0; k < s->L; k++) {
                    const int *off = &s->off[3 * k];
                    double w = s->w[k], fac;
                    if (irb < 4) {
                        fac = (irb <= 1) ? 1.0 : (irb == 2 ? c->inv_4pir[jc] : c->inv_4pirsq[jc]);
                        fac *= 1.0;
                    } else {
                        int idim = (irb < 4 + nd) ? irb - 4 : irb - 4 - nd;
                        fac = (irb < 4 + nd) ? 1.0 : c->inv_4pir[jc];
                        fac *= 1.0 * ((double)off[idim] * c->p.dx[idim] * c->inv_rad[jc]);
                    }
                    int jn = o2n(c, ijk, off, refl);
                    if (jn >= 0 && !c->zero[(size_t)jn * ncomp + jc])
                        sum += fac * w * X(c->iu_rho + jc, jn);
                    else if (jn < 0)
                        sum += fac * w * (jn == -1 ? c->p.rho_b[jc] : 0.0);
                }
            }
            if (irb == 0 && sum > 1.0) sum = 0.98;
            X(c->iu_rb + irb, inode) = sum;
        }
        for (int ic = 0; ic < c->nmf; ic++) {
            double sum = 0.0;
            for (int jc = 0; jc < ncomp; jc++) {
                const sten_t *s = &c->sten[TMO_UATT][ic + ncomp * jc];
                for (int k = 0; k < s->L; k++) {
                    int jn = o2n(c, ijk, &s->off[3 * k], refl);
                    if (jn >= 0 && !c->zero[(size_t)jn * ncomp + jc])
                        sum += s->w[k] * X(c->iu_rho + jc, jn);
                    else if (jn < 0)
                        sum += s->w[k] * (jn == -1 ? c->p.rho_b[jc] : 0.0);
                }
            }
            X(c->iu_mf + ic, inode) = sum;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Newton-Schur solve.  The Jacobian is assembled in CSR; the linear solve is
 * GMRES on S = A22 - A21 A11^{-1} A12 with A11 = -I and A22 preconditioning
 * (dft_HardSphereLinProbMgr.cpp:345-428, dft_Schur_Epetra_Operator.cpp:86-147);
 * the update follows update_solution, dft_newton.c:266-380.
 * Parity unpinned at the Krylov level (no reference vectors exist); pinned end
 * to end by convergence to F = 0 and the Examples' key outputs.             */
typedef struct {
    int n;
    long *rp;
    int *ci;
    double *v;
} csr_t;

static int cmp_pair(const void *a, const void *b)
{
    const long *x = (const long *)a, *y = (const long *)b;
    return (x[0] > y[0]) - (x[0] < y[0]);
}

static void build_csr(const tmo_ctx *c, const double *x, csr_t *A)
{
    long nnz = tmo_jacobian_coo(c, x, NULL, NULL, NULL, 0);
    int *r = (int *)malloc(sizeof(int) * nnz), *cc = (int *)malloc(sizeof(int) * nnz);
    double *v = (double *)malloc(sizeof(double) * nnz);
    tmo_jacobian_coo(c, x, r, cc, v, nnz);
    int n = c->nunk * c->N;
    long *key = (long *)malloc(sizeof(long) * 2 * nnz);
    for (long q = 0; q < nnz; q++) {
        key[2 * q] = (long)r[q] * n + cc[q];
        key[2 * q + 1] = q;
    }
    qsort(key, nnz, 2 * sizeof(long), cmp_pair);
    A->n = n;
    A->rp = (long *)calloc(n + 1, sizeof(long));
    A->ci = (int *)malloc(sizeof(int) * nnz);
    A->v = (double *)malloc(sizeof(double) * nnz);
    long m = 0;
    long q = 0;
    while (q < nnz) {
        long k = key[2 * q];
        /* += semantics in insertion order (stable sort keeps order of equal keys
           only if qsort were stable, so sum by original index order) */
        long q2 = q;
        while (q2 < nnz && key[2 * q2] == k) q2++;
        /* order duplicates by insertion index */
        for (long a = q; a < q2; a++)
            for (long b = a + 1; b < q2; b++)
                if (key[2 * b + 1] < key[2 * a + 1]) {
                    long t = key[2 * a + 1];
                    key[2 * a + 1] = key[2 * b + 1];
                    key[2 * b + 1] = t;
                }
        double s = 0.0;
        for (long a = q; a < q2; a++) s += v[key[2 * a + 1]];
        int row = (int)(k / n), col = (int)(k % n);
        A->ci[m] = col;
        A->v[m] = s;
        A->rp[row + 1]++;
        m++;
        q = q2;
    }
    for (int i = 0; i < n; i++) A->rp[i + 1] += A->rp[i];
    free(key);
    free(r);
    free(cc);
    free(v);
}

static void csr_free(csr_t *A)
{
    free(A->rp);
    free(A->ci);
    free(A->v);
}

static void spmv(const csr_t *A, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < A->n; i++) {
        double s = 0.0;
        for (long q = A->rp[i]; q < A->rp[i + 1]; q++) s += A->v[q] * x[A->ci[q]];
        y[i] = s;
    }
}

typedef struct {
    const csr_t *A;
    int n1, n2; /* block sizes: block 2 = densities (first ncomp*N GIDs) */
    double *t, *u;
    const tmo_ctx *c;
} schur_t;

/* One particle size with hard surfaces (Lhard_surf with Nlists_HW == 2, one component and walls):
 * the reference fills n1, n0 and nv1 as multiples of n2 and nv2 (dft_fill_FMT.c:215-236, 305-320)
 * instead of stencil sums, so A11 = [-I 0; B -I] with those multiples in B, and its A11 operator
 * applies the exact inverse [-I 0; -B -I] (dft_HardSphereA11_Epetra_Operator.cpp:127-172). */
static int one_size(const tmo_ctx *c) { return c->p.lhard_surf && c->p.nwall > 0 && c->p.ncomp == 1; }

/* v_dep += B v_ind over the rho-bar block v[nrb][N] */
static void one_size_lift(const tmo_ctx *c, double *v)
{
    if (!one_size(c)) return;
    const int N = c->N, nd = c->p.ndim;
    for (int i = 0; i < N; i++) {
        v[2 * N + i] += c->inv_4pir[0] * v[N + i];
        v[3 * N + i] += c->inv_4pirsq[0] * v[N + i];
        for (int d = 0; d < nd; d++) v[(size_t)(4 + nd + d) * N + i] += c->inv_4pir[0] * v[(size_t)(4 + d) * N + i];
    }
}

/* S x2 = A22 x2 + A21 (A12 x2) */
static void schur_apply(schur_t *S, const double *x2, double *y2)
{
    int n2 = S->n2, n = S->A->n;
    memset(S->t, 0, sizeof(double) * n);
    memcpy(S->t, x2, sizeof(double) * n2);
    spmv(S->A, S->t, S->u);                /* u = J [x2; 0] */
    for (int i = 0; i < n2; i++) y2[i] = S->u[i];   /* A22 x2 */
    memset(S->t, 0, sizeof(double) * n2);
    memcpy(S->t + n2, S->u + n2, sizeof(double) * (n - n2));  /* t = [0; -A11^{-1} A12 x2] */
    one_size_lift(S->c, S->t + n2);
    spmv(S->A, S->t, S->u);
    for (int i = 0; i < n2; i++) y2[i] += S->u[i];
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* right-preconditioned restarted GMRES (M = diag(A22)); returns iterations */
static int gmres(schur_t *S, const double *dinv, const double *b, double *x, int restart, int maxit, double rtol)
{
    int n = S->n2, m = restart, its = 0;
    double *V = (double *)malloc(sizeof(double) * (size_t)n * (m + 1));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
    double *g = (double *)malloc(sizeof(double) * (m + 1)), *z = (double *)malloc(sizeof(double) * n);
    double *w = (double *)malloc(sizeof(double) * n);
    double bn = sqrt(dot(n, b, b));
    if (bn == 0.0) bn = 1.0;
    memset(x, 0, sizeof(double) * n);
    while (its < maxit) {
        schur_apply(S, x, w);
        for (int i = 0; i < n; i++) V[i] = b[i] - w[i];
        double beta = sqrt(dot(n, V, V));
        if (beta / bn <= rtol) break;
        for (int i = 0; i < n; i++) V[i] /= beta;
        for (int i = 0; i <= m; i++) g[i] = 0.0;
        g[0] = beta;
        int k;
        for (k = 0; k < m && its < maxit; k++, its++) {
            for (int i = 0; i < n; i++) z[i] = dinv[i] * V[(size_t)k * n + i];
            schur_apply(S, z, w);
            for (int j = 0; j <= k; j++) {
                double h = dot(n, w, &V[(size_t)j * n]);
                H[j * m + k] = h;
                for (int i = 0; i < n; i++) w[i] -= h * V[(size_t)j * n + i];
            }
            double hn = sqrt(dot(n, w, w));
            H[(k + 1) * m + k] = hn;
            if (hn != 0.0)
                for (int i = 0; i < n; i++) V[(size_t)(k + 1) * n + i] = w[i] / hn;
            for (int j = 0; j < k; j++) {
                double a = H[j * m + k], bb = H[(j + 1) * m + k];
                H[j * m + k] = cs[j] * a + sn[j] * bb;
                H[(j + 1) * m + k] = -sn[j] * a + cs[j] * bb;
            }
            double a = H[k * m + k], bb = H[(k + 1) * m + k];
            double den = sqrt(a * a + bb * bb);
            cs[k] = a / den;
            sn[k] = bb / den;
            H[k * m + k] = den;
            H[(k + 1) * m + k] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            if (fabs(g[k + 1]) / bn <= rtol) {
                k++;
                its++;
                break;
            }
        }
        /* solve H y = g */
        double y[1024];
        for (int i = k - 1; i >= 0; i--) {
            double s = g[i];
            for (int j = i + 1; j < k; j++) s -= H[i * m + j] * y[j];
            y[i] = s / H[i * m + i];
        }
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int j = 0; j < k; j++) s += V[(size_t)j * n + i] * y[j];
            x[i] += dinv[i] * s;
        }
        if (fabs(g[k]) / bn <= rtol) break;
    }
    free(V);
    free(H);
    free(cs);
    free(sn);
    free(g);
    free(z);
    free(w);
    return its;
}

int tmo_newton(const tmo_ctx *c, double *x, const tmo_newton_opts *o, tmo_newton_stats *st)
{
    const int N = c->N, n = c->nunk * N, n2 = c->p.ncomp * N;
    double *F = (double *)malloc(sizeof(double) * n), *dx = (double *)malloc(sizeof(double) * n);
    double *b = (double *)malloc(sizeof(double) * n), *dinv = (double *)malloc(sizeof(double) * n2);
    double *bs = (double *)malloc(sizeof(double) * n2), *tmp = (double *)malloc(sizeof(double) * n);
    schur_t S;
    S.c = c;
    S.n2 = n2;
    S.n1 = n - n2;
    S.t = (double *)malloc(sizeof(double) * n);
    S.u = (double *)malloc(sizeof(double) * n);
    memset(st, 0, sizeof(*st));
    int iter = 0, converged = 0;
    if (o->resid_tol > 0.0 && o->max_iter > 0) {
        double r2 = tmo_residual(c, x, F);
        st->resid_norm = sqrt(r2);
        if (st->resid_norm < o->resid_tol) converged = 1;
    }
    while (!converged && iter < o->max_iter) {
        iter++;
        tmo_residual(c, x, F);
        csr_t A;
        build_csr(c, x, &A);
        S.A = &A;
        for (int i = 0; i < n; i++) b[i] = -F[i];
        for (int i = 0; i < n2; i++) {
            double d = 0.0;
            for (long q = A.rp[i]; q < A.rp[i + 1]; q++)
                if (A.ci[q] == i) d = A.v[q];
            dinv[i] = (d != 0.0) ? 1.0 / d : 1.0;
        }
        /* b_S = b2 - A21 A11^{-1} b1 */
        memset(tmp, 0, sizeof(double) * n2);
        memcpy(tmp + n2, b + n2, sizeof(double) * (n - n2));
        one_size_lift(c, tmp + n2);
        spmv(&A, tmp, S.u);
        for (int i = 0; i < n2; i++) bs[i] = b[i] + S.u[i];
        st->gmres_iters_total += gmres(&S, dinv, bs, dx, o->gmres_restart, o->gmres_max_iter, o->gmres_tol);
        /* x1 = A12 x2 - b1 */
        memset(tmp, 0, sizeof(double) * n);
        memcpy(tmp, dx, sizeof(double) * n2);
        spmv(&A, tmp, S.u);
        for (int i = n2; i < n; i++) dx[i] = S.u[i] - b[i];   /* x1 = A11^{-1} (b1 - A12 x2) */
        one_size_lift(c, dx + n2);
        csr_free(&A);
        /* update_solution, dft_newton.c:266-380 */
        double frac_min = 1.0;
        for (int iu = 0; iu < c->nunk; iu++) {
            if (iu >= c->iu_rb) continue;  /* DENSITY unknowns only */
            for (int i = 0; i < N; i++) {
                double xv = X(iu, i), d = dx[(size_t)iu * N + i], frac;
                if (xv + d < 0.0) {
                    frac = fmin(1.0, xv / (-d));
                    frac = fmax(frac, o->update_scaling);
                } else {
                    frac = 1.0;
                }
                if (frac < frac_min) frac_min = frac;
            }
        }
        double unorm = 0.0;
        for (int iu = 0; iu < c->nunk; iu++)
            for (int i = 0; i < N; i++) {
                double t = (frac_min * dx[(size_t)iu * N + i]) / (o->nl_rel_tol * X(iu, i) + o->nl_abs_tol);
                unorm += t * t;
            }
        for (int iu = 0; iu < c->nunk; iu++)
            for (int i = 0; i < N; i++) {
                double *xp = &X(iu, i);
                double d = dx[(size_t)iu * N + i];
                /* x + frac_min d < 1e-99 (dft_newton.c:344) read in exact arithmetic: a density
                   whose own fraction x / (-d) is at or below frac_min reaches x + frac_min d <= 0,
                   where the rounded sum is a residue of either sign */
                int limited = iu < c->iu_rb && *xp + d < 0.0 && fmin(1.0, *xp / (-d)) <= frac_min;
                if ((*xp + frac_min * d < 1.e-99 || limited) && iu < c->iu_rb)
                    *xp = 0.1 * *xp;
                else if (iu == c->iu_rb && *xp + frac_min * d > 1.0)
                    *xp += 0.5 * (1.0 - *xp);
                else
                    *xp += frac_min * d;
            }
        unorm = sqrt(unorm);
        double r2 = tmo_residual(c, x, F);
        st->resid_norm = sqrt(r2);
        if (o->resid_tol > 0.0)
            converged = st->resid_norm < o->resid_tol;
        else
            converged = unorm <= 1.0;
    }
    st->iters = iter;
    st->converged = converged;
    free(F);
    free(dx);
    free(b);
    free(dinv);
    free(bs);
    free(tmp);
    free(S.t);
    free(S.u);
    return converged ? iter : -iter;
}

/* ------------------------------------------------------------------------- */
/* Integrated observables of a profile x: calc_free_energy (dft_energy.c:30-391)
 * with the integrands of dft_energy_id.c:22-56, dft_energy_mu.c:22-96,
 * dft_energy_vext.c:22-37, dft_energy_hs.c:22-58 and dft_energy_att.c:22-65
 * (int_stencil, dft_energy_CLSintsten.c:73-129), and calc_adsorption
 * (dft_out_ads.c:33-110, integrands :141-157), each integrated as
 * integrateInSpace (dft_out_utils.c:38-98): the sum over nodes of
 * integrand * nelhit[comp][node] * Vol_el / 2^D, times scale (Fac_vol, or
 * Fac_vol / (Fac_area Area) per unit area).  Profile terms take Nel_hit2, bulk
 * terms and the hard-sphere term Nel_hit (of component 0 for the latter).
 * out[0] omega_sum, [1] omega_s_sum (the reference's "energy"), [2..10] ideal,
 * ideal_b, mu, mu_b, vext, hs, hs_b, att, att_b, [11..14] ads, [15..18] ads_b.
 * Pinned by: the README "energy" and "adsorption" key outputs of the 1D
 * Examples (tests/test_oracle_golden.py).                                     */
void tmo_observables(const tmo_ctx *c, const double *x, const int *nel_hit, const int *nel_hit2, double scale,
                     double *out)
{
    const int N = c->N, C = c->p.ncomp, nd = c->p.ndim;
    const int att_on = c->p.type_attr >= 0;
    double vol = 1.0;
    for (int d = 0; d < nd; d++) vol *= c->p.dx[d];
    const double f = vol / (double)(1 << nd) * scale;
    double betamu[NC];
    for (int i = 0; i < C; i++)
        betamu[i] = log(c->p.rho_b[i]) + c->betamu_hs_ex[i] + (att_on ? c->betamu_att[i] : 0.0);
    double id = 0.0, mu = 0.0, vx = 0.0, hs = 0.0, att = 0.0, ads[NC] = {0}, sw1[NC] = {0};
    for (int node = 0; node < N; node++) {
        int ijk[3], refl[3];
        node_ijk(c, node, ijk);
        for (int i = 0; i < C; i++) {
            const double rho = X(c->iu_rho + i, node);
            const double w2 = nel_hit2[(size_t)i * N + node], w1 = nel_hit[(size_t)i * N + node];
            if (rho > 1.e-20) { /* DENSITY_MIN */
                id += rho * (log(rho) - 1.0) * w2;
                mu += -rho * betamu[i] * w2;
                if (c->p.nwall > 0) vx += rho * c->vext[(size_t)node * C + i] * w2;
            }
            if (att_on && rho > 0.0) {
                double mf = 0.0;
                for (int jc = 0; jc < C; jc++) {
                    const sten_t *s = &c->sten[TMO_UATT][i + C * jc];
                    for (int k = 0; k < s->L; k++) {
                        int jn = o2n(c, ijk, &s->off[3 * k], refl);
                        /* int_stencil with hard surfaces, dft_energy_CLSintsten.c:110-117 */
                        const double w = c->p.lhard_surf ? hw_weight(c, s, k, jc, jn, refl, s->w[k]) : s->w[k];
                        mf += w * rho_val(c, x, jc, jn);
                    }
                }
                att += 0.5 * rho * mf * w2;
            }
            ads[i] += rho * w2;
            sw1[i] += w1;
        }
        double rb[10], n[10];
        for (int q = 0; q < c->nrb; q++) rb[q] = X(c->iu_rb + q, node);
        sv_to_n(nd, rb, n);
        /* White Bear's third term is 0/0 at n3 = 0 (the reference would add a NaN): taken as 0 for
           n3 <= 1e-10, the derivatives' cut-off (dft_physics_FMT3.c:68) */
        if ((c->p.type_func == TMO_FMT3 || c->p.type_func == TMO_FMT4) && !(n[3] > 1.e-10)) {
            double d12 = 0.0;
            for (int d = 0; d < nd; d++) d12 += n[4 + d] * n[4 + nd + d];
            hs += (-n[0] * log(1.0 - n[3]) + (n[1] * n[2] - d12) / (1.0 - n[3])) * nel_hit[node];
        } else
            hs += tmo_fmt_phi(c, n) * nel_hit[node];
    }
    double nb[10] = {0};
    sv_to_n(nd, c->rhobar_b, nb);
    for (int d = 0; d < 2 * nd; d++) nb[4 + d] = 0.0;
    double id_b = 0.0, mu_b = 0.0, att_b = 0.0;
    for (int i = 0; i < C; i++) {
        const double rb = c->p.rho_b[i];
        id_b += rb * (log(rb) - 1.0) * sw1[i];
        mu_b += -rb * betamu[i] * sw1[i];
        if (att_on) att_b += 0.5 * rb * c->betamu_att[i] * sw1[i];
    }
    const double hs_b = tmo_fmt_phi(c, nb) * sw1[0];
    double t[9] = {id, id_b, mu, mu_b, vx, hs, hs_b, att, att_b};
    for (int q = 0; q < 9; q++) t[q] *= f;
    out[0] = t[0] + t[2] + t[4] + t[5] + t[7];
    out[1] = (t[0] - t[1]) + (t[2] - t[3]) + t[4] + (t[5] - t[6]) + (t[7] - t[8]);
    for (int q = 0; q < 9; q++) out[2 + q] = t[q];
    for (int i = 0; i < NC; i++) {
        out[11 + i] = i < C ? ads[i] * f : 0.0;
        out[15 + i] = i < C ? c->p.rho_b[i] * sw1[i] * f : 0.0;
    }
}
