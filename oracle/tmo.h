/*
 * tmo.h -- CPU oracle for the Tramonto nonlocal-density fill (TEST INFRASTRUCTURE).
 *
 * This is a plain-C restatement of the reference algorithm, used only as the
 * checker: by tests/, by __graft_entry__.smoke() and by bench.py's
 * cpu_baseline leg.  Nothing in the product library links or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the Tramonto source tree).  Pinning of each function (golden vectors, closed
 * forms, invariants, finite differences) lives in tests/test_oracle_*.py.
 */
#ifndef TMO_H
#define TMO_H

#ifdef __cplusplus
extern "C" {
#endif

#define TMO_NC 4          /* max components (reference NCOMP_MAX is 6) */

/* boundary condition codes, dft_globals_const.h:496-499 */
#define TMO_IN_WALL  (-1)
#define TMO_IN_BULK    0
#define TMO_PERIODIC   1
#define TMO_REFLECT    2

/* functionals, Type_func */
#define TMO_FMT1 0
#define TMO_FMT2 1
#define TMO_FMT3 2
#define TMO_FMT4 3

/* stencil types, dft_globals_const.h:154-156 */
#define TMO_DELTA 0
#define TMO_THETA 1
#define TMO_UATT  2

typedef struct {
    int ndim, ncomp;
    int nx[3];
    double dx[3];
    int bc[3][2];
    int type_func;      /* TMO_FMT1 .. TMO_FMT4 */
    int type_attr;      /* -1: none; 0: strict MF, u = u(r_min) for r < r_min */
    int nwall;          /* > 0: external field added to the Euler-Lagrange rows */
    int lhard_surf;     /* hard surfaces: boundary-node weights from tmo_set_hw_masks */
    double vext_max;
    double hs_diam[TMO_NC];
    double sigma_ff[TMO_NC][TMO_NC];
    double eps_ff[TMO_NC][TMO_NC];   /* already divided by kT */
    double cut_ff[TMO_NC][TMO_NC];
    double rho_b[TMO_NC];
} tmo_params;

typedef struct tmo_ctx tmo_ctx;

typedef struct {
    int max_iter;         /* Newton iterations */
    double nl_rel_tol, nl_abs_tol, update_scaling;  /* update_solution damping */
    double resid_tol;     /* stop when ||F||_2 < resid_tol (0 disables) */
    int gmres_restart, gmres_max_iter;
    double gmres_tol;     /* relative tolerance of the Schur GMRES */
} tmo_newton_opts;

typedef struct {
    int iters;
    int converged;
    double resid_norm;    /* ||F||_2 after the last update */
    int gmres_iters_total;
} tmo_newton_stats;

/* vext and zero_tf are node-major [N][ncomp] (reference Vext[loc_inode][icomp],
 * Zero_density_TF[inode_box][icomp]).  Both may be NULL (no walls). */
tmo_ctx *tmo_create(const tmo_params *p, const double *vext, const unsigned char *zero_tf);
void tmo_destroy(tmo_ctx *c);
/* hard surfaces (lhard_surf): per component and node, 0 for nodes that are not wall-fluid boundary
 * nodes of the component's element list, else bit ln set when the element that has the node as
 * its local node ln is a fluid element (Nodes_2_boundary_wall / Wall_elems, dft_mesh.c:1452-1470);
 * [ncomp][N], NULL to clear */
void tmo_set_hw_masks(tmo_ctx *c, const unsigned char *mask);
/* message of the last failed tmo_create (NULL return) */
const char *tmo_last_error(void);

int tmo_nunk(const tmo_ctx *c);
int tmo_nnodes(const tmo_ctx *c);
/* layout of the unknowns: first index of densities, rho-bars, mean-field unknowns */
void tmo_unk_layout(const tmo_ctx *c, int *iu_rho, int *iu_rb, int *n_rb, int *iu_mf, int *n_mf);

int tmo_stencil_length(const tmo_ctx *c, int isten, int icomp, int jcomp);
/* copies offsets [L][3] and weights [L]; returns L */
int tmo_stencil_get(const tmo_ctx *c, int isten, int icomp, int jcomp, int *off, double *w);

/* bulk constants: out[0..9]=Rhobar_b (solution order), out[10..19]=Dphi_Drhobar_b (n order),
 * out[20..23]=Betamu_hs_ex, out[24..27]=Betamu_att, out[28..43]=Avdw[i][j], out[44..59]=Rmin_ff */
void tmo_bulk(const tmo_ctx *c, double *out);

/* FMT pieces (n order: n0 n1 n2 n3 nv1[D] nv2[D]) */
void tmo_fmt_dphi(const tmo_ctx *c, const double *n, double *dphi);
double tmo_fmt_phi(const tmo_ctx *c, const double *n);
/* delta / theta "kernel-vector contracted" second derivatives, as RB_Struct in n order */
void tmo_fmt_d2_delta(const tmo_ctx *c, const double *n, const int *offset, const double *sign,
                      int icomp, double *out);
void tmo_fmt_d2_theta(const tmo_ctx *c, const double *n, double *out);

/* residual F (reference sign: the right-hand side receives -F); x and F are [nunk][N].
 * Returns sum of F^2. */
double tmo_residual(const tmo_ctx *c, const double *x, double *F);

/* residual rows of the listed nodes only: F is compact [nunk][count], F[u][k] the row u of node
 * nodes[k] (a bounded sample of a large grid).  Returns the sum of their F^2. */
double tmo_residual_nodes(const tmo_ctx *c, const double *x, const int *nodes, int count, double *F);

/* Jacobian in coordinate form with += semantics already merged; GID = iunk*N + node.
 * First call with rows == NULL to get the count. */
long tmo_jacobian_coo(const tmo_ctx *c, const double *x, int *rows, int *cols, double *vals, long cap);

/* seed rho-bar and mean-field unknowns from the densities (calc_init_rho_bar / calc_init_mf_attract) */
void tmo_init_fields(const tmo_ctx *c, double *x);

/* built-in Newton with the reference's damping and a Schur-complement GMRES */
int tmo_newton(const tmo_ctx *c, double *x, const tmo_newton_opts *o, tmo_newton_stats *st);

/* integrated observables (free-energy terms, adsorption); 19 doubles out */
void tmo_observables(const tmo_ctx *c, const double *x, const int *nel_hit, const int *nel_hit2, double scale,
                     double *out);

int tmo_num_threads(void);

/* the n-point Gauss-Legendre rule on [0, 1] the stencils use (n = 1, 3, 6, 12, 20, 40), in
 * set_gauss_quad's point order (src/dft_quadrature.c:40-159) */
void tmo_gauss_rule(int n, double *gp, double *gw);

/* test hook: stencils built after this call use interior Gauss order gauss_order, npt_boundary
 * mid-point sub-points per axis on straddling elements (delta and theta) and a 3D delta shell band
 * of band_factor * ok_distance; 0 restores the reference's rule (6 or 3; 150/40 or 40/20; 2) */
void tmo_set_stencil_rule(int gauss_order, int npt_boundary, double band_factor);

#ifdef __cplusplus
}
#endif
#endif
