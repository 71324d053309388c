/*
 * tramonto_mi355x.h -- C ABI of the MI355X nonlocal-density fill and
 * hard-sphere Schur operator library (libtramonto_mi355x.so).
 *
 * Two groups of entry points:
 *
 *  1. Field-level entry points (tmi_*).  They replace the per-entry fill of
 *     Tramonto's fill_resid_and_matrix_control (src/dft_fill_control.c:39)
 *     and the matrix-based Schur apply (src/dft_Schur_Epetra_Operator.cpp:86)
 *     with whole-field device kernels.  Arrays named *_d are device pointers
 *     holding unknowns in the reference's x[iunk][inode_box] order, one
 *     contiguous row of nnodes doubles per unknown (src/dft_main.c:323-331).
 *     `stream` is a hipStream_t (NULL = default stream).
 *
 *  2. The reference's LinProbMgr ABI (dft_linprobmgr_* and
 *     dft_hardsphere_lin_prob_mgr_*, src/dft_basic_lin_prob_mgr_wrapper.h:44-92,
 *     src/dft_hardsphere_lin_prob_mgr_wrapper.h:43-58) and the fill entry
 *     point fill_resid_and_matrix_control with the reference's signature, so
 *     the reference's C Newton driver (src/dft_newton.c:131) and NOX group
 *     (src/NOXLOCA_Tramonto_Group.cpp:176-313) link against this library.
 *
 * Every function returns 0 on success and a negative TMI_E* code on error
 * (tmi_last_error() gives the message), except where stated.
 */
#ifndef TRAMONTO_MI355X_H
#define TRAMONTO_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

#define TMI_NCOMP_MAX 4

#define TMI_OK 0
#define TMI_EINVAL (-1)   /* bad argument */
#define TMI_EUNSUP (-2)   /* feature outside the implemented set */
#define TMI_EHIP (-3)     /* HIP runtime error */
#define TMI_ESTATE (-4)   /* call sequence error (e.g. solve before fill) */

/* boundary conditions, Type_bc (src/dft_globals_const.h:496-499) */
#define TMI_IN_WALL (-1)
#define TMI_IN_BULK 0
#define TMI_PERIODIC 1
#define TMI_REFLECT 2

/* functionals, Type_func (src/dft_globals_const.h:247-250) */
#define TMI_FMT1 0   /* Rosenfeld */
#define TMI_FMT2 1   /* Rosenfeld with the (n2 - nv2.nv2/n2)^3 third term */
#define TMI_FMT3 2   /* White Bear */
#define TMI_FMT4 3   /* Oleksy and Hansen: White Bear's n3 dependence with FMT2's third term */

/* stencil types, src/dft_globals_const.h:154-156 */
#define TMI_DELTA 0
#define TMI_THETA 1
#define TMI_UATT 2

/* Problem description: the inputs the reference fill reads from its globals
 * (Ndim, Nodes_x, Esize_x, Type_bc, Type_func, Type_attr, Nwall, Lhard_surf,
 * VEXT_MAX, HS_diam, Sigma_ff, Eps_ff/kT, Cut_ff, Rho_b). */
typedef struct tmi_params {
    int ndim, ncomp;
    int nx[3];
    double dx[3];
    int bc[3][2];
    int type_func;
    int type_attr;      /* -1 none; 0 strict mean field with u(r<r_min)=u(r_min) */
    int nwall;          /* > 0: external field present */
    int lhard_surf;     /* 1: hard surfaces (Lhard_surf); the boundary-node weights of the rho-bar
                           and mean-field sums come from tmi_set_hard_surface_masks */
    double vext_max;
    double hs_diam[TMI_NCOMP_MAX];
    double sigma_ff[TMI_NCOMP_MAX][TMI_NCOMP_MAX];
    double eps_ff[TMI_NCOMP_MAX][TMI_NCOMP_MAX];
    double cut_ff[TMI_NCOMP_MAX][TMI_NCOMP_MAX];
    double rho_b[TMI_NCOMP_MAX];
    int att_in_a22;     /* 1: the reference's ATTInA22Block = TRUE layout: no mean-field unknowns, the
                           attraction sums enter the Euler-Lagrange rows directly (A22 not diagonal);
                           0: one mean-field unknown per component (ATTInA22Block = FALSE) */
} tmi_params;

typedef struct tmi_info {
    int nunk, nnodes;
    int iu_rho, iu_rb, n_rb, iu_mf, n_mf;
    int sten_len[3][TMI_NCOMP_MAX * TMI_NCOMP_MAX];
    int mf_fft;   /* 1: the mean-field sums run as FFT convolutions (wide 2D/3D attraction stencils) */
    int fmt_fft;  /* FMT sums (rho-bar and Euler-Lagrange rows): 0 direct, 1 FFT convolutions, 2 FFT with the x pass
                     fused with the spectral products (2D/3D, whichever fills faster) */
    int precond;  /* preconditioner of the Schur GMRES: TMI_PRECOND_DIAG or TMI_PRECOND_BULK */
    int fft_fold; /* 1: the stencil spectra of the FFT sums are stored for k_y, k_z >= 0 only (every stencil
                     symmetric under the reflection of y and of z), 0: for the whole FFT grid or no FFT sums */
    int fft_xreg; /* the fused x pass runs in registers (x length 32-256, few enough outputs): bit 0 for the
                     density product, bit 1 for the Euler-Lagrange product; 0 in LDS or not fused */
    int fft_regpass; /* strided FFT axes of length 32-256: 0 LDS Stockham passes (TMI_FFT_REGPASS=0), 1 the
                        one-wave register pass where a pass has at least 1024 waves, 2 wherever the length
                        allows (TMI_FFT_REGPASS=2); read at creation */
    int jacobian;    /* TMI_JAC_EXACT or TMI_JAC_REFERENCE (tmi_problem_set_jacobian) */
} tmi_info;

/* precision of the inner Krylov products in tmi_newton_solve */
#define TMI_PREC_FP64 0
#define TMI_PREC_MIXED 1   /* FP32 stencil sums and JVPs, FP64 residual/update/norm */

/* right preconditioners of the Schur-complement GMRES */
#define TMI_PRECOND_DIAG 0   /* the A22 diagonal, as the reference's HardSphere LinProbMgr */
#define TMI_PRECOND_BULK 1   /* A22 diagonal times the inverse bulk Schur symbol (periodic FFT grids) */

/* convergence tests of tmi_newton_solve */
#define TMI_CONV_RESID 0    /* ||F||_2 < resid_tol, tested before each step (the benchmark's criterion) */
#define TMI_CONV_UPDATE 1   /* the reference's: the weighted update norm ||frac dx / (nl_rel_tol |x| + nl_abs_tol)||_2
                               of the step just taken is <= 1 (update_solution, src/dft_newton.c:266-380) */

typedef struct tmi_newton_opts {
    int max_iter;
    double nl_rel_tol, nl_abs_tol, update_scaling;   /* update_solution damping, src/dft_newton.c:266 */
    double resid_tol;        /* converged when ||F||_2 < resid_tol */
    int gmres_restart, gmres_max_iter;
    double gmres_tol;        /* relative tolerance of each Schur solve */
    int precision;           /* TMI_PREC_* */
    int verbose;
    int convergence;         /* TMI_CONV_* (zero-initialised options keep TMI_CONV_RESID) */
} tmi_newton_opts;

typedef struct tmi_newton_stats {
    int iters, converged, gmres_iters_total;
    double resid_norm;
    double seconds;
} tmi_newton_stats;

const char *tmi_last_error(void);
int tmi_version(void);

/* Builds stencils (src/dft_stencil.c:40 calc_stencils), bulk constants
 * (src/dft_thermo_hs.c:42, src/dft_thermo_att.c:44) and uploads tables.
 * vext and zero_tf are host arrays [nnodes][ncomp] (Vext[loc_inode][icomp],
 * Zero_density_TF[inode_box][icomp]); either may be NULL. */
int tmi_problem_create(const tmi_params *p, const double *vext, const unsigned char *zero_tf, void **handle);
int tmi_problem_destroy(void *handle);
int tmi_problem_info(void *handle, tmi_info *info);
/* copies stencil offsets [L][3] and weights [L]; *len receives L */
int tmi_stencil_get(void *handle, int isten, int icomp, int jcomp, int *len, int *off, double *w);
/* bulk constants, same packing as the oracle (60 doubles) */
int tmi_bulk_get(void *handle, double *out60);

/* Host-only variants (no GPU needed): the stencil table and the bulk
 * constants a problem description yields.  Same outputs as tmi_stencil_get /
 * tmi_bulk_get on a created problem. */
int tmi_host_stencil(const tmi_params *p, int isten, int icomp, int jcomp, int *len, int *off, double *w);
int tmi_host_bulk(const tmi_params *p, double *out60);

/* Parameter updates in place (no counterpart in the reference, which rebuilds
 * its globals between continuation steps, src/dft_continuation.c; these serve
 * continuation in a bulk or wall parameter and the adjoint harness).
 * set_rho_b: new bulk densities rho_b[ncomp]; the chemical potentials and the
 * bulk boundary values follow, stencils and tables are kept; the bulk-symbol
 * preconditioner is rebuilt for the new rho_b by the next Schur solve (or
 * tmi_problem_info / set_preconditioner call), not here.  set_vext: new
 * V_ext [nnodes][ncomp] (host, node-major as at creation); Zero_density_TF is
 * kept and rows with V_ext = VEXT_MAX become fixed rows as at creation.  Both
 * wait for the device to be idle first. */
int tmi_problem_set_rho_b(void *handle, const double *rho_b);
int tmi_problem_set_vext(void *handle, const double *vext);
/* Preconditioner of the Schur GMRES (no counterpart in the reference, whose
 * HardSphere LinProbMgr preconditions with the A22 diagonal,
 * src/dft_HardSphereLinProbMgr.cpp:345-428): TMI_PRECOND_DIAG keeps that;
 * TMI_PRECOND_BULK, the default where available, multiplies the diagonal by
 * the inverse symbol of the Schur operator at the uniform bulk state,
 * M^-1 v = F^-1[S_b(k)^-1 diag(1/rho_b) F(diag(rho) v)], exact for a uniform
 * fluid on a periodic FFT grid (every length 2^a 3^b 5^c).  On 2D/3D grids
 * with walls or reflecting boundaries the FFT grid pads those axes, and the
 * symbol of a periodic bulk fluid on that padded grid serves as an
 * approximation.  It exists where an FFT grid exists and the bulk is stable
 * (S_b(k) positive definite); TMI_PRECOND=diag in the environment at creation
 * selects the diagonal.  Fails with TMI_PRECOND_BULK where unavailable. */
int tmi_problem_set_preconditioner(void *handle, int kind);

/* Jacobian of the products (Jacobian-vector products, the Schur operator,
 * getmatrixvalue): TMI_JAC_EXACT (default) the exact derivative of the
 * residual; TMI_JAC_REFERENCE the reference's assembled entries
 * (load_nonlocal_hs_rosen_rb with d2phi_drb2_delta_rb_FMTk,
 * src/dft_fill_FMT.c:41-189, src/dft_physics_FMT1.c:79-126, FMT2.c:106-184,
 * FMT3.c:101-160, FMT4.c:101-191), which differ from the exact derivative in
 * the nv2 columns of FMT2 (2D/3D) and FMT4 and, for White Bear and FMT4,
 * across REFLECT boundaries.  The residual and the transpose product
 * (tmi_apply_jacobian_transpose, exact) are unaffected.  TMI_JAC_REFERENCE=1
 * in the environment at creation selects the reference entries. */
#define TMI_JAC_EXACT 0
#define TMI_JAC_REFERENCE 1
int tmi_problem_set_jacobian(void *handle, int mode);

/* Integrated observables of a profile, in the units of the reference's output. */
typedef struct tmi_observables {
    double omega;    /* grand potential, omega_sum of calc_free_energy */
    double omega_s;  /* its surface excess, omega_s_sum: the reference's "energy" key output */
    double ideal, ideal_b, mu, mu_b, vext, hs, hs_b, att, att_b;   /* terms and their bulk values */
    double ads[TMI_NCOMP_MAX], ads_b[TMI_NCOMP_MAX];              /* adsorption, its bulk value */
} tmi_observables;

/* Solvation force on the walls: out[k] = sum over components c and nodes n of
 * w_d[k][c][n] * rho_c(n), the reference's contact sums of sum_rho_wall
 * (src/dft_out_force.c:223-280, collated per linked wall and divided by the area per unit area at
 * :84-167) with the surface-element weights (-sign of the surface normal x Area_surf_el /
 * Nnodes_per_el_S per boundary-node surface element, src/dft_mesh.c:1760-2230) given by the
 * caller; w_d: device [nsum][ncomp][N], nsum 1..64 (links x directions); out: host [nsum]. */
int tmi_compute_force(void *handle, const double *x_d, const double *w_d, int nsum, double *out, void *stream);
/* The free-energy terms of calc_free_energy (src/dft_energy.c:30-391: ideal gas, chemical
 * potential, neutral external field, hard spheres, mean-field attraction; the functionals of this
 * library) and the adsorption of calc_adsorption (src/dft_out_ads.c:33-110), summed on the
 * device at the profile x_d (device pointer, x[iunk][inode] rows).  nel_hit, nel_hit2: host
 * arrays [ncomp][nnodes], the reference's Nel_hit and Nel_hit2 (setup_integrals,
 * src/dft_out_utils.c:182-270), the elements counted per node for the bulk and hard-sphere
 * terms and for the profile terms.  scale: the factor integrateInSpace applies
 * (src/dft_out_utils.c:55-59: Fac_vol, or Fac_vol / (Fac_area Area) per unit area).  No
 * counterpart in the reference's fill; it replaces the post-processing integrals for this path.
 * Not for batches of pores. */
int tmi_compute_observables(void *handle, const double *x_d, const int *nel_hit, const int *nel_hit2, double scale,
                            tmi_observables *out, void *stream);

/* F(x): the residual of every equation (the reference inserts -F into the
 * right-hand side).  *sumsq (host, may be NULL) receives sum F^2, the value
 * fill_resid_and_matrix_control returns. */
int tmi_fill_residual(void *handle, const double *x_d, double *F_d, double *sumsq, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation): tmi_fill_residual with
 * sum F^2 written to device memory (*sumsq_d, one double) instead of the host, without a host
 * synchronisation and without any allocation.  Its reduction scratch (the per-workgroup partial
 * sums of F^2) is a fixed library-owned device array of 131072 doubles (1 MiB) per GPU, held until
 * the process ends; problems whose fill needs more partial sums are rejected with TMI_EINVAL.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (several streams or host
 * threads) race.  F and the sum equal tmi_fill_residual's bit for bit. */
int tmi_fill_residual_noalloc(void *handle, const double *x_d, double *F_d, double *sumsq_d, void *stream);
/* The residual fill with HIP events recorded on `stream` between its stages;
 * stage_ms[0..4] receive the mean over `reps` fills of: density halo,
 * rho-bar rows, mean-field rows, pointwise FMT + adjoint halo, EL rows. */
int tmi_fill_residual_timed(void *handle, const double *x_d, double *F_d, int reps, double *stage_ms, void *stream);
/* Measurement hook, no counterpart in the reference: when the FMT sums run in
 * FFT form (tmi_info.fmt_fft), times one kernel family of a residual fill on
 * the problem's work buffers, in the fill's order and batch sizes, with HIP
 * events on `stream`: which = 0 the plain FFT axis passes, which = 1 the
 * fused x passes (tmi_info.fmt_fft == 2), which = 2 the density product's last
 * pass writing the rho-bar and mean-field rows (wrapping grids).  *ms receives
 * their mean time per fill over `reps` fills, *bytes their algorithmic traffic
 * per fill (axis pass: every point of every pair field read and written once,
 * 16 B each way; fused x pass: input pairs and distinct spectrum tables read
 * once, output pairs written once; rows pass: the output pairs read once, each
 * row's unknowns read and its residual written, 8 B each) and *launches their
 * launches per fill. */
int tmi_fft_timed(void *handle, int which, int reps, double *ms, double *bytes, int *launches, void *stream);
/* Jv = J(x) v, the assembled-Jacobian product, src/dft_Schur_Epetra_Operator.cpp:149 (ApplyGlobal) */
int tmi_apply_jacobian(void *handle, const double *x_d, const double *v_d, double *Jv_d, void *stream);
/* same product with FP32 stencil sums and FP32 accumulation (FP64 pointwise
 * free-energy derivatives); the inner product of the mixed-precision solve */
int tmi_apply_jacobian_fp32(void *handle, const double *x_d, const double *v_d, double *Jv_d, void *stream);
/* JTw = J(x)^T w, adjoint (transpose-Jacobian-vector) product */
int tmi_apply_jacobian_transpose(void *handle, const double *x_d, const double *w_d, double *JTw_d, void *stream);
/* Hard surfaces (lhard_surf = 1): the wall-fluid boundary nodes of each component's element list
 * and which of their elements are fluid, the inputs of HW_boundary_weight (src/dft_utils.c:41-62)
 * that resid_and_Jac_sten_fill_sum_Ncomp applies to the rho-bar and mean-field rows
 * (src/dft_fill_CLSmf.c:140-160) and int_stencil to the attraction energy
 * (src/dft_energy_CLSintsten.c:110-117).
 * mask: host [ncomp][N]; 0 for a node that is not a boundary node of the component's list
 * (Nodes_2_boundary_wall == -1), else bit ln set when the element that has the node as its local
 * node ln (x bit 0, y bit 1, z bit 2 of the element's corner) is a fluid element (Wall_elems == -1)
 * inside the domain or by reflection.  A stencil point that reaches the node through a REFLECT
 * boundary in dimension d reads the mirror-side element, bit ln ^ (1 << d), as the reference's
 * node_box_to_elem_box_reflect does (src/dft_mesh_lib.c:896-937).  Entries at a component's
 * Zero_density_TF nodes are ignored: those columns carry 0 in the reference's sums
 * (src/dft_fill_CLSmf.c:156).  NULL clears.  Every later fill, JVP, transpose product and
 * Schur solve uses the weights; calc_init_rho_bar (tmi_init_fields) does not, as in the reference. */
int tmi_set_hard_surface_masks(void *handle, const unsigned char *mask);
/* rho-bar and mean-field unknowns from the densities, src/dft_guess_HSFMT.c:70, src/dft_guess_MFATT.c:66 */
int tmi_init_fields(void *handle, double *x_d, void *stream);
/* Newton-Krylov solve in place on the device (Schur GMRES with A22 preconditioning) */
int tmi_newton_solve(void *handle, double *x_d, const tmi_newton_opts *opts, tmi_newton_stats *stats, void *stream);
/* one Schur-complement solve J dx = -F at x (dx full length, device) */
int tmi_schur_solve(void *handle, const double *x_d, const double *F_d, double *dx_d, double rtol, int restart,
                    int maxit, int precision, int *iters, void *stream);

/* Launch shapes the autotuner chose for the 1D row kernels (1D problems and
 * batches; no counterpart in the reference): shape[prec][kind][2] = outputs
 * per lane and lanes per workgroup, ms[prec][kind] = measured time, prec 0
 * FP64 / 1 FP32, kind 0 rho-bar rows / 1 mean-field rows / 2 Euler-Lagrange
 * rows.  Zeros when the problem does not use the row kernels. */
int tmi_row_tuning(void *handle, int *shape, double *ms);
/* Sets one of those launch shapes (outputs per lane 4, 8 or 16; lanes per
 * workgroup 64, 128 or 256), replacing the autotuner's choice; every shape
 * computes the same sums.  TMI_EINVAL when the problem has no row kernels. */
int tmi_row_tuning_set(void *handle, int prec, int kind, int per_lane, int lanes);

/* --------------------------------------------------------------------------
 * Batched 1D or 2D problems.  No counterpart in the reference, which solves
 * one problem per run (src/dft_main.c): npores independent problems
 * ("pores") that share the fluid of *fluid and differ in node count nodes[p],
 * bulk densities rho_b[p][ncomp], V_ext and Zero_density_TF (vext, zero_tf:
 * [N][ncomp] node-major over the packed nodes, pore 0 first, N = sum of
 * nodes).  fluid->ndim = 1: 1D pores with IN_WALL at both ends
 * (fluid->nx[0], fluid->bc and fluid->rho_b are ignored).  fluid->ndim = 2:
 * 2D pieces of fluid->nx[0] columns and nodes[p] / nx[0] rows (x fastest),
 * stacked along y, each with IN_WALL at both y ends; fluid->bc[0] (both
 * ends PERIODIC, or both IN_WALL) is their common x boundary; their sums run
 * in the direct form.  Each pore is the problem the reference would set up
 * with those inputs (one fill_resid_and_matrix_control system per pore).
 * The handle works with tmi_problem_info (nnodes = N), tmi_fill_residual,
 * tmi_apply_jacobian[_fp32], tmi_init_fields, tmi_schur_solve and
 * tmi_newton_solve; rows of x hold the packed nodes.  Transpose products
 * return TMI_EUNSUP.
 * ------------------------------------------------------------------------ */
int tmi_batch_create(const tmi_params *fluid, int npores, const int *nodes, const double *rho_b, const double *vext,
                     const unsigned char *zero_tf, void **handle);
/* number of pores and first node of each pore (pore_off[npores+1], may be NULL) */
int tmi_batch_info(void *handle, int *npores, int *pore_off);
/* Newton-Krylov per pore in lockstep (newton_solver, src/dft_newton.c:131-196,
 * applied to every pore): per-pore GMRES, damping and convergence; one
 * Newton step runs as replayed HIP graphs (TMI_BATCH_GRAPH=0 launches the
 * same kernels directly).  stats: iters = Newton steps taken, converged =
 * all pores converged, gmres_iters_total = sum over pores.  Per-pore outputs
 * ([npores], host, each may be NULL): Newton iterations, converged flag,
 * GMRES iterations, final ||F_p||_2. */
int tmi_batch_newton_solve(void *handle, double *x_d, const tmi_newton_opts *opts, tmi_newton_stats *stats,
                           int *pore_iters, int *pore_converged, int *pore_gmres, double *pore_resid, void *stream);

/* --------------------------------------------------------------------------
 * Reference ABI.  The handle is the LinProbMgr; the fill entry point uses
 * the problem registered with tmi_set_active_problem (the reference reads
 * its problem from globals).  Arrays are host double** in box layout,
 * exactly as in the reference.
 *
 * The communicator argument is the reference's MPI_Comm.  Where mpi.h is
 * included first it is declared as MPI_Comm; otherwise as int, the width of
 * MPICH's MPI_Comm (the MPI the reference builds with here).  Both travel in
 * one integer register of the x86-64 calling convention, and the library
 * never reads it: one GPU holds the whole domain.
 * ------------------------------------------------------------------------ */
#ifdef MPI_VERSION
typedef MPI_Comm tmi_comm_t;
#else
typedef int tmi_comm_t;
#endif

/* Registers the problem the fill entry point and the LinProbMgr serve.  When
 * the manager's block structure is already finalised, its equation IDs and
 * maps are checked against the problem's layout (TMI_EINVAL on mismatch). */
int tmi_set_active_problem(void *problem_handle, void *linprobmgr_handle);

/* src/dft_fill_control.c:39.  resid_only_flag 0 adds -F to the manager's
 * right-hand side and records x as the Jacobian's linearisation point; 1 and
 * 4 add -F only; 2 and 3 insert nothing.  Returns sum F^2 (-1 on error). */
double fill_resid_and_matrix_control(double **x, int iter, int resid_only_flag);

/* src/dft_parameterlist_wrapper.h:40-52.  A list of named solver options;
 * the managers read "Max_iter" and "Tol" at solve time
 * (src/dft_HardSphereLinProbMgr.cpp:401) and the basic manager "Kspace".
 * set_all_aztec stores the Aztec options/params arrays under the reference's
 * names (src/dft_parameterlist_wrapper.cpp:74-108). */
void *dft_parameterlist_create(void);
void dft_parameterlist_destruct(void *parameterlistptr);
void dft_parameterlist_set_int(void *parameterlistptr, char *name, int value);
void dft_parameterlist_set_double(void *parameterlistptr, char *name, double value);
void dft_parameterlist_set_int_array(void *parameterlistptr, char *name, int *values);
void dft_parameterlist_set_double_array(void *parameterlistptr, char *name, double *values);
void dft_parameterlist_set_all_aztec(void *parameterlistptr, int *options, double *params);

/* src/dft_hardsphere_lin_prob_mgr_wrapper.h:43-58.  finalizeblockstructure
 * checks the three equation-ID lists against the registered problem's
 * layout (linsolver_setup_HSTYPE, src/dft_linsolver_setup.c:182-239).  The
 * debug manager prints ||J dx - b||_2 after each solve, as the reference's
 * debug solve does (src/dft_HardSphereLinProbMgr.cpp:405-414), without
 * writing matrix files or aborting. */
void *dft_hardsphere_lin_prob_mgr_create(int numUnks, void *ParameterList, tmi_comm_t comm);
void *dft_hardsphere_lin_prob_mgr_create_debug(int numUnks, void *ParameterList, tmi_comm_t comm);
void dft_hardsphere_lin_prob_mgr_destruct(void *linprobmgr);
int dft_hardsphere_lin_prob_mgr_setindnonlocalequationids(void *linprobmgr, int numIndNonLocalEquations,
                                                          int *indNonLocalEquationIDs);
int dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids(void *linprobmgr, int numDepNonLocalEquations,
                                                          int *depNonLocalEquationIDs);
int dft_hardsphere_lin_prob_mgr_setdensityequationids(void *linprobmgr, int numDensityEquations,
                                                      int *densityEquationIDs);
int dft_hardsphere_lin_prob_mgr_seta22blockisdiagonal(void *linprobmgr, int isA22Diagonal);

/* src/dft_basic_lin_prob_mgr_wrapper.h:44-92.  The basic manager
 * (dft_linsolver_setup.c:48, problems without a Schur solver) takes no
 * equation IDs; its solve returns the solution of the same system J dx = b
 * through the device Schur complement. */
void *dft_basic_lin_prob_mgr_create(int numUnks, void *ParameterList, tmi_comm_t comm);
void dft_linprobmgr_destruct(void *linprobmgr);
int dft_linprobmgr_setnodalrowmap(void *linprobmgr, int numgids, int *gids);
int dft_linprobmgr_setnodalcolmap(void *linprobmgr, int numgids, int *gids);
int dft_linprobmgr_setcoarsenednodeslist(void *linprobmgr, int numgids, int *gids);
int dft_linprobmgr_finalizeblockstructure(void *linprobmgr);
int dft_linprobmgr_initializeproblemvalues(void *linprobmgr);
int dft_linprobmgr_insertrhsvalue(void *linprobmgr, int iunk, int inode, double value);
int dft_linprobmgr_insertonematrixvalue(void *linprobmgr, int iunk, int ownednode, int boxphysicsID, int boxnode,
                                        double value);
int dft_linprobmgr_insertmultinodematrixvalues(void *linprobmgr, int iunk, int ownednode, int boxphysicsID,
                                               int *boxnodes, double *values, int numEntries);
int dft_linprobmgr_insertmultiphysicsmatrixvalues(void *linprobmgr, int iunk, int ownednode, int *boxphysicsIDs,
                                                  int boxnode, double *values, int numEntries);
int dft_linprobmgr_finalizeproblemvalues(void *linprobmgr);
int dft_linprobmgr_setrhs(void *linprobmgr, double **b);
int dft_linprobmgr_getrhs(void *linprobmgr, double **b);
int dft_linprobmgr_getlhs(void *linprobmgr, double **x);
int dft_linprobmgr_setupsolver(void *linprobmgr);
int dft_linprobmgr_solve(void *linprobmgr);
int dft_linprobmgr_applymatrix(void *linprobmgr, double **x, double **b);
int dft_linprobmgr_importr2c(void *linprobmgr, double **xOwned, double **xBox);
int dft_linprobmgr_importnodalr2c(void *linprobmgr, double *xOwned, double *xBox);
int dft_linprobmgr_importsingleunknownr2c(void *linprobmgr, double *xOwned, double *xBox);
double dft_linprobmgr_getmatrixvalue(void *linprobmgr, int iunk, int ownednode, int junk, int boxnode);
int dft_linprobmgr_writeMatrix(void *linprobmgr, char *filename, char *matrixName, char *matrixDescription);

/* No counterpart in the reference (whose solver reports nothing through the
 * ABI): GMRES iterations of the manager's last solve and, for the debug
 * manager, its ||J dx - b||_2 (-1 otherwise). */
int tmi_linprobmgr_last_solve(void *linprobmgr, int *iters, double *debug_resid);

#ifdef __cplusplus
}
#endif
#endif
