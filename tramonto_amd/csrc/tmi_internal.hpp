// Internal declarations of libtramonto_mi355x.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "tramonto_mi355x.h"

namespace tmi {

constexpr double kPi = 3.141592653589793238462643383279502884197169399375;
constexpr int kNC = TMI_NCOMP_MAX;

// ---------------------------------------------------------------- host setup
struct HostStencil {
    int L = 0;
    std::vector<int> off;     // [L][3]
    std::vector<double> w;    // [L]
    std::vector<double> hw;   // [L][8]: the part of w from elements where the point is local node ln
};

// Sparse corrections of rows by sums over a few columns: out[rowid[r]] += sum val[k] in[col[k]],
// k in [rowptr[r], rowptr[r + 1]); device arrays, one thread per row (rows distinct).
struct HwCsr {
    int nrows = 0;
    int *rowptr = nullptr, *rowid = nullptr, *col = nullptr;
    double *val = nullptr;
};

struct BulkConstants {
    double inv_4pi = 0.0;
    double inv_rad[kNC] = {}, inv_4pir[kNC] = {}, inv_4pirsq[kNC] = {};
    double rhobar_b[10] = {};   // solution order
    double dphi_b[10] = {};     // n order
    double betamu_hs_ex[kNC] = {}, betamu_att[kNC] = {};
    double avdw[kNC][kNC] = {}, rmin[kNC][kNC] = {};
};

// stencils[type][icomp + ncomp*jcomp]
void build_all_stencils(const tmi_params &p, const BulkConstants &bc, HostStencil (*stencils)[kNC * kNC]);
void compute_bulk_constants(const tmi_params &p, BulkConstants &bc);
void compute_attraction_constants(const tmi_params &p, BulkConstants &bc, HostStencil (*stencils)[kNC * kNC]);
double lj_potential_minimum(double sigma, double eps, double rcut, double dx0);

// ---------------------------------------------------------------- device data
// Geometry of a halo-padded copy of the grid.  A node (x,y,z) sits at padded
// position (x+h0) + (y+h1)*P0 + (z+h2)*P0*P1; halo positions carry the value
// the reference's offset_to_node_box assigns to an out-of-box stencil point.
struct PadGeom {
    int h[3] = {0, 0, 0};
    int P[3] = {1, 1, 1};
    long long n = 0;   // P0*P1*P2
    // batched 1D problems only: explicit maps between nodes and padded positions
    int *d_node_pad = nullptr;   // [N] padded position of each node
    int *d_pad_node = nullptr;   // [n] node at each padded position, -1 in the wall gaps
};

template <typename T>
struct DevStencilT {
    int L = 0;
    int *lo = nullptr;   // padded linear offsets, [4][L]: geometries pad_u, pad_f, pad_u2, pad_f2
    T *w = nullptr;      // weights
    T *wv = nullptr;     // [L][ndim] w * offset_d*dx_d/R (delta stencils only)
};

// A 3D stencil regrouped into z-columns for register-blocked sums: for each
// distinct (dx, dy) the z-range [z0, z1] with dense weights, stored with
// kColB-1 leading zeros and zero-padded to a whole number of kColB chunks
// plus kColB-1 (so every chunk reads a full window of 2*kColB-1 weights).
// Delta stencils carry a second weight plane, w * dz*kz/R, and the column
// constants dx*kx/R, dy*ky/R of the x and y vector weights.
constexpr int kColB = 8;   // outputs per thread along z
template <typename T>
struct DevColumnsT {
    int ncol = 0;
    long long nw = 0;        // length of one weight plane
    int *lo = nullptr;       // [2][ncol] linear offset of (dx, dy, z0): density geometry, FMT geometry
    int *woff = nullptr;     // [ncol] start of the column's padded weights
    int *nchunk = nullptr;   // [ncol] number of kColB-chunks
    T *wpad = nullptr;       // [1 or 2][nw]
    T *fxy = nullptr;        // [ncol][2] (delta stencils)
};

// A 1D stencil as dense weights over its offset range [lo0, lo0 + L): kRowPad
// zeros, the L weights, then 4 kRowPad zeros, so that a kernel blocking B <=
// kRowPad outputs per lane reads whole windows of 2B-1 weights without
// bounds checks.  Delta stencils also carry the vector weights w k dx/R.
constexpr int kRowPad = 16;
template <typename T>
struct DevRow1DT {
    int lo0 = 0, L = 0;
    T *w = nullptr;    // [kRowPad + L + 4 kRowPad]
    T *wv = nullptr;   // same layout (delta stencils)
};

// Launch shape of a 1D row kernel chosen by the autotuner: outputs per lane
// and lanes per workgroup.
struct RowTune {
    int B = 8, wg = 256;
    double ms = 0.0;   // best measured time
};

struct Grid {
    int nd;
    int nx[3];
    int bc[3][2];
    int N;
};

struct Problem;
struct BatchWork;
struct FftPlan;

// kernel-side launch helpers (tmi_kernels.hip)
template <typename T>
struct Workspace {
    T *rho_pad = nullptr;    // [C][pad_u.n]   density (or direction) with halo
    T *adj_pad = nullptr;    // [nadj][pad_f.n] adjoint fields A_c, B_c,d, S3 with halo
    T *rb_part = nullptr;    // [C][5][N] per-component rho-bar sums (3D column kernels, C > 1)
    // transpose products (allocated on first use): zero-padded row weights
    T *corr_u = nullptr;     // [nrb + nmf][pad_u2.n]
    T *corr_f = nullptr;     // [C][pad_f2.n]
};

struct Problem {
    tmi_params p{};
    Grid g{};
    int N = 0, nunk = 0, nrb = 0, nmf = 0, iu_rho = 0, iu_rb = 0, iu_mf = 0;
    int nadj = 0;   // number of adjoint fields = C*(1+D) + 1
    BulkConstants bulk;
    HostStencil hs[3][kNC * kNC];
    PadGeom pad_u, pad_f;   // density halo (max reach of all stencils), FMT halo
    PadGeom pad_u2, pad_f2; // twice the halo: inputs of the transpose correlations
    // device tables
    DevStencilT<double> dsd[3][kNC * kNC];
    DevStencilT<float> dsf[3][kNC * kNC];
    // column form of the stencils (3D only), [type][idx] as the point tables
    DevColumnsT<double> cold[3][kNC * kNC];
    DevColumnsT<float> colf[3][kNC * kNC];
    bool use_columns = false;
    // dense row form of the stencils (1D only) and the tuned launch shapes
    DevRow1DT<double> rowd[3][kNC * kNC];
    DevRow1DT<float> rowf[3][kNC * kNC];
    bool use_rows1d = false;
    RowTune tune[2][3];   // [FP64, FP32][rb, mf, el]
    unsigned char *d_zero = nullptr;     // [C][N] Zero_density_TF
    unsigned char *d_elzero = nullptr;   // [C][N] EL row replaced by x (zero density or V_ext = VEXT_MAX)
    double *d_vext = nullptr;            // [C][N]
    std::vector<unsigned char> h_zero, h_elzero;
    std::vector<double> h_vext;
    Workspace<double> wsd;
    Workspace<float> wsf;
    // Batched 1D problems: npores independent pores (IN_WALL at both ends) that
    // share the fluid, packed node-contiguously, pore p owning nodes
    // [pore_off[p], pore_off[p+1]); in the padded copies consecutive pores are
    // separated by a zero gap of the halo width.  npores == 0: one problem.
    int npores = 0;
    std::vector<int> pore_off;
    int *d_pore_off = nullptr;     // [npores+1]
    int *d_pore_of = nullptr;      // [N] pore of each node
    double *d_mu = nullptr;        // [C][N] -ln rho_b - beta mu_hs,ex - beta mu_att of the node's pore
    std::vector<double> pore_rho_b;   // [npores][C]
    BatchWork *batch_work = nullptr;  // batched Newton buffers and graphs (tmi_batch.hip)
    FftPlan *fft = nullptr;           // FFT forms of the stencil sums (tmi_fft.hip), 2D/3D
    // solver scratch
    double *d_newton = nullptr;   // [2][nunk][N]: F and dx of tmi_newton_solve, from its first call on
    double *d_scratch = nullptr;
    size_t scratch_bytes = 0;
    double *d_red = nullptr;     // reduction buffer
    double *h_red = nullptr;     // pinned host mirror (1024 doubles), written by reduction kernels directly
    double *h_red_dev = nullptr; // its device address
    // external layout without the mean-field rows (attraction in the A22 block): the entry points
    // expand x to the internal layout, whose mean-field rows (the last nmf rows) are recomputed
    // from the densities, and return the first nunk_ext rows
    bool ext_mf = false;
    // hard surfaces: the rho-bar rows' weights at wall-fluid boundary nodes (hw_fwd: rows q*N + i
    // of the rho-bar block, columns c*N + j of the densities; hw_tr its transpose)
    HwCsr hw_fwd, hw_tr, hw_mf;   // hw_fwd rows cover the mean-field block too; hw_mf those rows alone
    // one particle size with hard surfaces (one component, walls): n1, n0, nv1 rows are multiples
    // of n2, nv2 as in the reference; d_w0 is the transpose product's copy of w
    bool one_size = false;
    // Jacobian-vector products with the reference's delta-stencil Hessian rows (TMI_JAC_REFERENCE:
    // its assembled entries, including the inexact ones) instead of the exact Hessian
    bool jac_ref = false;
    double *d_mfw2 = nullptr;   // 2D direct form: dense (2 hy + 1) x (2 hx + 1) attraction weights per component pair
    bool mf2d_mfma = true;      // FP64 dense 2D sums on the matrix cores (TMI_MF2D_MFMA=0: the FMA kernel)
    double *d_rbw2 = nullptr;   // 2D, one component: dense theta, delta, delta x and delta y weights [4][2 fy + 1][2 fx + 1]
    float *f_rbw2 = nullptr;    // the same in FP32 (mixed path)
    int rbw_h[2] = {0, 0};      // their reach fx, fy (nodes)
    bool rbw_el = false;        // the reach fits the adjoint halo too (Euler-Lagrange rows on the matrix cores)
    bool rb2d_mfma = true;      // FP64 rho-bar rows on the matrix cores (TMI_RB2D_MFMA=0: the gather kernel)
    float *f_mfw2 = nullptr;    // the same in FP32
    double *d_w0 = nullptr;
    int nunk_ext = 0;
    double *d_xi = nullptr, *d_vi = nullptr, *d_fi = nullptr, *d_gi = nullptr;   // [nunk][N] each
};

// ---------------------------------------------------------------- kernels
enum Mode { kResidual = 0, kLinear = 1 };

// F = residual (all rows); sumsq optional (host)
void launch_residual(Problem &P, const double *x, double *F, hipStream_t s, hipEvent_t *ev = nullptr);
// F(x) and sum F^2 (from the fused last passes' partial sums where every row comes from them)
double residual_sumsq(Problem &P, const double *x, double *F, hipStream_t s);
// out = u * rho (the mean-field sums, [C][N]) of the densities rho ([C][N] node rows; linear: zero
// boundary values, else the bulk values of the fill)
void launch_mf_conv(Problem &P, const double *rho, bool linear, double *out, hipStream_t s);
// weighted node sums of the free-energy and adsorption integrands of x (k_observables): sums[0..4]
// ideal, mu, vext, hs, att; sums[5 + c] adsorption; sums[5 + kNC + c] sum of Nel_hit of c.
// nh_d, nh2_d: device [C][N] element counts; mf_tmp: device [C][N] scratch
void observables_sums(Problem &P, const double *x, const int *nh_d, const int *nh2_d, double *mf_tmp, double *sums,
                      hipStream_t s);
// Jv for all rows, precision T for the stencil sums
template <typename T>
void launch_jvp(Problem &P, const double *x, const double *v, double *Jv, hipStream_t s);
template <typename T>
void launch_vjp(Problem &P, const double *x, const double *w, double *JTw, hipStream_t s);
void launch_init_fields(Problem &P, double *x, hipStream_t s);
// Schur pieces (T precision for stencil sums):
//   t1 = A12 x2 (block-1 rows of J [x2;0]);  y2 = A22 x2 + A21 t1
template <typename T>
void launch_schur_apply(Problem &P, const double *x, const double *x2, double *t1, double *y2, hipStream_t s);
// bS = b2 + A21 b1 (b full length)
void launch_schur_rhs(Problem &P, const double *x, const double *b, double *bS, hipStream_t s);
// x1 = A12 x2 - b1 written into dx (full length), x2 already in dx[0:C*N]
void launch_schur_backsub(Problem &P, const double *x, const double *b, double *dx, hipStream_t s);

// batched problems (tmi_batch.hip): per-pore Newton-Krylov; per-pore outputs may be NULL
int batch_newton_solve(Problem &P, double *x, const tmi_newton_opts &o, tmi_newton_stats &st, int *pore_iters,
                       int *pore_conv, int *pore_gmres, double *pore_resid, hipStream_t s);
void free_batch_work(Problem &P);

// FFT forms of the stencil sums (tmi_fft.hip)
void fft_setup(Problem &P, bool mf, bool fmt);
void fft_free(Problem &P);
bool fft_mf(const Problem &P);              // mean-field sums in FFT form
bool fft_fmt(const Problem &P);             // FMT sums (rho-bar, Euler-Lagrange) in FFT form
bool fft_fmt_available(const Problem &P);   // FMT spectra built
bool fft_wraps(const Problem &P);           // every axis periodic on the FFT grid (halos unused)
bool fft_folded(const Problem &P);          // spectrum tables stored for k_y, k_z >= 0 only
void fft_set_fmt(Problem &P, bool on);
bool fft_fused_available(const Problem &P);   // the fused x pass fits the LDS budget
bool fft_fused(const Problem &P);             // products through the fused x pass
int fft_xreg_mask(const Problem &P);          // fused x pass in registers: bit 0 density, bit 1 Euler-Lagrange product
int fft_regpass_mode(const Problem &P);       // strided register pass: 0 off, 1 where it has enough waves, 2 always
void fft_set_fused(Problem &P, int mask);     // fused x pass for the density (bit 0) / Euler-Lagrange (bit 1) product
// density products; rho_nodes (node arrays, fft_node_input) or, if null, workspace<T>(P).rho_pad
// ss (optional): partial sums of the squares of the rows written by a fused last pass, *nss of them
template <typename T>
void fft_rb_mf(Problem &P, const double *rho_nodes, bool with_rb, const double *in_rb, const double *sub_rb,
               double *out_rb, const double *in_mf, const double *sub_mf, double *out_mf, hipStream_t s,
               double *ss = nullptr, long long *nss = nullptr);
// pair buffer of a wrapping FFT grid: field f of node (x, y, z) at element 2 ((f / 2) ns + x + y S1 + z S2) + f % 2
struct FftPairs {
    void *buf;
    long long ns, S1, S2;
    int dform;   // the pairs carry the derivatives of Phi (rho-bar order) instead of A_c, B_c,d, S3
};
FftPairs fft_pairs(const Problem &P, bool linear);
bool fft_node_input(const Problem &P);   // every density product reads the node arrays (no density halo)
template <typename T>
void fft_el(Problem &P, bool linear, const double *x_rho, const double *in_rho, const double *in_mf, const double *add,
            double *out_el, hipStream_t s, double *ss = nullptr, long long *nss = nullptr);
double *fft_ss_buffer(const Problem &P, long long *cap);   // partial sums of F^2 of the fused last passes
void fft_time_kernels(Problem &P, int which, int reps, double *ms, double *bytes, int *launches, hipStream_t s);
// bulk-circulant Schur preconditioner on FFT grids (tmi_fft.hip)
void precond_setup(Problem &P);
Problem *create_bulk_problem(const tmi_params &p);   // tmi_capi.cpp
void destroy_bulk_problem(Problem *P);
bool precond_available(Problem &P);   // both rebuild a stale Q first (precond_invalidate)
bool precond_active(Problem &P);
void precond_invalidate(Problem &P);
void precond_enable(Problem &P, bool on);
void precond_apply(Problem &P, const double *x_rho, const double *v, double *z, hipStream_t s, bool fp32 = false);
// times the fill with the FMT sums in both forms and keeps the faster (tmi_kernels.hip)
void choose_fmt_form(Problem &P);

// 1D row-kernel autotuner (tmi_kernels.hip): times every launch shape on the
// problem's own buffers and stores the fastest in P.tune
void autotune_rows1d(Problem &P);

// vector helpers
double dev_dot(Problem &P, const double *a, const double *b, long long n, hipStream_t s);
// out_d[j] = V_j . w on the device (partial sums in P.d_red[0, k*1024)), no host synchronisation;
// mirror (optional, the device address of pinned host memory such as P.h_red_dev) gets a copy
void dev_multi_dot_async(Problem &P, const double *V, long long ld, int k, const double *w, long long n, double *out_d,
                         hipStream_t s, double *mirror = nullptr);
void dev_scale_by_norm(const double *nrm2_d, const double *x, double *y, long long n, hipStream_t s);
// classical Gram-Schmidt of w against V_0 .. V_k-1 with one reorthogonalisation pass, all on the
// device: coef_d[0, k) and coef_d[k, 2k) the two passes' projections, coef_d[2k] the squared
// norm of the result (copied to P.h_red[0, 2k] as well), next = the result normalised (0 if
// its norm is 0); w holds the result.  TV (double or float): the type of the basis V and next.
template <typename TV>
void dev_cgs2(Problem &P, const TV *V, long long ld, int k, double *w, long long n, TV *next, double *coef_d,
              hipStream_t s);
// y = (TV)(a x); b = (double) a
template <typename TV>
void dev_cast_scaled(double a, const double *x, TV *y, long long n, hipStream_t s);
void dev_widen(const float *a, double *b, long long n, hipStream_t s);
void dev_multi_dot(Problem &P, const double *V, long long ld, int k, const double *w, long long n, double *out,
                   hipStream_t s);
void dev_axpy(double a, const double *x, double *y, long long n, hipStream_t s);
// one particle size: out's n1, n0, nv1 rows = -in_dep + B in_ind (rows of F or of J v)
void one_size_rows(const Problem &P, const double *in_rb, double *out_rb, hipStream_t s);
// v_dep += B v_ind: -A11^{-1} of the reference's A11 = [-I 0; B -I] relative to A11 = -I
void one_size_lift(const Problem &P, double *v_rb, hipStream_t s);
// y += (rows n1, n0, nv1 of J)^T w_dep: -w_dep on the dependent, B^T w_dep on n2, nv2
void one_size_transpose(const Problem &P, const double *w_rb, double *y_rb, hipStream_t s);
// out += the rows of a sparse correction applied to in (no-op when empty)
void hw_apply(const HwCsr &h, const double *in, double *out, hipStream_t s);
// build / free the hard-surface corrections from per-component boundary-node masks [C][N]
void hw_build(Problem &P, const unsigned char *mask);
void hw_free(Problem &P);
// y[0..n) = v with a kernel (no memset node, so a captured graph replays it as launched)
void dev_fill(double v, double *y, long long n, hipStream_t s);
void dev_scale_copy(double a, const double *x, double *y, long long n, hipStream_t s);
void dev_mul(const double *a, const double *b, double *y, long long n, hipStream_t s);
void dev_gemv_sub(const double *V, long long ld, int k, const double *h, double *w, long long n, hipStream_t s);
template <typename TV>
void dev_gemv_add(const TV *V, long long ld, int k, const double *y, double *x, long long n, hipStream_t s);
void dev_a22_inverse_diag(Problem &P, const double *x, double *dinv, hipStream_t s);
void dev_newton_update(Problem &P, double *x, const double *dx, const tmi_newton_opts &o, double *frac_out,
                       double *unorm_out, hipStream_t s);

// allocation-free fill (tmi_noalloc.hip): the device address of the current GPU's fixed scratch array of
// kNoAllocScratchDoubles doubles, looked up once per device
constexpr long long kNoAllocScratchDoubles = 131072;   // 1 MiB
double *noalloc_scratch();
// partial sums the fill's Sigma F^2 needs: those of the fused last passes (FFT form), else the blocks
// of one dot product
long long residual_sumsq_partials(const Problem &P);
// F(x) and Sigma F^2 into sumsq_d (device), partial sums in `partials` (at least
// residual_sumsq_partials(P) doubles); no host synchronisation, no allocation
void residual_sumsq_device(Problem &P, const double *x, double *F, double *partials, double *sumsq_d, hipStream_t s);

// errors
void set_error(const std::string &msg);
#define TMI_HIP_CHECK(expr)                                                                            \
    do {                                                                                               \
        hipError_t _e = (expr);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            ::tmi::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " __FILE__ ":" + \
                             std::to_string(__LINE__));                                                \
            throw ::tmi::HipFailure();                                                                 \
        }                                                                                              \
    } while (0)

struct HipFailure {};

}  // namespace tmi
