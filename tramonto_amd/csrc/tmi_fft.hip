// The stencil sums of 2D/3D problems as FFT convolutions.
//
// Every stencil sum of the fill is a correlation of a halo-padded field with a
// stencil, out(i) = sum_k w(k) in(i + o_k):
//   rho-bar rows         n_q(i)  = sum_c a_qc sum_k w_c(k) rho_c(i + o_k) over
//                        the theta, delta and vector-delta stencils
//                        (load_rho_bar_s/_v, src/dft_fill_FMT.c:195-347);
//   mean-field rows      MF_c(i) = sum_c' sum_k u_cc'(k) rho_c'(i + o_k)
//                        (load_mean_field, src/dft_fill_mf.c:38-63, through
//                        resid_and_Jac_sten_fill_sum_Ncomp, src/dft_fill_CLSmf.c:38-171);
//   Euler-Lagrange sums  E_c(i)  = sum_k [w_d(k) A_c - wv(k).B_c + w_t(k) S3](i + o_k)
//                        (load_nonlocal_hs_rosen_rb, src/dft_fill_FMT.c:41-189).
// With hundreds (FMT) to tens of thousands (attraction) of stencil points per
// node the direct sums are bound by FP64 FMAs; on an FFT grid each sum is a
// product of spectra and the fill costs a few HBM-bound passes instead.
//
// FFT grid, per dimension: a periodic axis whose node count factors into 2,
// 3 and 5 is used as is (a circular convolution is the periodic image sum);
// any other axis uses the halo-padded copy of the field (the halo holds the
// boundary values offset_to_node_box gives) extended to a 2^a 3^b 5^c length.
// Interior outputs of a padded axis never wrap, because the halo is the
// stencil reach.  The transforms are mixed-radix Stockham passes in LDS, one
// launch per axis for a batch of fields.
//
// The flipped stencils are even (theta, delta, attraction) or odd (vector
// delta), so their spectra are real or imaginary: set-up keeps one real table
// per stencil, scaled by 1/M, after checking that the other part vanishes.
// Real fields are transformed two at a time, one as the real and one as the
// imaginary part (split again in the product kernels from Z(k) and Z(-k)),
// and pairs of real outputs share one inverse transform the same way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "tmi_device.hpp"
#include "tmi_internal.hpp"

namespace tmi {

template <typename T>
struct alignas(2 * sizeof(T)) Cx {
    T x, y;
};

constexpr int kMaxOut = 10 + kNC;   // outputs of the density product: 4 + 2*3 rho-bar rows and kNC mean-field rows
constexpr int kKindMf = 8;          // output kinds: 0 theta, 1 delta, 2 + d vector delta d, kKindMf + c mean-field row c

// Outputs of the product of the density spectra: Y_o = sum_c coef[o][c] K_kind[o],c rho_c
// (FMT kinds) or Y_o = sum_c' U_c,c' rho_c' (mean-field row c = kind - kKindMf)
struct FwdDesc {
    int C = 0, nout = 0;
    int kind[kMaxOut] = {};
    double coef[kMaxOut][kNC] = {};
};

// Rows written from the outputs: rho-bar row q (q < nrb) = scale[q] Y_src[q],
// then nmf mean-field rows
struct RowMap {
    int nrb = 0, nmf = 0;
    int src[kMaxOut] = {};
    double scale[kMaxOut] = {};
};

// Terms of a fused x-axis product: output o += coef (times i when imag) K_t(k) X_f(k), K_t a real
// spectrum table (t >= 0: FMT table t; t < 0: attraction table -t-1), grouped by input field f,
// terms first[f] .. first[f+1]-1
constexpr int kMaxTerms = 64;
constexpr int kMaxIn = kNC * 4 + 1;   // adjoint fields C(1+D)+1
struct Term {
    int o, t, imag;
    int odd;   // axis in which the table is odd (vector delta d: d), -1 if even
    double coef;
};
struct TermList {
    int nin = 0, nout = 0;
    int first[kMaxIn + 2] = {};   // one past the last field is the running end while a list is built
    Term term[kMaxTerms] = {};
};

struct FftPlan {
    int nd = 0, C = 0;
    int M[3] = {1, 1, 1};       // FFT grid
    int off[3] = {0, 0, 0};     // FFT-grid coordinate of node 0
    // per geometry (0: density halo pad_u, 1: FMT halo pad_f): FFT-grid positions
    // lo <= m < lim map onto the padded copy at m + src; the rest are zero
    int lo[2][3] = {}, lim[2][3] = {{1, 1, 1}, {1, 1, 1}}, src[2][3] = {};
    bool wraps = true;          // every axis periodic with M = n: halos are not read
    int rc[3][4] = {};          // Stockham stages per axis: counts of radix 4, 2, 3 and 5, run in that order
    long long n = 1;            // grid points M0 M1 M2
    long long S1 = 1, S2 = 1;   // storage pitches of y rows and z planes (row_pitch; one row of padding
                                // per plane), so that no stride of the strided passes is a power of two
    long long ns = 1;           // storage elements per field
    bool mf = false;            // mean-field sums in FFT form
    bool fmt_ok = false;        // FMT spectra built
    bool fmt = false;           // FMT sums (rho-bar, Euler-Lagrange) in FFT form
    double *ud = nullptr;       // [C*C][tn] attraction spectra (real), index icomp + C*jcomp as the stencils
    float *uf = nullptr;
    double *td = nullptr;       // [C][2+nd][tn] theta, delta (real) and vector-delta d (imaginary part) spectra
    float *tf = nullptr;
    // layout of the spectrum tables (ud, uf, td, tf): with fold set (every stencil symmetric under the
    // reflection of y and of z) only 0 <= k_y <= M_y/2, 0 <= k_z <= M_z/2 is stored, pitches tp1 and tp2;
    // otherwise the storage layout of the transforms (pitches S1, S2)
    bool fold = false;
    long long tp1 = 1, tp2 = 1, tn = 1;   // pitch of y rows and z planes, elements per table
    Cx<double> *twd[3] = {nullptr, nullptr, nullptr};   // exp(-2 pi i j / M_d), j < M_d
    Cx<float> *twf[3] = {nullptr, nullptr, nullptr};
    void *zbuf = nullptr;       // transforms of the input pairs
    void *obuf = nullptr;       // products, then the output pairs
    FwdDesc fwd[2];             // [0] mean-field rows only, [1] rho-bar rows then mean-field rows
    RowMap rows[2];
    TermList tfwd[2], tel;      // the same products as terms of the fused x pass (tel: adjoint fields
                                // A_c, B_c,d, S3; tel_d: the derivative form, when dform)
    TermList tel_d;
    // rows fed by each output of the density product (fused last pass): [w][output][k]
    int onrow[2][kMaxOut] = {}, orow[2][kMaxOut][3] = {};
    double oscale[2][kMaxOut][3] = {};
    double *ss = nullptr;       // partial sums of F^2 from the fused last passes (density, then Euler-Lagrange rows)
    long long ss_cap = 0;
    int xfit[2] = {0, 0};       // lines per workgroup of the fused x pass that fit, density / adjoint product
    int xlpb[2] = {0, 0};       // in use (0: separate x passes and product kernel)
    // bulk-circulant Schur preconditioner (wrapping grids): Q(k) = S_b(k)^-1 diag(1/rho_b) / n
    Cx<double> *pq = nullptr;   // [C*C][ns], entry (i, j) at i + C j
    // Q real and even in k_y, k_z (the bulk operator is real and reflection-symmetric): its real part
    // in the spectrum-table layout, applied by the fused x pass between fused-end passes
    double *pqr = nullptr;      // [C*C][tn]
    float *pqr_f = nullptr;     // the same in FP32, for the mixed path's GMRES
    TermList tpc;               // output i += Q_ij X_j
    int pc_lpb = 0;             // lines per workgroup of its LDS form
    bool pc_fused = false;
    bool pc_ok = false;         // Q built
    int pc_want = -1;           // wanted (TMI_PRECOND=diag: no; -1 until the first set-up reads it)
    bool pc_on = false;         // GMRES uses it: built and wanted
    bool pc_stale = false;      // rho_b changed since Q was built: rebuilt when next asked for
    // kernel-family switches, read from the environment when the plan is built: regpass 0 keeps the
    // LDS Stockham pass on strided axes (TMI_FFT_REGPASS=0), 1 uses the one-wave register pass where it
    // has enough waves, 2 wherever it has radices (TMI_FFT_REGPASS=2); xreg false keeps the LDS form of
    // the fused x pass (TMI_FFT_XREG=0); xcdmap 0 keeps the plain workgroup order of the register
    // passes instead of the XCD-contiguous one (TMI_FFT_XCDMAP=0)
    int regpass = 1;
    bool xreg = true;
    int xcdmap = 1;
    // adjoint fields of the Euler-Lagrange product on a wrapping grid: false the per-component fields
    // A_c, B_c,d and S3 (C (1 + D) + 1), true the derivatives of Phi themselves in the rho-bar order
    // (n3 n2 n1 n0 nv2 nv1: 4 + 2 D), whichever makes fewer pairs; the product then forms
    // A_c = D2 + D1/(4 pi R_c) + D0/(4 pi R_c^2) and B_c,d = Dv2_d + Dv1_d/(4 pi R_c) per wave number
    bool dform = false;
};

namespace {

using namespace dev;

constexpr int kFftBlock = 256;
constexpr size_t kFftLds = 48 * 1024;
constexpr int kFftMaxLen = 768;   // longest FFT axis: its twiddles and 2 lines of complex doubles fit kFftLds

// Stockham stages per axis in the order they run: radix 4, then 2, then 3, then 5
struct RadixCounts {
    int n4, n2, n3, n5;
};

template <typename T>
__device__ inline Cx<T> cmul(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ inline Cx<T> cadd(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ inline Cx<T> csub(Cx<T> a, Cx<T> b)
{
    return Cx<T>{a.x - b.x, a.y - b.y};
}
template <typename T>
__device__ inline Cx<T> cscale(Cx<T> a, T s)
{
    return Cx<T>{a.x * s, a.y * s};
}
// i * a * SIGN
template <int SIGN, typename T>
__device__ inline Cx<T> cis(Cx<T> a)
{
    if constexpr (SIGN > 0)
        return Cx<T>{-a.y, a.x};
    else
        return Cx<T>{a.y, -a.x};
}

// X_p = sum_q a_q exp(SIGN 2 pi i p q / R), in place, in registers
template <int R, int SIGN, typename T>
__device__ inline void dft_small(Cx<T> *a)
{
    if constexpr (R == 2) {
        const Cx<T> t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if constexpr (R == 4) {
        const Cx<T> t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        const Cx<T> t2 = cadd(a[1], a[3]), t3 = cis<SIGN>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    } else if constexpr (R == 3) {
        const T h = (T)0.86602540378443864676;   // sqrt(3)/2
        const Cx<T> s = cadd(a[1], a[2]), d = cis<SIGN>(cscale(csub(a[1], a[2]), h));
        const Cx<T> m = csub(a[0], cscale(s, (T)0.5));
        a[0] = cadd(a[0], s);
        a[1] = cadd(m, d);
        a[2] = csub(m, d);
    } else {
        static_assert(R == 5, "radix 2, 3, 4 or 5");
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        const Cx<T> p1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
        const Cx<T> p2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
        const Cx<T> b1 = cadd(a[0], cadd(cscale(p1, c1), cscale(p2, c2)));
        const Cx<T> b2 = cadd(a[0], cadd(cscale(p1, c2), cscale(p2, c1)));
        const Cx<T> e1 = cis<SIGN>(cadd(cscale(d1, s1), cscale(d2, s2)));
        const Cx<T> e2 = cis<SIGN>(csub(cscale(d1, s2), cscale(d2, s1)));
        a[0] = cadd(a[0], cadd(p1, p2));
        a[1] = cadd(b1, e1);
        a[4] = csub(b1, e1);
        a[2] = cadd(b2, e2);
        a[3] = csub(b2, e2);
    }
}

// q = x / d for 0 <= x < 2^22 through a float reciprocal, corrected to exact
__device__ inline int small_div(int x, int d, float inv)
{
    int q = (int)((float)x * inv);
    if ((q + 1) * d <= x) ++q;
    if (q * d > x) --q;
    return q;
}

// One radix-R Stockham stage over nl lines of length M held in LDS (row pitch
// ld): src -> dst, with Ns the length of the sub-transforms already done.
// The butterfly is a compile-time radix, so its operands stay in registers.
template <int R, int SIGN, typename T>
__device__ inline void stockham_stage(const Cx<T> *__restrict__ src, Cx<T> *__restrict__ dst, int ld, int nl, int M,
                                      int Ns, const Cx<T> *__restrict__ tw)
{
    const int nb = M / R;
    const int tstep = M / (Ns * R);   // twiddle exp(SIGN 2 pi i r k / (Ns R)) = tw[r k tstep]
    const float inv_nb = 1.0f / (float)nb, inv_ns = 1.0f / (float)Ns;
    for (int t = threadIdx.x; t < nl * nb; t += kFftBlock) {
        const int i = small_div(t, nb, inv_nb), jj = t - i * nb;
        const int q = small_div(jj, Ns, inv_ns), k = jj - q * Ns;
        const Cx<T> *a = src + (size_t)i * ld + jj;
        Cx<T> v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = a[r * nb];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                Cx<T> w = tw[r * k * tstep];
                if constexpr (SIGN > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        dft_small<R, SIGN>(v);
        Cx<T> *b = dst + (size_t)i * ld + q * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) b[r * Ns] = v[r];
    }
}

// Every Stockham stage of nl lines of length M held in LDS (row pitch ld),
// ping-ponging between a and b; returns the buffer holding the transform.
template <int SIGN, typename T>
__device__ inline Cx<T> *fft_lines(Cx<T> *a, Cx<T> *b, int ld, int nl, int M, RadixCounts rc, const Cx<T> *tw)
{
    int Ns = 1;
    auto run = [&](auto radix, int count) {
        constexpr int R = decltype(radix)::value;
        for (int s = 0; s < count; ++s) {
            stockham_stage<R, SIGN>(a, b, ld, nl, M, Ns, tw);
            __syncthreads();
            Cx<T> *tmp = a;
            a = b;
            b = tmp;
            Ns *= R;
        }
    };
    run(std::integral_constant<int, 4>{}, rc.n4);
    run(std::integral_constant<int, 2>{}, rc.n2);
    run(std::integral_constant<int, 3>{}, rc.n3);
    run(std::integral_constant<int, 5>{}, rc.n5);
    return a;
}

// non-temporal 16-byte store of a complex value (written once, read by a later kernel)
template <typename T>
__device__ __forceinline__ void nt_store(Cx<T> *p, Cx<T> v)
{
    typedef T V2 __attribute__((ext_vector_type(2)));
    V2 t;
    t.x = v.x;
    t.y = v.y;
    __builtin_nontemporal_store(t, reinterpret_cast<V2 *>(p));
}

// Fused ends of a product on a wrapping grid (grid point = node), for the
// strided passes: kPassGather reads the density pairs of the first forward
// pass straight from the node arrays (0 where Zero_density_TF); kPassRows
// writes the rows of the last inverse pass straight from the output pairs:
// output o feeds rows row[o][0..nrow[o]) with factors scale[o][.], rho-bar
// rows (< nrb) minus in_rb and sub_rb, mean-field rows minus in_mf and sub_mf.
// kPassEl writes the Euler-Lagrange rows of components 2 field and 2 field + 1
// with their local terms, as k_el_extract.
constexpr int kPassPlain = 0, kPassGather = 1, kPassRows = 2, kPassEl = 3, kPassPc = 4;
struct PassIO {
    const double *rho = nullptr;
    const unsigned char *zero = nullptr;
    int C = 0, nx0 = 1, nx1 = 1, N = 0;
    int nout = 0, nrb = 0;
    int nrow[kMaxOut] = {};
    int row[kMaxOut][3] = {};
    double scale[kMaxOut][3] = {};
    const double *in_rb = nullptr, *sub_rb = nullptr, *in_mf = nullptr, *sub_mf = nullptr;
    double *out_rb = nullptr, *out_mf = nullptr;
    // kPassEl
    const double *x_rho = nullptr, *el_in_rho = nullptr, *el_in_mf = nullptr, *vext = nullptr, *add = nullptr;
    const unsigned char *elzero = nullptr;
    double *out_el = nullptr;
    double mu[kNC] = {};
    int linear = 0, nwall = 0;
    // kPassRows, kPassEl: per-workgroup partial sums of the squares of the rows written
    double *ss = nullptr;
    // kPassGather: optional factor of the gathered values; kPassPc: the rows kept where elzero is set
    const double *mul = nullptr, *pv = nullptr;
    int xcd = 1;   // register passes: XCD-contiguous workgroup order (xcd_item)
};

// sum over the workgroup (kFftBlock threads), valid in thread 0
__device__ inline double fft_block_sum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kFftBlock / 64; ++w) t += sh[w];
    return t;
}

// Element j of line i of a strided-axis group (lines = consecutive x from x0
// at the other strided coordinate `other`): storage index base0 + i + j step;
// on a wrapping grid also node (x0 + i, y, z).
struct StridedGeo {
    long long base0, step;
    int x0, other, axis;
};

__device__ inline int strided_node(const StridedGeo &g, const PassIO &io, int i, int j)
{
    const int x = g.x0 + i, y = g.axis == 1 ? j : g.other, z = g.axis == 1 ? g.other : j;
    return x + io.nx0 * (y + io.nx1 * z);
}

template <typename T, int MODE>
__device__ inline Cx<T> strided_load(const Cx<T> *__restrict__ data, const StridedGeo &g, const PassIO &io, int field,
                                     int i, int j)
{
    if constexpr (MODE == kPassGather) {   // every load unconditional: none waits for another
        const size_t a = (size_t)(2 * field) * io.N + strided_node(g, io, i, j), b = a + io.N;
        const bool has_b = 2 * field + 1 < io.C;
        const double fa = io.mul ? io.mul[a] : 1.0, fb = (io.mul && has_b) ? io.mul[b] : 1.0;
        const double ra = io.rho[a], rb = has_b ? io.rho[b] : 0.0;
        const unsigned char za = io.zero[a], zb = has_b ? io.zero[b] : 1;
        const T re = za ? (T)0 : (T)(ra * fa);
        const T im = zb ? (T)0 : (T)(rb * fb);
        return Cx<T>{re, im};
    } else {
        return data[g.base0 + i + (long long)j * g.step];
    }
}

// store element (i, j) of pair field `field`; kPassRows / kPassEl add the squares of the rows written to ssq
template <typename T, int MODE>
__device__ inline void strided_store(Cx<T> *__restrict__ data, const StridedGeo &g, const PassIO &io, int field, int i,
                                     int j, Cx<T> v, double &ssq)
{
    if constexpr (MODE == kPassEl) {
        const int nd = strided_node(g, io, i, j);
        for (int h = 0; h < 2; ++h) {
            const int c = 2 * field + h;
            if (c >= io.C) break;
            const size_t row = (size_t)c * io.N + nd;
            double r;
            if (io.elzero[row]) {
                r = io.el_in_rho ? io.el_in_rho[row] : 0.0;
            } else {
                r = (double)(h ? v.y : v.x);
                const double xv = io.x_rho[row];
                if (!io.linear) {
                    r += log(xv) + io.mu[c];
                    if (io.nwall > 0) r += io.vext[row];
                    if (io.el_in_mf) r += io.el_in_mf[row];
                } else {
                    if (io.el_in_rho) r += io.el_in_rho[row] / xv;
                    if (io.el_in_mf) r += io.el_in_mf[row];
                }
            }
            if (io.add) r += io.add[row];
            io.out_el[row] = r;
            ssq += r * r;
        }
    } else if constexpr (MODE == kPassPc) {
        const int nd = strided_node(g, io, i, j);
        for (int h = 0; h < 2; ++h) {
            const int c = 2 * field + h;
            if (c >= io.C) break;
            const size_t row = (size_t)c * io.N + nd;
            io.out_el[row] = io.elzero[row] ? io.pv[row] : (double)(h ? v.y : v.x);
        }
    } else if constexpr (MODE == kPassRows) {
        const int nd = strided_node(g, io, i, j);
        for (int h = 0; h < 2; ++h) {
            const int o = 2 * field + h;
            if (o >= io.nout) break;
            const double val = (double)(h ? v.y : v.x);
            for (int r = 0; r < io.nrow[o]; ++r) {
                const int row = io.row[o][r];
                double out = val * io.scale[o][r];
                if (row < io.nrb) {
                    const size_t at = (size_t)row * io.N + nd;
                    if (io.in_rb) out -= io.in_rb[at];
                    if (io.sub_rb) out -= io.sub_rb[at];
                    io.out_rb[at] = out;
                } else {
                    const size_t at = (size_t)(row - io.nrb) * io.N + nd;
                    if (io.in_mf) out -= io.in_mf[at];
                    if (io.sub_mf) out -= io.sub_mf[at];
                    io.out_mf[at] = out;
                }
                ssq += out * out;
            }
        }
    } else {
        data[g.base0 + i + (long long)j * g.step] = v;
    }
}

// The R2 values u[k2] (elements k1 + R1 k2 of line i) of one phase of a register pass, stored as
// strided_store does; for the row-writing ends the node-side inputs of all R2 values of an output
// (or component) are loaded before any of them is stored, so R2 loads are in flight at once (the
// compiler cannot move a load past a store to a pointer it may alias).
template <typename T, int MODE, int R1, int R2>
__device__ inline void strided_store_phase(Cx<T> *__restrict__ data, const StridedGeo &g, const PassIO &io, int field,
                                           int i, int k1, const Cx<T> (&u)[R2], double &ssq)
{
    if constexpr (MODE == kPassRows) {
        int nd[R2];
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) nd[k2] = strided_node(g, io, i, k1 + R1 * k2);
#pragma unroll
        for (int h = 0; h < 2; ++h) {   // compile-time h: u[k2].x / .y stay in registers
            const int o = 2 * field + h;
            if (o >= io.nout) break;
            for (int r = 0; r < io.nrow[o]; ++r) {
                const int row = io.row[o][r];
                const bool rb = row < io.nrb;
                const size_t base = rb ? (size_t)row * io.N : (size_t)(row - io.nrb) * io.N;
                const double *in = rb ? io.in_rb : io.in_mf, *sub = rb ? io.sub_rb : io.sub_mf;
                double *out = rb ? io.out_rb : io.out_mf;
                double pin[R2], psub[R2];
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {   // read once, written rows never re-read here: non-temporal
                    pin[k2] = in ? __builtin_nontemporal_load(in + base + nd[k2]) : 0.0;
                    psub[k2] = sub ? __builtin_nontemporal_load(sub + base + nd[k2]) : 0.0;
                }
                const double sc = io.scale[o][r];
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) {   // the operations and their order of strided_store
                    double val = (double)(h ? u[k2].y : u[k2].x) * sc;
                    if (in) val -= pin[k2];
                    if (sub) val -= psub[k2];
                    __builtin_nontemporal_store(val, out + base + nd[k2]);
                    ssq += val * val;
                }
            }
        }
    } else if constexpr (MODE == kPassEl) {
        int nd[R2];
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) nd[k2] = strided_node(g, io, i, k1 + R1 * k2);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = 2 * field + h;
            if (c >= io.C) break;
            const size_t base = (size_t)c * io.N;
            unsigned char zr[R2];
            double xv[R2], inr[R2], vx[R2], mf[R2], ad[R2];
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {
                const size_t row = base + nd[k2];
                zr[k2] = io.elzero[row];
                xv[k2] = io.x_rho[row];
                inr[k2] = io.el_in_rho ? io.el_in_rho[row] : 0.0;
                vx[k2] = (!io.linear && io.nwall > 0) ? io.vext[row] : 0.0;
                mf[k2] = io.el_in_mf ? io.el_in_mf[row] : 0.0;
                ad[k2] = io.add ? io.add[row] : 0.0;
            }
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) {   // the operations and their order of strided_store
                double r;
                if (zr[k2]) {
                    r = inr[k2];
                } else {
                    r = (double)(h ? u[k2].y : u[k2].x);
                    if (!io.linear) {
                        r += log(xv[k2]) + io.mu[c];
                        if (io.nwall > 0) r += vx[k2];
                        if (io.el_in_mf) r += mf[k2];
                    } else {
                        if (io.el_in_rho) r += inr[k2] / xv[k2];
                        if (io.el_in_mf) r += mf[k2];
                    }
                }
                if (io.add) r += ad[k2];
                io.out_el[base + nd[k2]] = r;
                ssq += r * r;
            }
        }
    } else {
#pragma unroll
        for (int k2 = 0; k2 < R2; ++k2) strided_store<T, MODE>(data, g, io, field, i, k1 + R1 * k2, u[k2], ssq);
    }
}

// One axis of a 3D transform of a batch of fields, in place.  A workgroup
// takes lpb lines of the axis (consecutive lines along x for the strided
// axes, so that the loads are coalesced), runs every Stockham stage in LDS and
// writes back.
template <typename T, int SIGN, int MODE>
__global__ __launch_bounds__(kFftBlock) void k_fft_axis(Cx<T> *__restrict__ data, long long fstride, int M0, int M1,
                                                        int M2, long long S1, long long S2, int axis, RadixCounts rc,
                                                        int lpb, const Cx<T> *__restrict__ twtab, PassIO io)
{
    extern __shared__ char smem_raw[];
    const int field = blockIdx.y;
    data += (long long)field * fstride;   // one field of the batch per grid row
    const int M = axis == 0 ? M0 : (axis == 1 ? M1 : M2);
    const int ld = M + 1;
    Cx<T> *tw = reinterpret_cast<Cx<T> *>(smem_raw);   // [M] twiddles exp(-2 pi i j / M)
    Cx<T> *bufA = tw + M;
    Cx<T> *bufB = bufA + (size_t)lpb * ld;
    for (int j = threadIdx.x; j < M; j += kFftBlock) tw[j] = twtab[j];
    // line geometry
    long long base0 = 0;
    long long step = 1;   // between elements of a line (strided axes)
    int nl;               // lines in this group
    int l0 = 0;           // first line of the group (axis 0: line = y + M1 z)
    int x0 = 0, other = 0;
    if (axis == 0) {
        const int nlines = M1 * M2;
        l0 = blockIdx.x * lpb;
        nl = (nlines - l0) < lpb ? (nlines - l0) : lpb;
    } else {
        const int xg = (M0 + lpb - 1) / lpb;
        other = blockIdx.x / xg;
        x0 = (blockIdx.x % xg) * lpb;
        nl = (M0 - x0) < lpb ? (M0 - x0) : lpb;
        if (axis == 1) {
            base0 = x0 + S2 * other;
            step = S1;
        } else {
            base0 = x0 + S1 * other;
            step = S2;
        }
    }
    const float inv_m = 1.0f / (float)M, inv_nl = 1.0f / (float)nl, inv_m1 = 1.0f / (float)M1;
    // storage index of element j of line i of an axis-0 group
    auto row0 = [&](int i) {
        const int l = l0 + i, z = small_div(l, M1, inv_m1);
        return (long long)(l - z * M1) * S1 + (long long)z * S2;
    };
    const StridedGeo sg{base0, step, x0, other, axis};
    // load
    if (axis == 0) {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int i = small_div(e, M, inv_m), j = e - i * M;
            bufA[i * ld + j] = data[row0(i) + j];
        }
    } else {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int j = small_div(e, nl, inv_nl), i = e - j * nl;
            bufA[i * ld + j] = strided_load<T, MODE>(data, sg, io, field, i, j);
        }
    }
    __syncthreads();
    bufA = fft_lines<SIGN>(bufA, bufB, ld, nl, M, rc, tw);
    // store
    double ssq = 0.0;   // kPassRows, kPassEl: squares of the rows written
    if (axis == 0) {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int i = small_div(e, M, inv_m), j = e - i * M;
            data[row0(i) + j] = bufA[i * ld + j];
        }
    } else {
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int j = small_div(e, nl, inv_nl), i = e - j * nl;
            strided_store<T, MODE>(data, sg, io, field, i, j, bufA[i * ld + j], ssq);
        }
    }
    if constexpr (MODE == kPassRows || MODE == kPassEl) {
        if (io.ss) {
            __shared__ double sh[kFftBlock / 64];
            const double t = fft_block_sum(ssq, sh);
            if (threadIdx.x == 0) io.ss[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
        }
    }
}

// exp(SIGN 2 pi i k / R) for R dividing 16 (k a compile-time index after unrolling)
template <int R, int SIGN, typename T>
__device__ __forceinline__ Cx<T> unit_root(int k)
{
    constexpr double c16[16] = {1.0,  0.92387953251128675613,  0.70710678118654752440,  0.38268343236508977173,
                                0.0,  -0.38268343236508977173, -0.70710678118654752440, -0.92387953251128675613,
                                -1.0, -0.92387953251128675613, -0.70710678118654752440, -0.38268343236508977173,
                                0.0,  0.38268343236508977173,  0.70710678118654752440,  0.92387953251128675613};
    const int q = (k * (16 / R)) & 15;
    return Cx<T>{(T)c16[q], (T)(SIGN * c16[(q + 12) & 15])};   // sin(a) = cos(a - pi/2)
}

// X_p = sum_q a_q exp(SIGN 2 pi i p q / R), R = 2, 4, 8, 16, in registers (radix-2 splits down to 4)
template <int R, int SIGN, typename T>
__device__ __forceinline__ void dft_reg(Cx<T> *a)
{
    if constexpr (R <= 4) {
        dft_small<R, SIGN>(a);
    } else {
        constexpr int H = R / 2;
        Cx<T> e[H], o[H];
#pragma unroll
        for (int k = 0; k < H; ++k) {
            e[k] = a[2 * k];
            o[k] = a[2 * k + 1];
        }
        dft_reg<H, SIGN>(e);
        dft_reg<H, SIGN>(o);
#pragma unroll
        for (int k = 0; k < H; ++k) {
            const Cx<T> t = cmul(o[k], unit_root<R, SIGN, T>(k));
            a[k] = cadd(e[k], t);
            a[k + H] = csub(e[k], t);
        }
    }
}

// Workgroup b of nb -> its work item.  Workgroups are dispatched to the 8 XCDs round robin (b mod
// 8), each XCD with its own L2; here the workgroups of one XCD take one contiguous range of items,
// so neighbouring items (adjacent x blocks whose real rows share 128-byte lines, the lines that
// share a spectrum-table row) meet in one L2.
__device__ inline int xcd_item(int b, int nb)
{
    const int x = b & 7, q = nb >> 3, r = nb & 7;
    return x * q + min(x, r) + (b >> 3);
}

// Strided axis pass of length M = R1 R2 (R2 <= R1 <= 16, both powers of two) with one wave per
// workgroup: lane (i, n2) takes line i of L = 64 / R2 consecutive x and loads elements
// R2 n1 + n2 straight into registers; an R1-point DFT, the twiddles W_M^(n2 k1), a transpose
// through LDS, R1/R2 R2-point DFTs per lane and the store.  One LDS round trip per pass instead
// of one per Stockham stage.  The transpose runs in R1/R2 phases of R2 values of k1 each
// (phase q: k1 = R2 q .. R2 q + R2 - 1, read back by lane n2 as k1 = n2 + R2 q), so the LDS
// buffer holds L R2^2 values whatever R1: more resident waves per CU.
template <typename T, int SIGN, int MODE, int R1, int R2>
__global__ __launch_bounds__(64) void k_fft_strided(Cx<T> *__restrict__ data, long long fstride, int M0, int M1, int M2,
                                                    long long S1, long long S2, int axis,
                                                    const Cx<T> *__restrict__ twtab, PassIO io)
{
    constexpr int M = R1 * R2, L = 64 / R2, LD = R2 * R2 + 1;
    __shared__ Cx<T> tw[M];
    __shared__ Cx<T> buf[L * LD];
    // XCD-contiguous order within each field (every XCD gets an equal share of every field)
    const int field = blockIdx.y, bx = io.xcd ? xcd_item(blockIdx.x, gridDim.x) : (int)blockIdx.x;
    const int item = field * gridDim.x + bx;
    data += (long long)field * fstride;
    for (int j = threadIdx.x; j < M; j += 64) tw[j] = twtab[j];
    const int xg = (M0 + L - 1) / L;
    const int other = bx / xg, x0 = (bx % xg) * L;
    const StridedGeo sg{axis == 1 ? x0 + S2 * other : x0 + S1 * other, axis == 1 ? S1 : S2, x0, other, axis};
    const int i = threadIdx.x % L, n2 = threadIdx.x / L;
    const bool live = x0 + i < M0;
    __syncthreads();
    Cx<T> v[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1)
        v[n1] = live ? strided_load<T, MODE>(data, sg, io, field, i, R2 * n1 + n2) : Cx<T>{(T)0, (T)0};
    dft_reg<R1, SIGN>(v);
    double ssq = 0.0;
#pragma unroll
    for (int q = 0; q < R1 / R2; ++q) {
        if (q > 0) __syncthreads();   // the previous phase's reads are done
#pragma unroll
        for (int t = 0; t < R2; ++t) {   // twiddles W_M^(n2 k1)
            const int k1 = R2 * q + t;
            Cx<T> w = tw[n2 * k1];
            if constexpr (SIGN > 0) w.y = -w.y;
            buf[i * LD + t * R2 + n2] = cmul(v[k1], w);
        }
        __syncthreads();
        const int k1 = n2 + R2 * q;
        Cx<T> u[R2];
#pragma unroll
        for (int m = 0; m < R2; ++m) u[m] = buf[i * LD + n2 * R2 + m];
        dft_reg<R2, SIGN>(u);
        if (live) strided_store_phase<T, MODE, R1, R2>(data, sg, io, field, i, k1, u, ssq);
    }
    if constexpr (MODE == kPassRows || MODE == kPassEl) {
        if (io.ss) {
            for (int o = 32; o > 0; o >>= 1) ssq += __shfl_down(ssq, o, 64);
            if (threadIdx.x == 0) io.ss[item] = ssq;
        }
    }
}

// The x axis (contiguous lines) as a register pass of length M = R1 R2 with one wave per
// workgroup: lane (i, n2) takes line i of L = 64 / R2 consecutive x-lines, loads elements
// R2 n1 + n2 (R2 lanes read R2 consecutive elements), an R1-point DFT, the twiddles W_M^(n2 k1),
// and, in R1/R2 phases of R2 values of k1 through an LDS buffer of L R2^2 values (as
// k_fft_strided), the R2-point DFTs; stores k1 + R1 k2.  Every load of the workgroup precedes
// its first barrier, so the pass runs in place.
template <typename T, int SIGN, int R1, int R2>
__global__ __launch_bounds__(64) void k_fft_xline(Cx<T> *__restrict__ data, long long fstride, int M1, int M2,
                                                  long long S1, long long S2, const Cx<T> *__restrict__ twtab)
{
    constexpr int M = R1 * R2, L = 64 / R2, LD = R2 * R2 + 1;
    __shared__ Cx<T> tw[M];
    __shared__ Cx<T> buf[L * LD];
    data += (long long)blockIdx.y * fstride;
    for (int j = threadIdx.x; j < M; j += 64) tw[j] = twtab[j];
    const int i = threadIdx.x / R2, n2 = threadIdx.x - i * R2;
    const int l = blockIdx.x * L + i;
    const bool live = l < M1 * M2;
    const int z = live ? l / M1 : 0, y = live ? l - z * M1 : 0;
    const long long row = (long long)y * S1 + (long long)z * S2;
    Cx<T> v[R1];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) v[n1] = live ? data[row + R2 * n1 + n2] : Cx<T>{(T)0, (T)0};
    __syncthreads();   // tw
    dft_reg<R1, SIGN>(v);
#pragma unroll
    for (int q = 0; q < R1 / R2; ++q) {
        if (q > 0) __syncthreads();   // the previous phase's reads are done
#pragma unroll
        for (int t = 0; t < R2; ++t) {
            const int k1 = R2 * q + t;
            Cx<T> w = tw[n2 * k1];
            if constexpr (SIGN > 0) w.y = -w.y;
            buf[i * LD + t * R2 + n2] = cmul(v[k1], w);
        }
        __syncthreads();
        const int k1 = n2 + R2 * q;
        Cx<T> u[R2];
#pragma unroll
        for (int m = 0; m < R2; ++m) u[m] = buf[i * LD + n2 * R2 + m];
        dft_reg<R2, SIGN>(u);
        if (live) {
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) data[row + k1 + R1 * k2] = u[k2];
        }
    }
}

// Layout of the spectrum tables.  Folded (fold = 1): the table of a stencil that is symmetric
// under the reflection of each axis is even in every k_a (theta, delta, attraction) or odd in
// k_d (vector delta d), so it is stored for 0 <= k_y <= M_y/2 and 0 <= k_z <= M_z/2 only (whole
// x rows, which keep the x index affine in the fused x pass) and T(k) = s T(fold(k)) with
// fold(k_a) = min(k_a, M_a - k_a) for a = y, z and s = -1 when the table is odd in y or z and
// that k_a > M_a/2.  Unfolded: the transforms' layout.
struct TabGeo {
    int fold, M[3];
    long long p1, p2, tn;   // pitches of y rows and z planes, elements per table
};
__device__ inline bool tab_neg(const TabGeo &t, int a, int k) { return a > 0 && t.fold && 2 * k > t.M[a]; }
__device__ inline int tab_fold(const TabGeo &t, int a, int k) { return tab_neg(t, a, k) ? t.M[a] - k : k; }

struct FftGeo {
    int M[3], off[3];
    int lo[3], lim[3], src[3];
    long long n, S1, S2, ns;
    TabGeo tg;
};

// storage index of FFT-grid point m and of its negative -m; table index of m and the signs of
// an odd table's value there (bit d of neg: the table odd in axis d changes sign)
struct KIdx {
    long long q, qn, t;
    int neg;
};
__device__ inline KIdx k_index(const FftGeo &f, long long m)
{
    const int m0 = (int)(m % f.M[0]);
    const long long r = m / f.M[0];
    const int m1 = (int)(r % f.M[1]), m2 = (int)(r / f.M[1]);
    const int n0 = m0 ? f.M[0] - m0 : 0, n1 = m1 ? f.M[1] - m1 : 0, n2 = m2 ? f.M[2] - m2 : 0;
    const TabGeo &t = f.tg;
    return KIdx{m0 + m1 * f.S1 + m2 * f.S2, n0 + n1 * f.S1 + n2 * f.S2,
                tab_fold(t, 0, m0) + tab_fold(t, 1, m1) * t.p1 + tab_fold(t, 2, m2) * t.p2,
                (tab_neg(t, 0, m0) ? 1 : 0) | (tab_neg(t, 1, m1) ? 2 : 0) | (tab_neg(t, 2, m2) ? 4 : 0)};
}

// spectrum of the real field carried as the real (odd = 0) or imaginary
// (odd = 1) part of a pair transform, from Z(k) = a and Z(-k) = b
template <typename T>
__device__ inline Cx<T> unpack(Cx<T> a, Cx<T> b, int odd)
{
    return odd ? Cx<T>{(T)0.5 * (a.y + b.y), (T)0.5 * (b.x - a.x)} : Cx<T>{(T)0.5 * (a.x + b.x), (T)0.5 * (a.y - b.y)};
}

// Z_j = f_2j + i f_2j+1 at every FFT-grid position (0 off the padded copy)
template <typename T>
__global__ void k_gather_pairs(FftGeo f, Geo g, const T *__restrict__ pad, int nfield, Cx<T> *__restrict__ Z)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= f.n) return;
    const int m0 = (int)(m % f.M[0]);
    const long long r = m / f.M[0];
    const int m1 = (int)(r % f.M[1]), m2 = (int)(r / f.M[1]);
    const bool in = m0 >= f.lo[0] && m0 < f.lim[0] && m1 >= f.lo[1] && m1 < f.lim[1] && m2 >= f.lo[2] && m2 < f.lim[2];
    const long long p = in ? (long long)(m0 + f.src[0]) + (long long)(m1 + f.src[1]) * g.P[0] +
                                 (long long)(m2 + f.src[2]) * g.P[0] * g.P[1]
                           : 0;
    const long long q = m0 + m1 * f.S1 + m2 * f.S2;
    for (int j = 0; 2 * j < nfield; ++j) {
        const T re = in ? pad[(size_t)(2 * j) * g.np + p] : (T)0;
        const T im = (in && 2 * j + 1 < nfield) ? pad[(size_t)(2 * j + 1) * g.np + p] : (T)0;
        Z[(size_t)j * f.ns + q] = Cx<T>{re, im};
    }
}

// Z_j = rho_2j + i rho_2j+1 from the node arrays (0 where Zero_density_TF) on a
// wrapping FFT grid (grid point = node): the density halo is never read there
template <typename T>
__global__ void k_gather_nodes(FftGeo f, Geo g, const double *__restrict__ in, const unsigned char *__restrict__ zero,
                               Cx<T> *__restrict__ Z)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.N) return;
    const int x = i % g.nx[0];
    const int r = i / g.nx[0];
    const int y = r % g.nx[1], z = r / g.nx[1];
    const long long q = x + y * f.S1 + z * f.S2;
    for (int j = 0; 2 * j < g.C; ++j) {
        const size_t a = (size_t)(2 * j) * g.N + i, b = a + g.N;
        const T re = zero[a] ? (T)0 : (T)in[a];
        const T im = (2 * j + 1 < g.C && !zero[b]) ? (T)in[b] : (T)0;
        Z[(size_t)j * f.ns + q] = Cx<T>{re, im};
    }
}

// Products of the density spectra (pairs Z) with the stencil spectra, the
// outputs packed in pairs O_j = Y_2j + i Y_2j+1
template <typename T, int ND>
__global__ void k_fwd_mul(FftGeo f, FwdDesc d, const Cx<T> *__restrict__ Z, const T *__restrict__ tk,
                          const T *__restrict__ uk, Cx<T> *__restrict__ O)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= f.n) return;
    const KIdx ix = k_index(f, m);
    Cx<T> X[kNC];
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        X[c] = Cx<T>{(T)0, (T)0};
        if (c < d.C) {
            const Cx<T> *z = Z + (size_t)(c >> 1) * f.ns;
            X[c] = unpack(z[ix.q], z[ix.qn], c & 1);
        }
    }
    Cx<T> acc{(T)0, (T)0};
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o) {
        if (o < d.nout) {
            const int kind = d.kind[o];
            Cx<T> y{(T)0, (T)0};
            if (kind < kKindMf) {
#pragma unroll
                for (int c = 0; c < kNC; ++c) {
                    if (c < d.C) {
                        T w = (T)d.coef[o][c] * tk[((size_t)c * (2 + ND) + kind) * f.tg.tn + ix.t];
                        if (kind >= 2 && ((ix.neg >> (kind - 2)) & 1)) w = -w;
                        if (kind < 2)
                            y = cadd(y, cscale(X[c], w));
                        else   // imaginary spectrum: (i w) X
                            y = cadd(y, Cx<T>{-w * X[c].y, w * X[c].x});
                    }
                }
            } else {
                const int row = kind - kKindMf;
#pragma unroll
                for (int c = 0; c < kNC; ++c)
                    if (c < d.C) y = cadd(y, cscale(X[c], uk[(size_t)(row + d.C * c) * f.tg.tn + ix.t]));
            }
            if ((o & 1) == 0) {
                acc = y;
                if (o + 1 == d.nout) O[(size_t)(o >> 1) * f.ns + ix.q] = acc;
            } else {
                O[(size_t)(o >> 1) * f.ns + ix.q] = Cx<T>{acc.x - y.y, acc.y + y.x};
            }
        }
    }
}

// The products in groups of 8 wave numbers {(+-kx, +-ky, +-kz)} (folded tables only): thread
// (kx, ky, kz), 0 <= k_a <= M_a/2, takes the whole group, which holds the negative of each of its
// points (point s, bit a set where axis a is negated; -point(s) = point(s ^ 7)) and whose table
// values sit in one folded row (ky, kz) at the columns kx and -kx.  Every input point, table
// element and output point is then read or written once (points repeat where k_a = 0 or M_a/2:
// the same values are computed and stored twice by one thread).
struct Group8 {
    int q[8];         // storage index of point s (a field holds fewer than 2^31 elements when tables fold)
    long long trow;   // folded table row
    int tx[2];        // table columns of +kx and -kx
    int nyz;          // bit 1: the y of the points with s bit 1 set folds from above M_y/2; bit 2: the same for z
    // bit d of the result: an odd table in axis d changes sign at point s (d = 1, 2; x is never folded)
    __device__ int neg(int s) const { return nyz & (((s >> 1) & 1) << 1 | ((s >> 2) & 1) << 2); }
};
__device__ inline bool group8(const FftGeo &f, long long t, Group8 &g)
{
    const int hx = f.M[0] / 2, hy = f.M[1] / 2, hz = f.M[2] / 2;
    const long long nxr = hx + 1, nyr = hy + 1;
    if (t >= nxr * nyr * (hz + 1)) return false;
    const int kx = (int)(t % nxr);
    const long long r = t / nxr;
    const int ky = (int)(r % nyr), kz = (int)(r / nyr);
    const int xs[2] = {kx, kx ? f.M[0] - kx : 0}, ys[2] = {ky, ky ? f.M[1] - ky : 0}, zs[2] = {kz, kz ? f.M[2] - kz : 0};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const int y = ys[(s >> 1) & 1], z = zs[s >> 2];
        g.q[s] = (int)(xs[s & 1] + (long long)y * f.S1 + (long long)z * f.S2);
    }
    // ky, kz <= M/2 never fold; their mirrors fold when they exceed M/2
    g.nyz = (tab_neg(f.tg, 1, ys[1]) ? 2 : 0) | (tab_neg(f.tg, 2, zs[1]) ? 4 : 0);
    g.trow = (long long)ky * f.tg.p1 + (long long)kz * f.tg.p2;
    g.tx[0] = xs[0];
    g.tx[1] = xs[1];
    return true;
}

// A group of 8 split over two adjacent lanes: lane half h takes the 4 points s_j = 2h ^ S[j],
// S = {0, 7, 6, 1} (bits x, y, z of s: the axes negated).  Entries j and j ^ 1 are each other's
// negatives, and the x bit of s_j is j & 1, so entry j reads table column j & 1.  The two lanes read the same table elements in one
// instruction, so the tables still come from memory once.
struct Group4 {
    int q[4];         // storage index of point s_j
    long long trow;   // folded table row
    int tx[2];        // table columns of +kx and -kx
    int sneg[4];      // bit d: an odd table in axis d changes sign at point s_j
};
__device__ inline bool group4(const FftGeo &f, long long t, Group4 &g)
{
    const int hx = f.M[0] / 2, hy = f.M[1] / 2, hz = f.M[2] / 2;
    const long long nxr = hx + 1, nyr = hy + 1, grp = t >> 1;
    if (grp >= nxr * nyr * (hz + 1)) return false;
    const int h = (int)(t & 1);
    const int kx = (int)(grp % nxr);
    const long long r = grp / nxr;
    const int ky = (int)(r % nyr), kz = (int)(r / nyr);
    const int x0 = kx, x1 = kx ? f.M[0] - kx : 0, y1 = ky ? f.M[1] - ky : 0, z1 = kz ? f.M[2] - kz : 0;
    // the y of entry j is mirrored when bit 1 of s_j = 2h ^ S[j] is set
    const int ya = h ? y1 : ky, yb = h ? ky : y1;   // entries with S[j] bit 1 clear / set
    const bool ny = tab_neg(f.tg, 1, y1), nz = tab_neg(f.tg, 2, z1);
    const int na = (h && ny) ? 2 : 0, nb = (!h && ny) ? 2 : 0, nzb = nz ? 4 : 0;
    // S[j] as (x, y, z) bits: j = 0: (0,0,0), 1: (1,1,1), 2: (0,1,1), 3: (1,0,0); h flips y
    g.q[0] = (int)(x0 + (long long)ya * f.S1 + (long long)kz * f.S2);
    g.q[1] = (int)(x1 + (long long)yb * f.S1 + (long long)z1 * f.S2);
    g.q[2] = (int)(x0 + (long long)yb * f.S1 + (long long)z1 * f.S2);
    g.q[3] = (int)(x1 + (long long)ya * f.S1 + (long long)kz * f.S2);
    g.sneg[0] = na;
    g.sneg[1] = nb | nzb;
    g.sneg[2] = nb | nzb;
    g.sneg[3] = na;
    g.trow = (long long)ky * f.tg.p1 + (long long)kz * f.tg.p2;
    g.tx[0] = x0;
    g.tx[1] = x1;
    return true;
}
__host__ __device__ constexpr long long group4_threads(long long ngroups) { return 2 * ngroups; }

// The products over Group4 (folded tables only), one output pair (outputs o = 2 p, 2 p + 1) at a
// time: its table values for every component are loaded once, then each of the lane's 4 points
// is formed and stored.
template <typename T, int ND>
__global__ __launch_bounds__(kBlock) void k_fwd_mul4(FftGeo f, FwdDesc d, const Cx<T> *__restrict__ Z,
                                                     const T *__restrict__ tk, const T *__restrict__ uk,
                                                     Cx<T> *__restrict__ O)
{
    Group4 g;
    if (!group4(f, (long long)blockIdx.x * kBlock + threadIdx.x, g)) return;
    constexpr int NP = (kNC + 1) / 2;
    Cx<T> z[NP][4];
#pragma unroll
    for (int j = 0; j < NP; ++j)
#pragma unroll
        for (int e = 0; e < 4; ++e) z[j][e] = (2 * j < d.C) ? Z[(size_t)j * f.ns + g.q[e]] : Cx<T>{(T)0, (T)0};
    auto X = [&](int c, int e) { return unpack(z[c >> 1][e], z[c >> 1][e ^ 1], c & 1); };
    // table values of output o for component c at +kx and -kx (0 past the components)
    auto load_w = [&](int o, T (&w)[kNC][2]) {
        const int kind = d.kind[o];
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            w[c][0] = w[c][1] = (T)0;
            if (c < d.C) {
                const T *tb = kind < kKindMf ? tk + ((size_t)c * (2 + ND) + kind) * f.tg.tn + g.trow
                                             : uk + (size_t)(kind - kKindMf + d.C * c) * f.tg.tn + g.trow;
                const T cf = kind < kKindMf ? (T)d.coef[o][c] : (T)1;
                w[c][0] = cf * tb[g.tx[0]];
                w[c][1] = cf * tb[g.tx[1]];
            }
        }
    };
    // output o at entry e: the sum over components in ascending order
    auto term = [&](int kind, const T (&w)[kNC][2], int e) {
        Cx<T> y{(T)0, (T)0};
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            if (c < d.C) {
                T wv = w[c][e & 1];
                const Cx<T> xv = X(c, e);
                if (kind >= 2 && kind < kKindMf) {   // imaginary spectrum: (i w) X
                    if ((g.sneg[e] >> (kind - 2)) & 1) wv = -wv;
                    y = cadd(y, Cx<T>{-wv * xv.y, wv * xv.x});
                } else {
                    y = cadd(y, cscale(xv, wv));
                }
            }
        }
        return y;
    };
#pragma unroll 1
    for (int p = 0; 2 * p < d.nout; ++p) {
        const int oa = 2 * p, ob = 2 * p + 1;
        const bool has_b = ob < d.nout;
        T wa[kNC][2], wb[kNC][2];
        load_w(oa, wa);
        if (has_b)
            load_w(ob, wb);
        const int ka = d.kind[oa], kb = has_b ? d.kind[ob] : 0;
        Cx<T> *out = O + (size_t)p * f.ns;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const Cx<T> ya = term(ka, wa, e);
            if (has_b) {
                const Cx<T> yb = term(kb, wb, e);
                out[g.q[e]] = Cx<T>{ya.x - yb.y, ya.y + yb.x};
            } else {
                out[g.q[e]] = ya;
            }
        }
    }
}

// rho-bar and mean-field rows at the nodes: row = Y - in - sub
template <typename T>
__global__ void k_fwd_extract(FftGeo f, Geo g, RowMap rm, const Cx<T> *__restrict__ O, const double *__restrict__ in_rb,
                              const double *__restrict__ sub_rb, double *__restrict__ out_rb,
                              const double *__restrict__ in_mf, const double *__restrict__ sub_mf,
                              double *__restrict__ out_mf)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int row = blockIdx.y;
    if (i >= g.N) return;
    const int x = i % g.nx[0];
    const int r = i / g.nx[0];
    const int y = r % g.nx[1], z = r / g.nx[1];
    const long long m = (long long)(x + f.off[0]) + (long long)(y + f.off[1]) * f.S1 + (long long)(z + f.off[2]) * f.S2;
    const int o = rm.src[row];
    const Cx<T> v = O[(size_t)(o >> 1) * f.ns + m];
    double val = (double)((o & 1) ? v.y : v.x) * rm.scale[row];
    if (row < rm.nrb) {
        const size_t at = (size_t)row * g.N + i;
        if (in_rb) val -= in_rb[at];
        if (sub_rb) val -= sub_rb[at];
        out_rb[at] = val;
    } else {
        const size_t at = (size_t)(row - rm.nrb) * g.N + i;
        if (in_mf) val -= in_mf[at];
        if (sub_mf) val -= sub_mf[at];
        out_mf[at] = val;
    }
}

// Euler-Lagrange sums of the adjoint fields (pairs Z, fields A_c, B_c,d in
// the order c (1 + ND) + {0, 1 + d}, then S3):
// E_c = K_delta,c A_c - sum_d (i K'_c,d) B_c,d + K_theta,c S3, packed in pairs.
// DF: the fields are the derivatives of Phi in the rho-bar order (D3 D2 D1 D0 Dv2_d Dv1_d) and
// A_c = D2 + D1 i4pir_c + D0 i4pirsq_c, B_c,d = Dv2_d + Dv1_d i4pir_c, S3 = D3 are formed here.
struct ElConsts {
    double i4pir[kNC], i4pirsq[kNC];
};
template <typename T, int ND, bool DF>
__global__ void k_el_mul(FftGeo f, int C, ElConsts ec, const Cx<T> *__restrict__ Z, const T *__restrict__ tk,
                         Cx<T> *__restrict__ O)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= f.n) return;
    const KIdx ix = k_index(f, m);
    auto field = [&](int j) {
        const Cx<T> *z = Z + (size_t)(j >> 1) * f.ns;
        return unpack(z[ix.q], z[ix.qn], j & 1);
    };
    Cx<T> S3, D[4 + 2 * ND];
    if constexpr (DF) {
#pragma unroll
        for (int j = 0; j < 4 + 2 * ND; ++j) D[j] = field(j);
        S3 = D[0];
    } else {
        S3 = field(C * (1 + ND));
    }
    Cx<T> acc{(T)0, (T)0};
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        if (c < C) {
            const T *t = tk + (size_t)c * (2 + ND) * f.tg.tn + ix.t;
            Cx<T> A;
            if constexpr (DF)
                A = cadd(D[1], cadd(cscale(D[2], (T)ec.i4pir[c]), cscale(D[3], (T)ec.i4pirsq[c])));
            else
                A = field(c * (1 + ND));
            Cx<T> e = cadd(cscale(S3, t[0]), cscale(A, t[f.tg.tn]));
#pragma unroll
            for (int dd = 0; dd < ND; ++dd) {
                T kv = t[(size_t)(2 + dd) * f.tg.tn];
                if ((ix.neg >> dd) & 1) kv = -kv;
                Cx<T> B;
                if constexpr (DF)
                    B = cadd(D[4 + dd], cscale(D[4 + ND + dd], (T)ec.i4pir[c]));
                else
                    B = field(c * (1 + ND) + 1 + dd);
                e = cadd(e, Cx<T>{kv * B.y, -kv * B.x});   // -(i kv) B
            }
            if ((c & 1) == 0) {
                acc = e;
                if (c + 1 == C) O[(size_t)(c >> 1) * f.ns + ix.q] = acc;
            } else {
                O[(size_t)(c >> 1) * f.ns + ix.q] = Cx<T>{acc.x - e.y, acc.y + e.x};
            }
        }
    }
}

// k_el_mul over groups of 8 wave numbers (Group8; folded tables only): the table values of the
// group's row are loaded once, then the four point pairs (s, s ^ 7) are formed in turn.
template <typename T, int ND, bool DF>
__global__ __launch_bounds__(kBlock) void k_el_mul8(FftGeo f, int C, ElConsts ec, const Cx<T> *__restrict__ Z,
                                                    const T *__restrict__ tk, Cx<T> *__restrict__ O)
{
    Group8 g;
    if (!group8(f, (long long)blockIdx.x * kBlock + threadIdx.x, g)) return;
    constexpr int NT = 2 + ND;
    T tv[kNC][NT][2];
#pragma unroll
    for (int c = 0; c < kNC; ++c)
#pragma unroll
        for (int k = 0; k < NT; ++k) {
            const T *tb = tk + ((size_t)c * NT + k) * f.tg.tn + g.trow;
            tv[c][k][0] = c < C ? tb[g.tx[0]] : (T)0;
            tv[c][k][1] = c < C ? tb[g.tx[1]] : (T)0;
        }
    constexpr int NF = DF ? 4 + 2 * ND : kNC * (1 + ND) + 1;
    const int nf = DF ? 4 + 2 * ND : C * (1 + ND) + 1;
#pragma unroll
    for (int s0 = 0; s0 < 4; ++s0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {   // point s and its negative
            const int s = h ? (s0 ^ 7) : s0;
            Cx<T> fld[NF];
#pragma unroll
            for (int j = 0; j < NF; j += 2) {
                if (j < nf) {
                    const Cx<T> *zp = Z + (size_t)(j >> 1) * f.ns;
                    const Cx<T> a = zp[g.q[s]], b = zp[g.q[s ^ 7]];
                    fld[j] = unpack(a, b, 0);
                    if (j + 1 < NF) fld[j + 1] = unpack(a, b, 1);
                }
            }
            const T sg[2] = {(T)1, (T)-1};
            Cx<T> S3 = fld[0];
            if constexpr (!DF) {   // field C (1 + ND), selected with compile-time indices
#pragma unroll
                for (int j = 0; j < NF; ++j)
                    if (j == C * (1 + ND)) S3 = fld[j];
            }
            Cx<T> acc{(T)0, (T)0};
#pragma unroll
            for (int c = 0; c < kNC; ++c) {
                if (c < C) {
                    Cx<T> A;
                    if constexpr (DF)
                        A = cadd(fld[1], cadd(cscale(fld[2], (T)ec.i4pir[c]), cscale(fld[3], (T)ec.i4pirsq[c])));
                    else
                        A = fld[c * (1 + ND)];
                    Cx<T> e = cadd(cscale(S3, tv[c][0][s & 1]), cscale(A, tv[c][1][s & 1]));
#pragma unroll
                    for (int dd = 0; dd < ND; ++dd) {
                        const T kv = tv[c][2 + dd][s & 1] * sg[(g.neg(s) >> dd) & 1];
                        Cx<T> B;
                        if constexpr (DF)
                            B = cadd(fld[4 + dd], cscale(fld[4 + ND + dd], (T)ec.i4pir[c]));
                        else
                            B = fld[c * (1 + ND) + 1 + dd];
                        e = cadd(e, Cx<T>{kv * B.y, -kv * B.x});   // -(i kv) B
                    }
                    if ((c & 1) == 0) {
                        acc = e;
                        if (c + 1 == C) nt_store(O + (size_t)(c >> 1) * f.ns + g.q[s], acc);
                    } else {
                        nt_store(O + (size_t)(c >> 1) * f.ns + g.q[s], Cx<T>{acc.x - e.y, acc.y + e.x});
                    }
                }
            }
        }
    }
}

// Fused x pass of a product, after the y and z passes of the inputs and before those of the
// outputs: a workgroup takes lpb x-lines (y, z) of every input pair and their mirror lines
// (-y, -z), transforms them along x, splits each pair into its two real fields' spectra
// (X(k) from Z(k) and Z(-k)), accumulates the terms into the output pairs in LDS, transforms
// those back along x and stores them.
template <typename T>
__global__ __launch_bounds__(kFftBlock) void k_fft_xprod(const Cx<T> *__restrict__ Zin, Cx<T> *__restrict__ Oout,
                                                         long long ns, int M0, int M1, int M2, long long S1,
                                                         long long S2, RadixCounts rc, int lpb,
                                                         const Cx<T> *__restrict__ twtab, const T *__restrict__ tk,
                                                         const T *__restrict__ uk, TabGeo tg, TermList tl)
{
    extern __shared__ char smem_raw[];
    const int M = M0, ld = M + 1;
    Cx<T> *tw = reinterpret_cast<Cx<T> *>(smem_raw);
    // lines, then their mirror lines, transformed as one batch of 2 lpb lines; scratch of the same size
    Cx<T> *bA = tw + M, *bB = bA + (size_t)lpb * ld, *bS = bB + (size_t)lpb * ld;
    Cx<T> *out = bS + (size_t)2 * lpb * ld;   // [output pair][lpb][ld]
    const int nop = (tl.nout + 1) / 2, nip = (tl.nin + 1) / 2;
    for (int j = threadIdx.x; j < M; j += kFftBlock) tw[j] = twtab[j];
    for (int e = threadIdx.x; e < nop * lpb * ld; e += kFftBlock) out[e] = Cx<T>{(T)0, (T)0};
    const int l0 = blockIdx.x * lpb;
    const int nl = min(lpb, M1 * M2 - l0);
    const float inv_m = 1.0f / (float)M, inv_m1 = 1.0f / (float)M1;
    auto yz = [&](int i, int &y, int &z) {
        const int l = l0 + i;
        z = small_div(l, M1, inv_m1);
        y = l - z * M1;
    };
    for (int jp = 0; jp < nip; ++jp) {
        const Cx<T> *src = Zin + (size_t)jp * ns;
        __syncthreads();
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int i = small_div(e, M, inv_m), j = e - i * M;
            int y, z;
            yz(i, y, z);
            const int ym = y ? M1 - y : 0, zm = z ? M2 - z : 0;
            bA[i * ld + j] = src[y * S1 + z * S2 + j];
            bB[i * ld + j] = src[ym * S1 + zm * S2 + j];
        }
        __syncthreads();
        // all 2 lpb lines (rows past nl hold stale values that are never read)
        const Cx<T> *fa = fft_lines<-1>(bA, bS, ld, 2 * lpb, M, rc, tw);
        const Cx<T> *fb = fa + (size_t)lpb * ld;
        for (int e = threadIdx.x; e < nl * M; e += kFftBlock) {
            const int i = small_div(e, M, inv_m), k = e - i * M;
            int y, z;
            yz(i, y, z);
            const long long q = tab_fold(tg, 0, k) + tab_fold(tg, 1, y) * tg.p1 + tab_fold(tg, 2, z) * tg.p2;
            const int neg = (tab_neg(tg, 0, k) ? 1 : 0) | (tab_neg(tg, 1, y) ? 2 : 0) | (tab_neg(tg, 2, z) ? 4 : 0);
            const Cx<T> a = fa[i * ld + k], b = fb[i * ld + (k ? M - k : 0)];
            for (int h = 0; h < 2; ++h) {
                const int f = 2 * jp + h;
                if (f >= tl.nin) break;
                const Cx<T> X = unpack(a, b, h);
                int last = 0x7fffffff;   // terms are sorted by table: each table is read once per point
                T kt = (T)0;
                for (int t = tl.first[f]; t < tl.first[f + 1]; ++t) {
                    const Term tm = tl.term[t];
                    if (tm.t != last) {
                        kt = tm.t >= 0 ? tk[(size_t)tm.t * tg.tn + q] : uk[(size_t)(-tm.t - 1) * tg.tn + q];
                        if (tm.odd >= 0 && ((neg >> tm.odd) & 1)) kt = -kt;
                        last = tm.t;
                    }
                    const T kv = (T)tm.coef * kt;
                    const Cx<T> yv = tm.imag ? Cx<T>{-kv * X.y, kv * X.x} : Cx<T>{kv * X.x, kv * X.y};
                    Cx<T> &acc = out[((size_t)(tm.o >> 1) * lpb + i) * ld + k];
                    acc = (tm.o & 1) ? Cx<T>{acc.x - yv.y, acc.y + yv.x} : cadd(acc, yv);
                }
            }
        }
    }
    __syncthreads();
    // output pairs back along x, up to 4 pairs (4 lpb lines, the size of the input region) at a time
    for (int p0 = 0; p0 < nop; p0 += 4) {
        const int np = min(4, nop - p0);
        const Cx<T> *r = fft_lines<1>(out + (size_t)p0 * lpb * ld, bA, ld, np * lpb, M, rc, tw);
        for (int e = threadIdx.x; e < np * nl * M; e += kFftBlock) {
            const int pl = small_div(e, nl * M, 1.0f / (float)(nl * M)), r2 = e - pl * nl * M;
            const int i = small_div(r2, M, inv_m), j = r2 - i * M;
            int y, z;
            yz(i, y, z);
            Oout[(size_t)(p0 + pl) * ns + y * S1 + z * S2 + j] = r[((size_t)pl * lpb + i) * ld + j];
        }
        __syncthreads();
    }
}

// Primary x-lines of the fused register x pass, numbered l = 0, 1, ...: every line (y, z) of the
// planes 0 < z < M2 - z (z-major, y fastest), then the self-mirror planes (z = 0, and z = M2/2
// when M2 is even) with 0 <= y <= M1/2.  The mirror lines (-y, -z) of the primary lines are the
// other lines, each once, except the self-mirror lines (y, z) = (-y, -z).  False past the last.
__device__ inline bool xprod_line(int l, int M1, int M2, int &y, int &z)
{
    const int zr = (M2 - 1) / 2;
    if (l < zr * M1) {
        z = 1 + l / M1;
        y = l - (z - 1) * M1;
        return true;
    }
    l -= zr * M1;
    const int ms = M1 / 2 + 1, nself = (M2 > 1 && M2 % 2 == 0) ? 2 : 1;
    const int sp = l / ms;
    if (sp >= nself) return false;
    z = sp ? M2 / 2 : 0;
    y = l - sp * ms;
    return true;
}
inline long long xprod_primary_lines(int M1, int M2)
{
    const int zr = (M2 - 1) / 2, nself = (M2 > 1 && M2 % 2 == 0) ? 2 : 1;
    return (long long)zr * M1 + (long long)nself * (M1 / 2 + 1);
}

// Fused x pass with the x transforms in registers, for x-lines of length M = R1 R2 (R1 = 8 or 16,
// R2 = 8 or 4) and NOP output pairs whose NOP x R1 spectrum values fit in registers.  The
// workgroup has 2 L R2 threads: lanes (i, n2) of the first half take primary line i of L
// consecutive primary x-lines (y, z) (xprod_line), the second half their mirror lines (-y, -z),
// so every line is loaded once.  Per input pair every lane loads elements R2 n1 + n2 of its
// line, runs an R1-point DFT, the twiddles W_M^(n2 k1), a transpose through LDS and R1/R2
// R2-point DFTs, leaving spectrum values k = n2 + R2 q + R1 k2; those go back to LDS in natural
// order, every lane reads the partner line's spectrum at -k (the spectrum of its line's
// negative), splits the pair and accumulates the terms of its own line into output pairs held
// in registers (the table row of a line and of its mirror is the same folded row).  The outputs
// return through LDS one pair at a time, through the same two-stage transform with SIGN +1,
// and both halves store their lines (a self-mirror line once) for the y and z passes.
template <typename T, int R1, int R2, int NOP>
__global__ __launch_bounds__(2 * (64 / R2) * R2) void k_fft_xprod_reg(const Cx<T> *__restrict__ Zin,
                                                                    Cx<T> *__restrict__ Oout, long long ns, int M1,
                                                                    int M2, long long S1, long long S2, int L,
                                                                    const Cx<T> *__restrict__ twtab,
                                                                    const T *__restrict__ tk, const T *__restrict__ uk,
                                                                    TabGeo tg, TermList tl, int xcd)
{
    constexpr int M = R1 * R2, LD = M + 1;
    extern __shared__ char smem_raw[];
    Cx<T> *tw = reinterpret_cast<Cx<T> *>(smem_raw);   // [M]
    Cx<T> *buf = tw + M;                               // [2 L][LD]: primary lines, then mirror lines
    const int tid = threadIdx.x, half = tid / (L * R2), r0 = tid - half * L * R2;
    const int i = r0 / R2, n2 = r0 - i * R2;
    const int nth = 2 * L * R2;
    for (int j = tid; j < M; j += nth) tw[j] = twtab[j];
    int yp = 0, zp = 0;
    const bool live = xprod_line((xcd ? xcd_item(blockIdx.x, gridDim.x) : (int)blockIdx.x) * L + i, M1, M2, yp, zp);
    const int ym = yp ? M1 - yp : 0, zm = zp ? M2 - zp : 0;
    const bool self = ym == yp && zm == zp;
    const int y = half ? ym : yp, z = half ? zm : zp;   // this lane's line
    const long long row = (long long)y * S1 + (long long)z * S2;
    const long long trow = tab_fold(tg, 1, y) * tg.p1 + tab_fold(tg, 2, z) * tg.p2;   // the line's table row
    const bool neg_y = tab_neg(tg, 1, y), neg_z = tab_neg(tg, 2, z);
    Cx<T> *mine = buf + (size_t)(half * L + i) * LD;                 // this lane's line in LDS
    const Cx<T> *partner = buf + (size_t)((1 - half) * L + i) * LD;  // the line of the negative spectrum
    // spectrum position of register r (= q R2 + k2)
    auto kpos = [&](int r) { return n2 + R2 * (r / R2) + R1 * (r % R2); };
    Cx<T> acc[NOP][R1];
#pragma unroll
    for (int p = 0; p < NOP; ++p)
#pragma unroll
        for (int r = 0; r < R1; ++r) acc[p][r] = Cx<T>{(T)0, (T)0};
    __syncthreads();
    // forward transform of the lane's line from registers v (elements R2 n1 + n2) into X (positions kpos)
    auto forward = [&](Cx<T> (&v)[R1], Cx<T> (&X)[R1]) {
        dft_reg<R1, -1>(v);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) mine[k1 * R2 + n2] = cmul(v[k1], tw[n2 * k1]);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < R1 / R2; ++q) {
            const int k1 = n2 + R2 * q;
            Cx<T> u[R2];
#pragma unroll
            for (int m = 0; m < R2; ++m) u[m] = mine[k1 * R2 + m];
            dft_reg<R2, -1>(u);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) X[q * R2 + k2] = u[k2];
        }
        __syncthreads();
    };
    const int nip = (tl.nin + 1) / 2;
    for (int jp = 0; jp < nip; ++jp) {
        const Cx<T> *src = Zin + (size_t)jp * ns;
        Cx<T> v[R1], X[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = live ? src[row + R2 * n1 + n2] : Cx<T>{(T)0, (T)0};
        forward(v, X);
        // spectra in natural order: every half writes its own line
#pragma unroll
        for (int r = 0; r < R1; ++r) mine[kpos(r)] = X[r];
        __syncthreads();
        if (live) {
            Cx<T> Yr[R1];   // the spectrum at -k
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const int k = kpos(r);
                Yr[r] = partner[k ? M - k : 0];
            }
            for (int h = 0; h < 2; ++h) {
                const int f = 2 * jp + h;
                if (f >= tl.nin) break;
                int last = 0x7fffffff;
                T kt[R1];
                for (int t = tl.first[f]; t < tl.first[f + 1]; ++t) {
                    const Term tm = tl.term[t];
                    if (tm.t != last) {
                        const T *tab = tm.t >= 0 ? tk + (size_t)tm.t * tg.tn : uk + (size_t)(-tm.t - 1) * tg.tn;
#pragma unroll
                        for (int r = 0; r < R1; ++r) kt[r] = tab[trow + kpos(r)];
                        if ((tm.odd == 1 && neg_y) || (tm.odd == 2 && neg_z)) {
#pragma unroll
                            for (int r = 0; r < R1; ++r) kt[r] = -kt[r];
                        }
                        last = tm.t;
                    }
                    const int p = tm.o >> 1, odd = tm.o & 1;
                    const T cf = (T)tm.coef;
#pragma unroll
                    for (int r = 0; r < R1; ++r) {
                        const T kv = cf * kt[r];
                        const Cx<T> xh = unpack(X[r], Yr[r], h);
                        const Cx<T> yv = tm.imag ? Cx<T>{-kv * xh.y, kv * xh.x} : Cx<T>{kv * xh.x, kv * xh.y};
                        const Cx<T> add = odd ? Cx<T>{-yv.y, yv.x} : yv;
#pragma unroll
                        for (int pp = 0; pp < NOP; ++pp)
                            if (pp == p) acc[pp][r] = cadd(acc[pp][r], add);
                    }
                }
            }
        }
        __syncthreads();
    }
    // output pairs back along x, one at a time, every lane its own line
    const int nop = (tl.nout + 1) / 2;
    const bool store = live && (half == 0 || !self);
#pragma unroll
    for (int p = 0; p < NOP; ++p) {
        if (p >= nop) break;
#pragma unroll
        for (int r = 0; r < R1; ++r) mine[kpos(r)] = acc[p][r];
        __syncthreads();
        Cx<T> v[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) v[n1] = mine[R2 * n1 + n2];
        __syncthreads();
        dft_reg<R1, 1>(v);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            Cx<T> w = tw[n2 * k1];
            w.y = -w.y;
            mine[k1 * R2 + n2] = cmul(v[k1], w);
        }
        __syncthreads();
        if (store) {
#pragma unroll
            for (int q = 0; q < R1 / R2; ++q) {
                const int k1 = n2 + R2 * q;
                Cx<T> u[R2];
#pragma unroll
                for (int m = 0; m < R2; ++m) u[m] = mine[k1 * R2 + m];
                dft_reg<R2, 1>(u);
#pragma unroll
                for (int k2 = 0; k2 < R2; ++k2) Oout[(size_t)p * ns + row + k1 + R1 * k2] = u[k2];
            }
        }
        __syncthreads();
    }
}

// Euler-Lagrange rows from the sums E_c, as k_el_rows:
//  residual: E + ln x + mu [+ V_ext] [+ x_MF]      ; zero rows: x   (in_rho)
//  linear  : E + v/x + v_MF                        ; zero rows: v
template <typename T>
__global__ void k_el_extract(FftGeo f, Geo g, CompConsts cc, int linear, int nwall, const Cx<T> *__restrict__ O,
                             const double *__restrict__ x_rho, const double *__restrict__ in_rho,
                             const double *__restrict__ in_mf, const double *__restrict__ vext,
                             const unsigned char *__restrict__ elzero, const double *__restrict__ add,
                             double *__restrict__ out_el)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g.N) return;
    const size_t row = (size_t)c * g.N + i;
    double r;
    if (elzero[row]) {
        r = in_rho ? in_rho[row] : 0.0;
    } else {
        const int x = i % g.nx[0];
        const int rr = i / g.nx[0];
        const int y = rr % g.nx[1], z = rr / g.nx[1];
        const long long m =
            (long long)(x + f.off[0]) + (long long)(y + f.off[1]) * f.S1 + (long long)(z + f.off[2]) * f.S2;
        const Cx<T> v = O[(size_t)(c >> 1) * f.ns + m];
        r = (double)((c & 1) ? v.y : v.x);
        const double xv = x_rho[row];
        if (!linear) {
            r += log(xv) + (g.mu ? g.mu[row] : cc.mu[c]);
            if (nwall > 0) r += vext[row];
            if (in_mf) r += in_mf[row];
        } else {
            if (in_rho) r += in_rho[row] / xv;
            if (in_mf) r += in_mf[row];
        }
    }
    if (add) r += add[row];
    out_el[row] = r;
}

// Preconditioner input u_c = rho_c v_c on the free rows (0 on the zero rows), as pairs at the
// nodes' FFT-grid positions (node + off; the rest of a padded grid is zeroed beforehand)
__global__ void k_pc_gather(FftGeo f, Geo g, const double *__restrict__ x_rho, const double *__restrict__ v,
                            const unsigned char *__restrict__ elzero, Cx<double> *__restrict__ Z)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.N) return;
    const int x = i % g.nx[0];
    const int r = i / g.nx[0];
    const int y = r % g.nx[1], z = r / g.nx[1];
    const long long q = (x + f.off[0]) + (y + f.off[1]) * f.S1 + (long long)(z + f.off[2]) * f.S2;
    for (int j = 0; 2 * j < g.C; ++j) {
        const size_t a = (size_t)(2 * j) * g.N + i, b = a + g.N;
        const double re = elzero[a] ? 0.0 : x_rho[a] * v[a];
        const double im = (2 * j + 1 < g.C && !elzero[b]) ? x_rho[b] * v[b] : 0.0;
        Z[(size_t)j * f.ns + q] = Cx<double>{re, im};
    }
}

// Y = Q(k) U(k), outputs packed in pairs
__global__ void k_pc_mul(FftGeo f, int C, const Cx<double> *__restrict__ Z, const Cx<double> *__restrict__ Q,
                         Cx<double> *__restrict__ O)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m >= f.n) return;
    const KIdx ix = k_index(f, m);
    Cx<double> X[kNC];
#pragma unroll
    for (int c = 0; c < kNC; ++c) {
        X[c] = Cx<double>{0.0, 0.0};
        if (c < C) {
            const Cx<double> *z = Z + (size_t)(c >> 1) * f.ns;
            X[c] = unpack(z[ix.q], z[ix.qn], c & 1);
        }
    }
    Cx<double> acc{0.0, 0.0};
#pragma unroll
    for (int i = 0; i < kNC; ++i) {
        if (i < C) {
            Cx<double> y{0.0, 0.0};
#pragma unroll
            for (int j = 0; j < kNC; ++j)
                if (j < C) y = cadd(y, cmul(Q[(size_t)(i + C * j) * f.ns + ix.q], X[j]));
            if ((i & 1) == 0) {
                acc = y;
                if (i + 1 == C) O[(size_t)(i >> 1) * f.ns + ix.q] = acc;
            } else {
                O[(size_t)(i >> 1) * f.ns + ix.q] = Cx<double>{acc.x - y.y, acc.y + y.x};
            }
        }
    }
}

// z_c = M^-1 v on the free rows, v on the zero rows
__global__ void k_pc_extract(FftGeo f, Geo g, const Cx<double> *__restrict__ O, const double *__restrict__ v,
                             const unsigned char *__restrict__ elzero, double *__restrict__ zout)
{
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g.N) return;
    const size_t row = (size_t)c * g.N + i;
    if (elzero[row]) {
        zout[row] = v[row];
        return;
    }
    const int x = i % g.nx[0];
    const int r = i / g.nx[0];
    const int y = r % g.nx[1], z = r / g.nx[1];
    const Cx<double> o =
        O[(size_t)(c >> 1) * f.ns + (x + f.off[0]) + (y + f.off[1]) * f.S1 + (long long)(z + f.off[2]) * f.S2];
    zout[row] = (c & 1) ? o.y : o.x;
}

__global__ void k_zero_cx(Cx<double> *__restrict__ Z, long long n)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m < n) Z[m] = Cx<double>{0.0, 0.0};
}

__global__ void k_to_float(const double *__restrict__ a, float *__restrict__ b, long long n)
{
    const long long m = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (m < n) b[m] = (float)a[m];
}

// Row pitch of the FFT storage: a multiple of 16 elements (rows start on 128-byte lines in complex
// float and complex double), at least one 16-element pad past the row, and not a power of two.
long long row_pitch(int M0)
{
    long long s = ((M0 + 15) / 16) * 16 + 16;
    if ((s & (s - 1)) == 0) s += 16;
    return s;
}

int nice_size(int n)
{
    for (int m = std::max(n, 1);; ++m) {
        int r = m;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
        if (r == 1) return m;
    }
}

// M = 4^c[0] 2^c[1] 3^c[2] 5^c[3] with c[1] <= 1
void factor_radices(int M, int *c)
{
    int r = M;
    const int radix[4] = {4, 2, 3, 5};
    for (int i = 0; i < 4; ++i) {
        c[i] = 0;
        while (r % radix[i] == 0) {
            ++c[i];
            r /= radix[i];
        }
    }
    if (r != 1) throw std::logic_error("FFT length must factor into 2, 3 and 5");
}

template <typename T>
inline const Cx<T> *twiddles(const FftPlan &F, int axis);
template <>
inline const Cx<double> *twiddles<double>(const FftPlan &F, int axis)
{
    return F.twd[axis];
}
template <>
inline const Cx<float> *twiddles<float>(const FftPlan &F, int axis)
{
    return F.twf[axis];
}

// Lines per workgroup: as many as the LDS budget allows, but few enough that
// the pass has about kFftGroups workgroups (the whole chip busy).
constexpr long long kFftGroups = 1024;

// one axis pass over nfield fields ns elements apart; sign -1 forward, +1 inverse (unscaled);
// mode kPassGather / kPassRows (strided axes of a wrapping grid) with io
// lines per workgroup and workgroups per field of an axis pass
struct PassShape {
    long long lpb, groups;
};
template <typename T>
PassShape pass_shape(const FftPlan &F, int axis, int nfield)
{
    const int M = F.M[axis];
    const long long lines = F.n / M;
    const long long lds_max = (long long)((kFftLds / sizeof(Cx<T>) - M) / (2 * (M + 1)));
    long long lpb = std::max<long long>(1, std::min<long long>(lds_max, lines * nfield / kFftGroups));
    lpb = std::min<long long>(lpb, 64);
    if (axis > 0) {   // strided axes: whole 128-byte lines of consecutive x when the budget allows
        const long long line_elems = 128 / (long long)sizeof(Cx<T>);
        if (lds_max >= line_elems) lpb = std::max(line_elems, (lpb / line_elems) * line_elems);
    }
    long long groups;
    if (axis == 0)
        groups = (lines + lpb - 1) / lpb;
    else
        groups = (long long)((F.M[0] + lpb - 1) / lpb) * (axis == 1 ? F.M[2] : F.M[1]);
    return PassShape{lpb, groups};
}

// register radices (R1, R2) of the register passes for length M, or 0 when M has none
inline int reg_radices(int M, int &R1, int &R2)
{
    switch (M) {
    case 32: R1 = 8, R2 = 4; return 1;
    case 64: R1 = 8, R2 = 8; return 1;
    case 128: R1 = 16, R2 = 8; return 1;
    case 256: R1 = 16, R2 = 16; return 1;
    default: return 0;
    }
}

// The one-wave register pass where it has radices and at least kRegPassWaves waves (one per
// workgroup); fewer would leave SIMDs idle, and the LDS pass runs 4 waves per workgroup.
constexpr long long kRegPassWaves = 1024;
inline long long reg_pass_groups(const FftPlan &F, int axis, int R2)
{
    const int L = 64 / R2;
    if (axis == 0) return (F.n / F.M[0] + L - 1) / L;
    return (long long)((F.M[0] + L - 1) / L) * (axis == 1 ? F.M[2] : F.M[1]);
}
inline bool use_reg_pass(const FftPlan &F, int axis, int nfield, int &R1, int &R2)
{
    if (F.regpass == 0 || !reg_radices(F.M[axis], R1, R2)) return false;
    if (F.regpass == 2) return true;
    return reg_pass_groups(F, axis, R2) * nfield >= kRegPassWaves;
}

// workgroups of a pass over nfield fields
template <typename T>
long long pass_blocks(const FftPlan &F, int axis, int nfield)
{
    int R1 = 0, R2 = 0;
    if (use_reg_pass(F, axis, nfield, R1, R2)) return reg_pass_groups(F, axis, R2) * nfield;
    return pass_shape<T>(F, axis, nfield).groups * nfield;
}

template <typename T, int SIGN, int MODE>
void launch_strided_reg(const FftPlan &F, Cx<T> *data, int nfield, int axis, int R1, int R2, hipStream_t s,
                        const PassIO &io)
{
    const dim3 grid((unsigned)reg_pass_groups(F, axis, R2), (unsigned)nfield);
    PassIO iox = io;
    iox.xcd = F.xcdmap;
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(64), 0, s, data, F.ns, F.M[0], F.M[1], F.M[2], F.S1, F.S2, axis,
                           twiddles<T>(F, axis), iox);
    };
    if (R1 == 8 && R2 == 4)
        go(k_fft_strided<T, SIGN, MODE, 8, 4>);
    else if (R1 == 8 && R2 == 8)
        go(k_fft_strided<T, SIGN, MODE, 8, 8>);
    else if (R1 == 16 && R2 == 8)
        go(k_fft_strided<T, SIGN, MODE, 16, 8>);
    else
        go(k_fft_strided<T, SIGN, MODE, 16, 16>);
}

// returns the number of workgroups launched
template <typename T>
long long fft_pass(const FftPlan &F, Cx<T> *data, int nfield, int sign, int axis, hipStream_t s, int mode = kPassPlain,
                   const PassIO &io = PassIO())
{
    const int M = F.M[axis];
    if (M == 1 || nfield <= 0) return 0;
    if (mode != kPassPlain && (axis == 0 || !F.wraps)) throw std::logic_error("fused pass ends need a strided axis of a wrapping grid");
    int R1 = 0, R2 = 0;
    if (use_reg_pass(F, axis, nfield, R1, R2) && axis == 0) {
        const dim3 grid((unsigned)reg_pass_groups(F, 0, R2), (unsigned)nfield);
        auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, grid, dim3(64), 0, s, data, F.ns, F.M[1], F.M[2], F.S1, F.S2, twiddles<T>(F, 0));
        };
        auto by_sign = [&](auto r1, auto r2) {
            constexpr int A = decltype(r1)::value, B = decltype(r2)::value;
            if (sign < 0)
                go(k_fft_xline<T, -1, A, B>);
            else
                go(k_fft_xline<T, 1, A, B>);
        };
        if (R1 == 8 && R2 == 4)
            by_sign(std::integral_constant<int, 8>{}, std::integral_constant<int, 4>{});
        else if (R1 == 8 && R2 == 8)
            by_sign(std::integral_constant<int, 8>{}, std::integral_constant<int, 8>{});
        else if (R1 == 16 && R2 == 8)
            by_sign(std::integral_constant<int, 16>{}, std::integral_constant<int, 8>{});
        else
            by_sign(std::integral_constant<int, 16>{}, std::integral_constant<int, 16>{});
        return pass_blocks<T>(F, axis, nfield);
    }
    if (use_reg_pass(F, axis, nfield, R1, R2)) {
        if (sign < 0) {
            if (mode == kPassGather)
                launch_strided_reg<T, -1, kPassGather>(F, data, nfield, axis, R1, R2, s, io);
            else
                launch_strided_reg<T, -1, kPassPlain>(F, data, nfield, axis, R1, R2, s, io);
        } else {
            if (mode == kPassRows)
                launch_strided_reg<T, 1, kPassRows>(F, data, nfield, axis, R1, R2, s, io);
            else if (mode == kPassEl)
                launch_strided_reg<T, 1, kPassEl>(F, data, nfield, axis, R1, R2, s, io);
            else if (mode == kPassPc)
                launch_strided_reg<T, 1, kPassPc>(F, data, nfield, axis, R1, R2, s, io);
            else
                launch_strided_reg<T, 1, kPassPlain>(F, data, nfield, axis, R1, R2, s, io);
        }
        return pass_blocks<T>(F, axis, nfield);
    }
    const PassShape ps = pass_shape<T>(F, axis, nfield);
    const long long lpb = ps.lpb, groups = ps.groups;
    const RadixCounts rc{F.rc[axis][0], F.rc[axis][1], F.rc[axis][2], F.rc[axis][3]};
    const size_t lds = sizeof(Cx<T>) * ((size_t)M + 2 * (size_t)lpb * (M + 1));
    const dim3 grid((unsigned)groups, (unsigned)nfield);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, dim3(kFftBlock), lds, s, data, F.ns, F.M[0], F.M[1], F.M[2], F.S1, F.S2, axis, rc,
                           (int)lpb, twiddles<T>(F, axis), io);
    };
    if (sign < 0) {
        if (mode == kPassGather)
            go(k_fft_axis<T, -1, kPassGather>);
        else
            go(k_fft_axis<T, -1, kPassPlain>);
    } else {
        if (mode == kPassRows)
            go(k_fft_axis<T, 1, kPassRows>);
        else if (mode == kPassEl)
            go(k_fft_axis<T, 1, kPassEl>);
        else if (mode == kPassPc)
            go(k_fft_axis<T, 1, kPassPc>);
        else
            go(k_fft_axis<T, 1, kPassPlain>);
    }
    return groups * nfield;
}

// in-place 3D transform of nfield fields
template <typename T>
void fft3d(const FftPlan &F, Cx<T> *data, int nfield, int sign, hipStream_t s)
{
    for (int axis = 0; axis < 3; ++axis) fft_pass<T>(F, data, nfield, sign, axis, s);
}

// LDS budget of the fused x pass
constexpr size_t kXprodLds = 64 * 1024;

// Lines per workgroup of the fused x pass for nop output pairs: the largest power of two up to 8
// whose LDS fits kXprodLds and that leaves about kFftGroups workgroups; 0 when one line does not fit.
TabGeo tab_geo(const FftPlan &F)
{
    return TabGeo{F.fold ? 1 : 0, {F.M[0], F.M[1], F.M[2]}, F.tp1, F.tp2, F.tn};
}

int xprod_lpb(const FftPlan &F, int nop)
{
    const size_t M = (size_t)F.M[0], per_line = (4 + (size_t)nop) * (M + 1) * sizeof(Cx<double>);
    const long long lines = F.n / F.M[0];
    int lpb = 8;
    while (lpb > 1 && (M * sizeof(Cx<double>) + lpb * per_line > kXprodLds || lines / lpb < kFftGroups)) lpb /= 2;
    return (M * sizeof(Cx<double>) + lpb * per_line <= kXprodLds) ? lpb : 0;
}

// the register form of the fused x pass where the x length has register radices and the output
// pairs' spectrum values fit in registers (NOP R1 <= 32); false otherwise
inline bool xreg_fits(const FftPlan &F, const TermList &tl, int &R1, int &R2)
{
    if (!F.xreg || !reg_radices(F.M[0], R1, R2)) return false;
    const int nop = (tl.nout + 1) / 2;
    return nop <= 4 && nop * R1 <= 32;
}

template <typename T>
bool xprod_reg(const FftPlan &F, const Cx<T> *Z, Cx<T> *O, const TermList &tl, const T *tk, const T *uk, hipStream_t s)
{
    int R1 = 0, R2 = 0;
    if (!xreg_fits(F, tl, R1, R2)) return false;
    const int nop = (tl.nout + 1) / 2;
    const int L = 64 / R2;
    const long long lines = xprod_primary_lines(F.M[1], F.M[2]);
    const dim3 grid((unsigned)((lines + L - 1) / L)), block(2 * L * R2);
    const size_t lds = sizeof(Cx<T>) * ((size_t)F.M[0] + 2 * (size_t)L * (F.M[0] + 1));
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, grid, block, lds, s, Z, O, F.ns, F.M[1], F.M[2], F.S1, F.S2, L, twiddles<T>(F, 0), tk,
                           uk, tab_geo(F), tl, F.xcdmap);
    };
    auto by_nop = [&](auto r1, auto r2) {
        constexpr int A = decltype(r1)::value, B = decltype(r2)::value;
        if constexpr (A * 1 <= 32) {
            if (nop == 1) go(k_fft_xprod_reg<T, A, B, 1>);
        }
        if constexpr (A * 2 <= 32) {
            if (nop == 2) go(k_fft_xprod_reg<T, A, B, 2>);
        }
        if constexpr (A * 3 <= 32) {
            if (nop == 3) go(k_fft_xprod_reg<T, A, B, 3>);
        }
        if constexpr (A * 4 <= 32) {
            if (nop == 4) go(k_fft_xprod_reg<T, A, B, 4>);
        }
    };
    if (R1 == 8 && R2 == 4)
        by_nop(std::integral_constant<int, 8>{}, std::integral_constant<int, 4>{});
    else if (R1 == 8 && R2 == 8)
        by_nop(std::integral_constant<int, 8>{}, std::integral_constant<int, 8>{});
    else if (R1 == 16 && R2 == 8)
        by_nop(std::integral_constant<int, 16>{}, std::integral_constant<int, 8>{});
    else
        by_nop(std::integral_constant<int, 16>{}, std::integral_constant<int, 16>{});
    return true;
}

// product of input pairs Z (y and z already transformed) into output pairs O (y and z still to invert)
template <typename T>
void fft_xprod(const FftPlan &F, const Cx<T> *Z, Cx<T> *O, const TermList &tl, int lpb, const T *tk, const T *uk,
               hipStream_t s)
{
    if (xprod_reg<T>(F, Z, O, tl, tk, uk, s)) return;
    const int M = F.M[0], nop = (tl.nout + 1) / 2;
    const long long lines = F.n / M;
    const RadixCounts rc{F.rc[0][0], F.rc[0][1], F.rc[0][2], F.rc[0][3]};
    const size_t lds = sizeof(Cx<T>) * ((size_t)M + (4 + (size_t)nop) * lpb * (M + 1));
    hipLaunchKernelGGL(k_fft_xprod<T>, dim3((unsigned)((lines + lpb - 1) / lpb)), dim3(kFftBlock), lds, s, Z, O, F.ns,
                       F.M[0], F.M[1], F.M[2], F.S1, F.S2, rc, lpb, twiddles<T>(F, 0), tk, uk, tab_geo(F), tl);
}

// threads of the group-of-8 product kernels: one per (kx, ky, kz) with 0 <= k_a <= M_a / 2
long long group8_count(const FftPlan &F)
{
    return (long long)(F.M[0] / 2 + 1) * (F.M[1] / 2 + 1) * (F.M[2] / 2 + 1);
}

FftGeo fft_geo(const FftPlan &F, int which)
{
    FftGeo f;
    f.tg = tab_geo(F);
    for (int d = 0; d < 3; ++d) {
        f.M[d] = F.M[d];
        f.off[d] = F.off[d];
        f.lo[d] = F.lo[which][d];
        f.lim[d] = F.lim[which][d];
        f.src[d] = F.src[which][d];
    }
    f.n = F.n;
    f.S1 = F.S1;
    f.S2 = F.S2;
    f.ns = F.ns;
    return f;
}

template <typename T>
inline const T *attr_spectra(const FftPlan &F)
{
    if constexpr (sizeof(T) == 8)
        return reinterpret_cast<const T *>(F.ud);
    else
        return reinterpret_cast<const T *>(F.uf);
}
template <typename T>
inline const T *fmt_spectra(const FftPlan &F)
{
    if constexpr (sizeof(T) == 8)
        return reinterpret_cast<const T *>(F.td);
    else
        return reinterpret_cast<const T *>(F.tf);
}

void free_plan(FftPlan *F)
{
    if (!F) return;
    (void)hipFree(F->zbuf);
    (void)hipFree(F->obuf);
    (void)hipFree(F->ud);
    (void)hipFree(F->uf);
    (void)hipFree(F->td);
    (void)hipFree(F->tf);
    (void)hipFree(F->pq);
    (void)hipFree(F->pqr);
    (void)hipFree(F->pqr_f);
    (void)hipFree(F->ss);
    for (int d = 0; d < 3; ++d) {
        (void)hipFree(F->twd[d]);
        (void)hipFree(F->twf[d]);
    }
    delete F;
}

// Relative size of the part of a flipped-stencil spectrum that must vanish
// (imaginary for even stencils, real for odd ones) above which the FFT form
// is refused.
constexpr double kSpectrumParityTol = 1e-9;
constexpr double kFoldTol = 1e-13;   // reflection symmetry of the stencil weights (relative)

}  // namespace

// Plan, spectra and buffers of the FFT forms.  mf: the mean-field sums;
// fmt: build the FMT spectra (the FMT sums switch to the FFT form with
// fft_set_fmt).  P.fft stays null when an axis would be longer than
// kFftMaxLen or a stencil spectrum fails its parity check.
void fft_setup(Problem &P, bool mf, bool fmt)
{
    const int nd = P.p.ndim, C = P.p.ncomp;
    mf = mf && P.nmf > 0;
    if (!mf && !fmt) return;
    auto *F = new FftPlan();
    try {
        F->nd = nd;
        F->C = C;
        if (const char *e = std::getenv("TMI_FFT_REGPASS")) F->regpass = std::string(e) == "0" ? 0 : (std::string(e) == "2" ? 2 : 1);
        if (const char *e = std::getenv("TMI_FFT_XREG")) F->xreg = std::string(e) != "0";
        if (const char *e = std::getenv("TMI_FFT_XCDMAP")) F->xcdmap = std::string(e) != "0";
        for (int d = 0; d < nd; ++d) {
            const int n = P.p.nx[d], h = P.pad_u.h[d], hf = P.pad_f.h[d];
            const bool per = P.p.bc[d][0] == TMI_PERIODIC;
            if (per && nice_size(n) == n) {
                F->M[d] = n;
                F->off[d] = 0;
                for (int w = 0; w < 2; ++w) {
                    F->lo[w][d] = 0;
                    F->lim[w][d] = n;
                }
                F->src[0][d] = h;
                F->src[1][d] = hf;
            } else {
                F->M[d] = nice_size(n + 2 * h);
                F->off[d] = h;
                F->lo[0][d] = 0;
                F->lim[0][d] = n + 2 * h;
                F->src[0][d] = 0;
                F->lo[1][d] = h - hf;
                F->lim[1][d] = h + n + hf;
                F->src[1][d] = hf - h;
                F->wraps = false;
            }
            factor_radices(F->M[d], F->rc[d]);
            if (F->M[d] > kFftMaxLen) {   // one line must fit the LDS budget: keep the direct sums
                free_plan(F);
                return;
            }
        }
        F->n = (long long)F->M[0] * F->M[1] * F->M[2];
        F->S1 = F->M[1] > 1 ? row_pitch(F->M[0]) : F->M[0];
        F->S2 = F->S1 * (F->M[1] + (F->M[2] > 1 ? 1 : 0));
        F->ns = F->S2 * F->M[2];
        const size_t ns = (size_t)F->ns;
        for (int d = 0; d < 3; ++d) {
            const int M = F->M[d];
            std::vector<Cx<double>> td(M);
            std::vector<Cx<float>> tf(M);
            for (int j = 0; j < M; ++j) {
                const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)j / (long double)M;
                td[j] = Cx<double>{(double)cosl(a), (double)sinl(a)};
                tf[j] = Cx<float>{(float)td[j].x, (float)td[j].y};
            }
            TMI_HIP_CHECK(hipMalloc((void **)&F->twd[d], M * sizeof(Cx<double>)));
            TMI_HIP_CHECK(hipMalloc((void **)&F->twf[d], M * sizeof(Cx<float>)));
            TMI_HIP_CHECK(hipMemcpy(F->twd[d], td.data(), M * sizeof(Cx<double>), hipMemcpyHostToDevice));
            TMI_HIP_CHECK(hipMemcpy(F->twf[d], tf.data(), M * sizeof(Cx<float>), hipMemcpyHostToDevice));
        }
        // output descriptors
        const int ND = nd;
        F->dform = F->wraps && fmt && (4 + 2 * ND + 1) / 2 < (C * (1 + ND) + 2) / 2;
        const auto &b = P.bulk;
        FwdDesc &f0 = F->fwd[0], &f1 = F->fwd[1];
        RowMap &r0 = F->rows[0], &r1 = F->rows[1];
        f0.C = f1.C = C;
        if (C == 1) {   // n1, n0 and v1 are multiples of n2 and v2
            f1.kind[f1.nout] = 0, f1.coef[f1.nout++][0] = 1.0;
            f1.kind[f1.nout] = 1, f1.coef[f1.nout++][0] = 1.0;
            for (int d = 0; d < ND; ++d) f1.kind[f1.nout] = 2 + d, f1.coef[f1.nout++][0] = 1.0;
            const int src[10] = {0, 1, 1, 1, 2, 3, 4, 2, 3, 4};
            const double sc[4] = {1.0, 1.0, b.inv_4pir[0], b.inv_4pirsq[0]};
            r1.nrb = 4 + 2 * ND;
            for (int q = 0; q < 4; ++q) r1.src[q] = src[q], r1.scale[q] = sc[q];
            for (int d = 0; d < ND; ++d) {
                r1.src[4 + d] = 2 + d, r1.scale[4 + d] = 1.0;
                r1.src[4 + ND + d] = 2 + d, r1.scale[4 + ND + d] = b.inv_4pir[0];
            }
        } else {
            auto add = [&](int kind, const double *coef) {
                f1.kind[f1.nout] = kind;
                for (int c = 0; c < C; ++c) f1.coef[f1.nout][c] = coef ? coef[c] : 1.0;
                r1.src[f1.nout] = f1.nout;
                r1.scale[f1.nout] = 1.0;
                ++f1.nout;
            };
            add(0, nullptr);
            add(1, nullptr);
            add(1, b.inv_4pir);
            add(1, b.inv_4pirsq);
            for (int d = 0; d < ND; ++d) add(2 + d, nullptr);
            for (int d = 0; d < ND; ++d) add(2 + d, b.inv_4pir);
            r1.nrb = f1.nout;
        }
        if (mf) {
            for (int c = 0; c < C; ++c) {
                f0.kind[f0.nout] = kKindMf + c;
                r0.src[r0.nmf] = f0.nout++;
                r0.scale[r0.nmf++] = 1.0;
                f1.kind[f1.nout] = kKindMf + c;
                r1.src[r1.nrb + r1.nmf] = f1.nout++;
                r1.scale[r1.nrb + r1.nmf++] = 1.0;
            }
        }
        // the same products as term lists of the fused x pass
        // the current field nin owns terms first[nin] .. first[nin+1]-1; first[nin+1] is the running end
        auto add_term = [&](TermList &tl, int o, int t, int imag, double coef) {
            if (tl.first[tl.nin + 1] >= kMaxTerms) throw std::logic_error("too many product terms");
            const int odd = (t >= 0 && imag) ? t % (2 + ND) - 2 : -1;   // vector delta d is odd in axis d
            tl.term[tl.first[tl.nin + 1]++] = Term{o, t, imag, odd, coef};
        };
        auto close_input = [&](TermList &tl) {   // sort the field's terms by table; the next field starts
            if (tl.nin + 2 > kMaxIn + 1) throw std::logic_error("too many product inputs");
            std::stable_sort(tl.term + tl.first[tl.nin], tl.term + tl.first[tl.nin + 1],
                             [](const Term &a, const Term &b) { return a.t < b.t; });
            ++tl.nin;
            tl.first[tl.nin + 1] = tl.first[tl.nin];
        };
        for (int w = 0; w < 2; ++w) {
            const FwdDesc &d = F->fwd[w];
            TermList &tl = F->tfwd[w];
            tl.nout = d.nout;
            for (int c = 0; c < C; ++c) {
                for (int o = 0; o < d.nout; ++o) {
                    const int kind = d.kind[o];
                    if (kind < kKindMf) {
                        if (d.coef[o][c] != 0.0) add_term(tl, o, c * (2 + ND) + kind, kind >= 2 ? 1 : 0, d.coef[o][c]);
                    } else {
                        add_term(tl, o, -((kind - kKindMf) + C * c) - 1, 0, 1.0);
                    }
                }
                close_input(tl);
            }
        }
        if (F->dform) {   // inputs D3, D2, D1, D0, Dv2_d, Dv1_d (the derivatives of Phi)
            TermList &tl = F->tel_d;
            tl.nout = C;
            for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND), 0, 1.0);   // D3: theta
            close_input(tl);
            for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND) + 1, 0, 1.0);   // D2: delta
            close_input(tl);
            for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND) + 1, 0, b.inv_4pir[c]);   // D1
            close_input(tl);
            for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND) + 1, 0, b.inv_4pirsq[c]);   // D0
            close_input(tl);
            for (int d = 0; d < ND; ++d) {   // Dv2_d: -(i K'_d)
                for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND) + 2 + d, 1, -1.0);
                close_input(tl);
            }
            for (int d = 0; d < ND; ++d) {   // Dv1_d
                for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND) + 2 + d, 1, -b.inv_4pir[c]);
                close_input(tl);
            }
        }
        {   // the adjoint fields A_c, B_c,d, S3 (without dform, and for reference-Jacobian products)
            TermList &tl = F->tel;
            tl.nout = C;
            for (int c = 0; c < C; ++c) {
                add_term(tl, c, c * (2 + ND) + 1, 0, 1.0);   // A_c: delta
                close_input(tl);
                for (int d = 0; d < ND; ++d) {
                    add_term(tl, c, c * (2 + ND) + 2 + d, 1, -1.0);   // B_c,d: -(i K'_d)
                    close_input(tl);
                }
            }
            for (int c = 0; c < C; ++c) add_term(tl, c, c * (2 + ND), 0, 1.0);   // S3: theta of every component
            close_input(tl);
        }
        for (int w = 0; w < 2; ++w) {
            const RowMap &rm = F->rows[w];
            for (int r = 0; r < rm.nrb + rm.nmf; ++r) {
                const int o = rm.src[r];
                if (F->onrow[w][o] >= 3) throw std::logic_error("more than 3 rows per product output");
                F->orow[w][o][F->onrow[w][o]] = r;
                F->oscale[w][o][F->onrow[w][o]++] = rm.scale[r];
            }
        }
        F->xfit[0] = F->xlpb[0] = xprod_lpb(*F, std::max((F->fwd[0].nout + 1) / 2, (F->fwd[1].nout + 1) / 2));
        F->xfit[1] = F->xlpb[1] = xprod_lpb(*F, (C + 1) / 2);
        const int zpairs = std::max((C + 1) / 2, fmt ? (C * (1 + ND) + 2) / 2 : 0);   // either adjoint form
        const int opairs = std::max((f0.nout + 1) / 2, fmt ? std::max((f1.nout + 1) / 2, (C + 1) / 2) : 0);
        TMI_HIP_CHECK(hipMalloc(&F->zbuf, (size_t)zpairs * ns * sizeof(Cx<double>)));
        TMI_HIP_CHECK(hipMalloc(&F->obuf, (size_t)opairs * ns * sizeof(Cx<double>)));
        if (F->wraps) {   // partial sums of F^2 of the fused last passes (FP64 fills)
            int last = -1;
            for (int a : {1, 2})
                if (F->M[a] > 1) last = a;
            auto most = [&](int pairs) {   // workgroups of the pass over at most `pairs` fields
                long long m = 0;
                for (int nf = 1; nf <= pairs; ++nf) m = std::max(m, pass_blocks<double>(*F, last, nf));
                return m;
            };
            F->ss_cap = most((std::max(F->fwd[0].nout, F->fwd[1].nout) + 1) / 2) + most((C + 1) / 2);
            TMI_HIP_CHECK(hipMalloc((void **)&F->ss, (size_t)F->ss_cap * sizeof(double)));
        }
        // the tables fold when every stencil in use is symmetric under the reflection of y and of z
        // (weights at o and at o with o_a negated agree to kFoldTol of the largest weight)
        auto reflective = [&](const HostStencil &hs) {
            std::unordered_map<long long, double> at;
            double wmax = 0.0;
            auto key = [](int a, int b, int c) {
                return (((long long)a + (1 << 19)) << 40) | (((long long)b + (1 << 19)) << 20) | ((long long)c + (1 << 19));
            };
            for (int k = 0; k < hs.L; ++k) {
                const int *o = &hs.off[3 * (size_t)k];
                at[key(o[0], o[1], o[2])] += hs.w[k];
                wmax = std::max(wmax, std::fabs(hs.w[k]));
            }
            for (int k = 0; k < hs.L; ++k) {
                const int *o = &hs.off[3 * (size_t)k];
                const double w = at[key(o[0], o[1], o[2])];
                for (int a = 1; a < nd; ++a) {
                    int r[3] = {o[0], o[1], o[2]};
                    r[a] = -r[a];
                    const auto it = at.find(key(r[0], r[1], r[2]));
                    if (it == at.end() || std::fabs(it->second - w) > kFoldTol * wmax) return false;
                }
            }
            return true;
        };
        bool fold = true;
        if (mf)
            for (int q = 0; q < C * C; ++q) fold = fold && reflective(P.hs[TMI_UATT][q]);
        if (fmt)
            for (int c = 0; c < C; ++c) fold = fold && reflective(P.hs[TMI_THETA][c]) && reflective(P.hs[TMI_DELTA][c]);
        const char *fold_env = std::getenv("TMI_FFT_FOLD");   // TMI_FFT_FOLD=0 keeps whole-grid tables
        F->fold = fold && F->ns <= INT32_MAX && !(fold_env && std::string(fold_env) == "0");
        if (F->fold) {
            const long long H0 = F->M[0], H1 = F->M[1] / 2 + 1, H2 = F->M[2] / 2 + 1;
            F->tp1 = H0;
            F->tp2 = H0 * H1;
            F->tn = H0 * H1 * H2;
        } else {
            F->tp1 = F->S1;
            F->tp2 = F->S2;
            F->tn = F->ns;
        }
        const size_t tn = (size_t)F->tn;
        // spectra: flipped stencil g(-o mod M) = w(o) on the FFT grid, transformed (in zbuf), the
        // real (even) or imaginary (odd) part kept, scaled by 1/n, folded or in the transforms' layout
        std::vector<Cx<double>> host(ns);
        std::vector<double> part(ns);
        Cx<double> *work = reinterpret_cast<Cx<double> *>(F->zbuf);
        bool parity_ok = true;
        // vd < 0: the weights w; vd = d: the vector weights w o_d dx_d / R_comp (odd in axis d)
        auto spectrum = [&](const HostStencil &hs, int vd, int comp, double *dst_d, float *dst_f) {
            std::fill(host.begin(), host.end(), Cx<double>{0.0, 0.0});
            std::fill(part.begin(), part.end(), 0.0);
            const long long pitch[3] = {1, F->S1, F->S2};
            for (int k = 0; k < hs.L; ++k) {
                long long idx = 0;
                for (int d = 0; d < 3; ++d) {
                    const int o = (d < nd) ? hs.off[3 * (size_t)k + d] : 0;
                    int j = (-o) % F->M[d];
                    if (j < 0) j += F->M[d];
                    idx += (long long)j * pitch[d];
                }
                double w = hs.w[k];
                if (vd >= 0) w = hs.w[k] * ((double)hs.off[3 * (size_t)k + vd] * P.p.dx[vd] * P.bulk.inv_rad[comp]);
                host[(size_t)idx].x += w;
            }
            TMI_HIP_CHECK(hipMemcpy(work, host.data(), ns * sizeof(Cx<double>), hipMemcpyHostToDevice));
            fft3d<double>(*F, work, 1, -1, 0);
            TMI_HIP_CHECK(hipMemcpy(host.data(), work, ns * sizeof(Cx<double>), hipMemcpyDeviceToHost));
            double keep = 0.0, drop = 0.0, asym = 0.0;
            const double inv = 1.0 / (double)F->n;
            auto kept = [&](long long q) { return vd < 0 ? host[q].x : host[q].y; };
            for (int m2 = 0; m2 < F->M[2]; ++m2)
                for (int m1 = 0; m1 < F->M[1]; ++m1)
                    for (int m0 = 0; m0 < F->M[0]; ++m0) {
                        const long long q = m0 + m1 * F->S1 + m2 * F->S2;
                        const double a = kept(q), z = vd < 0 ? host[q].y : host[q].x;
                        keep = std::max(keep, std::fabs(a));
                        drop = std::max(drop, std::fabs(z));
                        if (!F->fold) {
                            part[q] = a * inv;
                            continue;
                        }
                        const int m[3] = {m0, m1, m2};
                        int f[3];
                        double sgn = 1.0;
                        for (int d = 0; d < 3; ++d) {   // y and z fold; x rows are kept whole
                            const bool hi = d > 0 && 2 * m[d] > F->M[d];
                            f[d] = hi ? F->M[d] - m[d] : m[d];
                            if (hi && d == vd) sgn = -1.0;
                        }
                        asym = std::max(asym, std::fabs(a - sgn * kept(f[0] + f[1] * F->S1 + f[2] * F->S2)));
                        if (f[0] == m0 && f[1] == m1 && f[2] == m2) part[m0 + m1 * F->tp1 + m2 * F->tp2] = a * inv;
                    }
            if (drop > kSpectrumParityTol * keep || asym > kSpectrumParityTol * keep) parity_ok = false;
            TMI_HIP_CHECK(hipMemcpy(dst_d, part.data(), tn * sizeof(double), hipMemcpyHostToDevice));
            hipLaunchKernelGGL(k_to_float, dim3(grid_for((long long)tn)), dim3(kBlock), 0, 0, dst_d, dst_f,
                               (long long)tn);
        };
        if (mf) {
            TMI_HIP_CHECK(hipMalloc((void **)&F->ud, (size_t)C * C * tn * sizeof(double)));
            TMI_HIP_CHECK(hipMalloc((void **)&F->uf, (size_t)C * C * tn * sizeof(float)));
            for (int q = 0; q < C * C; ++q)
                spectrum(P.hs[TMI_UATT][q], -1, q % C, F->ud + (size_t)q * tn, F->uf + (size_t)q * tn);
            if (!parity_ok) {   // keep the direct sums
                free_plan(F);
                return;
            }
            F->mf = true;
        }
        if (fmt) {
            const size_t nk = (size_t)C * (2 + ND);
            TMI_HIP_CHECK(hipMalloc((void **)&F->td, nk * tn * sizeof(double)));
            TMI_HIP_CHECK(hipMalloc((void **)&F->tf, nk * tn * sizeof(float)));
            for (int c = 0; c < C; ++c) {
                const size_t base = (size_t)c * (2 + ND);
                spectrum(P.hs[TMI_THETA][c], -1, c, F->td + base * tn, F->tf + base * tn);
                spectrum(P.hs[TMI_DELTA][c], -1, c, F->td + (base + 1) * tn, F->tf + (base + 1) * tn);
                for (int d = 0; d < ND; ++d)
                    spectrum(P.hs[TMI_DELTA][c], d, c, F->td + (base + 2 + d) * tn, F->tf + (base + 2 + d) * tn);
            }
            F->fmt_ok = parity_ok;
        }
        TMI_HIP_CHECK(hipGetLastError());
        TMI_HIP_CHECK(hipDeviceSynchronize());
    } catch (...) {
        free_plan(F);
        throw;
    }
    if (!F->mf && !F->fmt_ok) {
        free_plan(F);
        return;
    }
    P.fft = F;
}

void fft_free(Problem &P)
{
    free_plan(P.fft);
    P.fft = nullptr;
}

// the Euler-Lagrange product of a residual (linear = false) or a Jacobian-vector product runs on the
// derivatives of Phi (dform) unless the reference-Jacobian mode asks for per-component fields
bool el_dform(const Problem &P, bool linear) { return P.fft && P.fft->dform && !(linear && P.jac_ref); }

FftPairs fft_pairs(const Problem &P, bool linear)
{
    return FftPairs{P.fft->zbuf, P.fft->ns, P.fft->S1, P.fft->S2, el_dform(P, linear) ? 1 : 0};
}
bool fft_node_input(const Problem &P)
{
    return P.fft && P.fft->fmt && P.fft->wraps && (P.nmf == 0 || P.fft->mf);
}
bool fft_mf(const Problem &P) { return P.fft && P.fft->mf; }
double *fft_ss_buffer(const Problem &P, long long *cap)
{
    if (cap) *cap = P.fft ? P.fft->ss_cap : 0;
    return P.fft ? P.fft->ss : nullptr;
}
bool fft_fmt(const Problem &P) { return P.fft && P.fft->fmt; }
bool fft_fused(const Problem &P) { return P.fft && (P.fft->xlpb[0] || P.fft->xlpb[1]); }
int fft_xreg_mask(const Problem &P)
{
    if (!P.fft || !P.fft->fmt) return 0;
    const FftPlan &F = *P.fft;
    int R1 = 0, R2 = 0, m = 0;
    if (F.xlpb[0] && xreg_fits(F, F.tfwd[1], R1, R2)) m |= 1;
    if (F.xlpb[1] && xreg_fits(F, F.tel, R1, R2)) m |= 2;
    return m;
}
int fft_regpass_mode(const Problem &P) { return P.fft ? P.fft->regpass : 0; }
bool fft_fmt_available(const Problem &P) { return P.fft && P.fft->fmt_ok; }
bool fft_wraps(const Problem &P) { return P.fft && P.fft->wraps; }
bool fft_folded(const Problem &P) { return P.fft && P.fft->fold; }
bool fft_fused_available(const Problem &P) { return P.fft && (P.fft->xfit[0] || P.fft->xfit[1]); }
void fft_set_fused(Problem &P, int mask)
{
    if (!P.fft) return;
    for (int w = 0; w < 2; ++w) P.fft->xlpb[w] = ((mask >> w) & 1) ? P.fft->xfit[w] : 0;
}
void fft_set_fmt(Problem &P, bool on)
{
    if (P.fft) P.fft->fmt = on && P.fft->fmt_ok;
}

// The fused ends of the density product on a wrapping grid (product w): the first pass gathers the
// density pairs from the node arrays, the last one writes the rho-bar and mean-field rows
PassIO rows_io(const Problem &P, int w, const double *rho_nodes, const double *in_rb, const double *sub_rb,
               double *out_rb, const double *in_mf, const double *sub_mf, double *out_mf)
{
    const FftPlan &F = *P.fft;
    const FwdDesc &d = F.fwd[w];
    PassIO io;
    io.C = P.p.ncomp;
    io.nx0 = P.p.nx[0];
    io.nx1 = P.p.ndim > 1 ? P.p.nx[1] : 1;
    io.N = P.N;
    io.rho = rho_nodes;
    io.zero = P.d_zero;
    io.nout = d.nout;
    io.nrb = F.rows[w].nrb;
    for (int o = 0; o < d.nout; ++o) {
        io.nrow[o] = F.onrow[w][o];
        for (int k = 0; k < 3; ++k) {
            io.row[o][k] = F.orow[w][o][k];
            io.scale[o][k] = F.oscale[w][o][k];
        }
    }
    io.in_rb = in_rb;
    io.sub_rb = sub_rb;
    io.out_rb = out_rb;
    io.in_mf = in_mf;
    io.sub_mf = sub_mf;
    io.out_mf = out_mf;
    return io;
}

// Density products from the halo-padded densities in workspace<T>(P).rho_pad:
// with_rb: the rho-bar rows (out_rb = sums - in_rb - sub_rb), then the
// mean-field rows when they are in FFT form; else the mean-field rows only.
template <typename T>
void fft_rb_mf(Problem &P, const double *rho_nodes, bool with_rb, const double *in_rb, const double *sub_rb,
               double *out_rb, const double *in_mf, const double *sub_mf, double *out_mf, hipStream_t s, double *ss,
               long long *nss)
{
    if (nss) *nss = 0;
    const FftPlan &F = *P.fft;
    const int w = with_rb ? 1 : 0;
    const FwdDesc &d = F.fwd[w];
    const RowMap &rm = F.rows[w];
    if (d.nout == 0) return;
    const int C = P.p.ncomp;
    Cx<T> *Z = reinterpret_cast<Cx<T> *>(F.zbuf);
    Cx<T> *O = reinterpret_cast<Cx<T> *>(F.obuf);
    const Geo gu = make_geo(P, P.pad_u);
    const FftGeo fg = fft_geo(F, 0);
    const int nip = (C + 1) / 2, nop = (d.nout + 1) / 2;
    {   // z, y, x passes; the product (fused with the x passes or not); x, y, z passes back
        // on a wrapping grid the first pass reads the node arrays and the last one writes the rows
        const bool ends = F.wraps && rho_nodes;
        PassIO io = ends ? rows_io(P, w, rho_nodes, in_rb, sub_rb, out_rb, in_mf, sub_mf, out_mf) : PassIO();
        io.C = C;
        io.nx0 = P.p.nx[0];
        io.nx1 = P.p.ndim > 1 ? P.p.nx[1] : 1;
        io.N = P.N;
        if (ends) {
        } else if (rho_nodes) {
            hipLaunchKernelGGL(k_gather_nodes<T>, dim3(grid_for(P.N)), dim3(kBlock), 0, s, fg, gu, rho_nodes, P.d_zero, Z);
        } else {
            hipLaunchKernelGGL(k_gather_pairs<T>, dim3(grid_for(F.n)), dim3(kBlock), 0, s, fg, gu,
                               workspace<T>(P).rho_pad, C, Z);
        }
        const int fwd_axes[2] = {2, 1}, inv_axes[2] = {1, 2};
        int first = -1, last = -1;
        for (int a : fwd_axes)
            if (F.M[a] > 1 && first < 0) first = a;
        for (int a : inv_axes)
            if (F.M[a] > 1) last = a;
        for (int a : fwd_axes) fft_pass<T>(F, Z, nip, -1, a, s, (ends && a == first) ? kPassGather : kPassPlain, io);
        if (F.xlpb[0]) {
            fft_xprod<T>(F, Z, O, F.tfwd[w], F.xlpb[0], fmt_spectra<T>(F), attr_spectra<T>(F), s);
        } else {
            fft_pass<T>(F, Z, nip, -1, 0, s);
            if (F.fold) {
                const dim3 g4((unsigned)grid_for(group4_threads(group8_count(F))));
                if (P.p.ndim == 2)
                    hipLaunchKernelGGL((k_fwd_mul4<T, 2>), g4, dim3(kBlock), 0, s, fg, d, Z, fmt_spectra<T>(F),
                                       attr_spectra<T>(F), O);
                else
                    hipLaunchKernelGGL((k_fwd_mul4<T, 3>), g4, dim3(kBlock), 0, s, fg, d, Z, fmt_spectra<T>(F),
                                       attr_spectra<T>(F), O);
            } else if (P.p.ndim == 2) {
                hipLaunchKernelGGL((k_fwd_mul<T, 2>), dim3(grid_for(F.n)), dim3(kBlock), 0, s, fg, d, Z,
                                   fmt_spectra<T>(F), attr_spectra<T>(F), O);
            } else {
                hipLaunchKernelGGL((k_fwd_mul<T, 3>), dim3(grid_for(F.n)), dim3(kBlock), 0, s, fg, d, Z,
                                   fmt_spectra<T>(F), attr_spectra<T>(F), O);
            }
            fft_pass<T>(F, O, nop, +1, 0, s);
        }
        if (ends) io.ss = ss;
        long long blocks = 0;
        for (int a : inv_axes)
            blocks = fft_pass<T>(F, O, nop, +1, a, s, (ends && a == last) ? kPassRows : kPassPlain, io);
        if (ends) {
            if (ss && nss) *nss = blocks;
            TMI_HIP_CHECK(hipGetLastError());
            return;
        }
    }
    hipLaunchKernelGGL(k_fwd_extract<T>, dim3(grid_for(P.N), rm.nrb + rm.nmf), dim3(kBlock), 0, s, fg, gu, rm, O, in_rb,
                       sub_rb, out_rb, in_mf, sub_mf, out_mf);
    TMI_HIP_CHECK(hipGetLastError());
}
template void fft_rb_mf<double>(Problem &, const double *, bool, const double *, const double *, double *,
                                const double *, const double *, double *, hipStream_t, double *, long long *);
template void fft_rb_mf<float>(Problem &, const double *, bool, const double *, const double *, double *,
                               const double *, const double *, double *, hipStream_t, double *, long long *);

// Euler-Lagrange rows from the adjoint fields (as el_rows): in workspace<T>(P).adj_pad,
// or, on wrapping grids, already paired in the transform buffer (fft_pairs)
template <typename T>
void fft_el(Problem &P, bool linear, const double *x_rho, const double *in_rho, const double *in_mf, const double *add,
            double *out_el, hipStream_t s, double *ss, long long *nss)
{
    if (nss) *nss = 0;
    const FftPlan &F = *P.fft;
    const bool df = el_dform(P, linear);
    const int C = P.p.ncomp, nd = P.p.ndim, nadj = df ? 4 + 2 * nd : C * (1 + nd) + 1;
    Cx<T> *Z = reinterpret_cast<Cx<T> *>(F.zbuf);
    Cx<T> *O = reinterpret_cast<Cx<T> *>(F.obuf);
    const Geo gf = make_geo(P, P.pad_f);
    const FftGeo fg = fft_geo(F, 1);
    if (!F.wraps)   // on wrapping grids fmt_point wrote the pairs (fft_pairs)
        hipLaunchKernelGGL(k_gather_pairs<T>, dim3(grid_for(F.n)), dim3(kBlock), 0, s, fg, gf, workspace<T>(P).adj_pad,
                           nadj, Z);
    const int nip = (nadj + 1) / 2, nop = (C + 1) / 2;
    {   // z, y, x passes; the product (fused with the x passes or not); x, y, z passes back
        // on a wrapping grid the last pass writes the Euler-Lagrange rows
        const bool ends = F.wraps;
        PassIO io;
        io.C = C;
        io.nx0 = P.p.nx[0];
        io.nx1 = nd > 1 ? P.p.nx[1] : 1;
        io.N = P.N;
        if (ends) {
            const CompConsts cc = make_consts(P);
            io.x_rho = x_rho;
            io.el_in_rho = in_rho;
            io.el_in_mf = in_mf;
            io.vext = P.d_vext;
            io.add = add;
            io.elzero = P.d_elzero;
            io.out_el = out_el;
            for (int c = 0; c < C; ++c) io.mu[c] = cc.mu[c];
            io.linear = linear ? 1 : 0;
            io.nwall = P.p.nwall;
        }
        int last = -1;
        for (int a : {1, 2})
            if (F.M[a] > 1) last = a;
        fft_pass<T>(F, Z, nip, -1, 2, s);
        fft_pass<T>(F, Z, nip, -1, 1, s);
        if (F.xlpb[1]) {
            fft_xprod<T>(F, Z, O, df ? F.tel_d : F.tel, F.xlpb[1], fmt_spectra<T>(F), attr_spectra<T>(F), s);
        } else {
            fft_pass<T>(F, Z, nip, -1, 0, s);
            ElConsts ec{};
            for (int c = 0; c < C; ++c) {
                ec.i4pir[c] = P.bulk.inv_4pir[c];
                ec.i4pirsq[c] = P.bulk.inv_4pirsq[c];
            }
            const long long nthr = F.fold ? group8_count(F) : F.n;
            auto go = [&](auto kernel) {
                hipLaunchKernelGGL(kernel, dim3(grid_for(nthr)), dim3(kBlock), 0, s, fg, C, ec, Z, fmt_spectra<T>(F), O);
            };
            if (F.fold) {
                if (nd == 2)
                    df ? go(k_el_mul8<T, 2, true>) : go(k_el_mul8<T, 2, false>);
                else
                    df ? go(k_el_mul8<T, 3, true>) : go(k_el_mul8<T, 3, false>);
            } else if (nd == 2) {
                df ? go(k_el_mul<T, 2, true>) : go(k_el_mul<T, 2, false>);
            } else {
                df ? go(k_el_mul<T, 3, true>) : go(k_el_mul<T, 3, false>);
            }
            fft_pass<T>(F, O, nop, +1, 0, s);
        }
        if (ends) io.ss = ss;
        long long blocks = 0;
        for (int a : {1, 2}) blocks = fft_pass<T>(F, O, nop, +1, a, s, (ends && a == last) ? kPassEl : kPassPlain, io);
        if (ends) {
            if (ss && nss) *nss = blocks;
            TMI_HIP_CHECK(hipGetLastError());
            return;
        }
    }
    hipLaunchKernelGGL(k_el_extract<T>, dim3(grid_for(P.N), C), dim3(kBlock), 0, s, fg, gf, make_consts(P),
                       linear ? 1 : 0, P.p.nwall, O, x_rho, in_rho, in_mf, P.d_vext, P.d_elzero, add, out_el);
    TMI_HIP_CHECK(hipGetLastError());
}
template void fft_el<double>(Problem &, bool, const double *, const double *, const double *, const double *, double *,
                             hipStream_t, double *, long long *);
template void fft_el<float>(Problem &, bool, const double *, const double *, const double *, const double *, double *,
                            hipStream_t, double *, long long *);

// Times one kernel family of a residual fill in FFT form on the plan's work
// buffers, in the fill's order and batch sizes, with HIP events on s.
// which = 0: the plain axis passes (without fused ends); which = 1: the fused
// x passes (density and Euler-Lagrange products); which = 2: the density
// product's rows pass (fused end, on scratch rows).  *ms is their mean time per
// fill over reps fills, *bytes their algorithmic traffic per fill and
// *launches their launches per fill.  Algorithmic bytes: an axis pass reads
// and writes each grid point of each pair field once (16 B each way); a fused
// x pass reads each input pair once (16 B per point), each distinct spectrum
// table once (8 B per stored table element: the folded quarter or the whole
// grid) and writes each output pair once (16 B per point).  The buffers'
// contents are scratch between products.
void fft_time_kernels(Problem &P, int which, int reps, double *ms, double *bytes, int *launches, hipStream_t s)
{
    if (!fft_fmt(P)) throw std::logic_error("the FMT sums are not in FFT form");
    const FftPlan &F = *P.fft;
    if (which == 1 && !(F.xlpb[0] || F.xlpb[1])) throw std::logic_error("the fused x pass is not in use");
    const int C = P.p.ncomp, nadj = F.dform ? 4 + 2 * P.p.ndim : C * (1 + P.p.ndim) + 1;
    Cx<double> *Z = reinterpret_cast<Cx<double> *>(F.zbuf);
    Cx<double> *O = reinterpret_cast<Cx<double> *>(F.obuf);
    std::function<void()> fill;
    int count = 0;
    double b = 0.0;
    double *tmp = nullptr;
    if (which == 2) {   // the density product's last pass writing the rows (fused end), as in a residual fill
        if (!(F.wraps && fft_node_input(P))) throw std::logic_error("the density product has no fused rows pass");
        const RowMap &rm = F.rows[1];
        const int nrows = rm.nrb + rm.nmf, nop = (F.fwd[1].nout + 1) / 2;
        int last = -1;
        for (int a : {1, 2})
            if (F.M[a] > 1) last = a;
        TMI_HIP_CHECK(hipMalloc((void **)&tmp, sizeof(double) * 2 * (size_t)nrows * P.N));
        TMI_HIP_CHECK(hipMemsetAsync(tmp, 0, sizeof(double) * 2 * (size_t)nrows * P.N, s));
        const size_t N = P.N;
        PassIO io = rows_io(P, 1, nullptr, tmp, nullptr, tmp + (size_t)nrows * N, rm.nmf ? tmp + rm.nrb * N : nullptr,
                            nullptr, tmp + (size_t)nrows * N + rm.nrb * N);
        long long cap = 0;
        io.ss = fft_ss_buffer(P, &cap);
        count = 1;
        b = 16.0 * (double)F.n * nop + 16.0 * (double)nrows * N;
        fill = [&, io, last, nop] { fft_pass<double>(F, O, nop, +1, last, s, kPassRows, io); };
    } else if (which == 1) {
        const TermList *tls[2] = {&F.tfwd[1], F.dform ? &F.tel_d : &F.tel};
        for (int w = 0; w < 2; ++w) {
            if (!F.xlpb[w]) continue;
            const TermList *tl = tls[w];
            int tables = 0;
            for (int f = 0; f < tl->nin; ++f)
                for (int t = tl->first[f]; t < tl->first[f + 1]; ++t)
                    if (t == tl->first[f] || tl->term[t].t != tl->term[t - 1].t) ++tables;
            b += (double)F.n * (16.0 * ((tl->nin + 1) / 2) + 16.0 * ((tl->nout + 1) / 2)) + 8.0 * (double)F.tn * tables;
            ++count;
        }
        fill = [&] {
            if (F.xlpb[0]) fft_xprod<double>(F, Z, O, F.tfwd[1], F.xlpb[0], F.td, F.ud, s);
            if (F.xlpb[1]) fft_xprod<double>(F, Z, O, F.dform ? F.tel_d : F.tel, F.xlpb[1], F.td, F.ud, s);
        };
    } else {
        // which passes run with fused ends (those are other kernels): the density product's first and last
        // strided passes when it reads the node arrays, the Euler-Lagrange product's last one on wrapping grids
        const bool rb_ends = fft_node_input(P), el_ends = F.wraps;
        int first_strided = -1, last_strided = -1;
        for (int a : {2, 1})
            if (F.M[a] > 1 && first_strided < 0) first_strided = a;
        for (int a : {1, 2})
            if (F.M[a] > 1) last_strided = a;
        struct Set {
            Cx<double> *buf;
            int nfield, sign;
            bool fused;
            int skip;   // axis run as a fused-end pass, -1 if none
        };
        const Set sets[4] = {{Z, (C + 1) / 2, -1, F.xlpb[0] > 0, rb_ends ? first_strided : -1},
                             {O, (F.fwd[1].nout + 1) / 2, +1, F.xlpb[0] > 0, rb_ends ? last_strided : -1},
                             {Z, (nadj + 1) / 2, -1, F.xlpb[1] > 0, -1},
                             {O, (C + 1) / 2, +1, F.xlpb[1] > 0, el_ends ? last_strided : -1}};
        auto runs = [&](const Set &t, int axis) {
            return F.M[axis] > 1 && !(t.fused && axis == 0) && axis != t.skip;
        };
        for (const Set &t : sets)
            for (int axis = 0; axis < 3; ++axis)
                if (runs(t, axis)) {
                    ++count;
                    b += 2.0 * sizeof(Cx<double>) * (double)F.n * t.nfield;
                }
        fill = [&, sets, runs] {
            for (const Set &t : sets)
                for (int a = 0; a < 3; ++a) {
                    const int axis = t.sign < 0 ? 2 - a : a;   // forward z, y, x; back x, y, z
                    if (runs(t, axis)) fft_pass<double>(F, t.buf, t.nfield, t.sign, axis, s);
                }
        };
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    float t = 0.f;
    try {
        TMI_HIP_CHECK(hipEventCreate(&e0));
        TMI_HIP_CHECK(hipEventCreate(&e1));
        fill();
        TMI_HIP_CHECK(hipEventRecord(e0, s));
        for (int r = 0; r < reps; ++r) fill();
        TMI_HIP_CHECK(hipEventRecord(e1, s));
        TMI_HIP_CHECK(hipEventSynchronize(e1));
        TMI_HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
    } catch (...) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        (void)hipFree(tmp);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    (void)hipFree(tmp);
    *ms = (double)t / reps;
    *bytes = b;
    *launches = count;
}

// ---------------------------------------------------------------- bulk-circulant Schur preconditioner
// On a wrapping FFT grid the Schur operator S = A22 + A21 A12 at a uniform
// bulk state is a block convolution: S_b v = sum_j r_j * v_j with r_j = S_b e_j
// (unit impulse of component j at node 0), symbol S_b(k)_ij = FFT(r_j,i)(k).
// With S ~ D^-1 + K and D = diag(rho) (the reference's A22 diagonal
// preconditioner), S^-1 ~ S_b^-1 diag(1/rho_b) D, exact at the bulk state;
// GMRES applies it on the right: M^-1 v = F^-1[S_b(k)^-1 diag(1/rho_b) F(D v)]
// on the free rows, v on the zero rows (whose Schur rows are the identity).
// Built when every axis wraps and the Hermitian part of S_b(k) is positive
// definite at every k (a stable bulk); else GMRES keeps the diagonal.
void precond_setup(Problem &P)
{
    if (!P.fft) return;
    FftPlan &F = *P.fft;
    if (F.pc_want < 0) {
        const char *env = std::getenv("TMI_PRECOND");
        F.pc_want = (env && std::string(env) == "diag") ? 0 : 1;
    }
    F.pc_ok = F.pc_on = F.pc_fused = false;
    (void)hipFree(F.pq);
    (void)hipFree(F.pqr);
    (void)hipFree(F.pqr_f);
    F.pq = nullptr;
    F.pqr = nullptr;
    F.pqr_f = nullptr;
    if (P.npores > 0) return;
    if (!F.wraps) {
        // A walled or reflecting grid: its FFT grid pads every such axis (M = a 2-3-5 length of
        // n + 2 h).  The bulk Schur operator's symbol on that grid comes from a periodic bulk fluid
        // of M nodes per axis (same functional, components, potentials and rho_b; no walls), whose
        // FFT grid is the same grid with the same storage layout.  Applied to the zero-padded rho v,
        // it is a preconditioner only (not the inverse at any state): GMRES is right-preconditioned,
        // so the solution is unchanged.
        tmi_params q = P.p;
        for (int d = 0; d < P.p.ndim; ++d) {
            q.nx[d] = F.M[d];
            q.bc[d][0] = q.bc[d][1] = TMI_PERIODIC;
        }
        q.nwall = 0;
        q.lhard_surf = 0;
        q.att_in_a22 = 0;
        Problem *B = create_bulk_problem(q);
        const FftPlan *FB = B->fft;
        const bool same = FB && FB->pc_ok && FB->pq && FB->ns == F.ns && FB->S1 == F.S1 && FB->S2 == F.S2 &&
                          FB->M[0] == F.M[0] && FB->M[1] == F.M[1] && FB->M[2] == F.M[2];
        if (same) {
            const size_t bytes = (size_t)P.p.ncomp * P.p.ncomp * (size_t)F.ns * sizeof(Cx<double>);
            try {
                TMI_HIP_CHECK(hipMalloc((void **)&F.pq, bytes));
                TMI_HIP_CHECK(hipMemcpy(F.pq, FB->pq, bytes, hipMemcpyDeviceToDevice));
            } catch (...) {
                destroy_bulk_problem(B);
                throw;
            }
            F.pc_ok = true;
            F.pc_on = F.pc_want != 0;
        }
        destroy_bulk_problem(B);
        return;
    }
    const int C = P.p.ncomp;
    const size_t N = P.N, ns = (size_t)F.ns, nq = (size_t)P.nrb + P.nmf;
    double *xb = nullptr, *e = nullptr, *t1 = nullptr, *y = nullptr;
    unsigned char *flags = nullptr;
    unsigned char *zero_save = P.d_zero, *elzero_save = P.d_elzero;
    std::vector<std::complex<double>> S((size_t)C * C * ns);
    try {
        TMI_HIP_CHECK(hipMalloc((void **)&xb, (size_t)P.nunk * N * sizeof(double)));
        TMI_HIP_CHECK(hipMalloc((void **)&e, (size_t)C * N * sizeof(double)));
        TMI_HIP_CHECK(hipMalloc((void **)&t1, std::max<size_t>(nq, 1) * N * sizeof(double)));
        TMI_HIP_CHECK(hipMalloc((void **)&y, (size_t)C * N * sizeof(double)));
        TMI_HIP_CHECK(hipMalloc((void **)&flags, (size_t)C * N));
        TMI_HIP_CHECK(hipMemset(flags, 0, (size_t)C * N));
        {
            std::vector<double> h((size_t)P.nunk * N, 0.0);
            for (int c = 0; c < C; ++c)
                std::fill(h.begin() + (size_t)(P.iu_rho + c) * N, h.begin() + (size_t)(P.iu_rho + c + 1) * N,
                          P.p.rho_b[c]);
            TMI_HIP_CHECK(hipMemcpy(xb, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        P.d_zero = P.d_elzero = flags;   // the bulk operator has no zero rows
        launch_init_fields(P, xb, 0);
        std::vector<double> hy((size_t)C * N);
        std::vector<Cx<double>> grid(ns);
        Cx<double> *work = reinterpret_cast<Cx<double> *>(F.zbuf);
        for (int j = 0; j < C; ++j) {
            TMI_HIP_CHECK(hipMemset(e, 0, (size_t)C * N * sizeof(double)));
            const double one = 1.0;
            TMI_HIP_CHECK(hipMemcpy(e + (size_t)j * N, &one, sizeof(double), hipMemcpyHostToDevice));
            launch_schur_apply<double>(P, xb, e, t1, y, 0);
            TMI_HIP_CHECK(hipMemcpy(hy.data(), y, hy.size() * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 0; i < C; ++i) {
                std::fill(grid.begin(), grid.end(), Cx<double>{0.0, 0.0});
                for (size_t node = 0; node < N; ++node) {
                    const size_t x = node % P.p.nx[0], r = node / P.p.nx[0];
                    const size_t yy = P.p.ndim > 1 ? r % P.p.nx[1] : 0, z = P.p.ndim > 1 ? r / P.p.nx[1] : 0;
                    grid[x + yy * F.S1 + z * F.S2].x = hy[(size_t)i * N + node];
                }
                TMI_HIP_CHECK(hipMemcpy(work, grid.data(), ns * sizeof(Cx<double>), hipMemcpyHostToDevice));
                fft3d<double>(F, work, 1, -1, 0);
                TMI_HIP_CHECK(hipMemcpy(grid.data(), work, ns * sizeof(Cx<double>), hipMemcpyDeviceToHost));
                for (size_t q = 0; q < ns; ++q) S[(size_t)(i + C * j) * ns + q] = {grid[q].x, grid[q].y};
            }
        }
        P.d_zero = zero_save;
        P.d_elzero = elzero_save;
    } catch (...) {
        P.d_zero = zero_save;
        P.d_elzero = elzero_save;
        (void)hipFree(xb);
        (void)hipFree(e);
        (void)hipFree(t1);
        (void)hipFree(y);
        (void)hipFree(flags);
        throw;
    }
    (void)hipFree(xb);
    (void)hipFree(e);
    (void)hipFree(t1);
    (void)hipFree(y);
    (void)hipFree(flags);
    // per grid point: Cholesky of the Hermitian part (positive definite), then Q = S^-1 diag(1/rho_b) / n
    std::vector<Cx<double>> hq((size_t)C * C * ns, Cx<double>{0.0, 0.0});
    const double inv_n = 1.0 / (double)F.n;
    using cd = std::complex<double>;
    for (int m2 = 0; m2 < F.M[2]; ++m2)
        for (int m1 = 0; m1 < F.M[1]; ++m1)
            for (int m0 = 0; m0 < F.M[0]; ++m0) {
                const size_t q = m0 + m1 * F.S1 + m2 * F.S2;
                cd A[kNC][kNC], Hm[kNC][kNC], Inv[kNC][kNC];
                for (int i = 0; i < C; ++i)
                    for (int j = 0; j < C; ++j) A[i][j] = S[(size_t)(i + C * j) * ns + q];
                for (int i = 0; i < C; ++i)
                    for (int j = 0; j < C; ++j) Hm[i][j] = 0.5 * (A[i][j] + std::conj(A[j][i]));
                for (int k = 0; k < C; ++k) {   // in-place Cholesky, lower triangle
                    double dkk = Hm[k][k].real();
                    for (int l = 0; l < k; ++l) dkk -= std::norm(Hm[k][l]);
                    if (!(dkk > 0.0)) return;   // not positive definite: keep the diagonal preconditioner
                    const double lkk = std::sqrt(dkk);
                    Hm[k][k] = lkk;
                    for (int i = k + 1; i < C; ++i) {
                        cd sum = Hm[i][k];
                        for (int l = 0; l < k; ++l) sum -= Hm[i][l] * std::conj(Hm[k][l]);
                        Hm[i][k] = sum / lkk;
                    }
                }
                // Gauss-Jordan with partial pivoting: Inv = A^-1
                for (int i = 0; i < C; ++i)
                    for (int j = 0; j < C; ++j) Inv[i][j] = (i == j) ? 1.0 : 0.0;
                for (int k = 0; k < C; ++k) {
                    int piv = k;
                    for (int i = k + 1; i < C; ++i)
                        if (std::abs(A[i][k]) > std::abs(A[piv][k])) piv = i;
                    if (std::abs(A[piv][k]) == 0.0) return;
                    for (int j = 0; j < C; ++j) {
                        std::swap(A[k][j], A[piv][j]);
                        std::swap(Inv[k][j], Inv[piv][j]);
                    }
                    const cd d = 1.0 / A[k][k];
                    for (int j = 0; j < C; ++j) {
                        A[k][j] *= d;
                        Inv[k][j] *= d;
                    }
                    for (int i = 0; i < C; ++i) {
                        if (i == k) continue;
                        const cd f = A[i][k];
                        for (int j = 0; j < C; ++j) {
                            A[i][j] -= f * A[k][j];
                            Inv[i][j] -= f * Inv[k][j];
                        }
                    }
                }
                for (int i = 0; i < C; ++i)
                    for (int j = 0; j < C; ++j) {
                        const cd v = Inv[i][j] * (inv_n / P.p.rho_b[j]);
                        hq[(size_t)(i + C * j) * ns + q] = Cx<double>{v.real(), v.imag()};
                    }
            }
    TMI_HIP_CHECK(hipMalloc((void **)&F.pq, hq.size() * sizeof(Cx<double>)));
    TMI_HIP_CHECK(hipMemcpy(F.pq, hq.data(), hq.size() * sizeof(Cx<double>), hipMemcpyHostToDevice));
    F.pc_ok = true;
    F.pc_on = F.pc_want != 0;
    // the fused form: Q real (imaginary parts below 1e-9 of the largest entry) and, with folded
    // tables, even in k_y and k_z to the same tolerance; the x pass fused with the product fits
    const size_t tn = (size_t)F.tn;
    std::vector<double> qr((size_t)C * C * tn, 0.0);
    double qmax = 0.0, bad = 0.0;
    for (int e = 0; e < C * C; ++e)
        for (int m2 = 0; m2 < F.M[2]; ++m2)
            for (int m1 = 0; m1 < F.M[1]; ++m1)
                for (int m0 = 0; m0 < F.M[0]; ++m0) {
                    const Cx<double> v = hq[(size_t)e * ns + m0 + m1 * F.S1 + m2 * F.S2];
                    qmax = std::max(qmax, std::fabs(v.x));
                    bad = std::max(bad, std::fabs(v.y));
                    if (!F.fold) {
                        qr[(size_t)e * tn + m0 + m1 * F.S1 + m2 * F.S2] = v.x;
                        continue;
                    }
                    const int f1 = 2 * m1 > F.M[1] ? F.M[1] - m1 : m1, f2 = 2 * m2 > F.M[2] ? F.M[2] - m2 : m2;
                    bad = std::max(bad, std::fabs(v.x - hq[(size_t)e * ns + m0 + f1 * F.S1 + f2 * F.S2].x));
                    if (f1 == m1 && f2 == m2) qr[(size_t)e * tn + m0 + m1 * F.tp1 + m2 * F.tp2] = v.x;
                }
    F.pc_lpb = xprod_lpb(F, (C + 1) / 2);
    const char *pf = std::getenv("TMI_PRECOND_FUSED");   // 0: the separate gather, product and extract kernels
    if (bad <= 1e-9 * qmax && F.pc_lpb > 0 && !(pf && std::string(pf) == "0")) {
        TMI_HIP_CHECK(hipMalloc((void **)&F.pqr, qr.size() * sizeof(double)));
        TMI_HIP_CHECK(hipMemcpy(F.pqr, qr.data(), qr.size() * sizeof(double), hipMemcpyHostToDevice));
        TMI_HIP_CHECK(hipMalloc((void **)&F.pqr_f, qr.size() * sizeof(float)));
        hipLaunchKernelGGL(k_to_float, dim3(grid_for((long long)qr.size())), dim3(kBlock), 0, 0, F.pqr, F.pqr_f,
                           (long long)qr.size());
        TMI_HIP_CHECK(hipStreamSynchronize(0));
        TermList &tl = F.tpc;
        tl = TermList();
        tl.nout = C;
        for (int j = 0; j < C; ++j) {   // input component j feeds every output i through table i + C j
            for (int i = 0; i < C; ++i) tl.term[tl.first[tl.nin + 1]++] = Term{i, i + C * j, 0, -1, 1.0};
            ++tl.nin;
            tl.first[tl.nin + 1] = tl.first[tl.nin];
        }
        F.pc_fused = true;
    }
}

// a new rho_b marks Q stale; the queries below rebuild it first, so parameter sweeps that only fill
// (finite differences in the bulk density) never build the bulk problem of a walled grid
void precond_invalidate(Problem &P)
{
    if (P.fft) P.fft->pc_stale = true;
}
static void precond_refresh(Problem &P)
{
    if (!P.fft || !P.fft->pc_stale) return;
    TMI_HIP_CHECK(hipDeviceSynchronize());   // the set-up runs on the null stream and reuses the transform buffers
    precond_setup(P);
    P.fft->pc_stale = false;
}
bool precond_available(Problem &P)
{
    precond_refresh(P);
    return P.fft && P.fft->pc_ok;
}
bool precond_active(Problem &P)
{
    precond_refresh(P);
    return P.fft && P.fft->pc_on;
}
void precond_enable(Problem &P, bool on)
{
    if (on && !precond_available(P))
        throw std::domain_error("the bulk preconditioner needs an FFT grid and a stable bulk");
    if (P.fft) {
        P.fft->pc_want = on ? 1 : 0;
        P.fft->pc_on = on;
    }
}

// The fused apply in precision T: z, y passes (the first gathers rho v); x pass fused with Q; y, z
// passes (the last writes z, FP64)
template <typename T>
static void precond_apply_fused(Problem &P, const double *x_rho, const double *v, double *z, const T *qr,
                                hipStream_t s)
{
    const FftPlan &F = *P.fft;
    const int C = P.p.ncomp;
    Cx<T> *Z = reinterpret_cast<Cx<T> *>(F.zbuf);
    Cx<T> *O = reinterpret_cast<Cx<T> *>(F.obuf);
    const int np = (C + 1) / 2;
    PassIO io;
    io.C = C;
    io.nx0 = P.p.nx[0];
    io.nx1 = P.p.ndim > 1 ? P.p.nx[1] : 1;
    io.N = P.N;
    io.rho = x_rho;
    io.mul = v;
    io.zero = P.d_elzero;
    io.pv = v;
    io.elzero = P.d_elzero;
    io.out_el = z;
    int first = -1, last = -1;
    for (int a : {2, 1})
        if (F.M[a] > 1 && first < 0) first = a;
    for (int a : {1, 2})
        if (F.M[a] > 1) last = a;
    for (int a : {2, 1}) fft_pass<T>(F, Z, np, -1, a, s, a == first ? kPassGather : kPassPlain, io);
    fft_xprod<T>(F, Z, O, F.tpc, F.pc_lpb, qr, nullptr, s);
    for (int a : {1, 2}) fft_pass<T>(F, O, np, +1, a, s, a == last ? kPassPc : kPassPlain, io);
    TMI_HIP_CHECK(hipGetLastError());
}

// z = M^-1 v (bulk-circulant preconditioner), x_rho the densities of the Newton iterate; fp32: the
// fused apply's transforms in FP32 (the mixed path)
void precond_apply(Problem &P, const double *x_rho, const double *v, double *z, hipStream_t s, bool fp32)
{
    const FftPlan &F = *P.fft;
    const int C = P.p.ncomp;
    Cx<double> *Z = reinterpret_cast<Cx<double> *>(F.zbuf);
    Cx<double> *O = reinterpret_cast<Cx<double> *>(F.obuf);
    if (F.pc_fused && fp32 && F.pqr_f) {
        precond_apply_fused<float>(P, x_rho, v, z, F.pqr_f, s);
        return;
    }
    if (F.pc_fused) {
        precond_apply_fused<double>(P, x_rho, v, z, F.pqr, s);
        return;
    }
    const Geo g = make_geo(P, P.pad_u);
    const FftGeo fg = fft_geo(F, 0);
    if (!F.wraps) {   // zero padding around the nodes (a fill kernel: no memset nodes in captured graphs)
        const long long nz = (long long)((C + 1) / 2) * F.ns;
        hipLaunchKernelGGL(k_zero_cx, dim3(grid_for(nz)), dim3(kBlock), 0, s, Z, nz);
    }
    hipLaunchKernelGGL(k_pc_gather, dim3(grid_for(P.N)), dim3(kBlock), 0, s, fg, g, x_rho, v, P.d_elzero, Z);
    fft3d<double>(F, Z, (C + 1) / 2, -1, s);
    hipLaunchKernelGGL(k_pc_mul, dim3(grid_for(F.n)), dim3(kBlock), 0, s, fg, C, Z, F.pq, O);
    fft3d<double>(F, O, (C + 1) / 2, +1, s);
    hipLaunchKernelGGL(k_pc_extract, dim3(grid_for(P.N), C), dim3(kBlock), 0, s, fg, g, O, v, P.d_elzero, z);
    TMI_HIP_CHECK(hipGetLastError());
}

}  // namespace tmi
