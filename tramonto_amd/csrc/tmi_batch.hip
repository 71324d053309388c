// Batched Newton-Krylov for many independent 1D problems ("pores") packed
// into one device problem (see Problem::npores).
//
// Every pore is its own nonlinear system: the reference solves one such
// problem with newton_solver (src/dft_newton.c:131-196), a Schur-complement
// GMRES per step (src/dft_HardSphereLinProbMgr.cpp:345-428) and the damped
// update of update_solution (:266-380).  Here all pores advance in lockstep:
//
//  * the fill, the Schur products and the Newton update run once over the
//    whole packed batch (the stencil kernels see each pore between wall gaps);
//  * GMRES keeps one Krylov basis over the packed vector but a separate
//    Arnoldi process per pore: the Gram-Schmidt inner products are segmented
//    by pore, each pore has its own Hessenberg matrix and Givens rotations
//    and stops (freezes its iteration count) at its own tolerance;
//  * damping (frac_min), the convergence test and the iteration counts are
//    per pore; converged pores take no further updates.
//
// One Newton step is three device-only stages with no host synchronisation
// inside -- prefix (fill, norms, Schur right-hand side), a GMRES restart
// cycle of m Arnoldi steps, and the update -- each captured once into a HIP
// graph on a library-owned stream and replayed every iteration.  The host
// reads two counters per step (pores still active, pores whose GMRES has not
// converged).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "tmi_device.hpp"
#include "tmi_internal.hpp"

namespace tmi {

namespace {

using namespace dev;

constexpr int kSegBlock = 512;
constexpr int kChunk = 8;   // Arnoldi steps per captured graph

__device__ inline double seg_block_sum(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x < 64) {
        t = (threadIdx.x < kSegBlock / 64) ? sh[threadIdx.x] : 0.0;
        for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
    }
    return t;   // valid in thread 0
}

// out[p*ldo + j] = sum over pore p of V_j . w   (rows x nodes of the pore): one workgroup per pore,
// kSegGroup vectors per sweep over the pore's elements (w re-read per group, from cache) and one
// reduction round per group
constexpr int kSegGroup = 8;
__global__ __launch_bounds__(kSegBlock) void k_seg_dots(const double *__restrict__ V, long long ldv, int k,
                                                        const double *__restrict__ w, int rows, int N,
                                                        const int *__restrict__ off, double *__restrict__ out, int ldo)
{
    __shared__ double sh[kSegBlock / 64][kSegGroup];
    const int p = blockIdx.x;
    const int i0 = off[p], n = off[p + 1] - i0;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int j0 = 0; j0 < k; j0 += kSegGroup) {
        double acc[kSegGroup];
#pragma unroll
        for (int g = 0; g < kSegGroup; ++g) acc[g] = 0.0;
        for (int r = 0; r < rows; ++r) {
            const size_t base = (size_t)r * N + i0;
            for (int i = threadIdx.x; i < n; i += kSegBlock) {
                const double wv = w[base + i];
#pragma unroll
                for (int g = 0; g < kSegGroup; ++g)
                    if (j0 + g < k) acc[g] += V[(size_t)(j0 + g) * ldv + base + i] * wv;
            }
        }
#pragma unroll
        for (int g = 0; g < kSegGroup; ++g) {
            double a = acc[g];
            for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
            if (lane == 0) sh[wid][g] = a;
        }
        __syncthreads();
        if (threadIdx.x < kSegGroup && j0 + (int)threadIdx.x < k) {
            double t = 0.0;
            for (int q = 0; q < kSegBlock / 64; ++q) t += sh[q][threadIdx.x];
            out[(size_t)p * ldo + j0 + threadIdx.x] = t;
        }
        __syncthreads();   // sh is reused by the next group
    }
}

// w[e] += alpha * sum_j V_j[e] coef[pore(e)][j]
__global__ void k_seg_gemv(const double *__restrict__ V, long long ldv, int k, const double *__restrict__ coef, int ldc,
                           double alpha, double *__restrict__ w, int rows, int N, const int *__restrict__ pore_of)
{
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (long long)rows * N) return;
    const int i = (int)(e % N);
    const double *c = coef + (size_t)pore_of[i] * ldc;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += V[(size_t)j * ldv + e] * c[j];
    w[e] += alpha * s;
}

__global__ void k_fill(double *__restrict__ y, double v, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = v;
}

// y = a - b
__global__ void k_sub(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ y, long long n)
{
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = a[i] - b[i];
}

// out[p] = sum over pore p of x^2 (rows x nodes)
__global__ __launch_bounds__(kSegBlock) void k_seg_sumsq(const double *__restrict__ x, int rows, int N,
                                                         const int *__restrict__ off, double *__restrict__ out)
{
    __shared__ double sh[kSegBlock / 64];
    const int p = blockIdx.x;
    const int i0 = off[p], n = off[p + 1] - i0;
    double s = 0.0;
    for (int e = threadIdx.x; e < rows * n; e += kSegBlock) {
        const int r = e / n, i = e - r * n;
        const double v = x[(size_t)r * N + i0 + i];
        s += v * v;
    }
    s = seg_block_sum(s, sh);
    if (threadIdx.x == 0) out[p] = s;
}

// Normalise pore p's segment of w into v and store the norm in hn[p].
__global__ __launch_bounds__(kSegBlock) void k_seg_normalize(const double *__restrict__ w, double *__restrict__ v,
                                                             int rows, int N, const int *__restrict__ off,
                                                             double *__restrict__ hn)
{
    __shared__ double sh[kSegBlock / 64];
    __shared__ double nrm;
    const int p = blockIdx.x;
    const int i0 = off[p], n = off[p + 1] - i0;
    double s = 0.0;
    for (int e = threadIdx.x; e < rows * n; e += kSegBlock) {
        const int r = e / n, i = e - r * n;
        const double t = w[(size_t)r * N + i0 + i];
        s += t * t;
    }
    s = seg_block_sum(s, sh);
    if (threadIdx.x == 0) {
        nrm = sqrt(s);
        hn[p] = nrm;
    }
    __syncthreads();
    const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
    for (int e = threadIdx.x; e < rows * n; e += kSegBlock) {
        const int r = e / n, i = e - r * n;
        const size_t q = (size_t)r * N + i0 + i;
        v[q] = w[q] * inv;
    }
}

// Per-pore GMRES bookkeeping, one thread per pore.
struct GmresState {
    double *H;       // [P][m+1][m] (rotated: upper triangular)
    double *cs, *sn; // [P][m]
    double *g;       // [P][m+1]
    double *y;       // [P][m]
    double *h1, *h2; // [P][m+1] Gram-Schmidt coefficients of the two passes
    double *hn;      // [P]
    double *bn;      // [P] norm of the right-hand side
    int *kdone;      // [P] Arnoldi steps taken in this cycle
    int *done;       // [P] converged (or inactive)
    int *its;        // [P] GMRES iterations over the whole solve
    int m;
};

// start of a restart cycle: hn[p] = beta_p (already normalised V0);
// done_p is set for inactive pores and pores already at tolerance
__global__ void k_gmres_start(int P, GmresState st, const int *__restrict__ active, double tol)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int m = st.m;
    const double beta = st.hn[p];
    double *g = st.g + (size_t)p * (m + 1);
    for (int j = 0; j <= m; ++j) g[j] = 0.0;
    g[0] = beta;
    st.kdone[p] = 0;
    st.done[p] = (!active[p] || beta <= tol * st.bn[p] || beta == 0.0) ? 1 : 0;
}

// Arnoldi step k: new column (h1 + h2, hn) of each active pore's Hessenberg
// matrix, previous rotations applied, new rotation formed, residual updated.
__global__ void k_gmres_givens(int P, int k, GmresState st, double tol)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P || st.done[p]) return;
    const int m = st.m;
    double *H = st.H + (size_t)p * (m + 1) * m;
    double *cs = st.cs + (size_t)p * m, *sn = st.sn + (size_t)p * m, *g = st.g + (size_t)p * (m + 1);
    const double *h1 = st.h1 + (size_t)p * (m + 1), *h2 = st.h2 + (size_t)p * (m + 1);
    for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = h1[j] + h2[j];
    const double hn = st.hn[p];
    H[(size_t)(k + 1) * m + k] = hn;
    for (int j = 0; j < k; ++j) {
        const double a = H[(size_t)j * m + k], b = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a + sn[j] * b;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * b;
    }
    const double a = H[(size_t)k * m + k], b = H[(size_t)(k + 1) * m + k];
    const double den = sqrt(a * a + b * b);
    cs[k] = den > 0.0 ? a / den : 1.0;
    sn[k] = den > 0.0 ? b / den : 0.0;
    H[(size_t)k * m + k] = den;
    H[(size_t)(k + 1) * m + k] = 0.0;
    g[k + 1] = -sn[k] * g[k];
    g[k] = cs[k] * g[k];
    st.kdone[p] = k + 1;
    if (fabs(g[k + 1]) <= tol * st.bn[p] || hn == 0.0) st.done[p] = 1;
}

// y = R^{-1} g over each pore's kdone steps (zero beyond); iteration counts
__global__ void k_gmres_solve(int P, GmresState st)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int m = st.m, k = st.kdone[p];
    const double *H = st.H + (size_t)p * (m + 1) * m, *g = st.g + (size_t)p * (m + 1);
    double *y = st.y + (size_t)p * m;
    for (int j = k; j < m; ++j) y[j] = 0.0;
    for (int i = k - 1; i >= 0; --i) {
        double s = g[i];
        for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * y[j];
        y[i] = s / H[(size_t)i * m + i];
    }
    st.its[p] += k;
}

// pores still running: count[0] = #(active), or #(active and GMRES not done)
__global__ void k_count(int P, const int *__restrict__ active, const int *__restrict__ done, int *__restrict__ count)
{
    __shared__ int sh[kBlock];
    int c = 0;
    for (int p = threadIdx.x; p < P; p += kBlock) c += (active[p] && (!done || !done[p])) ? 1 : 0;
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) count[0] = sh[0];
}

// Newton bookkeeping after a fill: pores whose ||F_p|| < tol stop.
__global__ void k_newton_flags(int P, const double *__restrict__ rn2, double tol, int iter, int *__restrict__ active,
                               int *__restrict__ conv, int *__restrict__ iters, double *__restrict__ rnorm)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const double rn = sqrt(rn2[p]);
    rnorm[p] = rn;
    if (!active[p]) return;
    if (rn < tol) {
        active[p] = 0;
        conv[p] = 1;
        iters[p] = iter;
    } else if (!(rn == rn)) {
        active[p] = 0;   // NaN: stop this pore, reported as not converged
        iters[p] = iter;
    }
}

// update_solution per pore: frac_p = min over the pore's densities of
// max(min(1, x/(-dx)), scaling) where x + dx < 0
__global__ __launch_bounds__(kSegBlock) void k_seg_frac_min(const double *__restrict__ x, const double *__restrict__ dx,
                                                            int C, int N, const int *__restrict__ off, double scaling,
                                                            const int *__restrict__ active, double *__restrict__ frac)
{
    __shared__ double sh[kSegBlock];
    const int p = blockIdx.x;
    const int i0 = off[p], n = off[p + 1] - i0;
    double mn = 1.0;
    for (int e = threadIdx.x; e < C * n; e += kSegBlock) {
        const int r = e / n, i = e - r * n;
        const size_t q = (size_t)r * N + i0 + i;
        const double xv = x[q], d = dx[q];
        if (xv + d < 0.0) mn = fmin(mn, fmax(fmin(1.0, xv / (-d)), scaling));
    }
    sh[threadIdx.x] = mn;
    __syncthreads();
    for (int s = kSegBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) frac[p] = active[p] ? sh[0] : 0.0;
}

// x += frac_p dx with the clamps of update_solution (src/dft_newton.c:266-380)
__global__ void k_seg_apply_update(double *__restrict__ x, const double *__restrict__ dx, long long n, int C, int N,
                                   int iu_n3, const int *__restrict__ pore_of, const double *__restrict__ frac)
{
    const long long e = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n) return;
    const int u = (int)(e / N), i = (int)(e % N);
    const double f = frac[pore_of[i]];
    if (f == 0.0) return;
    const double xv = x[e], d = dx[e];
    // the limiting density goes to 0.1 x as in exact arithmetic, not to the rounding residue of
    // x + f dx (k_apply_update)
    const bool limited = u < C && xv + d < 0.0 && fmin(1.0, xv / (-d)) <= f;
    if (u < C && (xv + f * d < 1.e-99 || limited))
        x[e] = 0.1 * xv;
    else if (u == iu_n3 && xv + f * d > 1.0)
        x[e] = xv + 0.5 * (1.0 - xv);
    else
        x[e] = xv + f * d;
}

inline unsigned nblk(long long n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------- work buffers
struct BatchWork {
    int P = 0, m = 0;
    size_t ntot = 0, n2 = 0;
    hipStream_t stream = nullptr;
    double *buf = nullptr;     // all doubles below
    int *ibuf = nullptr;       // all ints below
    int *h_counts = nullptr;   // pinned [2]
    // vectors
    double *F, *b, *dx, *bS, *dinv, *V, *z, *w, *t1, *r;
    // per pore
    double *rn2, *rnorm, *frac;
    GmresState gs;
    int *active, *conv, *niters, *counts;
    // graphs
    hipGraphExec_t g_pre = nullptr, g_start = nullptr, g_finish = nullptr, g_post = nullptr;
    std::vector<hipGraphExec_t> g_steps;   // Arnoldi steps [j*kChunk, (j+1)*kChunk)
    const double *graph_x = nullptr;
    int graph_prec = -1;
    double graph_tol[2] = {0, 0};
};

void free_batch_work(Problem &P)
{
    BatchWork *W = P.batch_work;
    if (!W) return;
    for (hipGraphExec_t g : {W->g_pre, W->g_start, W->g_finish, W->g_post})
        if (g) (void)hipGraphExecDestroy(g);
    for (hipGraphExec_t g : W->g_steps)
        if (g) (void)hipGraphExecDestroy(g);
    (void)hipFree(W->buf);
    (void)hipFree(W->ibuf);
    if (W->h_counts) (void)hipHostFree(W->h_counts);
    if (W->stream) (void)hipStreamDestroy(W->stream);
    delete W;
    P.batch_work = nullptr;
}

static BatchWork &batch_work(Problem &P, int m)
{
    if (P.batch_work && P.batch_work->m == m) return *P.batch_work;
    free_batch_work(P);
    auto *W = new BatchWork();
    P.batch_work = W;
    W->P = P.npores;
    W->m = m;
    const int Pn = P.npores;
    W->ntot = (size_t)P.nunk * P.N;
    W->n2 = (size_t)P.p.ncomp * P.N;
    const size_t nt = W->ntot, n2 = W->n2;
    const size_t nd = 3 * nt + n2 /*bS*/ + n2 /*dinv*/ + (size_t)(m + 1) * n2 /*V*/ + 3 * n2 /*z w r*/ +
                      (size_t)(P.nrb + P.nmf) * P.N /*t1*/ + 3 * (size_t)Pn +
                      (size_t)Pn * ((size_t)(m + 1) * m + 2 * m + (m + 1) + m + 2 * (m + 1) + 2);
    const size_t ni = (size_t)Pn * 6 + 2;
    TMI_HIP_CHECK(hipStreamCreateWithFlags(&W->stream, hipStreamNonBlocking));
    TMI_HIP_CHECK(hipMalloc((void **)&W->buf, nd * sizeof(double)));
    TMI_HIP_CHECK(hipMalloc((void **)&W->ibuf, ni * sizeof(int)));
    TMI_HIP_CHECK(hipHostMalloc((void **)&W->h_counts, 2 * sizeof(int), hipHostMallocDefault));
    double *q = W->buf;
    auto take = [&](size_t n) {
        double *r = q;
        q += n;
        return r;
    };
    W->F = take(nt);
    W->b = take(nt);
    W->dx = take(nt);
    W->bS = take(n2);
    W->dinv = take(n2);
    W->V = take((size_t)(m + 1) * n2);
    W->z = take(n2);
    W->w = take(n2);
    W->r = take(n2);
    W->t1 = take((size_t)(P.nrb + P.nmf) * P.N);
    W->rn2 = take(Pn);
    W->rnorm = take(Pn);
    W->frac = take(Pn);
    W->gs.m = m;
    W->gs.H = take((size_t)Pn * (m + 1) * m);
    W->gs.cs = take((size_t)Pn * m);
    W->gs.sn = take((size_t)Pn * m);
    W->gs.g = take((size_t)Pn * (m + 1));
    W->gs.y = take((size_t)Pn * m);
    W->gs.h1 = take((size_t)Pn * (m + 1));
    W->gs.h2 = take((size_t)Pn * (m + 1));
    W->gs.hn = take(Pn);
    W->gs.bn = take(Pn);
    int *iq = W->ibuf;
    W->gs.kdone = iq;
    iq += Pn;
    W->gs.done = iq;
    iq += Pn;
    W->gs.its = iq;
    iq += Pn;
    W->active = iq;
    iq += Pn;
    W->conv = iq;
    iq += Pn;
    W->niters = iq;
    iq += Pn;
    W->counts = iq;
    return *W;
}

// ---------------------------------------------------------------- stages
// Schur product S v (v: density rows) in the chosen precision
static void schur_apply(Problem &P, const double *x, const double *v, double *t1, double *y, int prec, hipStream_t s)
{
    if (prec == TMI_PREC_MIXED)
        launch_schur_apply<float>(P, x, v, t1, y, s);
    else
        launch_schur_apply<double>(P, x, v, t1, y, s);
}

// prefix: F(x) and its per-pore sums of squares, b = -F, Schur right-hand
// side bS and its per-pore norms, A22^{-1}, x2 = 0
static void stage_pre(Problem &P, BatchWork &W, const double *x, hipStream_t s)
{
    const int Pn = W.P, C = P.p.ncomp, N = P.N;
    launch_residual(P, x, W.F, s);
    hipLaunchKernelGGL(k_seg_sumsq, dim3(Pn), dim3(kSegBlock), 0, s, W.F, P.nunk, N, P.d_pore_off, W.rn2);
    dev_scale_copy(-1.0, W.F, W.b, (long long)W.ntot, s);
    launch_schur_rhs(P, x, W.b, W.bS, s);
    dev_a22_inverse_diag(P, x, W.dinv, s);
    // bn_p = |bS|_p (the normalised copy in r is not used)
    hipLaunchKernelGGL(k_seg_normalize, dim3(Pn), dim3(kSegBlock), 0, s, W.bS, W.r, C, N, P.d_pore_off, W.gs.bn);
    hipLaunchKernelGGL(k_fill, dim3(nblk((long long)W.n2)), dim3(kBlock), 0, s, W.dx, 0.0, (long long)W.n2);
    TMI_HIP_CHECK(hipGetLastError());
}

// One GMRES restart cycle on S x2 = bS per pore (x2 = dx[0 : C N]) in three
// parts: start (r = bS - S x2, V0, per-pore beta), Arnoldi steps [k0, k1),
// finish (per-pore least squares, x2 update).  Each part ends with the count
// of pores whose GMRES is still running, so the host can skip the remaining
// steps of a cycle once every pore has converged.
static void stage_cycle_start(Problem &P, BatchWork &W, const double *x, double tol, int prec, hipStream_t s)
{
    const int Pn = W.P, C = P.p.ncomp, N = P.N;
    const long long n2 = (long long)W.n2;
    schur_apply(P, x, W.dx, W.t1, W.w, prec, s);
    hipLaunchKernelGGL(k_sub, dim3(nblk(n2)), dim3(kBlock), 0, s, W.bS, W.w, W.r, n2);
    hipLaunchKernelGGL(k_seg_normalize, dim3(Pn), dim3(kSegBlock), 0, s, W.r, W.V, C, N, P.d_pore_off, W.gs.hn);
    hipLaunchKernelGGL(k_gmres_start, dim3(nblk(Pn)), dim3(kBlock), 0, s, Pn, W.gs, W.active, tol);
    hipLaunchKernelGGL(k_count, dim3(1), dim3(kBlock), 0, s, Pn, W.active, W.gs.done, W.counts + 1);
    TMI_HIP_CHECK(hipGetLastError());
}

static void stage_cycle_steps(Problem &P, BatchWork &W, const double *x, double tol, int prec, int k0, int k1,
                              hipStream_t s)
{
    const int Pn = W.P, C = P.p.ncomp, N = P.N, m = W.m;
    const long long n2 = (long long)W.n2;
    for (int k = k0; k < k1; ++k) {
        double *Vk = W.V + (size_t)k * n2;
        dev_mul(W.dinv, Vk, W.z, n2, s);
        schur_apply(P, x, W.z, W.t1, W.w, prec, s);
        hipLaunchKernelGGL(k_seg_dots, dim3(Pn), dim3(kSegBlock), 0, s, W.V, n2, k + 1, W.w, C, N, P.d_pore_off,
                           W.gs.h1, m + 1);
        hipLaunchKernelGGL(k_seg_gemv, dim3(nblk(n2)), dim3(kBlock), 0, s, W.V, n2, k + 1, W.gs.h1, m + 1, -1.0, W.w,
                           C, N, P.d_pore_of);
        hipLaunchKernelGGL(k_seg_dots, dim3(Pn), dim3(kSegBlock), 0, s, W.V, n2, k + 1, W.w, C, N, P.d_pore_off,
                           W.gs.h2, m + 1);
        hipLaunchKernelGGL(k_seg_gemv, dim3(nblk(n2)), dim3(kBlock), 0, s, W.V, n2, k + 1, W.gs.h2, m + 1, -1.0, W.w,
                           C, N, P.d_pore_of);
        hipLaunchKernelGGL(k_seg_normalize, dim3(Pn), dim3(kSegBlock), 0, s, W.w, W.V + (size_t)(k + 1) * n2, C, N,
                           P.d_pore_off, W.gs.hn);
        hipLaunchKernelGGL(k_gmres_givens, dim3(nblk(Pn)), dim3(kBlock), 0, s, Pn, k, W.gs, tol);
    }
    hipLaunchKernelGGL(k_count, dim3(1), dim3(kBlock), 0, s, Pn, W.active, W.gs.done, W.counts + 1);
    TMI_HIP_CHECK(hipGetLastError());
}

static void stage_cycle_finish(Problem &P, BatchWork &W, hipStream_t s)
{
    const int Pn = W.P, C = P.p.ncomp, N = P.N, m = W.m;
    const long long n2 = (long long)W.n2;
    hipLaunchKernelGGL(k_gmres_solve, dim3(nblk(Pn)), dim3(kBlock), 0, s, Pn, W.gs);
    // x2 += dinv * (V y); y is zero beyond each pore's steps
    hipLaunchKernelGGL(k_fill, dim3(nblk(n2)), dim3(kBlock), 0, s, W.z, 0.0, n2);
    hipLaunchKernelGGL(k_seg_gemv, dim3(nblk(n2)), dim3(kBlock), 0, s, W.V, n2, m, W.gs.y, m, 1.0, W.z, C, N,
                       P.d_pore_of);
    dev_mul(W.dinv, W.z, W.w, n2, s);
    dev_axpy(1.0, W.w, W.dx, n2, s);
    TMI_HIP_CHECK(hipGetLastError());
}

// back-substitution of the nonlocal unknowns, per-pore damping and update
static void stage_post(Problem &P, BatchWork &W, double *x, double scaling, hipStream_t s)
{
    const int Pn = W.P, C = P.p.ncomp, N = P.N;
    launch_schur_backsub(P, x, W.b, W.dx, s);
    hipLaunchKernelGGL(k_seg_frac_min, dim3(Pn), dim3(kSegBlock), 0, s, x, W.dx, C, N, P.d_pore_off, scaling,
                       W.active, W.frac);
    hipLaunchKernelGGL(k_seg_apply_update, dim3(nblk((long long)W.ntot)), dim3(kBlock), 0, s, x, W.dx,
                       (long long)W.ntot, C, N, P.iu_rb, P.d_pore_of, W.frac);
    TMI_HIP_CHECK(hipGetLastError());
}

template <typename F>
static hipGraphExec_t capture(hipStream_t s, F &&enqueue)
{
    hipGraph_t g = nullptr;
    TMI_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    try {
        enqueue();
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
    }
    TMI_HIP_CHECK(hipStreamEndCapture(s, &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    TMI_HIP_CHECK(e);
    return ex;
}

int batch_newton_solve(Problem &P, double *x, const tmi_newton_opts &o, tmi_newton_stats &st, int *pore_iters,
                       int *pore_conv, int *pore_gmres, double *pore_resid, hipStream_t caller)
{
    if (P.npores < 1) throw std::logic_error("not a batched problem");
    const int m = std::max(1, std::min(o.gmres_restart, 200));
    BatchWork &W = batch_work(P, m);
    const int Pn = W.P;
    const char *ge = std::getenv("TMI_BATCH_GRAPH");
    const bool use_graph = !(ge && std::string(ge) == "0");
    hipStream_t s = W.stream;
    TMI_HIP_CHECK(hipStreamSynchronize(caller));
    auto t0 = std::chrono::steady_clock::now();
    st = tmi_newton_stats{};
    // per-pore state
    {
        std::vector<int> ones(Pn, 1), zeros(Pn, 0);
        TMI_HIP_CHECK(hipMemcpyAsync(W.active, ones.data(), Pn * sizeof(int), hipMemcpyHostToDevice, s));
        TMI_HIP_CHECK(hipMemcpyAsync(W.conv, zeros.data(), Pn * sizeof(int), hipMemcpyHostToDevice, s));
        TMI_HIP_CHECK(hipMemcpyAsync(W.niters, zeros.data(), Pn * sizeof(int), hipMemcpyHostToDevice, s));
        TMI_HIP_CHECK(hipMemcpyAsync(W.gs.its, zeros.data(), Pn * sizeof(int), hipMemcpyHostToDevice, s));
        TMI_HIP_CHECK(hipStreamSynchronize(s));
    }
    const int maxcycles = std::max(1, (o.gmres_max_iter + m - 1) / m);
    const int nchunks = (m + kChunk - 1) / kChunk;
    // the Newton iteration number only enters k_newton_flags, launched outside the graphs
    auto enqueue_pre = [&] { stage_pre(P, W, x, s); };
    auto enqueue_start = [&] { stage_cycle_start(P, W, x, o.gmres_tol, o.precision, s); };
    auto enqueue_steps = [&](int j) {
        stage_cycle_steps(P, W, x, o.gmres_tol, o.precision, j * kChunk, std::min(m, (j + 1) * kChunk), s);
    };
    auto enqueue_finish = [&] { stage_cycle_finish(P, W, s); };
    auto enqueue_post = [&] { stage_post(P, W, x, o.update_scaling, s); };
    if (use_graph && (W.graph_x != x || W.graph_prec != o.precision || W.graph_tol[0] != o.gmres_tol ||
                      W.graph_tol[1] != o.update_scaling)) {
        for (hipGraphExec_t *g : {&W.g_pre, &W.g_start, &W.g_finish, &W.g_post})
            if (*g) {
                (void)hipGraphExecDestroy(*g);
                *g = nullptr;
            }
        for (hipGraphExec_t &g : W.g_steps)
            if (g) (void)hipGraphExecDestroy(g);
        W.g_steps.assign(nchunks, nullptr);
        W.g_pre = capture(s, enqueue_pre);
        W.g_start = capture(s, enqueue_start);
        for (int j = 0; j < nchunks; ++j) W.g_steps[j] = capture(s, [&] { enqueue_steps(j); });
        W.g_finish = capture(s, enqueue_finish);
        W.g_post = capture(s, enqueue_post);
        W.graph_x = x;
        W.graph_prec = o.precision;
        W.graph_tol[0] = o.gmres_tol;
        W.graph_tol[1] = o.update_scaling;
    }
    auto run = [&](hipGraphExec_t g, auto &&enqueue) {
        if (use_graph)
            TMI_HIP_CHECK(hipGraphLaunch(g, s));
        else
            enqueue();
    };
    auto running = [&] {   // pores whose GMRES has not converged (one host read)
        TMI_HIP_CHECK(hipMemcpyAsync(W.h_counts + 1, W.counts + 1, sizeof(int), hipMemcpyDeviceToHost, s));
        TMI_HIP_CHECK(hipStreamSynchronize(s));
        return W.h_counts[1];
    };
    int iter = 0;
    while (true) {
        run(W.g_pre, enqueue_pre);
        hipLaunchKernelGGL(k_newton_flags, dim3(nblk(Pn)), dim3(kBlock), 0, s, Pn, W.rn2, o.resid_tol, iter, W.active,
                           W.conv, W.niters, W.rnorm);
        hipLaunchKernelGGL(k_count, dim3(1), dim3(kBlock), 0, s, Pn, W.active, (const int *)nullptr, W.counts);
        TMI_HIP_CHECK(hipMemcpyAsync(W.h_counts, W.counts, sizeof(int), hipMemcpyDeviceToHost, s));
        TMI_HIP_CHECK(hipStreamSynchronize(s));
        const int nactive = W.h_counts[0];
        if (o.verbose) std::fprintf(stderr, "[tmi batch] iter %d: %d of %d pores active\n", iter, nactive, Pn);
        if (nactive == 0 || iter >= o.max_iter) break;
        for (int c = 0; c < maxcycles; ++c) {
            run(W.g_start, enqueue_start);
            int left = running();
            for (int j = 0; j < nchunks && left > 0; ++j) {
                if (use_graph)
                    TMI_HIP_CHECK(hipGraphLaunch(W.g_steps[j], s));
                else
                    enqueue_steps(j);
                left = running();
            }
            run(W.g_finish, enqueue_finish);
            if (left == 0) break;
        }
        run(W.g_post, enqueue_post);
        ++iter;
    }
    // pores still active at the iteration limit
    std::vector<int> act(Pn), conv(Pn), nit(Pn), gits(Pn);
    std::vector<double> rn(Pn);
    TMI_HIP_CHECK(hipMemcpyAsync(act.data(), W.active, Pn * sizeof(int), hipMemcpyDeviceToHost, s));
    TMI_HIP_CHECK(hipMemcpyAsync(conv.data(), W.conv, Pn * sizeof(int), hipMemcpyDeviceToHost, s));
    TMI_HIP_CHECK(hipMemcpyAsync(nit.data(), W.niters, Pn * sizeof(int), hipMemcpyDeviceToHost, s));
    TMI_HIP_CHECK(hipMemcpyAsync(gits.data(), W.gs.its, Pn * sizeof(int), hipMemcpyDeviceToHost, s));
    TMI_HIP_CHECK(hipMemcpyAsync(rn.data(), W.rnorm, Pn * sizeof(double), hipMemcpyDeviceToHost, s));
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    double rsum = 0.0;
    int nconv = 0, gtot = 0;
    for (int p = 0; p < Pn; ++p) {
        if (act[p]) nit[p] = iter;
        rsum += rn[p] * rn[p];
        nconv += conv[p];
        gtot += gits[p];
        if (pore_iters) pore_iters[p] = nit[p];
        if (pore_conv) pore_conv[p] = conv[p];
        if (pore_gmres) pore_gmres[p] = gits[p];
        if (pore_resid) pore_resid[p] = rn[p];
    }
    st.iters = iter;
    st.converged = (nconv == Pn) ? 1 : 0;
    st.gmres_iters_total = gtot;
    st.resid_norm = std::sqrt(rsum);
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // namespace tmi
