// Transpose-Jacobian product y = J(x)^T w (the adjoint of tmi_apply_jacobian).
//
// The Jacobian-vector product is  pad (halo by offset_to_node_box rules)
// -> stencil sum -> pointwise combination.  Its transpose runs the same
// pieces backwards:
//   1. the row weights w are written, zero-padded, into a grid with twice
//      the stencil halo;
//   2. a correlation with the same weights (offsets reversed) produces the
//      transpose sums at every position of the padded grid, halo included;
//   3. a fold gathers each node's halo images (periodic copies, mirror
//      copies with the vector components negated) back onto the node;
//      bulk and wall positions carry constants and drop out.
// The density columns collect the rho-bar and mean-field rows (forward
// stencils, src/dft_fill_FMT.c:195-347, src/dft_fill_CLSmf.c:38-171) and
// the A22 diagonal; the rho-bar columns collect the Euler-Lagrange rows
// (adjoint stencils, src/dft_fill_FMT.c:41-189) through the symmetric Hessian
// of Phi; the mean-field columns collect -I and the +1 of the EL rows.
// Zero-density columns and rows follow the reference's insertion rules.
#include <hip/hip_runtime.h>

#include "tmi_device.hpp"
#include "tmi_fmt.hpp"
#include "tmi_internal.hpp"

namespace tmi {
namespace {

using namespace dev;

// position in geometry `to` of a position of geometry `from` (same grid, larger halo)
__device__ inline long long regeom(const Geo &from, const Geo &to, long long pos)
{
    const int x = (int)(pos % from.P[0]);
    const long long r = pos / from.P[0];
    const int y = (int)(r % from.P[1]);
    const int z = (int)(r / from.P[1]);
    return (long long)(x + to.h[0] - from.h[0]) + (long long)(y + to.h[1] - from.h[1]) * to.P[0] +
           (long long)(z + to.h[2] - from.h[2]) * to.P[0] * to.P[1];
}

// combined rho-bar / mean-field row weights, zero-padded (interior only):
//   Theta = w_n3; Delta_c = w_n2 + w_n1/4piR_c + w_n0/4piR_c^2;
//   V_c,d = w_nv2d + w_nv1d/4piR_c; M_c' = w_mf_c'
template <typename T, int ND>
__global__ void k_pad_rows_rho(Geo g2, const double *__restrict__ w_rb, const double *__restrict__ w_mf, int nmf,
                               CompConsts cc, T *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g2.N) return;
    const long long p = pad_pos(g2, i);
    const size_t N = g2.N;
    const int C = g2.C;
    out[p] = (T)w_rb[i];
    for (int c = 0; c < C; ++c) {
        const size_t b = 1 + (size_t)c * (1 + ND);
        out[b * g2.np + p] = (T)(w_rb[N + i] + cc.i4pir[c] * w_rb[2 * N + i] + cc.i4pirsq[c] * w_rb[3 * N + i]);
#pragma unroll
        for (int d = 0; d < ND; ++d)
            out[(b + 1 + d) * g2.np + p] = (T)(w_rb[(4 + d) * N + i] + cc.i4pir[c] * w_rb[(4 + ND + d) * N + i]);
    }
    for (int c = 0; c < nmf; ++c) out[(1 + (size_t)C * (1 + ND) + c) * g2.np + p] = (T)w_mf[c * N + i];
}

struct CorrRhoArgs {
    StenPtrs th[kNC], de[kNC], u[kNC * kNC];
};

// rho_hat[c][p] for every position p of the density geometry
template <typename T, int ND>
__global__ void k_corr_rho(Geo gu, Geo g2, CorrRhoArgs a, int nmf, const T *__restrict__ in, T *__restrict__ out)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (p >= gu.np) return;
    const long long q = regeom(gu, g2, p);
    const int C = gu.C;
    T acc = 0;
    {
        const T *src = in + q;
        const StenPtrs s = a.th[c];
        const T *w = (const T *)s.w;
        for (int k = 0; k < s.L; ++k) acc += w[k] * src[-(long long)s.lo[k]];
    }
    {
        const size_t b = 1 + (size_t)c * (1 + ND);
        const T *src = in + b * g2.np + q;
        const StenPtrs s = a.de[c];
        const T *w = (const T *)s.w;
        const T *wv = (const T *)s.wv;
        for (int k = 0; k < s.L; ++k) {
            const long long o = -(long long)s.lo[k];
            T t = w[k] * src[o];
#pragma unroll
            for (int d = 0; d < ND; ++d) t += wv[k * ND + d] * src[(long long)(1 + d) * g2.np + o];
            acc += t;
        }
    }
    for (int cp = 0; cp < nmf; ++cp) {
        const T *src = in + (1 + (size_t)C * (1 + ND) + cp) * g2.np + q;
        const StenPtrs s = a.u[cp + C * c];
        const T *w = (const T *)s.w;
        for (int k = 0; k < s.L; ++k) acc += w[k] * src[-(long long)s.lo[k]];
    }
    out[(size_t)c * gu.np + p] = acc;
}

// enumerate the padded positions that map onto node j (itself, periodic and
// mirror images) and call f(position, reflect mask)
template <typename F>
__device__ inline void for_each_image(const Geo &g, int j, F &&f)
{
    int c[3];
    c[0] = j % g.nx[0];
    const int r = j / g.nx[0];
    c[1] = r % g.nx[1];
    c[2] = r / g.nx[1];
    int cand[3][3], nc[3];
    for (int d = 0; d < 3; ++d) {
        nc[d] = 1;
        cand[d][0] = c[d];
        if (d >= g.nd) continue;
        const int n = g.nx[d];
        if (g.bc[d][0] == TMI_PERIODIC) {
            cand[d][nc[d]++] = c[d] - n;
            cand[d][nc[d]++] = c[d] + n;
        } else {
            if (g.bc[d][0] == TMI_REFLECT && c[d] != 0) cand[d][nc[d]++] = -c[d];
            if (g.bc[d][1] == TMI_REFLECT && c[d] != n - 1) cand[d][nc[d]++] = 2 * (n - 1) - c[d];
        }
    }
    for (int a = 0; a < nc[0]; ++a)
        for (int b = 0; b < nc[1]; ++b)
            for (int e = 0; e < nc[2]; ++e) {
                const int q0 = cand[0][a] + g.h[0], q1 = cand[1][b] + g.h[1], q2 = cand[2][e] + g.h[2];
                if (q0 < 0 || q0 >= g.P[0] || q1 < 0 || q1 >= g.P[1] || q2 < 0 || q2 >= g.P[2]) continue;
                const long long pos = (long long)q0 + (long long)q1 * g.P[0] + (long long)q2 * g.P[0] * g.P[1];
                int src = -1, refl = 0;
                if (map_pad(g, pos, src, refl) != 0 || src != j) continue;
                f(pos, refl);
            }
}

// density columns: y_c(j) = diag * w_EL + [not zero] * sum over images of rho_hat
template <typename T>
__global__ void k_fold_rho(Geo gu, const T *__restrict__ rho_hat, const double *__restrict__ w_el,
                           const double *__restrict__ x_rho, const unsigned char *__restrict__ zero,
                           const unsigned char *__restrict__ elzero, double *__restrict__ y)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (j >= gu.N) return;
    const size_t row = (size_t)c * gu.N + j;
    double s = 0.0;
    if (!zero[row]) {
        const T *h = rho_hat + (size_t)c * gu.np;
        for_each_image(gu, j, [&](long long pos, int) { s += (double)h[pos]; });
    }
    const double diag = elzero[row] ? 1.0 : 1.0 / x_rho[row];
    y[row] = s + diag * w_el[row];
}

// Euler-Lagrange row weights (zero on replaced rows), zero-padded
template <typename T>
__global__ void k_pad_rows_el(Geo g2, const double *__restrict__ w_el, const unsigned char *__restrict__ elzero,
                              T *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g2.N) return;
    const size_t row = (size_t)c * g2.N + i;
    out[(size_t)c * g2.np + pad_pos(g2, i)] = elzero[row] ? (T)0 : (T)w_el[row];
}

struct CorrAdjArgs {
    StenPtrs th[kNC], de[kNC];
};

// adjoint fields' transpose: A_hat_c, B_hat_c,d (with the EL minus sign), S3_hat
template <typename T, int ND>
__global__ void k_corr_adj(Geo gf, Geo g2, CorrAdjArgs a, const T *__restrict__ in, T *__restrict__ out)
{
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= gf.np) return;
    const long long q = regeom(gf, g2, p);
    const int C = gf.C;
    T s3 = 0;
    for (int c = 0; c < C; ++c) {
        const T *src = in + (size_t)c * g2.np + q;
        T A = 0, B[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) B[d] = 0;
        {
            const StenPtrs s = a.de[c];
            const T *w = (const T *)s.w;
            const T *wv = (const T *)s.wv;
            for (int k = 0; k < s.L; ++k) {
                const T v = src[-(long long)s.lo[k]];
                A += w[k] * v;
#pragma unroll
                for (int d = 0; d < ND; ++d) B[d] -= wv[k * ND + d] * v;
            }
        }
        {
            const StenPtrs s = a.th[c];
            const T *w = (const T *)s.w;
            for (int k = 0; k < s.L; ++k) s3 += w[k] * src[-(long long)s.lo[k]];
        }
        const size_t b = (size_t)c * (1 + ND);
        out[b * gf.np + p] = A;
#pragma unroll
        for (int d = 0; d < ND; ++d) out[(b + 1 + d) * gf.np + p] = B[d];
    }
    out[(size_t)C * (1 + ND) * gf.np + p] = s3;
}

// rho-bar and mean-field columns
template <typename T, int ND>
__global__ void k_fold_adj(Geo gf, int func, const T *__restrict__ adj_hat, CompConsts cc, const double *__restrict__ x_rb,
                           const double *__restrict__ w_rb, const double *__restrict__ w_el,
                           const unsigned char *__restrict__ elzero, const double *__restrict__ w_mf, int nmf,
                           double *__restrict__ y_rb, double *__restrict__ y_mf)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= gf.N) return;
    constexpr int NR = 4 + 2 * ND;
    const int C = gf.C;
    const size_t N = gf.N;
    double g[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) g[q] = 0.0;
    for (int c = 0; c < C; ++c) {
        const size_t b = (size_t)c * (1 + ND);
        double A = 0.0, B[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) B[d] = 0.0;
        for_each_image(gf, j, [&](long long pos, int refl) {
            A += (double)adj_hat[b * gf.np + pos];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const double v = (double)adj_hat[(b + 1 + d) * gf.np + pos];
                B[d] += ((refl >> d) & 1) ? -v : v;
            }
        });
        g[0] += A * cc.i4pirsq[c];
        g[1] += A * cc.i4pir[c];
        g[2] += A;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            g[4 + d] += B[d] * cc.i4pir[c];
            g[4 + ND + d] += B[d];
        }
    }
    {
        double S3 = 0.0;
        const size_t b = (size_t)C * (1 + ND);
        for_each_image(gf, j, [&](long long pos, int) { S3 += (double)adj_hat[b * gf.np + pos]; });
        g[3] = S3;
    }
    // n at node j (solution order -> n order), then H(n) g
    double rb[NR], n[NR], out[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) rb[q] = x_rb[(size_t)q * N + j];
    n[0] = rb[3];
    n[1] = rb[2];
    n[2] = rb[1];
    n[3] = rb[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        n[4 + d] = rb[4 + ND + d];
        n[4 + ND + d] = rb[4 + d];
    }
    fmt_hvp<ND>(func, n, g, out);
    double sol[NR];
    sol[0] = out[3];
    sol[1] = out[2];
    sol[2] = out[1];
    sol[3] = out[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        sol[4 + d] = out[4 + ND + d];
        sol[4 + ND + d] = out[4 + d];
    }
#pragma unroll
    for (int q = 0; q < NR; ++q) y_rb[(size_t)q * N + j] = sol[q] - w_rb[(size_t)q * N + j];
    for (int c = 0; c < nmf; ++c) {
        const size_t row = (size_t)c * N + j;
        y_mf[row] = -w_mf[row] + (elzero[row] ? 0.0 : w_el[row]);
    }
}

template <typename T>
void ensure_corr(Problem &P)
{
    Workspace<T> &ws = workspace<T>(P);
    const int C = P.p.ncomp, nd = P.p.ndim;
    if (!ws.corr_u) {
        const size_t rows = 1 + (size_t)C * (1 + nd) + P.nmf;
        TMI_HIP_CHECK(hipMalloc((void **)&ws.corr_u, rows * P.pad_u2.n * sizeof(T)));
        TMI_HIP_CHECK(hipMemset(ws.corr_u, 0, rows * P.pad_u2.n * sizeof(T)));
    }
    if (!ws.corr_f) {
        TMI_HIP_CHECK(hipMalloc((void **)&ws.corr_f, (size_t)C * P.pad_f2.n * sizeof(T)));
        TMI_HIP_CHECK(hipMemset(ws.corr_f, 0, (size_t)C * P.pad_f2.n * sizeof(T)));
    }
}

template <typename T, int ND>
void vjp_nd(Problem &P, const double *x, const double *w, double *y, hipStream_t s)
{
    ensure_corr<T>(P);
    Workspace<T> &ws = workspace<T>(P);
    const size_t N = P.N;
    const int C = P.p.ncomp;
    const Geo gu = make_geo(P, P.pad_u), gu2 = make_geo(P, P.pad_u2);
    const Geo gf = make_geo(P, P.pad_f), gf2 = make_geo(P, P.pad_f2);
    const CompConsts cc = make_consts(P);
    const double *w_el = w + (size_t)P.iu_rho * N, *w_rb = w + (size_t)P.iu_rb * N;
    const double *w_mf = P.nmf ? w + (size_t)P.iu_mf * N : nullptr;
    // density columns
    hipLaunchKernelGGL((k_pad_rows_rho<T, ND>), dim3(grid_for(P.N)), dim3(kBlock), 0, s, gu2, w_rb, w_mf, P.nmf, cc,
                       ws.corr_u);
    CorrRhoArgs ca{};
    for (int c = 0; c < C; ++c) {
        ca.th[c] = sten_ptrs<T>(P, TMI_THETA, c, kGeomU2);
        ca.de[c] = sten_ptrs<T>(P, TMI_DELTA, c, kGeomU2);
    }
    for (int q = 0; q < C * C && P.nmf; ++q) ca.u[q] = sten_ptrs<T>(P, TMI_UATT, q, kGeomU2);
    hipLaunchKernelGGL((k_corr_rho<T, ND>), dim3(grid_for(gu.np), C), dim3(kBlock), 0, s, gu, gu2, ca, P.nmf,
                       ws.corr_u, ws.rho_pad);
    hipLaunchKernelGGL(k_fold_rho<T>, dim3(grid_for(P.N), C), dim3(kBlock), 0, s, gu, ws.rho_pad, w_el,
                       x + (size_t)P.iu_rho * N, P.d_zero, P.d_elzero, y + (size_t)P.iu_rho * N);
    // rho-bar and mean-field columns
    hipLaunchKernelGGL(k_pad_rows_el<T>, dim3(grid_for(P.N), C), dim3(kBlock), 0, s, gf2, w_el, P.d_elzero, ws.corr_f);
    CorrAdjArgs aa{};
    for (int c = 0; c < C; ++c) {
        aa.th[c] = sten_ptrs<T>(P, TMI_THETA, c, kGeomF2);
        aa.de[c] = sten_ptrs<T>(P, TMI_DELTA, c, kGeomF2);
    }
    hipLaunchKernelGGL((k_corr_adj<T, ND>), dim3(grid_for(gf.np)), dim3(kBlock), 0, s, gf, gf2, aa, ws.corr_f,
                       ws.adj_pad);
    hipLaunchKernelGGL((k_fold_adj<T, ND>), dim3(grid_for(P.N)), dim3(kBlock), 0, s, gf, P.p.type_func, ws.adj_pad, cc,
                       x + (size_t)P.iu_rb * N, w_rb, w_el, P.d_elzero, w_mf, P.nmf, y + (size_t)P.iu_rb * N,
                       P.nmf ? y + (size_t)P.iu_mf * N : nullptr);
    TMI_HIP_CHECK(hipGetLastError());
}

}  // namespace

template <typename T>
void launch_vjp(Problem &P, const double *x, const double *w_in, double *y, hipStream_t s)
{
    const double *w = w_in;
    if (P.one_size) {   // the stencil transpose over w without the n1, n0, nv1 rows, which are algebraic
        const size_t N = P.N, nd = P.p.ndim;
        TMI_HIP_CHECK(hipMemcpyAsync(P.d_w0, w_in, sizeof(double) * P.nunk * N, hipMemcpyDeviceToDevice, s));
        double *rb = P.d_w0 + (size_t)P.iu_rb * N;
        dev_fill(0.0, rb + 2 * N, (long long)(2 * N), s);
        dev_fill(0.0, rb + (4 + nd) * N, (long long)(nd * N), s);
        w = P.d_w0;
    }
    switch (P.p.ndim) {
    case 1: vjp_nd<T, 1>(P, x, w, y, s); break;
    case 2: vjp_nd<T, 2>(P, x, w, y, s); break;
    default: vjp_nd<T, 3>(P, x, w, y, s); break;
    }
    hw_apply(P.hw_tr, w + (size_t)P.iu_rb * P.N, y + (size_t)P.iu_rho * P.N, s);   // hard-surface weights
    one_size_transpose(P, w_in + (size_t)P.iu_rb * P.N, y + (size_t)P.iu_rb * P.N, s);
}
template void launch_vjp<double>(Problem &, const double *, const double *, double *, hipStream_t);
template void launch_vjp<float>(Problem &, const double *, const double *, double *, hipStream_t);

}  // namespace tmi
