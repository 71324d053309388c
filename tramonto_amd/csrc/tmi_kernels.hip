// Device kernels of libtramonto_mi355x (gfx950).
//
// Data layout in HBM: every field is structure-of-arrays, one contiguous row
// of N doubles per unknown (the reference's x[iunk][inode_box]).  Stencil
// sums read halo-padded copies of their inputs: the halo holds exactly what
// offset_to_node_box (src/dft_mesh_lib.c:203-297) returns for out-of-box
// points -- the periodic image, the mirror image (REFLECT, with vector
// components negated), the bulk constant (IN_BULK) or zero (IN_WALL) -- so the
// sum kernels are branch-free: out[i] = sum_k w_k in_pad[p(i) + lo_k].
//
//   pad_rho    : densities rho_c (zero where Zero_density_TF), halo of the
//                widest stencil; read by the forward (rho-bar) and mean-field
//                (u) sums.
//   fmt_point  : dPhi/dn (residual) or H(n) dn (Jacobian-vector product) at
//                every node, folded per component into the adjoint fields
//                A_c = S0/4piR^2 + S1/4piR + S2, B_c,d = V1_d/4piR + V2_d, S3.
//   pad_adj    : halo of the adjoint fields.
//   rb_rows    : rho-bar definition rows  (load_rho_bar_s/_v, src/dft_fill_FMT.c:195-347)
//   mf_rows    : mean-field rows          (load_mean_field, src/dft_fill_mf.c:38-63)
//   el_rows    : Euler-Lagrange rows      (load_euler_lagrange, src/dft_fill_EL.c:38-197,
//                adjoint stencil load_nonlocal_hs_rosen_rb, src/dft_fill_FMT.c:41-189)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "tmi_device.hpp"
#include "tmi_fmt.hpp"
#include "tmi_internal.hpp"

namespace tmi {

namespace {

using namespace dev;

// ---------------------------------------------------------------- padding
// two values stored as one vector (a complex element of a pair transform buffer)
template <typename T>
struct alignas(2 * sizeof(T)) TPair {
    T a, b;
};

// rho_pad[c][pos] from in[c*N + node]; bulk value rho_b (residual) or 0 (linear)
template <typename T>
__global__ void k_pad_rho(Geo g, const double *__restrict__ in, const unsigned char *__restrict__ zero, CompConsts cc,
                          int linear, T *__restrict__ out)
{
    long long pos = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= g.np) return;
    int src = 0, refl = 0;
    int code = map_pad(g, pos, src, refl);
    for (int c = 0; c < g.C; ++c) {
        double v;
        if (code == 0)
            v = zero[(size_t)c * g.N + src] ? 0.0 : in[(size_t)c * g.N + src];
        else if (code == -1 && !linear)
            v = cc.rho_b[c];
        else
            v = 0.0;
        out[(size_t)c * g.np + pos] = (T)v;
    }
}

// pointwise FMT: residual (dphi) or linear (H dn); writes the adjoint fields
// at the interior positions of the FMT-halo geometry, or (pairs.buf set) as
// pairs of the transform buffer of a wrapping FFT grid.  MODE (kFmtResidual,
// kFmtLinear, kFmtLinearRef) is a template parameter so that each mode is
// compiled with only its own registers.
constexpr int kFmtResidual = 0, kFmtLinear = 1, kFmtLinearRef = 2;
template <typename T, int ND, int MODE>
__global__ __launch_bounds__(kBlock) void k_fmt_point(Geo g, int func, const double *__restrict__ xrb,
                                                      const double *__restrict__ dnrb, CompConsts cc,
                                                      T *__restrict__ adj, FftPairs pairs)
{
    constexpr bool linear = MODE != kFmtResidual, jref = MODE == kFmtLinearRef;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    constexpr int NR = 4 + 2 * ND;
    double rb[NR], n[NR], out[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) rb[q] = xrb[(size_t)q * g.N + i];
    // solution order (n3 n2 n1 n0 nv2 nv1) -> n order, src/dft_utils.c:205
    n[0] = rb[3];
    n[1] = rb[2];
    n[2] = rb[1];
    n[3] = rb[0];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        n[4 + d] = rb[4 + ND + d];
        n[4 + ND + d] = rb[4 + d];
    }
    if constexpr (!linear) {
        fmt_dphi<ND>(func, n, out);
    } else {   // both linear modes: the exact product (the reference mode keeps its S3 = out[3])
        double dr[NR], dn[NR];
#pragma unroll
        for (int q = 0; q < NR; ++q) dr[q] = dnrb[(size_t)q * g.N + i];
        dn[0] = dr[3];
        dn[1] = dr[2];
        dn[2] = dr[1];
        dn[3] = dr[0];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            dn[4 + d] = dr[4 + ND + d];
            dn[4 + ND + d] = dr[4 + d];
        }
        fmt_hvp<ND>(func, n, dn, out);
    }
    // the adjoint fields in v[0 .. nf): A_c, B_c,d (component c at c (1 + ND)), S3 at C (1 + ND), or in
    // the derivative form D3 D2 D1 D0 Dv2_d Dv1_d; every index below is a compile-time one
    const int C = g.C;
    constexpr int NV = 2 * ((kNC * (1 + ND) + 2) / 2);
    T v[NV];
#pragma unroll
    for (int f = 0; f < NV; ++f) v[f] = (T)0;
    int nf;
    bool dform_out = false;
    if constexpr (jref) {   // the reference's delta rows for A_c and B_c,d; S3 from the theta row (exact)
        double sg[ND > 0 ? ND : 1];
        for (int d = 0; d < ND; ++d) sg[d] = 1.0;
        double dn[NR];
        {
            double dr[NR];
#pragma unroll
            for (int q = 0; q < NR; ++q) dr[q] = dnrb[(size_t)q * g.N + i];
            dn[0] = dr[3];
            dn[1] = dr[2];
            dn[2] = dr[1];
            dn[3] = dr[0];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                dn[4 + d] = dr[4 + ND + d];
                dn[4 + ND + d] = dr[4 + d];
            }
        }
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            if (c < C) {
                double A, B[ND > 0 ? ND : 1];
                fmt_ref_adjoint<ND>(func, n, dn, sg, cc.i4pir[c], cc.i4pirsq[c], A, B);
                v[c * (1 + ND)] = (T)A;
#pragma unroll
                for (int d = 0; d < ND; ++d) v[c * (1 + ND) + 1 + d] = (T)B[d];
            }
        }
        nf = C * (1 + ND) + 1;
    } else if (pairs.buf && pairs.dform) {   // the derivatives themselves, rho-bar order
        v[0] = (T)out[3];
        v[1] = (T)out[2];
        v[2] = (T)out[1];
        v[3] = (T)out[0];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            v[4 + d] = (T)out[4 + ND + d];
            v[4 + ND + d] = (T)out[4 + d];
        }
        nf = 4 + 2 * ND;
        dform_out = true;
    } else {
#pragma unroll
        for (int c = 0; c < kNC; ++c) {   // compile-time component index: the constants stay in registers
            if (c < C) {
                v[c * (1 + ND)] = (T)(out[0] * cc.i4pirsq[c] + out[1] * cc.i4pir[c] + out[2]);
#pragma unroll
                for (int d = 0; d < ND; ++d) v[c * (1 + ND) + 1 + d] = (T)(out[4 + d] * cc.i4pir[c] + out[4 + ND + d]);
            }
        }
        nf = C * (1 + ND) + 1;
    }
    if (!dform_out) {   // S3 after the components (v[nf] stays 0: an odd count's empty partner)
#pragma unroll
        for (int f = 0; f < NV; ++f)
            if (f == C * (1 + ND)) v[f] = (T)out[3];
    }
    if (pairs.buf) {   // pair j = fields (2 j, 2 j + 1) at element j ns + q of the transform buffer: one 16-byte store
        const int x = i % g.nx[0];
        const int r = i / g.nx[0];
        const long long q = (long long)x + (r % g.nx[1]) * pairs.S1 + (r / g.nx[1]) * pairs.S2;
        TPair<T> *dst = reinterpret_cast<TPair<T> *>(pairs.buf);
#pragma unroll
        for (int j = 0; j < NV / 2; ++j)
            if (2 * j < nf) dst[(size_t)j * pairs.ns + q] = TPair<T>{v[2 * j], v[2 * j + 1]};
    } else {
        const long long p = pad_pos(g, i);
#pragma unroll
        for (int f = 0; f < NV; ++f)
            if (f < nf) adj[(size_t)f * g.np + p] = v[f];
    }
}

template <typename T, int ND, int MODE>
__global__ __launch_bounds__(kBlock) void k_pad_adj(Geo g, CompConsts cc, T *__restrict__ adj, int func,
                                                    const double *__restrict__ xrb, const double *__restrict__ dnrb)
{
    constexpr bool linear = MODE != kFmtResidual;
    long long pos = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pos >= g.np) return;
    int src = 0, refl = 0;
    int code = map_pad(g, pos, src, refl);
    const int C = g.C;
    if (code == 0) {
        const long long ps = pad_pos(g, src);
        if (ps == pos) return;   // interior position
        if (MODE == kFmtLinearRef && refl) {
            // reference-Jacobian mode across a REFLECT boundary: the reference evaluates its delta row
            // at the node with the mirror signs (src/dft_fill_FMT.c:41-189 with sign[] of
            // offset_to_node_box), which is not the mirror image of the node's own value for every
            // functional; S3 (theta row, no signs) is copied
            constexpr int NR = 4 + 2 * ND;
            double n[NR], dn[NR], sg[ND > 0 ? ND : 1];
            {
                double r[NR], dr[NR];
#pragma unroll
                for (int q = 0; q < NR; ++q) {
                    r[q] = xrb[(size_t)q * g.N + src];
                    dr[q] = dnrb[(size_t)q * g.N + src];
                }
                n[0] = r[3], n[1] = r[2], n[2] = r[1], n[3] = r[0];
                dn[0] = dr[3], dn[1] = dr[2], dn[2] = dr[1], dn[3] = dr[0];
#pragma unroll
                for (int d = 0; d < ND; ++d) {
                    n[4 + d] = r[4 + ND + d];
                    n[4 + ND + d] = r[4 + d];
                    dn[4 + d] = dr[4 + ND + d];
                    dn[4 + ND + d] = dr[4 + d];
                }
            }
#pragma unroll
            for (int d = 0; d < ND; ++d) sg[d] = ((refl >> d) & 1) ? -1.0 : 1.0;
            for (int c = 0; c < C; ++c) {
                const size_t b = (size_t)(c * (1 + ND));
                double A, B[ND > 0 ? ND : 1];
                fmt_ref_adjoint<ND>(func, n, dn, sg, cc.i4pir[c], cc.i4pirsq[c], A, B);
                adj[b * g.np + pos] = (T)A;
#pragma unroll
                for (int d = 0; d < ND; ++d) adj[(b + 1 + d) * g.np + pos] = (T)B[d];
            }
            adj[(size_t)(C * (1 + ND)) * g.np + pos] = adj[(size_t)(C * (1 + ND)) * g.np + ps];
            return;
        }
        for (int c = 0; c < C; ++c) {
            const size_t b = (size_t)(c * (1 + ND));
            adj[b * g.np + pos] = adj[b * g.np + ps];
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                T v = adj[(b + 1 + d) * g.np + ps];
                adj[(b + 1 + d) * g.np + pos] = ((refl >> d) & 1) ? -v : v;
            }
        }
        adj[(size_t)(C * (1 + ND)) * g.np + pos] = adj[(size_t)(C * (1 + ND)) * g.np + ps];
        return;
    }
    const bool bulk = (code == -1) && !linear;
    for (int c = 0; c < C; ++c) {
        const size_t b = (size_t)(c * (1 + ND));
        adj[b * g.np + pos] = bulk ? (T)cc.Ab[c] : (T)0;
#pragma unroll
        for (int d = 0; d < ND; ++d) adj[(b + 1 + d) * g.np + pos] = (T)0;
    }
    adj[(size_t)(C * (1 + ND)) * g.np + pos] = bulk ? (T)cc.S3b : (T)0;
}

// ---------------------------------------------------------------- stencil rows
struct RbArgs {
    StenPtrs th[kNC], de[kNC];
};

// rho-bar rows: out_rb[q][i] = sum - in_rb[q][i] - sub[q][i]
template <typename T, int ND>
__global__ void k_rb_rows(Geo g, RbArgs sa, const T *__restrict__ rho_pad, CompConsts cc, const double *__restrict__ in_rb,
                          const double *__restrict__ sub, double *__restrict__ out_rb)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const long long p = pad_pos(g, i);
    T n3 = 0, n2 = 0, n1 = 0, n0 = 0, v2[ND], v1[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) v2[d] = v1[d] = 0;
    for (int c = 0; c < g.C; ++c) {
        const T *src = rho_pad + (size_t)c * g.np + p;
        T st = 0;
        {
            const int *lo = sa.th[c].lo;
            const T *w = (const T *)sa.th[c].w;
            const int L = sa.th[c].L;
            for (int k = 0; k < L; ++k) st += w[k] * src[lo[k]];
        }
        T sd = 0, vd[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) vd[d] = 0;
        {
            const int *lo = sa.de[c].lo;
            const T *w = (const T *)sa.de[c].w;
            const T *wv = (const T *)sa.de[c].wv;
            const int L = sa.de[c].L;
            for (int k = 0; k < L; ++k) {
                const T r = src[lo[k]];
                sd += w[k] * r;
#pragma unroll
                for (int d = 0; d < ND; ++d) vd[d] += wv[k * ND + d] * r;
            }
        }
        n3 += st;
        n2 += sd;
        n1 += (T)cc.i4pir[c] * sd;
        n0 += (T)cc.i4pirsq[c] * sd;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            v2[d] += vd[d];
            v1[d] += (T)cc.i4pir[c] * vd[d];
        }
    }
    double vals[4 + 2 * ND];
    vals[0] = n3;
    vals[1] = n2;
    vals[2] = n1;
    vals[3] = n0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        vals[4 + d] = v2[d];
        vals[4 + ND + d] = v1[d];
    }
#pragma unroll
    for (int q = 0; q < 4 + 2 * ND; ++q) {
        double r = vals[q];
        if (in_rb) r -= in_rb[(size_t)q * g.N + i];
        if (sub) r -= sub[(size_t)q * g.N + i];
        out_rb[(size_t)q * g.N + i] = r;
    }
}

struct UArgs {
    StenPtrs u[kNC * kNC];
};

// mean-field rows: out[c][i] = sum_c' sum_k u_cc'(k) rho_pad[c'] - in_mf - sub
template <typename T>
__global__ void k_mf_rows(Geo g, UArgs ua, const T *__restrict__ rho_pad, const double *__restrict__ in_mf,
                          const double *__restrict__ sub, double *__restrict__ out_mf)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g.N) return;
    const long long p = pad_pos(g, i);
    T acc = 0;
    for (int cj = 0; cj < g.C; ++cj) {
        const StenPtrs s = ua.u[c + g.C * cj];
        const T *src = rho_pad + (size_t)cj * g.np + p;
        const T *w = (const T *)s.w;
        for (int k = 0; k < s.L; ++k) acc += w[k] * src[s.lo[k]];
    }
    double r = acc;
    if (in_mf) r -= in_mf[(size_t)c * g.N + i];
    if (sub) r -= sub[(size_t)c * g.N + i];
    out_mf[(size_t)c * g.N + i] = r;
}

// 2D mean-field rows as a dense convolution of the padded densities.  A workgroup of kMf2dWaves waves
// takes 64 consecutive interior x of TY consecutive padded rows; wave w sums the stencil columns
// dx = -hx + w, -hx + w + kMf2dWaves, ... and the partial sums meet in LDS.  Each density a wave loads feeds up to TY
// outputs (register blocking along y), and the dense (2 hy + 1) x (2 hx + 1) weights of a component
// pair (zeros off the stencil) are wave-uniform scalars.  Padded rows holding no node (the gap rows of
// a batch) produce no output.  Sums run in another order than k_mf_rows's.
constexpr int kMf2dWaves = 8;
template <typename T, int TY>
__global__ __launch_bounds__(64 * kMf2dWaves) void k_mf_rows_2d(Geo g, const T *__restrict__ W,
                                                                const T *__restrict__ rho_pad,
                                                                const double *__restrict__ in_mf,
                                                                const double *__restrict__ sub,
                                                                double *__restrict__ out_mf)
{
    __shared__ T part[kMf2dWaves - 1][TY][64];
    const int c = blockIdx.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int hx = g.h[0], hy = g.h[1], P0 = g.P[0], P1 = g.P[1];
    const int wx = 2 * hx + 1, wy = 2 * hy + 1;
    const int xi = blockIdx.x * 64 + lane;   // interior column
    const bool col_ok = xi < g.nx[0];
    const int px = hx + (col_ok ? xi : 0);
    const int py0 = hy + blockIdx.y * TY;    // first padded row of the tile
    // input row py0 + yy feeds output j through weight row dy = yy - j, |dy| <= hy; rows at or past P1
    // feed only outputs past the last padded node row
    const int yend = min(hy + TY - 1, P1 - 1 - py0);
    T acc[TY];
#pragma unroll
    for (int j = 0; j < TY; ++j) acc[j] = (T)0;
    for (int cj = 0; cj < g.C; ++cj) {
        const T *Wq = W + (size_t)(c + g.C * cj) * wx * wy;
        const T *src = rho_pad + (size_t)cj * g.np;
        for (int dx = -hx + wave; dx <= hx; dx += kMf2dWaves) {
            const T *col = src + px + dx + (long long)py0 * P0;
            const T *wcol = Wq + (dx + hx) + (size_t)hy * wx;   // weight row dy at wcol[dy * wx]
            int yy = -hy;
            for (; yy < -hy + TY - 1 && yy <= yend; ++yy) {   // ramp: only outputs j <= yy + hy
                const T v = col[(long long)yy * P0];
#pragma unroll
                for (int j = 0; j < TY; ++j)
                    if (yy - j >= -hy && yy - j <= hy) acc[j] += wcol[(yy - j) * wx] * v;
            }
            const int ysteady = min(hy, yend);   // from yy >= -hy + TY - 1 on, every output's weight row exists
            if (yy <= ysteady) {   // every output; wr[j] = weight row yy - j, rotated: one new weight per row
                T wr[TY];
#pragma unroll
                for (int j = 1; j < TY; ++j) wr[j] = wcol[(yy - j) * wx];
#pragma unroll TY
                for (; yy <= ysteady; ++yy) {
                    wr[0] = wcol[yy * wx];
                    const T v = col[(long long)yy * P0];
#pragma unroll
                    for (int j = 0; j < TY; ++j) acc[j] += wr[j] * v;
#pragma unroll
                    for (int j = TY - 1; j > 0; --j) wr[j] = wr[j - 1];
                }
            }
            for (; yy <= yend; ++yy) {   // tail: only outputs j >= yy - hy
                const T v = col[(long long)yy * P0];
#pragma unroll
                for (int j = 0; j < TY; ++j)
                    if (yy - j >= -hy && yy - j <= hy) acc[j] += wcol[(yy - j) * wx] * v;
            }
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int j = 0; j < TY; ++j) part[wave - 1][j][lane] = acc[j];
    }
    __syncthreads();
    if (wave > 0 || !col_ok) return;
#pragma unroll
    for (int j = 0; j < TY; ++j) {
        const int py = py0 + j;
        if (py >= P1 - hy) break;
        T a = acc[j];
#pragma unroll
        for (int w = 0; w < kMf2dWaves - 1; ++w) a += part[w][j][lane];
        const long long pos = (long long)py * P0 + px;
        const int node = g.pad_node ? g.pad_node[pos] : xi + (py - hy) * g.nx[0];
        if (node < 0) continue;
        double r = (double)a;
        if (in_mf) r -= in_mf[(size_t)c * g.N + node];
        if (sub) r -= sub[(size_t)c * g.N + node];
        out_mf[(size_t)c * g.N + node] = r;
    }
}
constexpr int kMf2dRows = 4;

// The same sums on the matrix cores (FP64 v_mfma_f64_16x16x4f64; FP32 v_mfma_f32_16x16x4f32 for the
// mixed path's sums).  A wave owns a tile of 16 padded rows x 16 columns; for each dx the tile gains
// the Toeplitz product A_dx B_dx over the 16 + 2 hy input rows k of its window (padded row
// py0 - hy + k): A_dx[i][k] = W[k - i][dx] (weight row dy + hy = k - i, 0 outside [0, 2 hy]),
// B_dx[k][j] = rho[py0 - hy + k][column j + dx], 4 rows k per instruction.  Operand lanes (both
// types): A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k; D[i][j] at lane j + 16 (i % 4), register
// i / 4 in FP64 and at lane j + 16 (i / 4), register i % 4 in FP32.  The kMfmaWaves waves of a
// workgroup split the dx range and add their tiles in LDS.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 step(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int row(int kq, int r) { return kq + 4 * r; }   // tile row of register r
};
template <>
struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static v4 step(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int row(int kq, int r) { return 4 * kq + r; }
};
constexpr int kMfmaWaves = 16;
template <typename T>
__global__ __launch_bounds__(64 * kMfmaWaves) void k_mf_rows_2d_mfma(Geo g, const T *__restrict__ W,
                                                                    const T *__restrict__ rho_pad,
                                                                    const double *__restrict__ in_mf,
                                                                    const double *__restrict__ sub,
                                                                    double *__restrict__ out_mf)
{
    typedef typename Mfma16<T>::v4 d4;
    __shared__ T part[kMfmaWaves - 1][4][64];
    const int c = blockIdx.z;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int hx = g.h[0], hy = g.h[1], P0 = g.P[0], P1 = g.P[1];
    const int wx = 2 * hx + 1, wy = 2 * hy + 1;
    const int lo = lane & 15, kq = lane >> 4;   // A: row i = lo; B and D: column j = lo; k offset kq
    const int x0 = blockIdx.x * 16;             // first interior column of the tile
    const int py0 = hy + blockIdx.y * 16;       // first padded row of the tile's outputs
    const bool col_ok = x0 + lo < g.nx[0];
    const int nk = 16 + 2 * hy;
    d4 acc = {(T)0, (T)0, (T)0, (T)0};
    for (int cj = 0; cj < g.C; ++cj) {
        const T *Wq = W + (size_t)(c + g.C * cj) * wx * wy;
        const T *src = rho_pad + (size_t)cj * g.np;
        for (int dx = -hx + wave; dx <= hx; dx += kMfmaWaves) {
            const long long colx = hx + x0 + lo + dx;   // padded column read by this lane's B entries
            const T *wcol = Wq + (dx + hx);             // weight row r at wcol[r * wx]
#pragma unroll 4
            for (int k0 = 0; k0 < nk; k0 += 4) {
                const int k = k0 + kq;
                const int prow = py0 - hy + k;   // >= 0; rows at or past P1 feed no written output
                const T b = (col_ok && k < nk && prow < P1) ? src[(long long)prow * P0 + colx] : (T)0;
                const int wr = k - lo;
                const T a = (wr >= 0 && wr < wy) ? wcol[(size_t)wr * wx] : (T)0;
                acc = Mfma16<T>::step(a, b, acc);
            }
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) part[wave - 1][r][lane] = acc[r];
    }
    __syncthreads();
    if (wave > 0 || !col_ok) return;
    const int xi = x0 + lo;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int py = py0 + Mfma16<T>::row(kq, r);
        if (py >= P1 - hy) continue;
        T a = acc[r];
#pragma unroll
        for (int w = 0; w < kMfmaWaves - 1; ++w) a += part[w][r][lane];
        const long long pos = (long long)py * P0 + hx + xi;
        const int node = g.pad_node ? g.pad_node[pos] : xi + (py - hy) * g.nx[0];
        if (node < 0) continue;
        double v = (double)a;
        if (in_mf) v -= in_mf[(size_t)c * g.N + node];
        if (sub) v -= sub[(size_t)c * g.N + node];
        out_mf[(size_t)c * g.N + node] = v;
    }
}

// The rho-bar rows of one component on a 2D grid in FP64 on the matrix cores, as k_mf_rows_2d_mfma
// with the FMT stencils' own reach (fx, fy) inside the density halo: four Toeplitz products per dx
// and 4-row chunk share each density load (theta, delta, delta x, delta y weights), and the rows are
// those of k_rb_rows for C = 1: n3, n2, n1 = n2 / (4 pi R), n0 = n2 / (4 pi R^2), nv2, nv1 = nv2 / (4 pi R).
constexpr int kRbMfmaWaves = 8;
template <typename T>
__global__ __launch_bounds__(64 * kRbMfmaWaves) void k_rb_rows_2d_mfma(Geo g, int fx, int fy, const T *__restrict__ W,
                                                                       const T *__restrict__ rho_pad, double i4pir,
                                                                       double i4pirsq, const double *__restrict__ in_rb,
                                                                       const double *__restrict__ sub,
                                                                       double *__restrict__ out_rb)
{
    typedef typename Mfma16<T>::v4 d4;
    __shared__ T part[kRbMfmaWaves - 1][4][4][64];   // [wave - 1][sum][register][lane]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int hx = g.h[0], hy = g.h[1], P0 = g.P[0], P1 = g.P[1];
    const int wx = 2 * fx + 1, wy = 2 * fy + 1;
    const size_t wsz = (size_t)wx * wy;
    const int lo = lane & 15, kq = lane >> 4;
    const int x0 = blockIdx.x * 16;          // first interior column of the tile
    const int py0 = hy + blockIdx.y * 16;    // first padded row of the tile's outputs
    const bool col_ok = x0 + lo < g.nx[0];
    const int nk = 16 + 2 * fy;              // window rows: padded rows py0 - fy + k (>= 0 as fy <= hy)
    d4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = (d4){(T)0, (T)0, (T)0, (T)0};
    for (int dx = -fx + wave; dx <= fx; dx += kRbMfmaWaves) {
        const long long colx = hx + x0 + lo + dx;
        const T *wcol = W + (dx + fx);
#pragma unroll 2
        for (int k0 = 0; k0 < nk; k0 += 4) {
            const int k = k0 + kq;
            const int prow = py0 - fy + k;
            const T b = (col_ok && k < nk && prow < P1) ? rho_pad[(long long)prow * P0 + colx] : (T)0;
            const int wr = k - lo;
            const bool win = wr >= 0 && wr < wy;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const T a = win ? wcol[t * wsz + (size_t)wr * wx] : (T)0;
                acc[t] = Mfma16<T>::step(a, b, acc[t]);
            }
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) part[wave - 1][t][r][lane] = acc[t][r];
    }
    __syncthreads();
    if (wave > 0 || !col_ok) return;
    const int xi = x0 + lo;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int py = py0 + Mfma16<T>::row(kq, r);
        if (py >= P1 - hy) continue;
        T sum[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            T a = acc[t][r];
#pragma unroll
            for (int w = 0; w < kRbMfmaWaves - 1; ++w) a += part[w][t][r][lane];
            sum[t] = a;
        }
        const long long pos = (long long)py * P0 + hx + xi;
        const int node = g.pad_node ? g.pad_node[pos] : xi + (py - hy) * g.nx[0];
        if (node < 0) continue;
        const T f1 = (T)i4pir, f0 = (T)i4pirsq;   // the products in T, as k_rb_rows forms them
        const double vals[8] = {(double)sum[0], (double)sum[1], (double)(f1 * sum[1]), (double)(f0 * sum[1]),
                                (double)sum[2], (double)sum[3], (double)(f1 * sum[2]), (double)(f1 * sum[3])};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            double v = vals[q];
            if (in_rb) v -= in_rb[(size_t)q * g.N + node];
            if (sub) v -= sub[(size_t)q * g.N + node];
            out_rb[(size_t)q * g.N + node] = v;
        }
    }
}

// The Euler-Lagrange rows of one component on a 2D grid in FP64 on the matrix cores, as k_el_rows for
// C = 1: the adjoint fields (A, B_x, B_y, S3, halo-padded in the adjoint geometry g) correlated with the
// delta, -delta x, -delta y and theta weights of k_rb_rows_2d_mfma's table (reach fx, fy inside the
// adjoint halo), four products per dx and 4-row chunk into one sum, then the local terms of k_el_rows.
template <typename T>
__global__ __launch_bounds__(64 * kRbMfmaWaves) void k_el_rows_2d_mfma(Geo g, int fx, int fy, const T *__restrict__ W,
                                                                       const T *__restrict__ adj, CompConsts cc,
                                                                       int linear, int nwall,
                                                                       const double *__restrict__ x_rho,
                                                                       const double *__restrict__ in_rho,
                                                                       const double *__restrict__ in_mf,
                                                                       const double *__restrict__ vext,
                                                                       const unsigned char *__restrict__ elzero,
                                                                       const double *__restrict__ add,
                                                                       double *__restrict__ out_el)
{
    typedef typename Mfma16<T>::v4 d4;
    __shared__ T part[kRbMfmaWaves - 1][4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int hx = g.h[0], hy = g.h[1], P0 = g.P[0], P1 = g.P[1];
    const int wx = 2 * fx + 1, wy = 2 * fy + 1;
    const size_t wsz = (size_t)wx * wy;
    const int lo = lane & 15, kq = lane >> 4;
    const int x0 = blockIdx.x * 16, py0 = hy + blockIdx.y * 16;
    const bool col_ok = x0 + lo < g.nx[0];
    const int nk = 16 + 2 * fy;
    d4 acc = {(T)0, (T)0, (T)0, (T)0};
    for (int dx = -fx + wave; dx <= fx; dx += kRbMfmaWaves) {
        const long long colx = hx + x0 + lo + dx;
        const T *wcol = W + (dx + fx);
#pragma unroll 2
        for (int k0 = 0; k0 < nk; k0 += 4) {
            const int k = k0 + kq;
            const int prow = py0 - fy + k;
            const bool inb = col_ok && k < nk && prow < P1;
            const long long at = (long long)prow * P0 + colx;
            const int wr = k - lo;
            const bool win = wr >= 0 && wr < wy;
            const size_t wo = (size_t)wr * wx;
            const T bA = inb ? adj[at] : (T)0, bX = inb ? adj[(size_t)g.np + at] : (T)0;
            const T bY = inb ? adj[2 * (size_t)g.np + at] : (T)0, bS = inb ? adj[3 * (size_t)g.np + at] : (T)0;
            acc = Mfma16<T>::step(win ? wcol[wsz + wo] : (T)0, bA, acc);
            acc = Mfma16<T>::step(win ? -wcol[2 * wsz + wo] : (T)0, bX, acc);
            acc = Mfma16<T>::step(win ? -wcol[3 * wsz + wo] : (T)0, bY, acc);
            acc = Mfma16<T>::step(win ? wcol[wo] : (T)0, bS, acc);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int r = 0; r < 4; ++r) part[wave - 1][r][lane] = acc[r];
    }
    __syncthreads();
    if (wave > 0 || !col_ok) return;
    const int xi = x0 + lo;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int py = py0 + Mfma16<T>::row(kq, r);
        if (py >= P1 - hy) continue;
        T sum = acc[r];
#pragma unroll
        for (int w = 0; w < kRbMfmaWaves - 1; ++w) sum += part[w][r][lane];
        const long long pos = (long long)py * P0 + hx + xi;
        const int node = g.pad_node ? g.pad_node[pos] : xi + (py - hy) * g.nx[0];
        if (node < 0) continue;
        const size_t row = (size_t)node;   // component 0
        double v;
        if (elzero[row]) {
            v = in_rho ? in_rho[row] : 0.0;
        } else {
            v = (double)sum;
            const double xv = x_rho[row];
            if (!linear) {
                v += log(xv) + (g.mu ? g.mu[row] : cc.mu[0]);
                if (nwall > 0) v += vext[row];
                if (in_mf) v += in_mf[row];
            } else {
                if (in_rho) v += in_rho[row] / xv;
                if (in_mf) v += in_mf[row];
            }
        }
        if (add) v += add[row];
        out_el[row] = v;
    }
}

// ---------------------------------------------------------------- 3D column kernels
// A column pass streams one input field along z for one stencil column and
// adds, for each of the lane's B consecutive z outputs, the column's weights
// times the loaded values (see DevColumnsT for the padded weight layout).
template <typename T>
struct ColPtrs {
    const int *lo, *woff, *nchunk;
    const T *w, *wz, *fxy;
    int ncol;
};

template <typename T>
inline ColPtrs<T> col_ptrs(const DevColumnsT<T> &dc, int geom)
{
    ColPtrs<T> cp;
    cp.lo = dc.lo + (size_t)geom * dc.ncol;
    cp.woff = dc.woff;
    cp.nchunk = dc.nchunk;
    cp.w = dc.wpad;
    cp.wz = dc.fxy ? dc.wpad + dc.nw : nullptr;
    cp.fxy = dc.fxy;
    cp.ncol = dc.ncol;
    return cp;
}

template <typename T>
struct ColArgs {
    ColPtrs<T> u[kNC * kNC];
};

template <typename T>
struct FmtColArgs {
    ColPtrs<T> th[kNC], de[kNC];
};

// Raw buffer loads for the column kernels: the descriptor holds the
// wave-uniform column start, the per-lane byte offset stays fixed in a vector
// register and the z step goes into the scalar offset, so a load costs no
// vector address arithmetic.  Descriptor word 3 is the gfx9-family raw
// 32-bit format (0x00020000); the range is the full 32-bit offset space.
typedef int tmi_v2i __attribute__((ext_vector_type(2)));

__device__ inline __amdgpu_buffer_rsrc_t col_rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)0xffffffff, 0x00020000);
}

template <typename T>
__device__ inline T rsrc_load(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff);
template <>
__device__ inline double rsrc_load<double>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    const tmi_v2i v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    return __builtin_bit_cast(double, v);
}
template <>
__device__ inline float rsrc_load<float>(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    const int v = __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
    return __builtin_bit_cast(float, v);
}

// acc[b] += sum_z' w(z' - z_b) s(z').  s is wave-uniform and lane the lane's
// offset from it (see col_rsrc).
template <typename T, int B>
__device__ __forceinline__ void col_sum(const T *__restrict__ s, unsigned lane, const T *__restrict__ w, int nk,
                                        long long PXY, T (&acc)[B])
{
    const __amdgpu_buffer_rsrc_t r = col_rsrc(s);
    const unsigned voff = lane * (unsigned)sizeof(T), pstride = (unsigned)(PXY * (long long)sizeof(T));
    for (int k = 0; k < nk; ++k) {
        T v[B], wk[2 * B - 1];
#pragma unroll
        for (int i = 0; i < B; ++i) v[i] = rsrc_load<T>(r, voff, (unsigned)(k * B + i) * pstride);
#pragma unroll
        for (int j = 0; j < 2 * B - 1; ++j) wk[j] = w[k * B + j];
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] += wk[i - b + B - 1] * v[i];
    }
}

// the same loads against two weight planes: a += w s, az += wz s
template <typename T, int B>
__device__ __forceinline__ void col_sum2(const T *__restrict__ s, unsigned lane, const T *__restrict__ w,
                                         const T *__restrict__ wz, int nk, long long PXY, T (&a)[B], T (&az)[B])
{
    const __amdgpu_buffer_rsrc_t r = col_rsrc(s);
    const unsigned voff = lane * (unsigned)sizeof(T), pstride = (unsigned)(PXY * (long long)sizeof(T));
    for (int k = 0; k < nk; ++k) {
        T v[B], wk[2 * B - 1];
#pragma unroll
        for (int i = 0; i < B; ++i) v[i] = rsrc_load<T>(r, voff, (unsigned)(k * B + i) * pstride);
#pragma unroll
        for (int j = 0; j < 2 * B - 1; ++j) wk[j] = w[k * B + j];
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int b = 0; b < B; ++b) a[b] += wk[i - b + B - 1] * v[i];
#pragma unroll
        for (int j = 0; j < 2 * B - 1; ++j) wk[j] = wz[k * B + j];
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int b = 0; b < B; ++b) az[b] += wk[i - b + B - 1] * v[i];
    }
}

// Sum the per-wave partial results of a workgroup into wave 0 through LDS,
// one wave at a time (buf holds B*64 values).
template <typename T, int B, int S>
__device__ __forceinline__ void wave_reduce(T (&acc)[B], T *buf, int grp, int lane)
{
    for (int s = 1; s < S; ++s) {
        if (grp == s) {
#pragma unroll
            for (int b = 0; b < B; ++b) buf[b * 64 + lane] = acc[b];
        }
        __syncthreads();
        if (grp == 0) {
#pragma unroll
            for (int b = 0; b < B; ++b) acc[b] += buf[b * 64 + lane];
        }
        __syncthreads();
    }
}

// Column-kernel launch shape: workgroup = 64 x-lanes x S waves for one
// (x-block, y, z-block, component) tile; each lane owns kColB consecutive z
// outputs and wave s takes the stencil columns s, s+S, ...; the S partial
// sums meet in LDS.  The grid is 1D.  Workgroups are dealt round-robin to the
// 8 XCDs, each with its own L2, so the block id is remapped to give every XCD
// a contiguous range of tiles, and tiles are ordered x-block fastest, then
// groups of kColGZ z-blocks by all y, so that a range is a compact (y, z)
// region of the halo-padded input.
constexpr int kColS = 8;
constexpr int kXcd = 8;
constexpr int kColGZ = 4;

struct ColTile {
    int lane, grp, x, y, c, zb;
    bool active;
    long long PXY, ubase;   // ubase: padded position of the tile's lane 0 (wave-uniform)
};

inline int col_tiles(const Problem &P)
{
    return ((P.p.nx[0] + 63) / 64) * P.p.nx[1] * ((P.p.nx[2] + kColB - 1) / kColB) * P.p.ncomp;
}

template <int B>
__device__ __forceinline__ ColTile col_tile(const Geo &g)
{
    ColTile t;
    t.lane = threadIdx.x;
    // the wave index is uniform across the wave: readfirstlane lets the
    // compiler keep the column loops and their weights in scalar registers
    t.grp = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int nxb = (g.nx[0] + 63) / 64, ny = g.nx[1];
    const int zblocks = (g.nx[2] + B - 1) / B;
    const int nb = nxb * ny * zblocks * g.C;
    int L = blockIdx.x;
    if (nb % kXcd == 0) L = (L % kXcd) * (nb / kXcd) + L / kXcd;
    const int xb = L % nxb;
    const int r = L / nxb;
    t.c = r / (ny * zblocks);
    const int q = r % (ny * zblocks);
    const int zg = q / (ny * kColGZ);
    const int gz = min(kColGZ, zblocks - zg * kColGZ);
    const int q2 = q - zg * ny * kColGZ;
    t.y = q2 / gz;
    t.zb = (zg * kColGZ + q2 % gz) * B;
    t.x = xb * 64 + t.lane;
    // lanes past the last x read inside the padded array (a row is followed by
    // others or by the slack) and write nothing
    t.active = (t.x < g.nx[0]);
    t.PXY = (long long)g.P[0] * g.P[1];
    t.ubase = (long long)(xb * 64 + g.h[0]) + (long long)(t.y + g.h[1]) * g.P[0] + (long long)(t.zb + g.h[2]) * t.PXY;
    return t;
}

// Mean-field rows on a 3D grid with the attraction stencil in column form.
template <typename T, int B, int S>
__global__ __launch_bounds__(64 * S) void k_mf_rows_cols(Geo g, ColArgs<T> ca, const T *__restrict__ rho_pad,
                                                         const double *__restrict__ in_mf, const double *__restrict__ sub,
                                                         double *__restrict__ out_mf)
{
    __shared__ T red[B * 64];
    const ColTile t = col_tile<B>(g);
    T acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;
    for (int cj = 0; cj < g.C; ++cj) {
        const ColPtrs<T> cp = ca.u[t.c + g.C * cj];
        const T *src0 = rho_pad + (size_t)cj * g.np + t.ubase;
        for (int col = t.grp; col < cp.ncol; col += S)
            col_sum<T, B>(src0 + cp.lo[col], t.lane, cp.w + cp.woff[col], cp.nchunk[col], t.PXY, acc);
    }
    wave_reduce<T, B, S>(acc, red, t.grp, t.lane);
    if (t.grp > 0 || !t.active) return;
    const int nx = g.nx[0], ny = g.nx[1], nz = g.nx[2];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int z = t.zb + b;
        if (z >= nz) break;
        const size_t row = (size_t)t.c * g.N + t.x + (size_t)nx * (t.y + (size_t)ny * z);
        double r = acc[b];
        if (in_mf) r -= in_mf[row];
        if (sub) r -= sub[row];
        out_mf[row] = r;
    }
}

// rho-bar rows on a 3D grid, theta and delta stencils in column form.  One
// workgroup handles one component; with one component the rows are written
// directly (out_rb = sums - in_rb - sub), otherwise the five raw sums
// (theta, delta, delta*k_x, delta*k_y, delta*k_z) go to part[c][5][N] and
// k_rb_combine forms the rows.
template <typename T, int B, int S>
__global__ __launch_bounds__(64 * S) void k_rb_rows_cols(Geo g, FmtColArgs<T> fa, const T *__restrict__ rho_pad,
                                                         CompConsts cc, const double *__restrict__ in_rb,
                                                         const double *__restrict__ sub, double *__restrict__ out_rb,
                                                         T *__restrict__ part)
{
    __shared__ T red[B * 64];
    const ColTile t = col_tile<B>(g);
    T n3[B], s2[B], vx[B], vy[B], vz[B];
#pragma unroll
    for (int b = 0; b < B; ++b) n3[b] = s2[b] = vx[b] = vy[b] = vz[b] = 0;
    const T *src0 = rho_pad + (size_t)t.c * g.np + t.ubase;
    {
        const ColPtrs<T> cp = fa.th[t.c];
        for (int col = t.grp; col < cp.ncol; col += S)
            col_sum<T, B>(src0 + cp.lo[col], t.lane, cp.w + cp.woff[col], cp.nchunk[col], t.PXY, n3);
    }
    {
        const ColPtrs<T> cp = fa.de[t.c];
        for (int col = t.grp; col < cp.ncol; col += S) {
            T cs[B];
#pragma unroll
            for (int b = 0; b < B; ++b) cs[b] = 0;
            const int wo = cp.woff[col];
            col_sum2<T, B>(src0 + cp.lo[col], t.lane, cp.w + wo, cp.wz + wo, cp.nchunk[col], t.PXY, cs, vz);
            const T fx = cp.fxy[2 * col], fy = cp.fxy[2 * col + 1];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                s2[b] += cs[b];
                vx[b] += fx * cs[b];
                vy[b] += fy * cs[b];
            }
        }
    }
    wave_reduce<T, B, S>(n3, red, t.grp, t.lane);
    wave_reduce<T, B, S>(s2, red, t.grp, t.lane);
    wave_reduce<T, B, S>(vx, red, t.grp, t.lane);
    wave_reduce<T, B, S>(vy, red, t.grp, t.lane);
    wave_reduce<T, B, S>(vz, red, t.grp, t.lane);
    if (t.grp > 0 || !t.active) return;
    const int nx = g.nx[0], ny = g.nx[1], nz = g.nx[2];
    const int c = t.c;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int z = t.zb + b;
        if (z >= nz) break;
        const size_t i = t.x + (size_t)nx * (t.y + (size_t)ny * z);
        if (part) {
            T *pp = part + (size_t)c * 5 * g.N + i;
            pp[0] = n3[b];
            pp[(size_t)g.N] = s2[b];
            pp[(size_t)2 * g.N] = vx[b];
            pp[(size_t)3 * g.N] = vy[b];
            pp[(size_t)4 * g.N] = vz[b];
            continue;
        }
        const T i4r = (T)cc.i4pir[c], i4r2 = (T)cc.i4pirsq[c];
        const T v[3] = {vx[b], vy[b], vz[b]};
        double vals[10];
        vals[0] = n3[b];
        vals[1] = s2[b];
        vals[2] = i4r * s2[b];
        vals[3] = i4r2 * s2[b];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            vals[4 + d] = v[d];
            vals[7 + d] = i4r * v[d];
        }
#pragma unroll
        for (int q = 0; q < 10; ++q) {
            double r = vals[q];
            if (in_rb) r -= in_rb[(size_t)q * g.N + i];
            if (sub) r -= sub[(size_t)q * g.N + i];
            out_rb[(size_t)q * g.N + i] = r;
        }
    }
}

// rho-bar rows from the per-component sums of k_rb_rows_cols (3D)
template <typename T>
__global__ void k_rb_combine(Geo g, const T *__restrict__ part, CompConsts cc, const double *__restrict__ in_rb,
                             const double *__restrict__ sub, double *__restrict__ out_rb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    T n3 = 0, n2 = 0, n1 = 0, n0 = 0, v2[3] = {0, 0, 0}, v1[3] = {0, 0, 0};
    for (int c = 0; c < g.C; ++c) {
        const T *pp = part + (size_t)c * 5 * g.N + i;
        const T st = pp[0], sd = pp[(size_t)g.N];
        n3 += st;
        n2 += sd;
        n1 += (T)cc.i4pir[c] * sd;
        n0 += (T)cc.i4pirsq[c] * sd;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T vd = pp[(size_t)(2 + d) * g.N];
            v2[d] += vd;
            v1[d] += (T)cc.i4pir[c] * vd;
        }
    }
    double vals[10] = {n3, n2, n1, n0, v2[0], v2[1], v2[2], v1[0], v1[1], v1[2]};
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        double r = vals[q];
        if (in_rb) r -= in_rb[(size_t)q * g.N + i];
        if (sub) r -= sub[(size_t)q * g.N + i];
        out_rb[(size_t)q * g.N + i] = r;
    }
}

// Euler-Lagrange rows on a 3D grid, adjoint sums in column form (FMT-halo
// geometry); the pointwise terms are those of k_el_rows.
template <typename T, int B, int S>
__global__ __launch_bounds__(64 * S) void k_el_rows_cols(Geo g, FmtColArgs<T> fa, const T *__restrict__ adj,
                                                         CompConsts cc, int linear, int nwall,
                                                         const double *__restrict__ x_rho, const double *__restrict__ in_rho,
                                                         const double *__restrict__ in_mf, const double *__restrict__ vext,
                                                         const unsigned char *__restrict__ elzero,
                                                         const double *__restrict__ add, double *__restrict__ out_el)
{
    __shared__ T red[B * 64];
    const ColTile t = col_tile<B>(g);
    T acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;
    {
        const T *S3 = adj + (size_t)(g.C * 4) * g.np + t.ubase;
        const ColPtrs<T> cp = fa.th[t.c];
        for (int col = t.grp; col < cp.ncol; col += S)
            col_sum<T, B>(S3 + cp.lo[col], t.lane, cp.w + cp.woff[col], cp.nchunk[col], t.PXY, acc);
    }
    {
        const T *A0 = adj + (size_t)(t.c * 4) * g.np + t.ubase;
        const size_t np = g.np;
        const ColPtrs<T> cp = fa.de[t.c];
        for (int col = t.grp; col < cp.ncol; col += S) {
            const T *s = A0 + cp.lo[col];
            const int wo = cp.woff[col];
            const T *w = cp.w + wo, *wz = cp.wz + wo;
            const T fx = cp.fxy[2 * col], fy = cp.fxy[2 * col + 1];
            const int nk = cp.nchunk[col];
            const __amdgpu_buffer_rsrc_t rA = col_rsrc(s), rX = col_rsrc(s + np), rY = col_rsrc(s + 2 * np),
                                         rZ = col_rsrc(s + 3 * np);
            const unsigned voff = (unsigned)t.lane * (unsigned)sizeof(T);
            const unsigned pstride = (unsigned)(t.PXY * (long long)sizeof(T));
            for (int k = 0; k < nk; ++k) {
                T a[B], bz[B], wk[2 * B - 1];
#pragma unroll
                for (int i = 0; i < B; ++i) {
                    const unsigned so = (unsigned)(k * B + i) * pstride;
                    a[i] = rsrc_load<T>(rA, voff, so) - fx * rsrc_load<T>(rX, voff, so) -
                           fy * rsrc_load<T>(rY, voff, so);
                    bz[i] = rsrc_load<T>(rZ, voff, so);
                }
#pragma unroll
                for (int j = 0; j < 2 * B - 1; ++j) wk[j] = w[k * B + j];
#pragma unroll
                for (int i = 0; i < B; ++i)
#pragma unroll
                    for (int b = 0; b < B; ++b) acc[b] += wk[i - b + B - 1] * a[i];
#pragma unroll
                for (int j = 0; j < 2 * B - 1; ++j) wk[j] = wz[k * B + j];
#pragma unroll
                for (int i = 0; i < B; ++i)
#pragma unroll
                    for (int b = 0; b < B; ++b) acc[b] -= wk[i - b + B - 1] * bz[i];
            }
        }
    }
    wave_reduce<T, B, S>(acc, red, t.grp, t.lane);
    if (t.grp > 0 || !t.active) return;
    const int nx = g.nx[0], ny = g.nx[1], nz = g.nx[2];
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int z = t.zb + b;
        if (z >= nz) break;
        const size_t row = (size_t)t.c * g.N + t.x + (size_t)nx * (t.y + (size_t)ny * z);
        double r;
        if (elzero[row]) {
            r = in_rho ? in_rho[row] : 0.0;
        } else {
            r = (double)acc[b];
            const double xv = x_rho[row];
            if (!linear) {
                r += log(xv) + cc.mu[t.c];
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
}

// ---------------------------------------------------------------- 1D row kernels
// Each lane owns B consecutive padded positions q0 .. q0+B-1 of the span
// [qlo, qhi) and streams its inputs once through the dense weights of every
// stencil (DevRow1DT): every loaded value feeds B fused multiply-adds against
// weights in scalar registers.  Outputs go to node q - h (one problem) or to
// pad_node[q] (a batch; gap positions are skipped).
template <typename T>
struct RowPtrs {
    const T *w, *wv;
    int lo0, L;
};

template <typename T>
struct RowArgs {
    RowPtrs<T> th[kNC], de[kNC], u[kNC * kNC];
};

template <typename T>
inline RowPtrs<T> row_ptrs(const DevRow1DT<T> &d)
{
    return RowPtrs<T>{d.w, d.wv, d.lo0, d.L};
}

__device__ inline int row_node(const Geo &g, int q) { return g.pad_node ? g.pad_node[q] : q - g.h[0]; }

// acc[b] +=/-= sum_j w_j s[b + lo0 + j] over the stencil's dense range
template <typename T, int B, bool NEG = false>
__device__ __forceinline__ void row_sum(const T *__restrict__ s, const T *__restrict__ wbase, int L, T (&acc)[B])
{
    const T *w = wbase + (kRowPad - (B - 1));
    const int nk = (L + 2 * B - 2) / B;
    for (int k = 0; k < nk; ++k) {
        T v[B], wk[2 * B - 1];
#pragma unroll
        for (int i = 0; i < B; ++i) v[i] = s[k * B + i];
#pragma unroll
        for (int j = 0; j < 2 * B - 1; ++j) wk[j] = w[k * B + j];
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if (NEG)
                    acc[b] -= wk[i - b + B - 1] * v[i];
                else
                    acc[b] += wk[i - b + B - 1] * v[i];
            }
    }
}

// the same loads against the scalar and vector weights of a delta stencil
template <typename T, int B>
__device__ __forceinline__ void row_sum2(const T *__restrict__ s, const T *__restrict__ wbase,
                                         const T *__restrict__ wvbase, int L, T (&a)[B], T (&av)[B])
{
    const T *w = wbase + (kRowPad - (B - 1)), *wv = wvbase + (kRowPad - (B - 1));
    const int nk = (L + 2 * B - 2) / B;
    for (int k = 0; k < nk; ++k) {
        T v[B], wk[2 * B - 1];
#pragma unroll
        for (int i = 0; i < B; ++i) v[i] = s[k * B + i];
#pragma unroll
        for (int j = 0; j < 2 * B - 1; ++j) wk[j] = w[k * B + j];
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int b = 0; b < B; ++b) a[b] += wk[i - b + B - 1] * v[i];
#pragma unroll
        for (int j = 0; j < 2 * B - 1; ++j) wk[j] = wv[k * B + j];
#pragma unroll
        for (int i = 0; i < B; ++i)
#pragma unroll
            for (int b = 0; b < B; ++b) av[b] += wk[i - b + B - 1] * v[i];
    }
}

template <typename T, int B>
__global__ __launch_bounds__(256) void k_rb_rows_1d(Geo g, RowArgs<T> ra, const T *__restrict__ rho_pad, CompConsts cc,
                             const double *__restrict__ in_rb, const double *__restrict__ sub,
                             double *__restrict__ out_rb, int qlo, int qhi)
{
    const int q0 = qlo + (int)(blockIdx.x * blockDim.x + threadIdx.x) * B;
    if (q0 >= qhi) return;
    T n3[B], n2[B], n1[B], n0[B], v2[B], v1[B];
#pragma unroll
    for (int b = 0; b < B; ++b) n3[b] = n2[b] = n1[b] = n0[b] = v2[b] = v1[b] = 0;
    for (int c = 0; c < g.C; ++c) {
        const T *src = rho_pad + (size_t)c * g.np + q0;
        T st[B], sd[B], sv[B];
#pragma unroll
        for (int b = 0; b < B; ++b) st[b] = sd[b] = sv[b] = 0;
        const RowPtrs<T> th = ra.th[c], de = ra.de[c];
        row_sum<T, B>(src + th.lo0, th.w, th.L, st);
        row_sum2<T, B>(src + de.lo0, de.w, de.wv, de.L, sd, sv);
        const T i4r = (T)cc.i4pir[c], i4r2 = (T)cc.i4pirsq[c];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            n3[b] += st[b];
            n2[b] += sd[b];
            n1[b] += i4r * sd[b];
            n0[b] += i4r2 * sd[b];
            v2[b] += sv[b];
            v1[b] += i4r * sv[b];
        }
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int q = q0 + b;
        if (q >= qhi) break;
        const int i = row_node(g, q);
        if (i < 0) continue;
        const double vals[6] = {n3[b], n2[b], n1[b], n0[b], v2[b], v1[b]};
#pragma unroll
        for (int r = 0; r < 6; ++r) {
            double o = vals[r];
            if (in_rb) o -= in_rb[(size_t)r * g.N + i];
            if (sub) o -= sub[(size_t)r * g.N + i];
            out_rb[(size_t)r * g.N + i] = o;
        }
    }
}

template <typename T, int B>
__global__ __launch_bounds__(256) void k_mf_rows_1d(Geo g, RowArgs<T> ra, const T *__restrict__ rho_pad, const double *__restrict__ in_mf,
                             const double *__restrict__ sub, double *__restrict__ out_mf, int qlo, int qhi)
{
    const int q0 = qlo + (int)(blockIdx.x * blockDim.x + threadIdx.x) * B;
    const int c = blockIdx.y;
    if (q0 >= qhi) return;
    T acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;
    for (int cj = 0; cj < g.C; ++cj) {
        const RowPtrs<T> u = ra.u[c + g.C * cj];
        row_sum<T, B>(rho_pad + (size_t)cj * g.np + q0 + u.lo0, u.w, u.L, acc);
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int q = q0 + b;
        if (q >= qhi) break;
        const int i = row_node(g, q);
        if (i < 0) continue;
        const size_t row = (size_t)c * g.N + i;
        double o = acc[b];
        if (in_mf) o -= in_mf[row];
        if (sub) o -= sub[row];
        out_mf[row] = o;
    }
}

template <typename T, int B>
__global__ __launch_bounds__(256) void k_el_rows_1d(Geo g, RowArgs<T> ra, const T *__restrict__ adj, CompConsts cc, int linear, int nwall,
                             const double *__restrict__ x_rho, const double *__restrict__ in_rho,
                             const double *__restrict__ in_mf, const double *__restrict__ vext,
                             const unsigned char *__restrict__ elzero, const double *__restrict__ add,
                             double *__restrict__ out_el, int qlo, int qhi)
{
    const int q0 = qlo + (int)(blockIdx.x * blockDim.x + threadIdx.x) * B;
    const int c = blockIdx.y;
    if (q0 >= qhi) return;
    T acc[B];
#pragma unroll
    for (int b = 0; b < B; ++b) acc[b] = 0;
    const RowPtrs<T> th = ra.th[c], de = ra.de[c];
    const T *S3 = adj + (size_t)(g.C * 2) * g.np + q0;
    const T *A = adj + (size_t)(c * 2) * g.np + q0;
    row_sum<T, B>(S3 + th.lo0, th.w, th.L, acc);
    row_sum<T, B>(A + de.lo0, de.w, de.L, acc);
    row_sum<T, B, true>(A + g.np + de.lo0, de.wv, de.L, acc);
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int q = q0 + b;
        if (q >= qhi) break;
        const int i = row_node(g, q);
        if (i < 0) continue;
        const size_t row = (size_t)c * g.N + i;
        double r;
        if (elzero[row]) {
            r = in_rho ? in_rho[row] : 0.0;
        } else {
            r = (double)acc[b];
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
}

template <typename T>
RowArgs<T> row_args(const Problem &P)
{
    RowArgs<T> ra{};
    for (int c = 0; c < P.p.ncomp; ++c) {
        ra.th[c] = row_ptrs(dev_row<T>(P, TMI_THETA, c));
        ra.de[c] = row_ptrs(dev_row<T>(P, TMI_DELTA, c));
    }
    if (P.nmf)
        for (int q = 0; q < P.p.ncomp * P.p.ncomp; ++q) ra.u[q] = row_ptrs(dev_row<T>(P, TMI_UATT, q));
    return ra;
}

// kernel kinds of the 1D row autotuner
enum RowKind { kRowRb = 0, kRowMf = 1, kRowEl = 2 };

template <typename T>
inline const RowTune &row_tune(const Problem &P, int kind)
{
    return P.tune[sizeof(T) == 4 ? 1 : 0][kind];
}

inline dim3 row_grid(const PadGeom &pg, int B, int wg, int ny)
{
    const long long span = (long long)pg.P[0] - 2LL * pg.h[0];
    const long long lanes = (span + B - 1) / B;
    return dim3((unsigned)((lanes + wg - 1) / wg), (unsigned)ny);
}

#define TMI_ROW_DISPATCH(B_, LAUNCH) \
    switch (B_) {                    \
    case 4: LAUNCH(4); break;        \
    case 16: LAUNCH(16); break;      \
    default: LAUNCH(8); break;       \
    }

template <typename T>
void rb_rows_1d(Problem &P, const RowTune &t, const double *in_rb, const double *sub, double *out_rb, hipStream_t s)
{
    Geo gu = make_geo(P, P.pad_u);
    const RowArgs<T> ra = row_args<T>(P);
    const CompConsts cc = make_consts(P);
    const T *rp = workspace<T>(P).rho_pad;
    const int qlo = P.pad_u.h[0], qhi = P.pad_u.P[0] - P.pad_u.h[0];
#define L_(B) hipLaunchKernelGGL((k_rb_rows_1d<T, B>), row_grid(P.pad_u, B, t.wg, 1), dim3(t.wg), 0, s, gu, ra, rp, cc, \
                                 in_rb, sub, out_rb, qlo, qhi)
    TMI_ROW_DISPATCH(t.B, L_)
#undef L_
}

template <typename T>
void mf_rows_1d(Problem &P, const RowTune &t, const double *in_mf, const double *sub, double *out_mf, hipStream_t s)
{
    Geo gu = make_geo(P, P.pad_u);
    const RowArgs<T> ra = row_args<T>(P);
    const T *rp = workspace<T>(P).rho_pad;
    const int qlo = P.pad_u.h[0], qhi = P.pad_u.P[0] - P.pad_u.h[0];
#define L_(B) hipLaunchKernelGGL((k_mf_rows_1d<T, B>), row_grid(P.pad_u, B, t.wg, P.p.ncomp), dim3(t.wg), 0, s, gu, \
                                 ra, rp, in_mf, sub, out_mf, qlo, qhi)
    TMI_ROW_DISPATCH(t.B, L_)
#undef L_
}

template <typename T>
void el_rows_1d(Problem &P, const RowTune &t, bool linear, const double *x_rho, const double *in_rho,
                const double *in_mf, const double *add, double *out_el, hipStream_t s)
{
    Geo gf = make_geo(P, P.pad_f);
    const RowArgs<T> ra = row_args<T>(P);
    const CompConsts cc = make_consts(P);
    const T *adj = workspace<T>(P).adj_pad;
    const int qlo = P.pad_f.h[0], qhi = P.pad_f.P[0] - P.pad_f.h[0];
    const int lin = linear ? 1 : 0;
#define L_(B) hipLaunchKernelGGL((k_el_rows_1d<T, B>), row_grid(P.pad_f, B, t.wg, P.p.ncomp), dim3(t.wg), 0, s, gf, \
                                 ra, adj, cc, lin, P.p.nwall, x_rho, in_rho, in_mf, P.d_vext, P.d_elzero, add, out_el, \
                                 qlo, qhi)
    TMI_ROW_DISPATCH(t.B, L_)
#undef L_
}

struct ElArgs {
    StenPtrs th[kNC], de[kNC];
};

// Euler-Lagrange rows.
//  residual: F = ln x + mu_c + Vext + adjoint sums + x_MF ; zero rows: x
//  linear  : J v = v/x + adjoint sums + v_MF           ; zero rows: v
// in_rho may be NULL (treated as 0); add (may be NULL) is added to the result.
template <typename T, int ND>
__global__ void k_el_rows(Geo g, ElArgs ea, const T *__restrict__ adj, CompConsts cc, int linear, int nwall,
                          const double *__restrict__ x_rho, const double *__restrict__ in_rho,
                          const double *__restrict__ in_mf, const double *__restrict__ vext,
                          const unsigned char *__restrict__ elzero, const double *__restrict__ add,
                          double *__restrict__ out_el)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= g.N) return;
    const size_t row = (size_t)c * g.N + i;
    double r;
    if (elzero[row]) {
        r = in_rho ? in_rho[row] : 0.0;
    } else {
        const long long p = pad_pos(g, i);
        const T *A = adj + (size_t)(c * (1 + ND)) * g.np + p;
        T acc = 0;
        {
            const int *lo = ea.de[c].lo;
            const T *w = (const T *)ea.de[c].w;
            const T *wv = (const T *)ea.de[c].wv;
            const int L = ea.de[c].L;
            for (int k = 0; k < L; ++k) {
                const long long q = lo[k];
                T t = w[k] * A[q];
#pragma unroll
                for (int d = 0; d < ND; ++d) t -= wv[k * ND + d] * A[(size_t)(1 + d) * g.np + q];
                acc += t;
            }
        }
        {
            const T *S3 = adj + (size_t)(g.C * (1 + ND)) * g.np + p;
            const int *lo = ea.th[c].lo;
            const T *w = (const T *)ea.th[c].w;
            const int L = ea.th[c].L;
            for (int k = 0; k < L; ++k) acc += w[k] * S3[lo[k]];
        }
        r = (double)acc;
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

// ---------------------------------------------------------------- init fields
// calc_init_rho_bar (src/dft_guess_HSFMT.c:70-83): rho-bar = forward sums of the
// densities with bulk halo, n3 capped at 0.98 when above 1
__global__ void k_cap_n3(double *n3, int N)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N && n3[i] > 1.0) n3[i] = 0.98;
}

// ---------------------------------------------------------------- vector ops
template <int B>
__device__ inline double block_sum(double v)
{
    __shared__ double sh[B / 64];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (wid == 0) {
        v = (lane < B / 64) ? sh[lane] : 0.0;
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    }
    __syncthreads();
    return v;
}

// partial[j*gridDim.x + blk] = sum over the block's range of V_j . w, four vectors per sweep over the
// range (w is re-read once per four vectors, from cache): V is read once, with few registers.  Each
// V_j . w is summed in the same order as by a sweep per vector.
// V may be an FP32 basis (the mixed path's Krylov vectors); products and sums are FP64.
constexpr int kDotGroup = 4;
template <typename TV>
__global__ __launch_bounds__(kBlock) void k_multidot(const TV *__restrict__ V, long long ld, int k,
                                                     const double *__restrict__ w, long long n,
                                                     double *__restrict__ partial)
{
    for (int j0 = 0; j0 < k; j0 += kDotGroup) {
        double acc[kDotGroup];
#pragma unroll
        for (int g = 0; g < kDotGroup; ++g) acc[g] = 0.0;
        for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
            const double wi = w[i];
#pragma unroll
            for (int g = 0; g < kDotGroup; ++g)
                if (j0 + g < k) acc[g] += (double)V[(size_t)(j0 + g) * ld + i] * wi;
        }
#pragma unroll
        for (int g = 0; g < kDotGroup; ++g) {
            if (j0 + g < k) {
                const double t = block_sum<kBlock>(acc[g]);
                if (threadIdx.x == 0) partial[(size_t)(j0 + g) * gridDim.x + blockIdx.x] = t;
            }
        }
    }
}

// out[j] = sum_b partial[j*nb + b]; mirror (optional, pinned host memory) gets a copy, which the
// host reads after synchronising with the stream: no separate transfer
__global__ void k_reduce_rows(const double *__restrict__ partial, int nb, int k, double *__restrict__ out,
                              double *__restrict__ mirror)
{
    int j = blockIdx.x;
    if (j >= k) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += partial[(size_t)j * nb + b];
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) {
        if (out) out[j] = s;
        if (mirror) mirror[j] = s;
    }
}

// sum of nb partial sums by one workgroup of 1024 threads, four independent loads in flight per
// thread (a single wide block: the sum is one short dependent chain instead of nb / 256 loads);
// the order is fixed, so the result is the same on every call
constexpr int kSumBlock = 1024;
__global__ __launch_bounds__(kSumBlock) void k_sum_partials(const double *__restrict__ partial, int nb,
                                                            double *__restrict__ out, double *__restrict__ mirror)
{
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int b = threadIdx.x;
    for (; b + 3 * kSumBlock < nb; b += 4 * kSumBlock) {
        a0 += partial[b];
        a1 += partial[b + kSumBlock];
        a2 += partial[b + 2 * kSumBlock];
        a3 += partial[b + 3 * kSumBlock];
    }
    for (; b < nb; b += kSumBlock) a0 += partial[b];
    const double s = block_sum<kSumBlock>((a0 + a1) + (a2 + a3));
    if (threadIdx.x == 0) {
        if (out) out[0] = s;
        if (mirror) mirror[0] = s;
    }
}

__global__ void k_axpy(double a, const double *__restrict__ x, double *__restrict__ y, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] += a * x[i];
}

__global__ void k_hw_gather(int nrows, const int *__restrict__ rowptr, const int *__restrict__ rowid,
                            const int *__restrict__ col, const double *__restrict__ val, const double *__restrict__ in,
                            double *__restrict__ out)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    double sum = 0.0;
    for (int k = rowptr[r]; k < rowptr[r + 1]; ++k) sum += val[k] * in[col[k]];
    out[rowid[r]] += sum;
}
// one particle size (dft_fill_FMT.c:215-236, 305-320): rows n1 = 2, n0 = 3, nv1 = 4 + D + d of the
// rho-bar block depend on n2 = 1 and nv2 = 4 + d with factors c1 = 1/(4 pi R), c0 = 1/(4 pi R^2).
// mode 0: out_dep = -in_dep + B in_ind; 1: out_dep += B out_ind (in unused); 2: out_dep -= in_dep,
// out_ind += B^T in_dep
__global__ void k_one_size(int mode, int N, int nd, double c1, double c0, const double *__restrict__ in,
                           double *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const size_t n = (size_t)N;
    if (mode == 0) {
        out[2 * n + i] = -in[2 * n + i] + in[n + i] * c1;
        out[3 * n + i] = -in[3 * n + i] + in[n + i] * c0;
        for (int d = 0; d < nd; ++d)
            out[(4 + nd + d) * n + i] = -in[(4 + nd + d) * n + i] + in[(4 + d) * n + i] * c1;
    } else if (mode == 1) {
        out[2 * n + i] += c1 * out[n + i];
        out[3 * n + i] += c0 * out[n + i];
        for (int d = 0; d < nd; ++d) out[(4 + nd + d) * n + i] += c1 * out[(4 + d) * n + i];
    } else {
        out[2 * n + i] -= in[2 * n + i];
        out[3 * n + i] -= in[3 * n + i];
        out[n + i] += c1 * in[2 * n + i] + c0 * in[3 * n + i];
        for (int d = 0; d < nd; ++d) {
            out[(4 + nd + d) * n + i] -= in[(4 + nd + d) * n + i];
            out[(4 + d) * n + i] += c1 * in[(4 + nd + d) * n + i];
        }
    }
}
__global__ void k_fill(double v, double *__restrict__ y, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        y[i] = v;
}
__global__ void k_scale_copy(double a, const double *__restrict__ x, double *__restrict__ y, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

__global__ void k_mul(const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ y, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = a[i] * b[i];
}

// w -= V h (k columns)
// y = x / sqrt(*nrm2) (0 when *nrm2 is 0): the next Arnoldi vector without a host round trip, stored
// in the basis type
template <typename TV>
__global__ void k_scale_by_norm(const double *__restrict__ nrm2, const double *__restrict__ x, TV *__restrict__ y,
                                long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double q = *nrm2;
    y[i] = (TV)(q > 0.0 ? x[i] * (1.0 / sqrt(q)) : 0.0);
}

// y = (TV)(a x), and its widening b = (double) a
template <typename TV>
__global__ void k_cast_scaled(double a, const double *__restrict__ x, TV *__restrict__ y, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = (TV)(a * x[i]);
}
template <typename TV>
__global__ void k_widen(const TV *__restrict__ a, double *__restrict__ b, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

__global__ void k_gemv_sub(const double *__restrict__ V, long long ld, int k, const double *__restrict__ h,
                           double *__restrict__ w, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += V[(size_t)j * ld + i] * h[j];
    w[i] -= s;
}

// w -= V h (k columns); NORM: partial[blk] = sum over the block's elements of w^2 (updated w)
template <typename TV, bool NORM>
__global__ __launch_bounds__(kBlock) void k_gemv_sub_norm(const TV *__restrict__ V, long long ld, int k,
                                                          const double *__restrict__ h, double *__restrict__ w,
                                                          long long n, double *__restrict__ partial)
{
    double q = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += (double)V[(size_t)j * ld + i] * h[j];
        const double wi = w[i] - s;
        w[i] = wi;
        q += wi * wi;
    }
    if constexpr (NORM) {
        q = block_sum<kBlock>(q);
        if (threadIdx.x == 0) partial[blockIdx.x] = q;
    }
}

template <typename TV>
__global__ void k_gemv_add(const TV *__restrict__ V, long long ld, int k, const double *__restrict__ y,
                           double *__restrict__ x, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s += (double)V[(size_t)j * ld + i] * y[j];
    x[i] += s;
}

// inverse of the A22 diagonal: x (EL rows: dF/drho = 1/rho) or 1 on zero rows
__global__ void k_a22_dinv(const double *__restrict__ x_rho, const unsigned char *__restrict__ elzero,
                           double *__restrict__ dinv, long long n)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) dinv[i] = elzero[i] ? 1.0 : x_rho[i];
}

// update_solution (src/dft_newton.c:266-380): fraction limited by density positivity
__global__ void k_frac_min(const double *__restrict__ x, const double *__restrict__ dx, long long n, double scaling,
                           double *__restrict__ partial)
{
    double m = 1.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const double xv = x[i], d = dx[i];
        if (xv + d < 0.0) {
            double f = fmin(1.0, xv / (-d));
            f = fmax(f, scaling);
            m = fmin(m, f);
        }
    }
    __shared__ double sh[kBlock];
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_update_norm(const double *__restrict__ x, const double *__restrict__ dx, long long n, double frac,
                              double rel, double abs_tol, double *__restrict__ partial)
{
    double s = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n; i += (long long)gridDim.x * kBlock) {
        const double t = (frac * dx[i]) / (rel * x[i] + abs_tol);
        s += t * t;
    }
    s = block_sum<kBlock>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_apply_update(double *__restrict__ x, const double *__restrict__ dx, long long n, long long n_rho,
                               int N, int iu_n3, double frac)
{
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double xv = x[i], d = dx[i];
    const bool dens = i < n_rho;
    const bool is_n3 = (i / N) == iu_n3;
    // a density whose own fraction x / (-dx) is at or below frac (the node that set frac, or one
    // past it) reaches x + frac dx <= 0 in exact arithmetic; in floating point the limiting node's
    // x + frac dx is a residue of either sign (~1e-16 x), and a positive one would leave a density
    // of ~1e-20 whose 1/rho term swamps the next Jacobian, so it takes the 0.1 x branch by this test
    const bool limited = dens && xv + d < 0.0 && fmin(1.0, xv / (-d)) <= frac;
    if ((xv + frac * d < 1.e-99 || limited) && dens)
        x[i] = 0.1 * xv;
    else if (is_n3 && xv + frac * d > 1.0)
        x[i] = xv + 0.5 * (1.0 - xv);
    else
        x[i] = xv + frac * d;
}

template <typename T>
void pad_rho(Problem &P, const double *in, bool linear, hipStream_t s)
{
    Geo gu = make_geo(P, P.pad_u);
    hipLaunchKernelGGL(k_pad_rho<T>, dim3(grid_for(gu.np)), dim3(kBlock), 0, s, gu, in, P.d_zero, make_consts(P),
                       linear ? 1 : 0, workspace<T>(P).rho_pad);
}

template <typename T>
void rb_rows(Problem &P, const double *in_rb, const double *sub, double *out_rb, hipStream_t s)
{
    Geo gu = make_geo(P, P.pad_u);
    RbArgs ra{};
    for (int c = 0; c < P.p.ncomp; ++c) {
        ra.th[c] = sten_ptrs<T>(P, TMI_THETA, c, kGeomU);
        ra.de[c] = sten_ptrs<T>(P, TMI_DELTA, c, kGeomU);
    }
    const CompConsts cc = make_consts(P);
    const T *rp = workspace<T>(P).rho_pad;
    if (P.use_rows1d) {
        rb_rows_1d<T>(P, row_tune<T>(P, kRowRb), in_rb, sub, out_rb, s);
        return;
    }
    if (P.use_columns) {
        FmtColArgs<T> fa{};
        for (int c = 0; c < P.p.ncomp; ++c) {
            fa.th[c] = col_ptrs(dev_cols<T>(P, TMI_THETA, c), kGeomU);
            fa.de[c] = col_ptrs(dev_cols<T>(P, TMI_DELTA, c), kGeomU);
        }
        T *part = (P.p.ncomp > 1) ? workspace<T>(P).rb_part : nullptr;
        dim3 grid(col_tiles(P));
        hipLaunchKernelGGL((k_rb_rows_cols<T, kColB, kColS>), grid, dim3(64, kColS), 0, s, gu, fa, rp, cc, in_rb, sub,
                           out_rb, part);
        if (part)
            hipLaunchKernelGGL(k_rb_combine<T>, dim3(grid_for(P.N)), dim3(kBlock), 0, s, gu, part, cc, in_rb, sub, out_rb);
        return;
    }
    if (P.p.ndim == 2 && P.d_rbw2 && P.rb2d_mfma) {   // one component: 16 x 16 tiles on the matrix cores
        const T *Wr = std::is_same<T, float>::value ? (const T *)P.f_rbw2 : (const T *)P.d_rbw2;
        const dim3 grid((unsigned)((P.p.nx[0] + 15) / 16), (unsigned)((P.pad_u.P[1] - 2 * P.pad_u.h[1] + 15) / 16));
        hipLaunchKernelGGL(k_rb_rows_2d_mfma<T>, grid, dim3(64 * kRbMfmaWaves), 0, s, gu, P.rbw_h[0], P.rbw_h[1], Wr, rp,
                           cc.i4pir[0], cc.i4pirsq[0], in_rb, sub, out_rb);
        return;
    }
    switch (P.p.ndim) {
    case 1: hipLaunchKernelGGL((k_rb_rows<T, 1>), dim3(grid_for(P.N)), dim3(kBlock), 0, s, gu, ra, rp, cc, in_rb, sub, out_rb); break;
    case 2: hipLaunchKernelGGL((k_rb_rows<T, 2>), dim3(grid_for(P.N)), dim3(kBlock), 0, s, gu, ra, rp, cc, in_rb, sub, out_rb); break;
    default: hipLaunchKernelGGL((k_rb_rows<T, 3>), dim3(grid_for(P.N)), dim3(kBlock), 0, s, gu, ra, rp, cc, in_rb, sub, out_rb); break;
    }
}

template <typename T>
void mf_rows(Problem &P, const double *in_mf, const double *sub, double *out_mf, hipStream_t s)
{
    if (P.nmf == 0) return;
    if (fft_mf(P)) {
        fft_rb_mf<T>(P, nullptr, false, nullptr, nullptr, nullptr, in_mf, sub, out_mf, s);
        return;
    }
    if (P.use_rows1d) {
        mf_rows_1d<T>(P, row_tune<T>(P, kRowMf), in_mf, sub, out_mf, s);
        return;
    }
    Geo gu = make_geo(P, P.pad_u);
    if (P.use_columns) {
        ColArgs<T> ca{};
        for (int q = 0; q < P.p.ncomp * P.p.ncomp; ++q) ca.u[q] = col_ptrs(dev_cols<T>(P, TMI_UATT, q), kGeomU);
        dim3 grid(col_tiles(P));
        hipLaunchKernelGGL((k_mf_rows_cols<T, kColB, kColS>), grid, dim3(64, kColS), 0, s, gu, ca,
                           workspace<T>(P).rho_pad, in_mf, sub, out_mf);
        return;
    }
    if (P.p.ndim == 2 && P.d_mfw2) {   // dense 2D convolution
        const T *W = std::is_same<T, float>::value ? (const T *)P.f_mfw2 : (const T *)P.d_mfw2;
        if (P.mf2d_mfma) {   // 16 x 16 output tiles on the matrix cores
            const dim3 grid((unsigned)((P.p.nx[0] + 15) / 16),
                            (unsigned)((P.pad_u.P[1] - 2 * P.pad_u.h[1] + 15) / 16), (unsigned)P.p.ncomp);
            hipLaunchKernelGGL(k_mf_rows_2d_mfma<T>, grid, dim3(64 * kMfmaWaves), 0, s, gu, W, workspace<T>(P).rho_pad,
                               in_mf, sub, out_mf);
            return;
        }
        const dim3 grid((unsigned)((P.p.nx[0] + 63) / 64),
                        (unsigned)((P.pad_u.P[1] - 2 * P.pad_u.h[1] + kMf2dRows - 1) / kMf2dRows), (unsigned)P.p.ncomp);
        hipLaunchKernelGGL((k_mf_rows_2d<T, kMf2dRows>), grid, dim3(64 * kMf2dWaves), 0, s, gu, W,
                           workspace<T>(P).rho_pad, in_mf, sub, out_mf);
        return;
    }
    UArgs ua{};
    for (int q = 0; q < P.p.ncomp * P.p.ncomp; ++q) ua.u[q] = sten_ptrs<T>(P, TMI_UATT, q, kGeomU);
    hipLaunchKernelGGL(k_mf_rows<T>, dim3(grid_for(P.N), P.p.ncomp), dim3(kBlock), 0, s, gu, ua,
                       workspace<T>(P).rho_pad, in_mf, sub, out_mf);
}

// rho-bar rows and mean-field rows of the densities rho (node arrays [C][N];
// residual: bulk halo, linear: zero halo): one forward transform serves both
// when the FMT sums are in FFT form, and on wrapping grids it reads the node
// arrays without a density halo.  ev_pad and ev_mid (optional) are recorded
// after the density halo and between the two row blocks.
template <typename T>
void rb_mf_rows(Problem &P, const double *rho, bool linear, const double *in_rb, const double *sub_rb, double *out_rb,
                const double *in_mf, const double *sub_mf, double *out_mf, hipStream_t s, hipEvent_t ev_pad = nullptr,
                hipEvent_t ev_mid = nullptr, double *ss = nullptr, long long *nss = nullptr)
{
    if (nss) *nss = 0;
    const bool nodes = fft_node_input(P);
    if (!nodes) pad_rho<T>(P, rho, linear, s);
    if (ev_pad) TMI_HIP_CHECK(hipEventRecord(ev_pad, s));
    if (fft_fmt(P)) {
        fft_rb_mf<T>(P, nodes ? rho : nullptr, true, in_rb, sub_rb, out_rb, in_mf, sub_mf, out_mf, s, ss, nss);
        if (ev_mid) TMI_HIP_CHECK(hipEventRecord(ev_mid, s));
        if (!fft_mf(P)) mf_rows<T>(P, in_mf, sub_mf, out_mf, s);
        return;
    }
    rb_rows<T>(P, in_rb, sub_rb, out_rb, s);
    if (ev_mid) TMI_HIP_CHECK(hipEventRecord(ev_mid, s));
    mf_rows<T>(P, in_mf, sub_mf, out_mf, s);
}

template <typename T>
void fmt_point(Problem &P, const double *x_rb, const double *dn_rb, bool linear, hipStream_t s)
{
    Geo gf = make_geo(P, P.pad_f);
    const CompConsts cc = make_consts(P);
    T *adj = workspace<T>(P).adj_pad;
    const int func = P.p.type_func;
    const int mode = !linear ? kFmtResidual : (P.jac_ref ? kFmtLinearRef : kFmtLinear);
    const bool pairs = fft_fmt(P) && fft_wraps(P);   // straight into the transform buffer as pairs: no halo
    const FftPairs fp = pairs ? fft_pairs(P, linear) : FftPairs{nullptr, 0, 0, 0, 0};
    auto go = [&](auto nd, auto md) {
        constexpr int ND = decltype(nd)::value, MODE = decltype(md)::value;
        hipLaunchKernelGGL((k_fmt_point<T, ND, MODE>), dim3(grid_for(P.N)), dim3(kBlock), 0, s, gf, func, x_rb,
                           dn_rb, cc, adj, fp);
        if (!pairs)
            hipLaunchKernelGGL((k_pad_adj<T, ND, MODE>), dim3(grid_for(gf.np)), dim3(kBlock), 0, s, gf, cc, adj, func,
                               x_rb, dn_rb);
    };
    auto by_mode = [&](auto nd) {
        if (mode == kFmtResidual)
            go(nd, std::integral_constant<int, kFmtResidual>{});
        else if (mode == kFmtLinear)
            go(nd, std::integral_constant<int, kFmtLinear>{});
        else
            go(nd, std::integral_constant<int, kFmtLinearRef>{});
    };
    if (P.p.ndim == 1)
        by_mode(std::integral_constant<int, 1>{});
    else if (P.p.ndim == 2)
        by_mode(std::integral_constant<int, 2>{});
    else
        by_mode(std::integral_constant<int, 3>{});
}

template <typename T>
void el_rows(Problem &P, bool linear, const double *x_rho, const double *in_rho, const double *in_mf, const double *add,
             double *out_el, hipStream_t s, double *ss = nullptr, long long *nss = nullptr)
{
    if (nss) *nss = 0;
    if (fft_fmt(P)) {
        fft_el<T>(P, linear, x_rho, in_rho, in_mf, add, out_el, s, ss, nss);
        return;
    }
    Geo gf = make_geo(P, P.pad_f);
    ElArgs ea{};
    for (int c = 0; c < P.p.ncomp; ++c) {
        ea.th[c] = sten_ptrs<T>(P, TMI_THETA, c, kGeomF);
        ea.de[c] = sten_ptrs<T>(P, TMI_DELTA, c, kGeomF);
    }
    const CompConsts cc = make_consts(P);
    const T *adj = workspace<T>(P).adj_pad;
    const int lin = linear ? 1 : 0;
    if (P.use_rows1d) {
        el_rows_1d<T>(P, row_tune<T>(P, kRowEl), linear, x_rho, in_rho, in_mf, add, out_el, s);
        return;
    }
    if (P.use_columns) {
        FmtColArgs<T> fa{};
        for (int c = 0; c < P.p.ncomp; ++c) {
            fa.th[c] = col_ptrs(dev_cols<T>(P, TMI_THETA, c), kGeomF);
            fa.de[c] = col_ptrs(dev_cols<T>(P, TMI_DELTA, c), kGeomF);
        }
        dim3 cgrid(col_tiles(P));
        hipLaunchKernelGGL((k_el_rows_cols<T, kColB, kColS>), cgrid, dim3(64, kColS), 0, s, gf, fa, adj, cc, lin,
                           P.p.nwall, x_rho, in_rho, in_mf, P.d_vext, P.d_elzero, add, out_el);
        return;
    }
    if (P.p.ndim == 2 && P.d_rbw2 && P.rb2d_mfma && P.rbw_el) {   // one component: 16 x 16 tiles on the matrix cores
        const T *Wr = std::is_same<T, float>::value ? (const T *)P.f_rbw2 : (const T *)P.d_rbw2;
        const dim3 g2((unsigned)((P.p.nx[0] + 15) / 16), (unsigned)((P.pad_f.P[1] - 2 * P.pad_f.h[1] + 15) / 16));
        hipLaunchKernelGGL(k_el_rows_2d_mfma<T>, g2, dim3(64 * kRbMfmaWaves), 0, s, gf, P.rbw_h[0], P.rbw_h[1], Wr, adj, cc,
                           lin, P.p.nwall, x_rho, in_rho, in_mf, P.d_vext, P.d_elzero, add, out_el);
        return;
    }
    dim3 grid(grid_for(P.N), P.p.ncomp);
    switch (P.p.ndim) {
    case 1:
        hipLaunchKernelGGL((k_el_rows<T, 1>), grid, dim3(kBlock), 0, s, gf, ea, adj, cc, lin, P.p.nwall, x_rho, in_rho, in_mf,
                           P.d_vext, P.d_elzero, add, out_el);
        break;
    case 2:
        hipLaunchKernelGGL((k_el_rows<T, 2>), grid, dim3(kBlock), 0, s, gf, ea, adj, cc, lin, P.p.nwall, x_rho, in_rho, in_mf,
                           P.d_vext, P.d_elzero, add, out_el);
        break;
    default:
        hipLaunchKernelGGL((k_el_rows<T, 3>), grid, dim3(kBlock), 0, s, gf, ea, adj, cc, lin, P.p.nwall, x_rho, in_rho, in_mf,
                           P.d_vext, P.d_elzero, add, out_el);
        break;
    }
}

}  // namespace

// ---------------------------------------------------------------- 1D row autotuner
// Times every (outputs per lane, lanes per workgroup) shape of the three 1D
// row kernels, in both precisions, on the problem's own padded buffers and
// keeps the fastest (TMI_AUTOTUNE=0 keeps the defaults: 8 per lane, 256 lanes).
// Problems under kAutotuneMinNodes keep the defaults: their fills are bound by
// launch latency, and timing 54 shapes would cost more than it could save.
constexpr int kAutotuneMinNodes = 16384;
void autotune_rows1d(Problem &P)
{
    if (!P.use_rows1d || P.N < kAutotuneMinNodes) return;
    const char *env = std::getenv("TMI_AUTOTUNE");
    if (env && std::string(env) == "0") return;
    const size_t n = (size_t)P.nunk * P.N;
    double *out = nullptr, *xz = nullptr;
    TMI_HIP_CHECK(hipMalloc((void **)&out, n * sizeof(double)));
    hipError_t ea = hipMalloc((void **)&xz, (size_t)P.p.ncomp * P.N * sizeof(double));
    if (ea != hipSuccess) {
        (void)hipFree(out);
        TMI_HIP_CHECK(ea);
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    try {
        TMI_HIP_CHECK(hipMemset(xz, 0, (size_t)P.p.ncomp * P.N * sizeof(double)));
        TMI_HIP_CHECK(hipEventCreate(&e0));
        TMI_HIP_CHECK(hipEventCreate(&e1));
        const int Bs[3] = {4, 8, 16}, wgs[3] = {64, 128, 256};
        for (int prec = 0; prec < 2; ++prec)
            for (int kind = 0; kind < 3; ++kind) {
                if (kind == kRowMf && P.nmf == 0) continue;
                RowTune best;
                best.ms = 1e30;
                for (int B : Bs)
                    for (int wg : wgs) {
                        RowTune t;
                        t.B = B;
                        t.wg = wg;
                        auto launch = [&] {
                            if (prec == 0) {
                                if (kind == kRowRb) rb_rows_1d<double>(P, t, nullptr, nullptr, out + (size_t)P.iu_rb * P.N, 0);
                                if (kind == kRowMf) mf_rows_1d<double>(P, t, nullptr, nullptr, out + (size_t)P.iu_mf * P.N, 0);
                                if (kind == kRowEl) el_rows_1d<double>(P, t, true, xz, nullptr, nullptr, nullptr, out, 0);
                            } else {
                                if (kind == kRowRb) rb_rows_1d<float>(P, t, nullptr, nullptr, out + (size_t)P.iu_rb * P.N, 0);
                                if (kind == kRowMf) mf_rows_1d<float>(P, t, nullptr, nullptr, out + (size_t)P.iu_mf * P.N, 0);
                                if (kind == kRowEl) el_rows_1d<float>(P, t, true, xz, nullptr, nullptr, nullptr, out, 0);
                            }
                        };
                        launch();
                        TMI_HIP_CHECK(hipEventRecord(e0, 0));
                        for (int r = 0; r < 5; ++r) launch();
                        TMI_HIP_CHECK(hipEventRecord(e1, 0));
                        TMI_HIP_CHECK(hipEventSynchronize(e1));
                        float ms = 0.f;
                        TMI_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                        t.ms = ms / 5.0;
                        if (t.ms < best.ms) best = t;
                    }
                P.tune[prec][kind] = best;
            }
    } catch (...) {
        (void)hipFree(out);
        (void)hipFree(xz);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipFree(out);
    (void)hipFree(xz);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// ---------------------------------------------------------------- FMT form
// Times residual fills with the FMT sums as direct stencil sums, as FFT
// convolutions with separate x passes and product kernels, and with the fused
// x pass for the density product, the Euler-Lagrange product or both (state:
// every unknown 4.8e-4, a positive byte pattern), and keeps the fastest.
// TMI_FMT_FFT=1 / 0 forces the FFT / direct form (the fused x pass is then
// used for both products where it fits, unless TMI_FFT_XFUSE=0).
void choose_fmt_form(Problem &P)
{
    if (!fft_fmt_available(P)) return;
    const char *env = std::getenv("TMI_FMT_FFT"), *xf = std::getenv("TMI_FFT_XFUSE");
    if (env) {
        fft_set_fmt(P, std::string(env) == "1");
        fft_set_fused(P, (xf && std::string(xf) == "0") ? 0 : 3);
        return;
    }
    const size_t n = (size_t)P.nunk * P.N;
    double *x = nullptr, *F = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // forms: 0 direct; 1 + mask: FFT with the fused x pass for the products in mask (bit 0 density, bit 1 EL)
    const int nform = fft_fused_available(P) ? 5 : 2;
    float ms[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
    auto set_form = [&](int form) {
        fft_set_fmt(P, form > 0);
        fft_set_fused(P, form > 0 ? form - 1 : 0);
    };
    try {
        TMI_HIP_CHECK(hipMalloc((void **)&x, n * sizeof(double)));
        TMI_HIP_CHECK(hipMalloc((void **)&F, n * sizeof(double)));
        TMI_HIP_CHECK(hipMemset(x, 0x3f, n * sizeof(double)));
        TMI_HIP_CHECK(hipEventCreate(&e0));
        TMI_HIP_CHECK(hipEventCreate(&e1));
        for (int form = 0; form < nform; ++form) {
            set_form(form);
            launch_residual(P, x, F, 0);
            TMI_HIP_CHECK(hipEventRecord(e0, 0));
            for (int r = 0; r < 3; ++r) launch_residual(P, x, F, 0);
            TMI_HIP_CHECK(hipEventRecord(e1, 0));
            TMI_HIP_CHECK(hipEventSynchronize(e1));
            TMI_HIP_CHECK(hipEventElapsedTime(&ms[form], e0, e1));
        }
    } catch (...) {
        (void)hipFree(x);
        (void)hipFree(F);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipFree(x);
    (void)hipFree(F);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int best = 0;
    for (int form = 1; form < nform; ++form)
        if (ms[form] < ms[best]) best = form;
    set_form(best);
}

// ---------------------------------------------------------------- public launchers
// ev (optional): 6 events recorded between the stages pad_rho | rb_rows |
// mf_rows | fmt_point+pad_adj | el_rows
namespace {
// ss (optional): partial sums of F^2 from the fused last passes; *nss their number when they cover
// every row, else 0
void residual_impl(Problem &P, const double *x, double *F, hipStream_t s, hipEvent_t *ev, double *ss, long long *nss)
{
    const size_t N = P.N;
    long long n_rb = 0, n_el = 0;
    if (P.hw_fwd.nrows || P.one_size) ss = nullptr;   // the rho-bar rows change after the fused passes
    if (ev) TMI_HIP_CHECK(hipEventRecord(ev[0], s));
    rb_mf_rows<double>(P, x + P.iu_rho * N, false, x + P.iu_rb * N, nullptr, F + P.iu_rb * N,
                       P.nmf ? x + P.iu_mf * N : nullptr, nullptr, F + P.iu_mf * N, s, ev ? ev[1] : nullptr,
                       ev ? ev[2] : nullptr, ss, &n_rb);
    hw_apply(P.hw_fwd, x + P.iu_rho * N, F + P.iu_rb * N, s);
    one_size_rows(P, x + P.iu_rb * N, F + P.iu_rb * N, s);
    if (ev) TMI_HIP_CHECK(hipEventRecord(ev[3], s));
    fmt_point<double>(P, x + P.iu_rb * N, nullptr, false, s);
    if (ev) TMI_HIP_CHECK(hipEventRecord(ev[4], s));
    el_rows<double>(P, false, x + P.iu_rho * N, x + P.iu_rho * N, P.nmf ? x + P.iu_mf * N : nullptr, nullptr,
                    F + P.iu_rho * N, s, ss ? ss + n_rb : nullptr, &n_el);
    if (ev) TMI_HIP_CHECK(hipEventRecord(ev[5], s));
    TMI_HIP_CHECK(hipGetLastError());
    if (nss) *nss = (n_rb > 0 && n_el > 0) ? n_rb + n_el : 0;
}
}  // namespace

void launch_residual(Problem &P, const double *x, double *F, hipStream_t s, hipEvent_t *ev)
{
    residual_impl(P, x, F, s, ev, nullptr, nullptr);
}

// ---------------------------------------------------------------- integrated observables
// Integrands of calc_free_energy (src/dft_energy.c:30-391) and calc_adsorption
// (src/dft_out_ads.c:33-110) at each node, weighted by the element counts as integrateInSpace
// (src/dft_out_utils.c:38-98) weights them: partial[v][blk], v = 0 ideal gas
// rho (ln rho - 1) (src/dft_energy_id.c:22-34), 1 chemical potential -rho beta mu
// (src/dft_energy_mu.c:22-45), 2 external field rho V_ext (src/dft_energy_vext.c:22-37), 3 Phi(n)
// of the rho-bar unknowns (src/dft_energy_hs.c:22-42), 4 attraction rho (u * rho) / 2
// (src/dft_energy_att.c:22-31, int_stencil src/dft_energy_CLSintsten.c:73-129), 5 + c adsorption
// of component c, 5 + C + c the sum of its Nel_hit counts (the bulk terms' integrals).  The first
// three count only rho > 1e-20 (DENSITY_MIN), the attraction rho > 0.
constexpr int kObsVals = 5 + 2 * kNC;
struct ObsConsts {
    double betamu[kNC];
};
template <int ND>
__global__ __launch_bounds__(kBlock) void k_observables(int N, int C, int func, int nwall, int att_on,
                                                        const double *__restrict__ x_rho,
                                                        const double *__restrict__ x_rb,
                                                        const double *__restrict__ mf,
                                                        const double *__restrict__ vext, const int *__restrict__ nh,
                                                        const int *__restrict__ nh2, ObsConsts oc,
                                                        double *__restrict__ partial)
{
    double acc[kObsVals];
#pragma unroll
    for (int v = 0; v < kObsVals; ++v) acc[v] = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < N; i += gridDim.x * kBlock) {
#pragma unroll
        for (int c = 0; c < kNC; ++c) {
            if (c >= C) break;
            const size_t row = (size_t)c * N + i;
            const double rho = x_rho[row], w2 = (double)nh2[row], w1 = (double)nh[row];
            if (rho > 1.e-20) {
                acc[0] += rho * (log(rho) - 1.0) * w2;
                acc[1] += -rho * oc.betamu[c] * w2;
                if (nwall > 0) acc[2] += rho * vext[row] * w2;
            }
            if (att_on && rho > 0.0) acc[4] += 0.5 * rho * mf[row] * w2;
            acc[5 + c] += rho * w2;
            acc[5 + kNC + c] += w1;
        }
        double n[4 + 2 * ND];   // n order from the solution order n3 n2 n1 n0 nv2[D] nv1[D]
        n[0] = x_rb[(size_t)3 * N + i];
        n[1] = x_rb[(size_t)2 * N + i];
        n[2] = x_rb[(size_t)1 * N + i];
        n[3] = x_rb[i];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            n[4 + d] = x_rb[(size_t)(4 + ND + d) * N + i];
            n[4 + ND + d] = x_rb[(size_t)(4 + d) * N + i];
        }
        acc[3] += fmt_phi<ND>(func, n) * (double)nh[i];
    }
    __shared__ double sh[kBlock / 64][kObsVals];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < kObsVals; ++v) {
        double a = acc[v];
        for (int o = 32; o > 0; o >>= 1) a += __shfl_down(a, o, 64);
        if (lane == 0) sh[wid][v] = a;
    }
    __syncthreads();
    if (threadIdx.x < kObsVals) {
        double a = 0.0;
        for (int q = 0; q < kBlock / 64; ++q) a += sh[q][threadIdx.x];
        partial[(size_t)threadIdx.x * gridDim.x + blockIdx.x] = a;
    }
}

void launch_mf_conv(Problem &P, const double *rho, bool linear, double *out, hipStream_t s)
{
    if (P.nmf == 0) return;
    pad_rho<double>(P, rho, linear, s);
    mf_rows<double>(P, nullptr, nullptr, out, s);
    hw_apply(P.hw_mf, rho, out, s);
}

void observables_sums(Problem &P, const double *x, const int *nh_d, const int *nh2_d, double *mf_tmp, double *sums,
                      hipStream_t s)
{
    if (P.npores > 0) throw std::invalid_argument("observables of a batch of pores are not supported");
    const size_t N = P.N;
    const int C = P.p.ncomp, nd = P.p.ndim;
    const bool att_on = P.p.type_attr >= 0 && P.nmf > 0;
    if (att_on)   // u * rho with the boundary values and hard-surface weights of the fill (int_stencil)
        launch_mf_conv(P, x + (size_t)P.iu_rho * N, false, mf_tmp, s);
    ObsConsts oc{};
    for (int c = 0; c < C; ++c)
        oc.betamu[c] = std::log(P.p.rho_b[c]) + P.bulk.betamu_hs_ex[c] + (P.p.type_attr >= 0 ? P.bulk.betamu_att[c] : 0.0);
    const unsigned nb = (unsigned)std::min<long long>(1024, ((long long)N + kBlock - 1) / kBlock);
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(nb), dim3(kBlock), 0, s, (int)N, C, P.p.type_func, P.p.nwall, att_on ? 1 : 0,
                           x + (size_t)P.iu_rho * N, x + (size_t)P.iu_rb * N, mf_tmp, P.d_vext, nh_d, nh2_d, oc,
                           P.d_red);
    };
    if (nd == 1)
        go(k_observables<1>);
    else if (nd == 2)
        go(k_observables<2>);
    else
        go(k_observables<3>);
    hipLaunchKernelGGL(k_reduce_rows, dim3(kObsVals), dim3(kBlock), 0, s, P.d_red, (int)nb, kObsVals, nullptr,
                       P.h_red_dev);
    TMI_HIP_CHECK(hipGetLastError());
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    for (int v = 0; v < kObsVals; ++v) sums[v] = P.h_red[v];
}

long long residual_sumsq_partials(const Problem &P)
{
    long long cap = 0;
    fft_ss_buffer(P, &cap);
    return std::max<long long>(cap, 1024);
}

void residual_sumsq_device(Problem &P, const double *x, double *F, double *partials, double *sumsq_d, hipStream_t s)
{
    long long cnt = 0;
    residual_impl(P, x, F, s, nullptr, P.fft ? partials : nullptr, &cnt);
    if (cnt > 0) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kSumBlock), 0, s, partials, (int)cnt, sumsq_d, nullptr);
    } else {   // the blocks and order of dev_dot
        const long long n = (long long)P.nunk * P.N;
        const unsigned nb = (unsigned)std::min<long long>(1024, (n + kBlock - 1) / kBlock);
        hipLaunchKernelGGL(k_multidot<double>, dim3(nb), dim3(kBlock), 0, s, F, 0LL, 1, F, n, partials);
        hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kBlock), 0, s, partials, (int)nb, 1, sumsq_d, nullptr);
    }
    TMI_HIP_CHECK(hipGetLastError());
}

double residual_sumsq(Problem &P, const double *x, double *F, hipStream_t s)
{
    long long cap = 0, cnt = 0;
    double *ss = fft_ss_buffer(P, &cap);
    residual_impl(P, x, F, s, nullptr, ss, &cnt);
    if (cnt > cap) throw std::logic_error("partial sums past their buffer");
    if (cnt > 0) {
        hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kSumBlock), 0, s, ss, (int)cnt, nullptr, P.h_red_dev);
        TMI_HIP_CHECK(hipStreamSynchronize(s));
        return P.h_red[0];
    }
    return dev_dot(P, F, F, (long long)P.nunk * P.N, s);
}

template <typename T>
void launch_jvp(Problem &P, const double *x, const double *v, double *Jv, hipStream_t s)
{
    const size_t N = P.N;
    rb_mf_rows<T>(P, v + P.iu_rho * N, true, v + P.iu_rb * N, nullptr, Jv + P.iu_rb * N,
                  P.nmf ? v + P.iu_mf * N : nullptr, nullptr, Jv + P.iu_mf * N, s);
    hw_apply(P.hw_fwd, v + P.iu_rho * N, Jv + P.iu_rb * N, s);
    one_size_rows(P, v + P.iu_rb * N, Jv + P.iu_rb * N, s);
    fmt_point<T>(P, x + P.iu_rb * N, v + P.iu_rb * N, true, s);
    el_rows<T>(P, true, x + P.iu_rho * N, v + P.iu_rho * N, P.nmf ? v + P.iu_mf * N : nullptr, nullptr, Jv + P.iu_rho * N,
               s);
    TMI_HIP_CHECK(hipGetLastError());
}
template void launch_jvp<double>(Problem &, const double *, const double *, double *, hipStream_t);
template void launch_jvp<float>(Problem &, const double *, const double *, double *, hipStream_t);

void launch_init_fields(Problem &P, double *x, hipStream_t s)
{
    const size_t N = P.N;
    rb_mf_rows<double>(P, x + P.iu_rho * N, false, nullptr, nullptr, x + P.iu_rb * N, nullptr, nullptr,
                       x + P.iu_mf * N, s);
    hipLaunchKernelGGL(k_cap_n3, dim3(grid_for(P.N)), dim3(kBlock), 0, s, x + P.iu_rb * N, P.N);
    TMI_HIP_CHECK(hipGetLastError());
}

// t1 = A12 x2 (rho-bar and mean-field rows of J [x2; 0]); y2 = A22 x2 + A21 t1
template <typename T>
void launch_schur_apply(Problem &P, const double *x, const double *x2, double *t1, double *y2, hipStream_t s)
{
    const size_t N = P.N;
    rb_mf_rows<T>(P, x2, true, nullptr, nullptr, t1, nullptr, nullptr, t1 + (size_t)P.nrb * N, s);
    hw_apply(P.hw_fwd, x2, t1, s);
    fmt_point<T>(P, x + P.iu_rb * N, t1, true, s);
    el_rows<T>(P, true, x + P.iu_rho * N, x2, P.nmf ? t1 + (size_t)P.nrb * N : nullptr, nullptr, y2, s);
    TMI_HIP_CHECK(hipGetLastError());
}
template void launch_schur_apply<double>(Problem &, const double *, const double *, double *, double *, hipStream_t);
template void launch_schur_apply<float>(Problem &, const double *, const double *, double *, double *, hipStream_t);

// bS = b2 + A21 b1  (A11 = -I, so -A21 A11^{-1} b1 = A21 b1)
void launch_schur_rhs(Problem &P, const double *x, const double *b, double *bS, hipStream_t s)
{
    const size_t N = P.N;
    fmt_point<double>(P, x + P.iu_rb * N, b + P.iu_rb * N, true, s);
    el_rows<double>(P, true, x + P.iu_rho * N, nullptr, P.nmf ? b + P.iu_mf * N : nullptr, b + P.iu_rho * N, bS, s);
    TMI_HIP_CHECK(hipGetLastError());
}

// x1 = A11^{-1}(b1 - A12 x2) = A12 x2 - b1
void launch_schur_backsub(Problem &P, const double *x, const double *b, double *dx, hipStream_t s)
{
    (void)x;
    const size_t N = P.N;
    rb_mf_rows<double>(P, dx + P.iu_rho * N, true, nullptr, b + P.iu_rb * N, dx + P.iu_rb * N, nullptr,
                       P.nmf ? b + P.iu_mf * N : nullptr, P.nmf ? dx + P.iu_mf * N : nullptr, s);
    hw_apply(P.hw_fwd, dx + P.iu_rho * N, dx + P.iu_rb * N, s);
    TMI_HIP_CHECK(hipGetLastError());
}

// ---------------------------------------------------------------- vector ops
static unsigned red_blocks(long long n) { return (unsigned)std::min<long long>(1024, (n + kBlock - 1) / kBlock); }

void dev_multi_dot(Problem &P, const double *V, long long ld, int k, const double *w, long long n, double *out,
                   hipStream_t s)
{
    unsigned nb = red_blocks(n);
    double *partial = P.d_red;            // [k][nb]
    hipLaunchKernelGGL(k_multidot<double>, dim3(nb), dim3(kBlock), 0, s, V, ld, k, w, n, partial);
    hipLaunchKernelGGL(k_reduce_rows, dim3(k), dim3(kBlock), 0, s, partial, (int)nb, k, nullptr, P.h_red_dev);
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    for (int j = 0; j < k; ++j) out[j] = P.h_red[j];
}

void dev_multi_dot_async(Problem &P, const double *V, long long ld, int k, const double *w, long long n, double *out_d,
                         hipStream_t s, double *mirror)
{
    unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_multidot<double>, dim3(nb), dim3(kBlock), 0, s, V, ld, k, w, n, P.d_red);
    hipLaunchKernelGGL(k_reduce_rows, dim3(k), dim3(kBlock), 0, s, P.d_red, (int)nb, k, out_d, mirror);
}

// Workgroups of the Gram-Schmidt passes: 2048 (the streams of V then run near the copy rate), fewer
// when k rows of partial sums would not fit below the coefficient slot of P.d_red (512 Ki doubles)
static unsigned gs_blocks(long long n, int k)
{
    const long long cap = (512LL * 1024) / std::max(k, 1);
    return (unsigned)std::max<long long>(1, std::min<long long>({2048, cap, (n + kBlock - 1) / kBlock}));
}

// classical Gram-Schmidt with one reorthogonalisation pass: h1 = V^T w, w -= V h1, h2 = V^T w,
// w -= V h2 with |w|^2, next = w / |w|; the coefficients and |w|^2 land in coef_d and in the
// pinned host mirror (read after the caller synchronises).  TV: the basis type (FP32 on the mixed
// path); w and every sum stay FP64.
template <typename TV>
void dev_cgs2(Problem &P, const TV *V, long long ld, int k, double *w, long long n, TV *next, double *coef_d,
              hipStream_t s)
{
    const unsigned nb = gs_blocks(n, k);
    double *h1 = coef_d, *h2 = coef_d + k, *nrm = coef_d + 2 * k;
    double *m1 = P.h_red_dev, *m2 = m1 + k, *mn = m1 + 2 * k;   // host copies (P.h_red)
    hipLaunchKernelGGL(k_multidot<TV>, dim3(nb), dim3(kBlock), 0, s, V, ld, k, w, n, P.d_red);
    hipLaunchKernelGGL(k_reduce_rows, dim3(k), dim3(kBlock), 0, s, P.d_red, (int)nb, k, h1, m1);
    hipLaunchKernelGGL((k_gemv_sub_norm<TV, false>), dim3(nb), dim3(kBlock), 0, s, V, ld, k, h1, w, n, nullptr);
    hipLaunchKernelGGL(k_multidot<TV>, dim3(nb), dim3(kBlock), 0, s, V, ld, k, w, n, P.d_red);
    hipLaunchKernelGGL(k_reduce_rows, dim3(k), dim3(kBlock), 0, s, P.d_red, (int)nb, k, h2, m2);
    hipLaunchKernelGGL((k_gemv_sub_norm<TV, true>), dim3(nb), dim3(kBlock), 0, s, V, ld, k, h2, w, n, P.d_red);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kBlock), 0, s, P.d_red, (int)nb, 1, nrm, mn);
    hipLaunchKernelGGL(k_scale_by_norm<TV>, dim3(grid_for(n)), dim3(kBlock), 0, s, nrm, w, next, n);
}
template void dev_cgs2<double>(Problem &, const double *, long long, int, double *, long long, double *, double *,
                               hipStream_t);
template void dev_cgs2<float>(Problem &, const float *, long long, int, double *, long long, float *, double *,
                              hipStream_t);

void dev_scale_by_norm(const double *nrm2_d, const double *x, double *y, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_by_norm<double>, dim3(grid_for(n)), dim3(kBlock), 0, s, nrm2_d, x, y, n);
}

template <typename TV>
void dev_cast_scaled(double a, const double *x, TV *y, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_cast_scaled<TV>, dim3(grid_for(n)), dim3(kBlock), 0, s, a, x, y, n);
}
template void dev_cast_scaled<double>(double, const double *, double *, long long, hipStream_t);
template void dev_cast_scaled<float>(double, const double *, float *, long long, hipStream_t);
void dev_widen(const float *a, double *b, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_widen<float>, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, n);
}

double dev_dot(Problem &P, const double *a, const double *b, long long n, hipStream_t s)
{
    double r;
    dev_multi_dot(P, a, 0, 1, b, n, &r, s);
    return r;
}

void dev_axpy(double a, const double *x, double *y, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(kBlock), 0, s, a, x, y, n);
}
static void one_size_launch(const Problem &P, int mode, const double *in, double *out, hipStream_t s)
{
    if (!P.one_size) return;
    hipLaunchKernelGGL(k_one_size, dim3(grid_for(P.N)), dim3(kBlock), 0, s, mode, (int)P.N, P.p.ndim,
                       P.bulk.inv_4pir[0], P.bulk.inv_4pirsq[0], in, out);
}
void one_size_rows(const Problem &P, const double *in_rb, double *out_rb, hipStream_t s)
{
    one_size_launch(P, 0, in_rb, out_rb, s);
}
void one_size_lift(const Problem &P, double *v_rb, hipStream_t s) { one_size_launch(P, 1, nullptr, v_rb, s); }
void one_size_transpose(const Problem &P, const double *w_rb, double *y_rb, hipStream_t s)
{
    one_size_launch(P, 2, w_rb, y_rb, s);
}
void hw_apply(const HwCsr &h, const double *in, double *out, hipStream_t s)
{
    if (h.nrows == 0) return;
    hipLaunchKernelGGL(k_hw_gather, dim3((h.nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, s, h.nrows, h.rowptr,
                       h.rowid, h.col, h.val, in, out);
}
void dev_fill(double v, double *y, long long n, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(grid_for(n)), dim3(kBlock), 0, s, v, y, n);
}
void dev_scale_copy(double a, const double *x, double *y, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_copy, dim3(grid_for(n)), dim3(kBlock), 0, s, a, x, y, n);
}
void dev_mul(const double *a, const double *b, double *y, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, y, n);
}
void dev_gemv_sub(const double *V, long long ld, int k, const double *h, double *w, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gemv_sub, dim3(grid_for(n)), dim3(kBlock), 0, s, V, ld, k, h, w, n);
}
template <typename TV>
void dev_gemv_add(const TV *V, long long ld, int k, const double *y, double *x, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_gemv_add<TV>, dim3(grid_for(n)), dim3(kBlock), 0, s, V, ld, k, y, x, n);
}
template void dev_gemv_add<double>(const double *, long long, int, const double *, double *, long long, hipStream_t);
template void dev_gemv_add<float>(const float *, long long, int, const double *, double *, long long, hipStream_t);
void dev_a22_inverse_diag(Problem &P, const double *x, double *dinv, hipStream_t s)
{
    long long n = (long long)P.p.ncomp * P.N;
    hipLaunchKernelGGL(k_a22_dinv, dim3(grid_for(n)), dim3(kBlock), 0, s, x + (size_t)P.iu_rho * P.N, P.d_elzero, dinv, n);
}

void dev_newton_update(Problem &P, double *x, const double *dx, const tmi_newton_opts &o, double *frac_out,
                       double *unorm_out, hipStream_t s)
{
    const long long n = (long long)P.nunk * P.N, n_rho = (long long)P.p.ncomp * P.N;
    unsigned nb = red_blocks(n_rho);
    hipLaunchKernelGGL(k_frac_min, dim3(nb), dim3(kBlock), 0, s, x + (size_t)P.iu_rho * P.N, dx + (size_t)P.iu_rho * P.N,
                       n_rho, o.update_scaling, P.d_red);
    TMI_HIP_CHECK(hipMemcpyAsync(P.h_red, P.d_red, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    double frac = 1.0;
    for (unsigned b = 0; b < nb; ++b) frac = std::min(frac, P.h_red[b]);
    unsigned nb2 = red_blocks(n);
    hipLaunchKernelGGL(k_update_norm, dim3(nb2), dim3(kBlock), 0, s, x, dx, n, frac, o.nl_rel_tol, o.nl_abs_tol, P.d_red);
    hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(kBlock), 0, s, P.d_red, (int)nb2, 1, nullptr, P.h_red_dev);
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    const double unorm = std::sqrt(P.h_red[0]);
    hipLaunchKernelGGL(k_apply_update, dim3(grid_for(n)), dim3(kBlock), 0, s, x, dx, n, n_rho, P.N, P.iu_rb, frac);
    TMI_HIP_CHECK(hipGetLastError());
    *frac_out = frac;
    *unorm_out = unorm;
}

}  // namespace tmi
