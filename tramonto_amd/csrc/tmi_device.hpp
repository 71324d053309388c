// Device-side helpers shared by the kernel files of libtramonto_mi355x.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "tmi_internal.hpp"

namespace tmi {
namespace dev {

constexpr int kBlock = 256;

// grid + one halo-padded geometry, passed by value to kernels
struct Geo {
    int nd, C, N;
    int nx[3];
    int bc[3][2];
    int h[3], P[3];
    long long np;
    const int *node_pad;   // batched problems: node -> padded position (else null)
    const int *pad_node;   // batched problems: padded position -> node or -1 (else null)
    const double *mu;      // batched problems: [C][N] chemical-potential term (else null)
};

inline Geo make_geo(const Problem &P, const PadGeom &pg)
{
    Geo g;
    g.nd = P.p.ndim;
    g.C = P.p.ncomp;
    g.N = P.N;
    for (int d = 0; d < 3; ++d) {
        g.nx[d] = (d < g.nd) ? P.p.nx[d] : 1;
        g.bc[d][0] = P.p.bc[d][0];
        g.bc[d][1] = P.p.bc[d][1];
        g.h[d] = pg.h[d];
        g.P[d] = pg.P[d];
    }
    g.np = pg.n;
    g.node_pad = pg.d_node_pad;
    g.pad_node = pg.d_pad_node;
    g.mu = P.d_mu;
    return g;
}

struct CompConsts {
    double i4pir[kNC], i4pirsq[kNC], rho_b[kNC], mu[kNC];
    double Ab[kNC], S3b;    // bulk adjoint-field values
};

inline CompConsts make_consts(const Problem &P)
{
    CompConsts cc{};
    const auto &b = P.bulk;
    for (int c = 0; c < P.p.ncomp; ++c) {
        cc.i4pir[c] = b.inv_4pir[c];
        cc.i4pirsq[c] = b.inv_4pirsq[c];
        cc.rho_b[c] = P.p.rho_b[c];
        double mu = 0.0;
        mu -= std::log(P.p.rho_b[c]);
        mu -= b.betamu_hs_ex[c];
        if (P.p.type_attr >= 0) mu -= b.betamu_att[c];
        cc.mu[c] = mu;
        cc.Ab[c] = b.dphi_b[0] * b.inv_4pirsq[c] + b.dphi_b[1] * b.inv_4pir[c] + b.dphi_b[2];
    }
    cc.S3b = b.dphi_b[3];
    return cc;
}

inline unsigned grid_for(long long n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// padded position of node i
__device__ inline long long pad_pos(const Geo &g, int i)
{
    if (g.node_pad) return g.node_pad[i];
    int x = i % g.nx[0];
    int r = i / g.nx[0];
    int y = r % g.nx[1];
    int z = r / g.nx[1];
    return (long long)(x + g.h[0]) + (long long)(y + g.h[1]) * g.P[0] + (long long)(z + g.h[2]) * g.P[0] * g.P[1];
}

// map a padded position to (code, source node, reflect mask); code 0: node,
// -1: bulk, -2: wall.  Dimension order and codes as offset_to_node_box
// (src/dft_mesh_lib.c:203-297).
__device__ inline int map_pad(const Geo &g, long long pos, int &src, int &refl)
{
    refl = 0;
    if (g.pad_node) {   // batched: the gaps between pores are walls
        src = g.pad_node[pos];
        return src >= 0 ? 0 : -2;
    }
    int c[3];
    c[0] = (int)(pos % g.P[0]);
    long long r = pos / g.P[0];
    c[1] = (int)(r % g.P[1]);
    c[2] = (int)(r / g.P[1]);
    refl = 0;
    int s[3];
    for (int d = 0; d < 3; ++d) {
        int v = c[d] - g.h[d];
        const int n = g.nx[d];
        if (d < g.nd) {
            if (v < 0) {
                const int b = g.bc[d][0];
                if (b == TMI_IN_BULK) return -1;
                if (b == TMI_IN_WALL) return -2;
                if (b == TMI_PERIODIC) {
                    v += n;
                } else {
                    refl |= 1 << d;
                    v = -v;
                    if (v >= n) {
                        if (g.bc[d][1] == TMI_IN_BULK) return -1;
                        return -2;
                    }
                }
            } else if (v >= n) {
                const int b = g.bc[d][1];
                if (b == TMI_IN_BULK) return -1;
                if (b == TMI_IN_WALL) return -2;
                if (b == TMI_PERIODIC) {
                    v -= n;
                } else {
                    refl |= 1 << d;
                    v = 2 * (n - 1) - v;
                    if (v < 0) {
                        if (g.bc[d][0] == TMI_IN_BULK) return -1;
                        return -2;
                    }
                }
            }
        }
        s[d] = v;
    }
    src = s[0] + s[1] * g.nx[0] + s[2] * g.nx[0] * g.nx[1];
    return 0;
}

struct StenPtrs {
    const int *lo;
    const void *w;
    const void *wv;
    int L;
};

template <typename T>
inline const DevStencilT<T> &dev_sten(const Problem &P, int type, int idx);
template <>
inline const DevStencilT<double> &dev_sten<double>(const Problem &P, int type, int idx)
{
    return P.dsd[type][idx];
}
template <>
inline const DevStencilT<float> &dev_sten<float>(const Problem &P, int type, int idx)
{
    return P.dsf[type][idx];
}

template <typename T>
inline const DevColumnsT<T> &dev_cols(const Problem &P, int type, int idx);
template <>
inline const DevColumnsT<double> &dev_cols<double>(const Problem &P, int type, int idx)
{
    return P.cold[type][idx];
}
template <>
inline const DevColumnsT<float> &dev_cols<float>(const Problem &P, int type, int idx)
{
    return P.colf[type][idx];
}

template <typename T>
inline const DevRow1DT<T> &dev_row(const Problem &P, int type, int idx);
template <>
inline const DevRow1DT<double> &dev_row<double>(const Problem &P, int type, int idx)
{
    return P.rowd[type][idx];
}
template <>
inline const DevRow1DT<float> &dev_row<float>(const Problem &P, int type, int idx)
{
    return P.rowf[type][idx];
}

template <typename T>
inline Workspace<T> &workspace(Problem &P);
template <>
inline Workspace<double> &workspace<double>(Problem &P)
{
    return P.wsd;
}
template <>
inline Workspace<float> &workspace<float>(Problem &P)
{
    return P.wsf;
}

// linear-offset tables of a stencil for one padded geometry
enum GeomSel { kGeomU = 0, kGeomF = 1, kGeomU2 = 2, kGeomF2 = 3 };

template <typename T>
inline StenPtrs sten_ptrs(const Problem &P, int type, int idx, int geom)
{
    const DevStencilT<T> &s = dev_sten<T>(P, type, idx);
    StenPtrs sp;
    sp.L = s.L;
    sp.lo = s.lo + (size_t)geom * s.L;
    sp.w = s.w;
    sp.wv = s.wv;
    return sp;
}

}  // namespace dev
}  // namespace tmi
