// Hard surfaces (Lhard_surf): the weights of the rho-bar rows at wall-fluid boundary nodes.
//
// With hard surfaces the reference weights a stencil point that lands on a wall-fluid boundary node
// by the parts of its weight that come from fluid elements only (HW_boundary_weight,
// src/dft_utils.c:41-62, applied in resid_and_Jac_sten_fill_sum_Ncomp, src/dft_fill_CLSmf.c:140-160;
// the parts are the HW_Weight of merge_load_stencil, src/dft_stencil.c:597-646).  The device
// kernels sum every stencil with its full weights; the corrections built here add
// fac * (w_fluid - w) * rho(j) to each rho-bar row whose stencil reaches a boundary node j, as a
// sparse matrix over the density columns (and its transpose for the adjoint product).  The
// caller gives, per component and node, 0 for nodes that are not boundary nodes of the
// component's element list and otherwise the bits of the local nodes whose element is fluid.
#include <algorithm>
#include <stdexcept>
#include <tuple>
#include <vector>

#include "tmi_internal.hpp"

namespace tmi {
namespace {

// node reached from ijk by off with the boundary rules of offset_to_node_box
// (src/dft_mesh_lib.c:203-297); -1 bulk, -2 wall.  refl (optional) receives the reflect flags as
// bits, bit d when the point was mirrored in dimension d.
int node_of(const tmi_params &p, const int *ijk, const int *off, unsigned *refl = nullptr)
{
    int s[3] = {0, 0, 0};
    if (refl) *refl = 0;
    for (int d = 0; d < p.ndim; ++d) {
        int v = ijk[d] + off[d];
        const int n = p.nx[d];
        if (v < 0) {
            const int b = p.bc[d][0];
            if (b == TMI_IN_BULK) return -1;
            if (b == TMI_IN_WALL) return -2;
            if (b == TMI_PERIODIC) {
                v += n;
            } else {
                v = -v;
                if (refl) *refl |= 1u << d;
                if (v >= n) return p.bc[d][1] == TMI_IN_BULK ? -1 : -2;
            }
        } else if (v >= n) {
            const int b = p.bc[d][1];
            if (b == TMI_IN_BULK) return -1;
            if (b == TMI_IN_WALL) return -2;
            if (b == TMI_PERIODIC) {
                v -= n;
            } else {
                v = 2 * (n - 1) - v;
                if (refl) *refl |= 1u << d;
                if (v < 0) return p.bc[d][0] == TMI_IN_BULK ? -1 : -2;
            }
        }
        s[d] = v;
    }
    return s[0] + s[1] * p.nx[0] + s[2] * p.nx[0] * p.nx[1];
}

// prefactor_rho_bar_s / _v (src/dft_fill_FMT.c:242-252, 330-347) of rho-bar unknown q (solution
// order n3 n2 n1 n0 nv2[D] nv1[D]) for component c at stencil offset off
double rb_prefactor(const Problem &P, int q, int c, const int *off)
{
    const int nd = P.p.ndim;
    if (q < 4) return q <= 1 ? 1.0 : (q == 2 ? P.bulk.inv_4pir[c] : P.bulk.inv_4pirsq[c]);
    const int idim = q < 4 + nd ? q - 4 : q - 4 - nd;
    double fac = q < 4 + nd ? 1.0 : P.bulk.inv_4pir[c];
    return fac * ((double)off[idim] * P.p.dx[idim] * P.bulk.inv_rad[c]);
}

// HW_boundary_weight (src/dft_utils.c:41-62) of stencil point k at a boundary node with fluid-element
// bits m: the parts hw[ln] of the local nodes ln whose element is fluid.  The element is
// node_box_to_elem_box_reflect's (src/dft_mesh_lib.c:896-937): for a point mirrored in dimension d it
// lies on the mirror side, local node ln with bit d flipped in the node's own numbering.
double fluid_part(const HostStencil &hs, int k, unsigned m, unsigned flip, int nd)
{
    double wf = 0.0;
    for (int ln = 0; ln < (1 << nd); ++ln)
        if (m & (1u << (ln ^ flip))) wf += hs.hw[8 * (size_t)k + ln];
    return wf;
}

template <typename T>
T *to_device(const std::vector<T> &v)
{
    T *d = nullptr;
    TMI_HIP_CHECK(hipMalloc(&d, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) TMI_HIP_CHECK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

// rows sorted by row id: (row, col, val) triplets to a device CSR
HwCsr make_csr(std::vector<std::tuple<int, int, double>> &t)
{
    std::stable_sort(t.begin(), t.end(), [](const auto &a, const auto &b) { return std::get<0>(a) < std::get<0>(b); });
    std::vector<int> rowptr{0}, rowid, col;
    std::vector<double> val;
    for (size_t k = 0; k < t.size(); ++k) {
        if (k == 0 || std::get<0>(t[k]) != std::get<0>(t[k - 1])) {
            if (k) rowptr.push_back((int)col.size());
            rowid.push_back(std::get<0>(t[k]));
        }
        col.push_back(std::get<1>(t[k]));
        val.push_back(std::get<2>(t[k]));
    }
    rowptr.push_back((int)col.size());
    HwCsr h;
    h.nrows = (int)rowid.size();
    h.rowptr = to_device(rowptr);
    h.rowid = to_device(rowid);
    h.col = to_device(col);
    h.val = to_device(val);
    return h;
}

void free_csr(HwCsr &h)
{
    (void)hipFree(h.rowptr);
    (void)hipFree(h.rowid);
    (void)hipFree(h.col);
    (void)hipFree(h.val);
    h = HwCsr();
}

}  // namespace

void hw_free(Problem &P)
{
    free_csr(P.hw_fwd);
    free_csr(P.hw_tr);
    free_csr(P.hw_mf);
}

void hw_build(Problem &P, const unsigned char *mask)
{
    hw_free(P);
    if (!mask) return;
    const tmi_params &p = P.p;
    const int C = p.ncomp, nd = p.ndim, N = P.N;
    if (P.npores > 0) throw std::invalid_argument("hard surfaces in a batch of pores are not supported");
    const unsigned full = (1u << (1 << nd)) - 1u;
    bool any = false;
    for (size_t k = 0; k < (size_t)C * N; ++k) {
        if (mask[k] & ~full) throw std::invalid_argument("hard-surface mask bits past the element's local nodes");
        any = any || mask[k];
    }
    if (!any) return;
    // a row can reach a boundary node only if, in every dimension, its coordinate reaches one of
    // the boundary nodes' coordinates (a conservative filter before the exact stencil walk)
    int reach[3] = {0, 0, 0};
    for (int t = TMI_DELTA; t <= TMI_UATT; ++t)
        for (int c = 0; c < (t == TMI_UATT ? (P.nmf ? C * C : 0) : C); ++c) {
            const HostStencil &hs = P.hs[t][c];
            for (int k = 0; k < hs.L; ++k)
                for (int d = 0; d < nd; ++d) reach[d] = std::max(reach[d], std::abs(hs.off[3 * k + d]));
        }
    std::vector<std::vector<char>> hit(nd), near(nd);
    for (int d = 0; d < nd; ++d) {
        hit[d].assign(p.nx[d], 0);
        near[d].assign(p.nx[d], 0);
    }
    for (int j = 0; j < N; ++j) {
        bool b = false;
        for (int c = 0; c < C; ++c) b = b || mask[(size_t)c * N + j];
        if (!b) continue;
        const int ijk[3] = {j % p.nx[0], (j / p.nx[0]) % p.nx[1], j / (p.nx[0] * p.nx[1])};
        for (int d = 0; d < nd; ++d) hit[d][ijk[d]] = 1;
    }
    for (int d = 0; d < nd; ++d)
        for (int i = 0; i < p.nx[d]; ++i)
            for (int o = -reach[d]; o <= reach[d] && !near[d][i]; ++o) {
                int ijk[3] = {0, 0, 0}, off[3] = {0, 0, 0};
                ijk[d] = i;
                off[d] = o;
                tmi_params q = p;
                q.ndim = d + 1;   // map along dimension d alone (lower dimensions offset 0)
                int j = node_of(q, ijk, off);
                if (j >= 0) {
                    int cd = j;
                    for (int e = 0; e < d; ++e) cd /= p.nx[e];
                    near[d][i] = hit[d][cd % p.nx[d]];
                }
            }
    std::vector<std::tuple<int, int, double>> fwd, tr, mf;
    for (int i = 0; i < N; ++i) {
        const int ijk[3] = {i % p.nx[0], (i / p.nx[0]) % p.nx[1], i / (p.nx[0] * p.nx[1])};
        bool cand = true;
        for (int d = 0; d < nd; ++d) cand = cand && near[d][ijk[d]];
        if (!cand) continue;
        for (int q = 0; q < P.nrb; ++q) {
            const int type = q == 0 ? TMI_THETA : TMI_DELTA;
            for (int c = 0; c < C; ++c) {
                const HostStencil &hs = P.hs[type][c];
                for (int k = 0; k < hs.L; ++k) {
                    const int *off = &hs.off[3 * k];
                    unsigned flip = 0;
                    const int j = node_of(p, ijk, off, &flip);
                    if (j < 0 || P.h_zero[(size_t)c * N + j]) continue;   // zero-density columns carry 0
                    const unsigned m = mask[(size_t)c * N + j];
                    if (!m) continue;
                    const double wf = fluid_part(hs, k, m, flip, nd);
                    const double v = rb_prefactor(P, q, c, off) * (wf - hs.w[k]);
                    fwd.emplace_back(q * N + i, c * N + j, v);
                    tr.emplace_back(c * N + j, q * N + i, v);
                }
            }
        }
        // mean-field rows (load_mean_field through the same stencil-sum fill), right after the
        // rho-bar block: rows (nrb + ic) N + i
        for (int ic = 0; ic < P.nmf; ++ic)
            for (int c = 0; c < C; ++c) {
                const HostStencil &hs = P.hs[TMI_UATT][ic + C * c];
                for (int k = 0; k < hs.L; ++k) {
                    const int *off = &hs.off[3 * k];
                    unsigned flip = 0;
                    const int j = node_of(p, ijk, off, &flip);
                    if (j < 0 || P.h_zero[(size_t)c * N + j]) continue;
                    const unsigned m = mask[(size_t)c * N + j];
                    if (!m) continue;
                    const double wf = fluid_part(hs, k, m, flip, nd);
                    const double v = wf - hs.w[k];
                    fwd.emplace_back((P.nrb + ic) * N + i, c * N + j, v);
                    tr.emplace_back(c * N + j, (P.nrb + ic) * N + i, v);
                    mf.emplace_back(ic * N + i, c * N + j, v);
                }
            }
    }
    P.hw_fwd = make_csr(fwd);
    P.hw_tr = make_csr(tr);
    if (!mf.empty()) P.hw_mf = make_csr(mf);
}

}  // namespace tmi
