// C ABI of libtramonto_mi355x: problem lifecycle and the field-level entry
// points (fill, Jacobian products, Schur solve, Newton).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <cstdlib>
#include <map>
#include <type_traits>
#include <vector>

#include "tmi_fmt.hpp"
#include "tmi_internal.hpp"

namespace tmi {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

namespace {

template <typename F>
int guarded(F &&fn)
{
    try {
        return fn();
    } catch (const HipFailure &) {
        return TMI_EHIP;
    } catch (const std::invalid_argument &e) {
        set_error(e.what());
        return TMI_EINVAL;
    } catch (const std::domain_error &e) {
        set_error(e.what());
        return TMI_EUNSUP;
    } catch (const std::exception &e) {
        set_error(e.what());
        return TMI_EINVAL;
    }
}

template <typename T>
T *dev_alloc(size_t n)
{
    void *p = nullptr;
    if (n == 0) n = 1;
    TMI_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    return (T *)p;
}

template <typename T>
void upload(T *dst, const T *src, size_t n)
{
    if (n) TMI_HIP_CHECK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
}

void validate(const tmi_params &p)
{
    if (p.ndim < 1 || p.ndim > 3) throw std::invalid_argument("ndim must be 1, 2 or 3");
    if (p.ncomp < 1 || p.ncomp > TMI_NCOMP_MAX) throw std::invalid_argument("ncomp must be 1..4");
    if (p.type_func < TMI_FMT1 || p.type_func > TMI_FMT4)
        throw std::domain_error("type_func: 0 (FMT1), 1 (FMT2), 2 (FMT3) or 3 (FMT4)");
    if (p.type_attr != -1 && p.type_attr != 0) throw std::domain_error("type_attr: only -1 and 0 are implemented");

    for (int d = 0; d < p.ndim; ++d) {
        if (p.nx[d] < 1) throw std::invalid_argument("nx must be positive");
        if (!(p.dx[d] > 0.0)) throw std::invalid_argument("dx must be positive");
        for (int s = 0; s < 2; ++s) {
            int b = p.bc[d][s];
            if (b != TMI_IN_WALL && b != TMI_IN_BULK && b != TMI_PERIODIC && b != TMI_REFLECT)
                throw std::invalid_argument("bc code");
        }
        if ((p.bc[d][0] == TMI_PERIODIC) != (p.bc[d][1] == TMI_PERIODIC))
            throw std::invalid_argument("periodic boundaries must be paired");
    }
    for (int i = 0; i < p.ncomp; ++i) {
        if (!(p.hs_diam[i] > 0.0)) throw std::invalid_argument("hs_diam must be positive");
        if (!(p.rho_b[i] > 0.0)) throw std::invalid_argument("rho_b must be positive");
    }
}

void set_geometry(Problem &P, PadGeom &pg, const int *halo, int scale = 1)
{
    pg.n = 1;
    for (int d = 0; d < 3; ++d) {
        int n = (d < P.p.ndim) ? P.p.nx[d] : 1;
        pg.h[d] = (d < P.p.ndim) ? scale * halo[d] : 0;
        // periodic images are read from one neighbouring period only
        if (scale == 1 && d < P.p.ndim && P.p.bc[d][0] == TMI_PERIODIC && pg.h[d] > n)
            throw std::domain_error("stencil reach exceeds the periodic domain");
        pg.P[d] = n + 2 * pg.h[d];
        pg.n *= pg.P[d];
    }
}

template <typename T>
void upload_stencil(Problem &P, int type, int idx, int comp, DevStencilT<T> &ds)
{
    const HostStencil &hs = P.hs[type][idx];
    const int L = hs.L, nd = P.p.ndim;
    ds.L = L;
    std::vector<int> lo(4 * (size_t)L);
    std::vector<T> w(L), wv((size_t)L * nd);
    const PadGeom *geoms[4] = {&P.pad_u, &P.pad_f, &P.pad_u2, &P.pad_f2};
    for (int k = 0; k < L; ++k) {
        const int *o = &hs.off[3 * (size_t)k];
        for (int g = 0; g < 4; ++g) {
            const PadGeom &pg = *geoms[g];
            long long v = o[0] + (long long)o[1] * pg.P[0] + (long long)o[2] * pg.P[0] * pg.P[1];
            if (v > INT32_MAX || v < INT32_MIN) throw std::domain_error("grid too large for 32-bit stencil offsets");
            lo[(size_t)g * L + k] = (int)v;
        }
        w[k] = (T)hs.w[k];
        for (int d = 0; d < nd; ++d) wv[(size_t)k * nd + d] = (T)(hs.w[k] * ((double)o[d] * P.p.dx[d] * P.bulk.inv_rad[comp]));
    }
    ds.lo = dev_alloc<int>(lo.size());
    ds.w = dev_alloc<T>(w.size());
    ds.wv = dev_alloc<T>(wv.size());
    upload(ds.lo, lo.data(), lo.size());
    upload(ds.w, w.data(), w.size());
    upload(ds.wv, wv.data(), wv.size());
}

// z-column form of a 3D stencil (see DevColumnsT)
template <typename T>
void upload_columns(Problem &P, int type, int idx, int comp, DevColumnsT<T> &dc)
{
    const HostStencil &hs = P.hs[type][idx];
    std::map<std::pair<int, int>, std::vector<std::pair<int, double>>> cols;
    for (int k = 0; k < hs.L; ++k) {
        const int *o = &hs.off[3 * (size_t)k];
        cols[{o[1], o[0]}].push_back({o[2], hs.w[k]});
    }
    const int B = kColB;
    const bool delta = (type == TMI_DELTA);
    const double inv_r = P.bulk.inv_rad[comp];
    const PadGeom *geoms[2] = {&P.pad_u, &P.pad_f};
    std::vector<int> lo[2], woff, nchunk;
    std::vector<T> wpad, wzpad, fxy;
    for (auto &kv : cols) {
        int z0 = 1 << 30, z1 = -(1 << 30);
        for (auto &e : kv.second) {
            z0 = std::min(z0, e.first);
            z1 = std::max(z1, e.first);
        }
        const int len = z1 - z0 + 1;
        const int nk = (len + B - 1 + B - 1) / B;
        const int dy = kv.first.first, dx = kv.first.second;
        for (int g = 0; g < 2; ++g) {
            const PadGeom &pg = *geoms[g];
            lo[g].push_back(dx + dy * pg.P[0] + z0 * pg.P[0] * pg.P[1]);
        }
        if (wpad.size() > INT32_MAX) throw std::domain_error("stencil too large for 32-bit column offsets");
        woff.push_back((int)wpad.size());
        nchunk.push_back(nk);
        std::vector<T> w((size_t)nk * B + B - 1, (T)0), wz(w.size(), (T)0);
        for (auto &e : kv.second) {
            w[(size_t)(B - 1) + (e.first - z0)] = (T)e.second;
            wz[(size_t)(B - 1) + (e.first - z0)] = (T)(e.second * ((double)e.first * P.p.dx[2] * inv_r));
        }
        wpad.insert(wpad.end(), w.begin(), w.end());
        if (delta) {
            wzpad.insert(wzpad.end(), wz.begin(), wz.end());
            fxy.push_back((T)((double)dx * P.p.dx[0] * inv_r));
            fxy.push_back((T)((double)dy * P.p.dx[1] * inv_r));
        }
    }
    dc.ncol = (int)woff.size();
    dc.nw = (long long)wpad.size();
    if (delta) wpad.insert(wpad.end(), wzpad.begin(), wzpad.end());
    std::vector<int> lo2(lo[0]);
    lo2.insert(lo2.end(), lo[1].begin(), lo[1].end());
    dc.lo = dev_alloc<int>(lo2.size());
    dc.woff = dev_alloc<int>(woff.size());
    dc.nchunk = dev_alloc<int>(nchunk.size());
    dc.wpad = dev_alloc<T>(wpad.size());
    upload(dc.lo, lo2.data(), lo2.size());
    upload(dc.woff, woff.data(), woff.size());
    upload(dc.nchunk, nchunk.data(), nchunk.size());
    upload(dc.wpad, wpad.data(), wpad.size());
    if (delta) {
        dc.fxy = dev_alloc<T>(fxy.size());
        upload(dc.fxy, fxy.data(), fxy.size());
    }
}

// dense row form of a 1D stencil (see DevRow1DT)
template <typename T>
void upload_row1d(Problem &P, int type, int idx, int comp, DevRow1DT<T> &dr)
{
    const HostStencil &hs = P.hs[type][idx];
    if (hs.L == 0) return;
    int lo = 1 << 30, hi = -(1 << 30);
    for (int k = 0; k < hs.L; ++k) {
        lo = std::min(lo, hs.off[3 * (size_t)k]);
        hi = std::max(hi, hs.off[3 * (size_t)k]);
    }
    const int L = hi - lo + 1;
    std::vector<T> w((size_t)L + 5 * kRowPad, (T)0), wv(w.size(), (T)0);
    for (int k = 0; k < hs.L; ++k) {
        const int o = hs.off[3 * (size_t)k];
        w[(size_t)kRowPad + (o - lo)] = (T)hs.w[k];
        wv[(size_t)kRowPad + (o - lo)] = (T)(hs.w[k] * ((double)o * P.p.dx[0] * P.bulk.inv_rad[comp]));
    }
    dr.lo0 = lo;
    dr.L = L;
    dr.w = dev_alloc<T>(w.size());
    upload(dr.w, w.data(), w.size());
    if (type == TMI_DELTA) {
        dr.wv = dev_alloc<T>(wv.size());
        upload(dr.wv, wv.data(), wv.size());
    }
}

template <typename T>
void free_row1d(DevRow1DT<T> &dr)
{
    (void)hipFree(dr.w);
    (void)hipFree(dr.wv);
    dr = DevRow1DT<T>();
}

template <typename T>
void free_columns(DevColumnsT<T> &dc)
{
    (void)hipFree(dc.lo);
    (void)hipFree(dc.woff);
    (void)hipFree(dc.nchunk);
    (void)hipFree(dc.wpad);
    (void)hipFree(dc.fxy);
    dc = DevColumnsT<T>();
}

template <typename T>
void free_stencil(DevStencilT<T> &ds)
{
    (void)hipFree(ds.lo);
    (void)hipFree(ds.w);
    (void)hipFree(ds.wv);
    ds = DevStencilT<T>();
}

// attraction stencils with at least this many points in all (summed over the
// component pairs) use the FFT form of the mean-field sums
constexpr long long kFftMinStencil = 2000;

// Pores of a batched 1D problem: node counts and bulk densities [npores][C].
struct BatchLayout {
    int npores = 0;
    const int *nodes = nullptr;
    const double *rho_b = nullptr;
};

// The padded copy of a batched problem: pores back to back, each preceded
// (and the last also followed) by a wall gap of h zero positions.
void batch_geometry(Problem &P, PadGeom &pg, int h)
{
    const long long n = (long long)P.N + (long long)(P.npores + 1) * h;
    if (n > INT32_MAX) throw std::domain_error("batch too large for 32-bit padded positions");
    pg.h[0] = h;
    pg.h[1] = pg.h[2] = 0;
    pg.P[0] = (int)n;
    pg.P[1] = pg.P[2] = 1;
    pg.n = n;
    std::vector<int> node_pad(P.N), pad_node((size_t)n, -1);
    for (int q = 0; q < P.npores; ++q)
        for (int i = P.pore_off[q]; i < P.pore_off[q + 1]; ++i) {
            node_pad[i] = i + (q + 1) * h;
            pad_node[(size_t)i + (size_t)(q + 1) * h] = i;
        }
    pg.d_node_pad = dev_alloc<int>(node_pad.size());
    pg.d_pad_node = dev_alloc<int>(pad_node.size());
    upload(pg.d_node_pad, node_pad.data(), node_pad.size());
    upload(pg.d_pad_node, pad_node.data(), pad_node.size());
}

// The padded copy of a batch of 2D pieces stacked along y (piece q: rows pore_off[q] / nx ..): every
// piece is preceded by hy wall rows and the last one followed by hy more; along x the halo of hx
// columns holds the periodic image (periodic x) or wall (walled x).  Stencil offsets keep the
// padded grid's row pitch nx + 2 hx.
void batch_geometry_2d(Problem &P, PadGeom &pg, int hx, int hy)
{
    const int nx = P.p.nx[0];
    const bool per = P.p.bc[0][0] == TMI_PERIODIC;
    if (per && hx > nx) throw std::domain_error("stencil reach exceeds the periodic domain");
    const long long P0 = nx + 2LL * hx, rows = (long long)P.p.nx[1] + (long long)(P.npores + 1) * hy;
    const long long n = P0 * rows;
    if (n > INT32_MAX) throw std::domain_error("batch too large for 32-bit padded positions");
    pg.h[0] = hx;
    pg.h[1] = hy;
    pg.h[2] = 0;
    pg.P[0] = (int)P0;
    pg.P[1] = (int)rows;
    pg.P[2] = 1;
    pg.n = n;
    std::vector<int> node_pad(P.N), pad_node((size_t)n, -1);
    for (int q = 0; q < P.npores; ++q) {
        const int r0 = P.pore_off[q] / nx, r1 = P.pore_off[q + 1] / nx;
        for (int r = r0; r < r1; ++r) {
            const long long prow = r + (long long)(q + 1) * hy;
            for (int xp = 0; xp < P0; ++xp) {
                int x = xp - hx;
                if (x < 0 || x >= nx) {
                    if (!per) continue;
                    x = ((x % nx) + nx) % nx;
                }
                const int node = x + r * nx;
                pad_node[(size_t)(xp + prow * P0)] = node;
                if (xp - hx == x) node_pad[node] = (int)(xp + prow * P0);
            }
        }
    }
    pg.d_node_pad = dev_alloc<int>(node_pad.size());
    pg.d_pad_node = dev_alloc<int>(pad_node.size());
    upload(pg.d_node_pad, node_pad.data(), node_pad.size());
    upload(pg.d_pad_node, pad_node.data(), pad_node.size());
}

void free_geometry_maps(PadGeom &pg)
{
    (void)hipFree(pg.d_node_pad);
    (void)hipFree(pg.d_pad_node);
    pg.d_node_pad = pg.d_pad_node = nullptr;
}

void destroy_problem(Problem *P);

// symbol_only: a bulk fluid built only for its Schur symbol (FFT forms taken without timing the
// direct ones)
Problem *create_problem(const tmi_params &p_in, const double *vext, const unsigned char *zero_tf,
                        const BatchLayout *batch = nullptr, bool symbol_only = false)
{
    tmi_params p = p_in;
    std::vector<int> pore_off;
    if (batch) {
        if (p.ndim != 1 && p.ndim != 2) throw std::invalid_argument("batched problems are 1D or 2D");
        if (batch->npores < 1 || !batch->nodes || !batch->rho_b) throw std::invalid_argument("batch layout");
        if (p.ndim == 2) {
            if (p.nx[0] < 1) throw std::invalid_argument("2D pieces need a common nx[0]");
            if (p.bc[0][0] != p.bc[0][1] || (p.bc[0][0] != TMI_PERIODIC && p.bc[0][0] != TMI_IN_WALL))
                throw std::domain_error("2D pieces of a batch are periodic or walled along x");
        }
        pore_off.assign(1, 0);
        for (int q = 0; q < batch->npores; ++q) {
            if (batch->nodes[q] < 1) throw std::invalid_argument("pore node count must be positive");
            if (p.ndim == 2 && batch->nodes[q] % p.nx[0]) throw std::invalid_argument("2D piece node count not a multiple of nx");
            if ((long long)pore_off.back() + batch->nodes[q] > INT32_MAX) throw std::domain_error("batch too large");
            pore_off.push_back(pore_off.back() + batch->nodes[q]);
        }
        if (p.ndim == 1) {
            p.nx[0] = pore_off.back();
            p.bc[0][0] = p.bc[0][1] = TMI_IN_WALL;
        } else {   // pieces stacked along y, each between walls in y
            p.nx[1] = pore_off.back() / p.nx[0];
            p.bc[1][0] = p.bc[1][1] = TMI_IN_WALL;
        }
        for (int c = 0; c < p.ncomp; ++c) p.rho_b[c] = batch->rho_b[c];
        for (int q = 0; q < batch->npores; ++q)
            for (int c = 0; c < p.ncomp; ++c)
                if (!(batch->rho_b[(size_t)q * p.ncomp + c] > 0.0)) throw std::invalid_argument("rho_b must be positive");
    }
    validate(p);
    auto P = new Problem();
    P->p = p;
    const int C = p.ncomp, nd = p.ndim;
    P->N = 1;
    for (int d = 0; d < nd; ++d) P->N *= p.nx[d];
    if (batch) {
        P->npores = batch->npores;
        P->pore_off = pore_off;
        P->pore_rho_b.assign(batch->rho_b, batch->rho_b + (size_t)batch->npores * C);
    }
    P->nrb = 4 + 2 * nd;
    P->nmf = (p.type_attr >= 0) ? C : 0;
    P->iu_rho = 0;
    P->iu_rb = C;
    P->iu_mf = C + P->nrb;
    P->nunk = C + P->nrb + P->nmf;
    P->ext_mf = p.att_in_a22 != 0 && P->nmf > 0;
    if (p.lhard_surf && P->npores > 0) throw std::domain_error("lhard_surf in a batch of pores is not supported");
    // Nlists_HW == 2 with hard surfaces (src/dft_mesh.c:330-345): one component and walls
    P->one_size = p.lhard_surf && p.nwall > 0 && p.ncomp == 1;
    if (const char *e = std::getenv("TMI_JAC_REFERENCE")) P->jac_ref = std::string(e) == "1";
    if (P->one_size) TMI_HIP_CHECK(hipMalloc((void **)&P->d_w0, sizeof(double) * (size_t)P->nunk * P->N));
    P->nunk_ext = P->ext_mf ? C + P->nrb : P->nunk;
    if (P->ext_mf && batch) throw std::domain_error("att_in_a22 is not supported for batches of pores");
    P->nadj = C * (1 + nd) + 1;
    compute_bulk_constants(p, P->bulk);
    if (p.type_attr >= 0)
        for (int i = 0; i < C; ++i)
            for (int j = 0; j < C; ++j)
                P->bulk.rmin[i][j] = lj_potential_minimum(p.sigma_ff[i][j], p.eps_ff[i][j], p.cut_ff[i][j], p.dx[0]);
    build_all_stencils(p, P->bulk, P->hs);
    if (p.type_attr >= 0) compute_attraction_constants(p, P->bulk, P->hs);

    int hu[3] = {0, 0, 0}, hf[3] = {0, 0, 0};
    for (int t = 0; t < 3; ++t)
        for (int q = 0; q < C * C; ++q) {
            const HostStencil &s = P->hs[t][q];
            for (int k = 0; k < s.L; ++k)
                for (int d = 0; d < nd; ++d) {
                    int a = std::abs(s.off[3 * (size_t)k + d]);
                    hu[d] = std::max(hu[d], a);
                    if (t != TMI_UATT) hf[d] = std::max(hf[d], a);
                }
        }
    if (!batch) {
        set_geometry(*P, P->pad_u, hu);
        set_geometry(*P, P->pad_f, hf);
    }
    set_geometry(*P, P->pad_u2, hu, 2);
    set_geometry(*P, P->pad_f2, hf, 2);

    // node-major inputs -> component-major device rows
    const size_t N = P->N;
    P->h_zero.assign((size_t)C * N, 0);
    P->h_elzero.assign((size_t)C * N, 0);
    P->h_vext.assign((size_t)C * N, 0.0);
    for (size_t i = 0; i < N; ++i)
        for (int c = 0; c < C; ++c) {
            unsigned char z = zero_tf ? zero_tf[i * C + c] : 0;
            double v = vext ? vext[i * C + c] : 0.0;
            P->h_zero[c * N + i] = z ? 1 : 0;
            P->h_vext[c * N + i] = v;
            P->h_elzero[c * N + i] = (z || (p.nwall > 0 && v == p.vext_max)) ? 1 : 0;
        }
    try {
        if (batch) {
            if (nd == 1) {
                batch_geometry(*P, P->pad_u, hu[0]);
                batch_geometry(*P, P->pad_f, hf[0]);
            } else {
                batch_geometry_2d(*P, P->pad_u, hu[0], hu[1]);
                batch_geometry_2d(*P, P->pad_f, hf[0], hf[1]);
            }
            std::vector<int> pore_of(N);
            for (int q = 0; q < P->npores; ++q)
                for (int i = P->pore_off[q]; i < P->pore_off[q + 1]; ++i) pore_of[i] = q;
            // per-pore chemical-potential term, as make_consts forms it for one problem
            std::vector<double> mu((size_t)C * N);
            for (int q = 0; q < P->npores; ++q) {
                tmi_params pq = p;
                for (int c = 0; c < C; ++c) pq.rho_b[c] = P->pore_rho_b[(size_t)q * C + c];
                BulkConstants bq;
                compute_bulk_constants(pq, bq);
                if (p.type_attr >= 0) compute_attraction_constants(pq, bq, P->hs);
                for (int c = 0; c < C; ++c) {
                    double m = -std::log(pq.rho_b[c]) - bq.betamu_hs_ex[c];
                    if (p.type_attr >= 0) m -= bq.betamu_att[c];
                    for (int i = P->pore_off[q]; i < P->pore_off[q + 1]; ++i) mu[(size_t)c * N + i] = m;
                }
            }
            P->d_pore_off = dev_alloc<int>(P->pore_off.size());
            P->d_pore_of = dev_alloc<int>(N);
            P->d_mu = dev_alloc<double>(mu.size());
            upload(P->d_pore_off, P->pore_off.data(), P->pore_off.size());
            upload(P->d_pore_of, pore_of.data(), pore_of.size());
            upload(P->d_mu, mu.data(), mu.size());
        }
        P->d_zero = dev_alloc<unsigned char>(P->h_zero.size());
        P->d_elzero = dev_alloc<unsigned char>(P->h_elzero.size());
        P->d_vext = dev_alloc<double>(P->h_vext.size());
        upload(P->d_zero, P->h_zero.data(), P->h_zero.size());
        upload(P->d_elzero, P->h_elzero.data(), P->h_elzero.size());
        upload(P->d_vext, P->h_vext.data(), P->h_vext.size());
        for (int i = 0; i < C; ++i) {
            upload_stencil(*P, TMI_DELTA, i, i, P->dsd[TMI_DELTA][i]);
            upload_stencil(*P, TMI_THETA, i, i, P->dsd[TMI_THETA][i]);
            upload_stencil(*P, TMI_DELTA, i, i, P->dsf[TMI_DELTA][i]);
            upload_stencil(*P, TMI_THETA, i, i, P->dsf[TMI_THETA][i]);
        }
        if (p.type_attr >= 0)
            for (int q = 0; q < C * C; ++q) {
                upload_stencil(*P, TMI_UATT, q, q % C, P->dsd[TMI_UATT][q]);
                upload_stencil(*P, TMI_UATT, q, q % C, P->dsf[TMI_UATT][q]);
            }
        // column kernels on 3D grids (TMI_STENCIL_KERNEL=gather keeps the per-point gather kernels
        // for comparison)
        const char *kern = std::getenv("TMI_STENCIL_KERNEL");
        // the column kernels address a column with 32-bit byte offsets: (2h + 1 + 2 kColB) planes
        // of the padded grid must span less than 4 GiB
        auto columns_fit = [&](const PadGeom &pg) {
            return (double)(2 * pg.h[2] + 1 + 2 * kColB) * pg.P[0] * pg.P[1] * sizeof(double) < 4294967296.0;
        };
        if (nd == 3 && !(kern && std::string(kern) == "gather") && columns_fit(P->pad_u) && columns_fit(P->pad_f)) {
            for (int c = 0; c < C; ++c)
                for (int t : {TMI_DELTA, TMI_THETA}) {
                    upload_columns(*P, t, c, c, P->cold[t][c]);
                    upload_columns(*P, t, c, c, P->colf[t][c]);
                }
            if (p.type_attr >= 0)
                for (int q = 0; q < C * C; ++q) {
                    upload_columns(*P, TMI_UATT, q, q % C, P->cold[TMI_UATT][q]);
                    upload_columns(*P, TMI_UATT, q, q % C, P->colf[TMI_UATT][q]);
                }
            P->use_columns = true;
        }
        // dense 2D convolution weights of the attraction sums (TMI_STENCIL_KERNEL=gather keeps the
        // per-point gather kernel)
        if (nd == 2 && p.type_attr >= 0 && !(kern && std::string(kern) == "gather")) {
            const int hx = P->pad_u.h[0], hy = P->pad_u.h[1], wx = 2 * hx + 1, wy = 2 * hy + 1;
            std::vector<double> wd((size_t)C * C * wx * wy, 0.0);
            bool fits = true;
            for (int q = 0; q < C * C; ++q) {
                const HostStencil &hs = P->hs[TMI_UATT][q];
                for (int k = 0; k < hs.L; ++k) {
                    const int dx = hs.off[3 * (size_t)k], dy = hs.off[3 * (size_t)k + 1];
                    if (dx < -hx || dx > hx || dy < -hy || dy > hy || hs.off[3 * (size_t)k + 2] != 0) {
                        fits = false;
                        break;
                    }
                    wd[(size_t)q * wx * wy + (size_t)(dy + hy) * wx + (dx + hx)] += hs.w[k];
                }
            }
            if (fits) {
                std::vector<float> wf(wd.begin(), wd.end());
                P->d_mfw2 = dev_alloc<double>(wd.size());
                P->f_mfw2 = dev_alloc<float>(wf.size());
                upload(P->d_mfw2, wd.data(), wd.size());
                upload(P->f_mfw2, wf.data(), wf.size());
                const char *mm = std::getenv("TMI_MF2D_MFMA");
                P->mf2d_mfma = !(mm && std::string(mm) == "0");
            }
        }
        // dense 2D weights of the rho-bar sums of one component for the matrix-core kernel: theta,
        // delta and the delta vector weights w (o_d dx_d / R) along x and y, as the gather kernel's
        // w and wv (TMI_STENCIL_KERNEL=gather keeps the gather kernel)
        if (nd == 2 && C == 1 && !P->use_columns && !(kern && std::string(kern) == "gather")) {
            const HostStencil &th = P->hs[TMI_THETA][0], &de = P->hs[TMI_DELTA][0];
            int fx = 0, fy = 0;
            bool flat = true;
            for (const HostStencil *hs : {&th, &de})
                for (int k = 0; k < hs->L; ++k) {
                    fx = std::max(fx, std::abs(hs->off[3 * (size_t)k]));
                    fy = std::max(fy, std::abs(hs->off[3 * (size_t)k + 1]));
                    flat = flat && hs->off[3 * (size_t)k + 2] == 0;
                }
            if (flat && fx <= P->pad_u.h[0] && fy <= P->pad_u.h[1]) {
                const int wx = 2 * fx + 1, wy = 2 * fy + 1;
                std::vector<double> wr((size_t)4 * wx * wy, 0.0);
                auto at = [&](int t, const int *o) -> double & {
                    return wr[((size_t)t * wy + (o[1] + fy)) * wx + (o[0] + fx)];
                };
                for (int k = 0; k < th.L; ++k) at(0, &th.off[3 * (size_t)k]) += th.w[k];
                for (int k = 0; k < de.L; ++k) {
                    const int *o = &de.off[3 * (size_t)k];
                    at(1, o) += de.w[k];
                    for (int d = 0; d < 2; ++d)
                        at(2 + d, o) += de.w[k] * ((double)o[d] * p.dx[d] * P->bulk.inv_rad[0]);
                }
                std::vector<float> wrf(wr.begin(), wr.end());
                P->d_rbw2 = dev_alloc<double>(wr.size());
                P->f_rbw2 = dev_alloc<float>(wrf.size());
                upload(P->d_rbw2, wr.data(), wr.size());
                upload(P->f_rbw2, wrf.data(), wrf.size());
                P->rbw_h[0] = fx;
                P->rbw_h[1] = fy;
                P->rbw_el = fx <= P->pad_f.h[0] && fy <= P->pad_f.h[1];
                const char *rm = std::getenv("TMI_RB2D_MFMA");
                P->rb2d_mfma = !(rm && std::string(rm) == "0");
            }
        }
        // dense row kernels on 1D grids (TMI_STENCIL_KERNEL=gather keeps the gather kernels)
        if (nd == 1 && !(kern && std::string(kern) == "gather")) {
            for (int c = 0; c < C; ++c)
                for (int t : {TMI_DELTA, TMI_THETA}) {
                    upload_row1d(*P, t, c, c, P->rowd[t][c]);
                    upload_row1d(*P, t, c, c, P->rowf[t][c]);
                }
            if (p.type_attr >= 0)
                for (int q = 0; q < C * C; ++q) {
                    upload_row1d(*P, TMI_UATT, q, q % C, P->rowd[TMI_UATT][q]);
                    upload_row1d(*P, TMI_UATT, q, q % C, P->rowf[TMI_UATT][q]);
                }
            P->use_rows1d = true;
        }
        // blocked reads run past the last field's halo (times zero weights): up to 2 kColB
        // planes (3D columns) or rows (2D), or 4 kRowPad positions (1D rows)
        auto slack = [&](const PadGeom &pg) -> size_t {
            if (nd == 3) return (size_t)2 * kColB * pg.P[0] * pg.P[1];
            if (nd == 2) return (size_t)2 * kColB * pg.P[0];
            return (size_t)4 * kRowPad;
        };
        const size_t slack_u = slack(P->pad_u), slack_f = slack(P->pad_f);
        const size_t nu = (size_t)C * P->pad_u.n + slack_u, nf = (size_t)P->nadj * P->pad_f.n + slack_f;
        P->wsd.rho_pad = dev_alloc<double>(nu);
        P->wsd.adj_pad = dev_alloc<double>(nf);
        P->wsf.rho_pad = dev_alloc<float>(nu);
        P->wsf.adj_pad = dev_alloc<float>(nf);
        TMI_HIP_CHECK(hipMemset(P->wsd.rho_pad, 0, nu * sizeof(double)));
        TMI_HIP_CHECK(hipMemset(P->wsd.adj_pad, 0, nf * sizeof(double)));
        TMI_HIP_CHECK(hipMemset(P->wsf.rho_pad, 0, nu * sizeof(float)));
        TMI_HIP_CHECK(hipMemset(P->wsf.adj_pad, 0, nf * sizeof(float)));
        if (P->use_columns && C > 1) {
            P->wsd.rb_part = dev_alloc<double>((size_t)5 * C * N);
            P->wsf.rb_part = dev_alloc<float>((size_t)5 * C * N);
        }
        P->d_red = dev_alloc<double>(1024 * 2 + 512 * 1024);
        if (P->ext_mf)   // the entry points' expansion buffers, so that no fill allocates
            for (double **b : {&P->d_xi, &P->d_vi, &P->d_fi, &P->d_gi}) *b = dev_alloc<double>((size_t)P->nunk * N);
        noalloc_scratch();
        autotune_rows1d(*P);
        TMI_HIP_CHECK(hipHostMalloc((void **)&P->h_red, sizeof(double) * 1024, hipHostMallocDefault));
        TMI_HIP_CHECK(hipHostGetDevicePointer((void **)&P->h_red_dev, P->h_red, 0));
        // FFT forms of the stencil sums on 2D/3D grids: the mean-field sums when the attraction
        // stencils are wide (TMI_MF_FFT=1 forces it, 0 keeps the direct sums), the FMT sums when
        // that form fills faster (TMI_FMT_FFT=1 / 0 forces the FFT / direct form)
        if (nd >= 2 && !batch) {   // batches of 2D pieces keep the direct sums (their padded copy has gap rows)
            const char *fe = std::getenv("TMI_MF_FFT"), *fm = std::getenv("TMI_FMT_FFT");
            long long lsum = 0;
            if (p.type_attr >= 0)
                for (int q = 0; q < C * C; ++q) lsum += P->hs[TMI_UATT][q].L;
            const bool mf = p.type_attr >= 0 && (fe ? std::string(fe) == "1" : lsum >= kFftMinStencil);
            const bool fmt = !(fm && std::string(fm) == "0");
            fft_setup(*P, mf, fmt);
            if (!symbol_only) {
                choose_fmt_form(*P);
            } else if (fft_fmt_available(*P)) {
                fft_set_fmt(*P, true);
                fft_set_fused(*P, 0);
            }
            precond_setup(*P);
            // an unused plan is kept only for the bulk preconditioner
            if (P->fft && !fft_mf(*P) && !fft_fmt(*P) && !precond_available(*P)) fft_free(*P);
        }
    } catch (...) {
        destroy_problem(P);
        throw;
    }
    return P;
}

void destroy_problem(Problem *P)
{
    hw_free(*P);
    (void)hipFree(P->d_w0);
    for (int t = 0; t < 3; ++t)
        for (int q = 0; q < kNC * kNC; ++q) {
            free_stencil(P->dsd[t][q]);
            free_stencil(P->dsf[t][q]);
            free_columns(P->cold[t][q]);
            free_columns(P->colf[t][q]);
            free_row1d(P->rowd[t][q]);
            free_row1d(P->rowf[t][q]);
        }
    free_batch_work(*P);
    fft_free(*P);
    free_geometry_maps(P->pad_u);
    free_geometry_maps(P->pad_f);
    (void)hipFree(P->d_pore_off);
    (void)hipFree(P->d_pore_of);
    (void)hipFree(P->d_mu);
    (void)hipFree(P->d_zero);
    (void)hipFree(P->d_elzero);
    (void)hipFree(P->d_vext);
    (void)hipFree(P->wsd.rho_pad);
    (void)hipFree(P->wsd.adj_pad);
    (void)hipFree(P->wsf.rho_pad);
    (void)hipFree(P->wsf.adj_pad);
    (void)hipFree(P->wsd.rb_part);
    (void)hipFree(P->wsf.rb_part);
    (void)hipFree(P->wsd.corr_u);
    (void)hipFree(P->wsd.corr_f);
    (void)hipFree(P->wsf.corr_u);
    (void)hipFree(P->wsf.corr_f);
    (void)hipFree(P->d_red);
    (void)hipFree(P->d_mfw2);
    (void)hipFree(P->f_mfw2);
    (void)hipFree(P->d_rbw2);
    (void)hipFree(P->f_rbw2);
    (void)hipFree(P->d_scratch);
    (void)hipFree(P->d_newton);
    (void)hipFree(P->d_xi);
    (void)hipFree(P->d_vi);
    (void)hipFree(P->d_fi);
    (void)hipFree(P->d_gi);
    if (P->h_red) (void)hipHostFree(P->h_red);
    delete P;
}

Problem *as_problem(void *h)
{
    if (!h) throw std::invalid_argument("null problem handle");
    return (Problem *)h;
}

// ---- external layout without mean-field rows (att_in_a22): adapters of the entry points
double *ibuf(Problem &P, double *&b)
{
    if (!b) b = dev_alloc<double>((size_t)P.nunk * P.N);
    return b;
}
// x in the internal layout: the external rows, then u * rho of its densities (linear: a direction,
// zero boundary values); x itself when the layouts agree
const double *expand(Problem &P, const double *x, double *&buf, bool linear, hipStream_t s)
{
    if (!P.ext_mf) return x;
    double *b = ibuf(P, buf);
    const size_t N = P.N;
    if (b != x) TMI_HIP_CHECK(hipMemcpyAsync(b, x, sizeof(double) * P.nunk_ext * N, hipMemcpyDeviceToDevice, s));
    launch_mf_conv(P, b + (size_t)P.iu_rho * N, linear, b + (size_t)P.iu_mf * N, s);
    return b;
}
// the external rows of an internal-layout result
void shrink(Problem &P, const double *in, double *out, hipStream_t s)
{
    if (in != out)
        TMI_HIP_CHECK(hipMemcpyAsync(out, in, sizeof(double) * P.nunk_ext * P.N, hipMemcpyDeviceToDevice, s));
}

double *scratch(Problem &P, size_t doubles)
{
    size_t bytes = doubles * sizeof(double);
    if (bytes > P.scratch_bytes) {
        (void)hipFree(P.d_scratch);
        P.d_scratch = nullptr;
        P.scratch_bytes = 0;
        TMI_HIP_CHECK(hipMalloc((void **)&P.d_scratch, bytes));
        P.scratch_bytes = bytes;
    }
    return P.d_scratch;
}

// right-preconditioned restarted GMRES on the Schur complement
// S = A22 - A21 A11^{-1} A12 (dft_HardSphereLinProbMgr.cpp:345-428);
// classical Gram-Schmidt with one reorthogonalisation pass.  TV is the Krylov basis type: FP32 on
// the mixed path (half the Gram-Schmidt traffic; each basis vector is widened to FP64 before its
// products, and the fused bulk preconditioner transforms in FP32), FP64 otherwise; the residual, w,
// the Arnoldi coefficients and x2 are FP64 throughout.
template <typename TV>
int gmres_schur_t(Problem &P, const double *x, const double *bS, double *x2, double rtol, int m, int maxit, int prec,
                  double *work, hipStream_t s)
{
    const long long n = (long long)P.p.ncomp * P.N;
    m = std::max(1, std::min(m, 500));
    TV *V = reinterpret_cast<TV *>(work);                                      // (m+1) n
    double *z = work + ((size_t)(m + 1) * n * sizeof(TV) + sizeof(double) - 1) / sizeof(double);   // n
    double *w = z + n;                                 // n
    double *dinv = w + n;                              // n
    double *vk = dinv + n;                             // n: V_j widened to FP64 (FP32 basis only)
    double *t1 = vk + (std::is_same<TV, float>::value ? n : 0);   // (nrb + nmf) N
    // right preconditioner: the A22 diagonal (dinv), or its bulk-circulant correction on wrapping FFT grids
    const bool bulk_pc = precond_active(P);
    const double *x_rho = x + (size_t)P.iu_rho * P.N;
    if (!bulk_pc) dev_a22_inverse_diag(P, x, dinv, s);
    dev_fill(0.0, x2, (long long)n, s);
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), h(m + 1);
    // device slot of the Arnoldi coefficients (past the partial sums of the dot products)
    double *dh = P.d_red + 1024 * 1024 / 2;
    // w = S M^-1 V_j, the products of an Arnoldi step
    auto products = [&](int j) {
        const double *vj;
        if constexpr (std::is_same<TV, float>::value) {
            dev_widen(V + (size_t)j * n, vk, n, s);
            vj = vk;
        } else {
            vj = V + (size_t)j * n;
        }
        if (bulk_pc)
            precond_apply(P, x_rho, vj, z, s, std::is_same<TV, float>::value);
        else
            dev_mul(dinv, vj, z, n, s);
        if (prec == TMI_PREC_MIXED)
            launch_schur_apply<float>(P, x, z, t1, w, s);
        else
            launch_schur_apply<double>(P, x, z, t1, w, s);
    };
    const double bn0 = std::sqrt(dev_dot(P, bS, bS, n, s));
    const double bn = bn0 > 0.0 ? bn0 : 1.0;
    int its = 0;
    bool first = true;
    while (its < maxit) {
        // r = bS - S x2 (in z), V_0 = r / |r|
        const double *r = bS;
        if (!first) {
            if (prec == TMI_PREC_MIXED)
                launch_schur_apply<float>(P, x, x2, t1, w, s);
            else
                launch_schur_apply<double>(P, x, x2, t1, w, s);
            dev_scale_copy(1.0, bS, z, n, s);
            dev_axpy(-1.0, w, z, n, s);
            r = z;
        }
        const double beta = first ? bn0 : std::sqrt(dev_dot(P, r, r, n, s));
        first = false;
        if (beta / bn <= rtol || beta == 0.0) break;
        dev_cast_scaled<TV>(1.0 / beta, r, V, n, s);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        bool done = false;
        for (; k < m && its < maxit; ++k) {
            products(k);
            // classical Gram-Schmidt and one reorthogonalisation pass with the coefficients kept
            // on the device; one host synchronisation per step, for the Hessenberg column
            dev_cgs2<TV>(P, V, n, k + 1, w, n, V + (size_t)(k + 1) * n, dh, s);
            TMI_HIP_CHECK(hipStreamSynchronize(s));
            for (int j = 0; j <= k; ++j) h[j] = P.h_red[j] + P.h_red[k + 1 + j];
            const double hn = std::sqrt(P.h_red[2 * (k + 1)]);
            for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = h[j];
            H[(size_t)(k + 1) * m + k] = hn;
            for (int j = 0; j < k; ++j) {
                double a = H[(size_t)j * m + k], b = H[(size_t)(j + 1) * m + k];
                H[(size_t)j * m + k] = cs[j] * a + sn[j] * b;
                H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * b;
            }
            double a = H[(size_t)k * m + k], b = H[(size_t)(k + 1) * m + k];
            double den = std::sqrt(a * a + b * b);
            cs[k] = den > 0 ? a / den : 1.0;
            sn[k] = den > 0 ? b / den : 0.0;
            H[(size_t)k * m + k] = den;
            H[(size_t)(k + 1) * m + k] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            ++its;
            if (std::fabs(g[k + 1]) / bn <= rtol || hn == 0.0) {
                ++k;
                done = true;
                break;
            }
        }
        for (int i = k - 1; i >= 0; --i) {
            double sum = g[i];
            for (int j = i + 1; j < k; ++j) sum -= H[(size_t)i * m + j] * y[j];
            y[i] = sum / H[(size_t)i * m + i];
        }
        TMI_HIP_CHECK(hipMemcpyAsync(P.d_red + 1024 * 1024 / 2, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, s));
        dev_fill(0.0, z, (long long)n, s);
        dev_gemv_add<TV>(V, n, k, P.d_red + 1024 * 1024 / 2, z, n, s);
        if (bulk_pc)   // the same preconditioner as in the products
            precond_apply(P, x_rho, z, w, s, std::is_same<TV, float>::value);
        else
            dev_mul(dinv, z, w, n, s);
        dev_axpy(1.0, w, x2, n, s);
        TMI_HIP_CHECK(hipStreamSynchronize(s));
        if (done) break;
    }
    return its;
}

int gmres_schur(Problem &P, const double *x, const double *bS, double *x2, double rtol, int m, int maxit, int prec,
                double *work, hipStream_t s)
{
    if (prec == TMI_PREC_MIXED) return gmres_schur_t<float>(P, x, bS, x2, rtol, m, maxit, prec, work, s);
    return gmres_schur_t<double>(P, x, bS, x2, rtol, m, maxit, prec, work, s);
}

size_t gmres_work_doubles(const Problem &P, int m)
{
    const size_t n = (size_t)P.p.ncomp * P.N, t1 = (size_t)(P.nrb + P.nmf) * P.N;
    const size_t fp64 = (size_t)(m + 1) * n + 3 * n + t1;           // FP64 basis, z, w, dinv, t1
    const size_t fp32 = ((size_t)(m + 1) * n + 1) / 2 + 4 * n + t1;   // FP32 basis, z, w, dinv, the widened V_j, t1
    return std::max(fp64, fp32);
}

int schur_solve(Problem &P, const double *x, const double *F, double *dx, double rtol, int restart, int maxit, int prec,
                int *iters, hipStream_t s)
{
    const size_t N = P.N, ntot = (size_t)P.nunk * N, n2 = (size_t)P.p.ncomp * N;
    restart = std::max(1, std::min(restart, 500));
    double *work = scratch(P, 2 * ntot + n2 + gmres_work_doubles(P, restart));
    double *b = work, *bS = b + ntot, *gw = bS + n2;
    dev_scale_copy(-1.0, F, b, (long long)ntot, s);
    // with algebraic n1, n0, nv1 rows the reference's A11 is [-I 0; B -I]: -A11^{-1} b1 is the lifted
    // b1 under A11 = -I, and the back-substitution A11^{-1} (b1 - A12 x2) is the same lift of
    // A12 x2 - b1 with the stencil rows of A12 (n1 = n2 / (4 pi R) holds for the stencil sums too)
    one_size_lift(P, b + (size_t)P.iu_rb * N, s);
    launch_schur_rhs(P, x, b, bS, s);
    int its = gmres_schur(P, x, bS, dx, rtol, restart, maxit, prec, gw, s);
    launch_schur_backsub(P, x, b, dx, s);
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    if (iters) *iters = its;
    return 0;
}

}  // namespace

// a periodic bulk fluid of the given parameters (no walls, no zero nodes), for the bulk Schur
// symbol of a walled grid's preconditioner (tmi_fft.hip); free with destroy_bulk_problem
Problem *create_bulk_problem(const tmi_params &p) { return create_problem(p, nullptr, nullptr, nullptr, true); }
void destroy_bulk_problem(Problem *P) { destroy_problem(P); }

// exported for the LinProbMgr layer
int schur_solve_export(Problem &P, const double *x, const double *F, double *dx, double rtol, int restart, int maxit,
                       int prec, int *iters, hipStream_t s)
{
    return schur_solve(P, x, F, dx, rtol, restart, maxit, prec, iters, s);
}

int newton_solve(Problem &P, double *x, const tmi_newton_opts &o, tmi_newton_stats &st, hipStream_t s)
{
    if (o.convergence != TMI_CONV_RESID && o.convergence != TMI_CONV_UPDATE)
        throw std::invalid_argument("convergence must be TMI_CONV_RESID or TMI_CONV_UPDATE");
    const size_t N = P.N, ntot = (size_t)P.nunk * N;
    // F and dx: allocated at the first solve and kept with the problem (no allocation in later solves)
    if (!P.d_newton) P.d_newton = dev_alloc<double>(2 * ntot);
    double *F = P.d_newton, *dx = P.d_newton + ntot;
    auto t0 = std::chrono::steady_clock::now();
    st = tmi_newton_stats{};
    int iter = 0;
    while (true) {
        const double rn = std::sqrt(residual_sumsq(P, x, F, s));
        st.resid_norm = rn;
        if (o.verbose) std::fprintf(stderr, "[tmi newton] iter %d ||F|| = %.6e\n", iter, rn);
        if (!(rn == rn)) throw std::runtime_error("residual is NaN");
        if (o.convergence == TMI_CONV_RESID && rn < o.resid_tol) {
            st.converged = 1;
            break;
        }
        if (iter >= o.max_iter) break;
        int its = 0;
        schur_solve(P, x, F, dx, o.gmres_tol, o.gmres_restart, o.gmres_max_iter, o.precision, &its, s);
        st.gmres_iters_total += its;
        double frac = 1.0, unorm = 0.0;
        dev_newton_update(P, x, dx, o, &frac, &unorm, s);
        if (o.verbose) std::fprintf(stderr, "[tmi newton]   gmres %d its, frac %.3g, weighted update %.3e\n", its, frac, unorm);
        ++iter;
        if (o.convergence == TMI_CONV_UPDATE && unorm <= 1.0) {   // src/dft_newton.c:378
            st.converged = 1;
            st.resid_norm = std::sqrt(residual_sumsq(P, x, F, s));
            break;
        }
    }
    st.iters = iter;
    TMI_HIP_CHECK(hipStreamSynchronize(s));
    st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

Problem *problem_from_handle(void *h) { return as_problem(h); }

}  // namespace tmi

using namespace tmi;

extern "C" {

const char *tmi_last_error(void) { return g_last_error.c_str(); }
int tmi_version(void) { return 1; }

int tmi_problem_create(const tmi_params *p, const double *vext, const unsigned char *zero_tf, void **handle)
{
    return guarded([&] {
        if (!p || !handle) throw std::invalid_argument("null argument");
        *handle = create_problem(*p, vext, zero_tf);
        return 0;
    });
}

int tmi_problem_destroy(void *handle)
{
    return guarded([&] {
        if (handle) destroy_problem((Problem *)handle);
        return 0;
    });
}

int tmi_problem_info(void *handle, tmi_info *info)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        std::memset(info, 0, sizeof(*info));
        info->nunk = P->nunk_ext;
        info->nnodes = P->N;
        info->iu_rho = P->iu_rho;
        info->iu_rb = P->iu_rb;
        info->n_rb = P->nrb;
        info->iu_mf = P->ext_mf ? P->nunk_ext : P->iu_mf;   // no mean-field rows in the att_in_a22 layout
        info->n_mf = P->ext_mf ? 0 : P->nmf;
        for (int t = 0; t < 3; ++t)
            for (int q = 0; q < kNC * kNC; ++q) info->sten_len[t][q] = P->hs[t][q].L;
        info->mf_fft = fft_mf(*P) ? 1 : 0;
        info->fmt_fft = fft_fmt(*P) ? (fft_fused(*P) ? 2 : 1) : 0;
        info->precond = precond_active(*P) ? 1 : 0;
        info->fft_fold = fft_folded(*P) ? 1 : 0;
        info->fft_xreg = fft_xreg_mask(*P);
        info->fft_regpass = fft_regpass_mode(*P);
        info->jacobian = P->jac_ref ? TMI_JAC_REFERENCE : TMI_JAC_EXACT;
        return 0;
    });
}

int tmi_stencil_get(void *handle, int isten, int icomp, int jcomp, int *len, int *off, double *w)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (isten < 0 || isten > 2 || icomp < 0 || icomp >= P->p.ncomp || jcomp < 0 || jcomp >= P->p.ncomp)
            throw std::invalid_argument("stencil index");
        const HostStencil &s = P->hs[isten][icomp + P->p.ncomp * jcomp];
        if (len) *len = s.L;
        if (off) std::memcpy(off, s.off.data(), sizeof(int) * s.off.size());
        if (w) std::memcpy(w, s.w.data(), sizeof(double) * s.w.size());
        return 0;
    });
}

static void pack_bulk(const tmi_params &p, const BulkConstants &b, double *out)
{
    std::memset(out, 0, 60 * sizeof(double));
    for (int q = 0; q < 10; ++q) out[q] = b.rhobar_b[q];
    for (int q = 0; q < 10; ++q) out[10 + q] = b.dphi_b[q];
    for (int i = 0; i < p.ncomp; ++i) {
        out[20 + i] = b.betamu_hs_ex[i];
        out[24 + i] = b.betamu_att[i];
        for (int j = 0; j < p.ncomp; ++j) {
            out[28 + 4 * i + j] = b.avdw[i][j];
            out[44 + 4 * i + j] = b.rmin[i][j];
        }
    }
}

// host-side tables without touching the GPU
static void host_tables(const tmi_params &p, BulkConstants &b, HostStencil (*hs)[kNC * kNC])
{
    validate(p);
    compute_bulk_constants(p, b);
    if (p.type_attr >= 0)
        for (int i = 0; i < p.ncomp; ++i)
            for (int j = 0; j < p.ncomp; ++j)
                b.rmin[i][j] = lj_potential_minimum(p.sigma_ff[i][j], p.eps_ff[i][j], p.cut_ff[i][j], p.dx[0]);
    build_all_stencils(p, b, hs);
    if (p.type_attr >= 0) compute_attraction_constants(p, b, hs);
}

int tmi_host_stencil(const tmi_params *p, int isten, int icomp, int jcomp, int *len, int *off, double *w)
{
    return guarded([&] {
        if (!p) throw std::invalid_argument("null params");
        if (isten < 0 || isten > 2 || icomp < 0 || icomp >= p->ncomp || jcomp < 0 || jcomp >= p->ncomp)
            throw std::invalid_argument("stencil index");
        BulkConstants b;
        static thread_local HostStencil hs[3][kNC * kNC];
        for (auto &row : hs)
            for (auto &s : row) s = HostStencil();
        host_tables(*p, b, hs);
        const HostStencil &s = hs[isten][icomp + p->ncomp * jcomp];
        if (len) *len = s.L;
        if (off) std::memcpy(off, s.off.data(), sizeof(int) * s.off.size());
        if (w) std::memcpy(w, s.w.data(), sizeof(double) * s.w.size());
        return 0;
    });
}

int tmi_host_bulk(const tmi_params *p, double *out)
{
    return guarded([&] {
        if (!p || !out) throw std::invalid_argument("null argument");
        BulkConstants b;
        static thread_local HostStencil hs[3][kNC * kNC];
        for (auto &row : hs)
            for (auto &s : row) s = HostStencil();
        host_tables(*p, b, hs);
        pack_bulk(*p, b, out);
        return 0;
    });
}

int tmi_bulk_get(void *handle, double *out)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        pack_bulk(P->p, P->bulk, out);
        return 0;
    });
}

int tmi_fill_residual(void *handle, const double *x_d, double *F_d, double *sumsq, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !F_d) throw std::invalid_argument("null device pointer");
        hipStream_t s = (hipStream_t)stream;
        const double *xi = expand(*P, x_d, P->d_xi, false, s);
        double *Fi = P->ext_mf ? ibuf(*P, P->d_fi) : F_d;
        if (sumsq)
            *sumsq = residual_sumsq(*P, xi, Fi, s);
        else
            launch_residual(*P, xi, Fi, s);
        if (P->ext_mf) shrink(*P, Fi, F_d, s);
        return 0;
    });
}

int tmi_fill_residual_noalloc(void *handle, const double *x_d, double *F_d, double *sumsq_d, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !F_d || !sumsq_d) throw std::invalid_argument("null device pointer");
        const long long need = residual_sumsq_partials(*P);
        if (need > kNoAllocScratchDoubles)
            throw std::invalid_argument("tmi_fill_residual_noalloc: the fill needs " + std::to_string(need) +
                                        " partial sums, more than the fixed scratch of " +
                                        std::to_string(kNoAllocScratchDoubles) + " doubles");
        double *partials = noalloc_scratch();
        hipStream_t s = (hipStream_t)stream;
        const double *xi = expand(*P, x_d, P->d_xi, false, s);   // buffers of the external layout exist from creation
        double *Fi = P->ext_mf ? P->d_fi : F_d;
        residual_sumsq_device(*P, xi, Fi, partials, sumsq_d, s);
        if (P->ext_mf) shrink(*P, Fi, F_d, s);
        return 0;
    });
}

int tmi_fill_residual_timed(void *handle, const double *x_d, double *F_d, int reps, double *stage_ms, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !F_d || !stage_ms || reps < 1) throw std::invalid_argument("bad argument");
        hipStream_t s = (hipStream_t)stream;
        const double *xi = expand(*P, x_d, P->d_xi, false, s);
        double *Fi = P->ext_mf ? ibuf(*P, P->d_fi) : F_d;
        hipEvent_t ev[6];
        for (auto &e : ev) TMI_HIP_CHECK(hipEventCreate(&e));
        for (int q = 0; q < 5; ++q) stage_ms[q] = 0.0;
        for (int r = 0; r < reps; ++r) {
            launch_residual(*P, xi, Fi, s, ev);
            TMI_HIP_CHECK(hipEventSynchronize(ev[5]));
            for (int q = 0; q < 5; ++q) {
                float ms = 0.f;
                TMI_HIP_CHECK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
                stage_ms[q] += ms / reps;
            }
        }
        for (auto &e : ev) (void)hipEventDestroy(e);
        return 0;
    });
}

int tmi_fft_timed(void *handle, int which, int reps, double *ms, double *bytes, int *launches, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!ms || !bytes || !launches || reps < 1 || which < 0 || which > 2) throw std::invalid_argument("bad argument");
        fft_time_kernels(*P, which, reps, ms, bytes, launches, (hipStream_t)stream);
        return 0;
    });
}

int tmi_apply_jacobian(void *handle, const double *x_d, const double *v_d, double *Jv_d, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !v_d || !Jv_d) throw std::invalid_argument("null device pointer");
        hipStream_t s = (hipStream_t)stream;
        const double *xi = expand(*P, x_d, P->d_xi, false, s), *vi = expand(*P, v_d, P->d_vi, true, s);
        double *Ji = P->ext_mf ? ibuf(*P, P->d_fi) : Jv_d;
        launch_jvp<double>(*P, xi, vi, Ji, s);
        if (P->ext_mf) shrink(*P, Ji, Jv_d, s);
        return 0;
    });
}

int tmi_apply_jacobian_fp32(void *handle, const double *x_d, const double *v_d, double *Jv_d, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !v_d || !Jv_d) throw std::invalid_argument("null device pointer");
        hipStream_t s = (hipStream_t)stream;
        const double *xi = expand(*P, x_d, P->d_xi, false, s), *vi = expand(*P, v_d, P->d_vi, true, s);
        double *Ji = P->ext_mf ? ibuf(*P, P->d_fi) : Jv_d;
        launch_jvp<float>(*P, xi, vi, Ji, s);
        if (P->ext_mf) shrink(*P, Ji, Jv_d, s);
        return 0;
    });
}

int tmi_apply_jacobian_transpose(void *handle, const double *x_d, const double *w_d, double *JTw_d, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !w_d || !JTw_d) throw std::invalid_argument("null device pointer");
        if (P->npores > 0) throw std::domain_error("transpose products of batched problems are not implemented");
        hipStream_t s = (hipStream_t)stream;
        if (!P->ext_mf) {
            launch_vjp<double>(*P, x_d, w_d, JTw_d, s);
            return 0;
        }
        // J_ext = J[e,e] + J[e,mf] J[mf,e]: J_ext^T w = (J^T [w; 0])_e + (J^T [0; (J^T [w; 0])_mf])_e, since
        // the mean-field rows are x_mf - u * rho (J[mf,mf] = -I): two internal transpose products
        const size_t N = P->N, ne = (size_t)P->nunk_ext * N, ni = (size_t)P->nunk * N;
        const double *xi = expand(*P, x_d, P->d_xi, false, s);
        double *wi = ibuf(*P, P->d_vi), *y = ibuf(*P, P->d_fi), *z = ibuf(*P, P->d_gi);
        TMI_HIP_CHECK(hipMemcpyAsync(wi, w_d, sizeof(double) * ne, hipMemcpyDeviceToDevice, s));
        dev_fill(0.0, wi + ne, (long long)(ni - ne), s);
        launch_vjp<double>(*P, xi, wi, y, s);
        dev_fill(0.0, wi, (long long)ne, s);
        TMI_HIP_CHECK(hipMemcpyAsync(wi + ne, y + ne, sizeof(double) * (ni - ne), hipMemcpyDeviceToDevice, s));
        launch_vjp<double>(*P, xi, wi, z, s);
        dev_axpy(1.0, z, y, (long long)ne, s);
        shrink(*P, y, JTw_d, s);
        return 0;
    });
}

int tmi_problem_set_rho_b(void *handle, const double *rho_b)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!rho_b) throw std::invalid_argument("null rho_b");
        if (P->npores > 0) throw std::domain_error("a batch holds one set of bulk densities per pore");
        for (int c = 0; c < P->p.ncomp; ++c)
            if (!(rho_b[c] > 0.0)) throw std::invalid_argument("rho_b must be positive");
        TMI_HIP_CHECK(hipDeviceSynchronize());   // launches in flight read the old constants by value only
        for (int c = 0; c < P->p.ncomp; ++c) P->p.rho_b[c] = rho_b[c];
        compute_bulk_constants(P->p, P->bulk);
        if (P->p.type_attr >= 0) compute_attraction_constants(P->p, P->bulk, P->hs);
        precond_invalidate(*P);   // the bulk Schur symbol depends on rho_b: rebuilt when GMRES next needs it
        return 0;
    });
}

int tmi_problem_set_vext(void *handle, const double *vext)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!vext) throw std::invalid_argument("null vext");
        if (P->p.nwall <= 0) throw std::invalid_argument("the problem has no external field (nwall = 0)");
        const size_t N = P->N;
        const int C = P->p.ncomp;
        for (size_t i = 0; i < N; ++i)
            for (int c = 0; c < C; ++c) {
                const double v = vext[i * C + c];
                P->h_vext[c * N + i] = v;
                P->h_elzero[c * N + i] = (P->h_zero[c * N + i] || v == P->p.vext_max) ? 1 : 0;
            }
        TMI_HIP_CHECK(hipDeviceSynchronize());
        upload(P->d_vext, P->h_vext.data(), P->h_vext.size());
        upload(P->d_elzero, P->h_elzero.data(), P->h_elzero.size());
        return 0;
    });
}

int tmi_problem_set_preconditioner(void *handle, int kind)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (kind != TMI_PRECOND_DIAG && kind != TMI_PRECOND_BULK) throw std::invalid_argument("preconditioner kind");
        if (kind == TMI_PRECOND_BULK && !precond_available(*P))
            throw std::domain_error("the bulk preconditioner needs an FFT grid and a stable bulk");
        precond_enable(*P, kind == TMI_PRECOND_BULK);
        return 0;
    });
}

int tmi_problem_set_jacobian(void *handle, int mode)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (mode != TMI_JAC_EXACT && mode != TMI_JAC_REFERENCE) throw std::invalid_argument("Jacobian mode");
        P->jac_ref = mode == TMI_JAC_REFERENCE;
        return 0;
    });
}

int tmi_compute_force(void *handle, const double *x_d, const double *w_d, int nsum, double *out, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !w_d || !out) throw std::invalid_argument("null pointer");
        if (nsum < 1 || nsum > 64) throw std::invalid_argument("nsum must be 1..64");
        if (P->npores > 0) throw std::invalid_argument("surface sums of a batch of pores are not supported");
        const long long n = (long long)P->p.ncomp * P->N;
        dev_multi_dot(*P, w_d, n, nsum, x_d + (size_t)P->iu_rho * P->N, n, out, (hipStream_t)stream);
        return 0;
    });
}

int tmi_compute_observables(void *handle, const double *x_d, const int *nel_hit, const int *nel_hit2, double scale,
                            tmi_observables *out, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!x_d || !nel_hit || !nel_hit2 || !out) throw std::invalid_argument("null pointer");
        hipStream_t s = (hipStream_t)stream;
        const int C = P->p.ncomp, nd = P->p.ndim;
        const size_t N = P->N, cn = (size_t)C * N;
        // scratch: u * rho [C][N], then the two count arrays (ints, two to a double)
        double *mf = scratch(*P, cn + cn);
        int *nh = reinterpret_cast<int *>(mf + cn), *nh2 = nh + cn;
        TMI_HIP_CHECK(hipMemcpyAsync(nh, nel_hit, cn * sizeof(int), hipMemcpyHostToDevice, s));
        TMI_HIP_CHECK(hipMemcpyAsync(nh2, nel_hit2, cn * sizeof(int), hipMemcpyHostToDevice, s));
        double sums[5 + 2 * TMI_NCOMP_MAX];
        observables_sums(*P, x_d, nh, nh2, mf, sums, s);
        double vol = 1.0;
        for (int d = 0; d < nd; ++d) vol *= P->p.dx[d];
        const double f = vol / (double)(1 << nd) * scale;
        const bool att_on = P->p.type_attr >= 0 && P->nmf > 0;
        double nb[10] = {0.0};   // bulk rho-bar values in n order, vector parts 0
        nb[0] = P->bulk.rhobar_b[3];
        nb[1] = P->bulk.rhobar_b[2];
        nb[2] = P->bulk.rhobar_b[1];
        nb[3] = P->bulk.rhobar_b[0];
        const double phi_b = nd == 1 ? fmt_phi<1>(P->p.type_func, nb)
                                     : (nd == 2 ? fmt_phi<2>(P->p.type_func, nb) : fmt_phi<3>(P->p.type_func, nb));
        tmi_observables o{};
        double id_b = 0.0, mu_b = 0.0, att_b = 0.0;
        for (int c = 0; c < C; ++c) {
            const double rb = P->p.rho_b[c], sw1 = sums[5 + TMI_NCOMP_MAX + c];
            const double betamu = std::log(rb) + P->bulk.betamu_hs_ex[c] + (P->p.type_attr >= 0 ? P->bulk.betamu_att[c] : 0.0);
            id_b += rb * (std::log(rb) - 1.0) * sw1;
            mu_b += -rb * betamu * sw1;
            if (att_on) att_b += 0.5 * rb * P->bulk.betamu_att[c] * sw1;
            o.ads[c] = sums[5 + c] * f;
            o.ads_b[c] = rb * sw1 * f;
        }
        o.ideal = sums[0] * f;
        o.ideal_b = id_b * f;
        o.mu = sums[1] * f;
        o.mu_b = mu_b * f;
        o.vext = sums[2] * f;
        o.hs = sums[3] * f;
        o.hs_b = phi_b * sums[5 + TMI_NCOMP_MAX] * f;
        o.att = sums[4] * f;
        o.att_b = att_b * f;
        o.omega = o.ideal + o.mu + o.vext + o.hs + o.att;
        o.omega_s = (o.ideal - o.ideal_b) + (o.mu - o.mu_b) + o.vext + (o.hs - o.hs_b) + (o.att - o.att_b);
        *out = o;
        return 0;
    });
}

int tmi_set_hard_surface_masks(void *handle, const unsigned char *mask)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!P->p.lhard_surf) throw std::invalid_argument("hard-surface masks need lhard_surf = 1");
        TMI_HIP_CHECK(hipDeviceSynchronize());
        hw_build(*P, mask);
        return 0;
    });
}

int tmi_init_fields(void *handle, double *x_d, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        hipStream_t s = (hipStream_t)stream;
        if (!P->ext_mf) {
            launch_init_fields(*P, x_d, s);
            return 0;
        }
        double *xi = ibuf(*P, P->d_xi);
        TMI_HIP_CHECK(hipMemcpyAsync(xi, x_d, sizeof(double) * P->nunk_ext * P->N, hipMemcpyDeviceToDevice, s));
        launch_init_fields(*P, xi, s);
        shrink(*P, xi, x_d, s);
        return 0;
    });
}

int tmi_schur_solve(void *handle, const double *x_d, const double *F_d, double *dx_d, double rtol, int restart,
                    int maxit, int precision, int *iters, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        hipStream_t s = (hipStream_t)stream;
        if (!P->ext_mf) return schur_solve(*P, x_d, F_d, dx_d, rtol, restart, maxit, precision, iters, s);
        // internal F: the external rows, mean-field rows 0 (x_mf = u * rho holds); dx's mean-field rows
        // (u * drho) are dropped
        const double *xi = expand(*P, x_d, P->d_xi, false, s);
        double *Fi = ibuf(*P, P->d_fi), *dxi = ibuf(*P, P->d_gi);
        const size_t ne = (size_t)P->nunk_ext * P->N;
        TMI_HIP_CHECK(hipMemcpyAsync(Fi, F_d, sizeof(double) * ne, hipMemcpyDeviceToDevice, s));
        dev_fill(0.0, Fi + ne, (long long)((size_t)P->nunk * P->N - ne), s);
        const int rc = schur_solve(*P, xi, Fi, dxi, rtol, restart, maxit, precision, iters, s);
        shrink(*P, dxi, dx_d, s);
        TMI_HIP_CHECK(hipStreamSynchronize(s));
        return rc;
    });
}

int tmi_newton_solve(void *handle, double *x_d, const tmi_newton_opts *opts, tmi_newton_stats *stats, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!opts || !stats) throw std::invalid_argument("null argument");
        if (P->npores > 0) {
            if (opts->convergence != TMI_CONV_RESID) throw std::domain_error("batches of pores converge on ||F_p||_2");
            return batch_newton_solve(*P, x_d, *opts, *stats, nullptr, nullptr, nullptr, nullptr, (hipStream_t)stream);
        }
        if (!P->ext_mf) return newton_solve(*P, x_d, *opts, *stats, (hipStream_t)stream);
        hipStream_t s = (hipStream_t)stream;
        double *xi = const_cast<double *>(expand(*P, x_d, P->d_xi, false, s));
        const int rc = newton_solve(*P, xi, *opts, *stats, s);
        shrink(*P, xi, x_d, s);
        TMI_HIP_CHECK(hipStreamSynchronize(s));
        return rc;
    });
}

int tmi_row_tuning(void *handle, int *shape, double *ms)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        for (int prec = 0; prec < 2; ++prec)
            for (int k = 0; k < 3; ++k) {
                const RowTune &t = P->tune[prec][k];
                if (shape) {
                    shape[(prec * 3 + k) * 2] = P->use_rows1d ? t.B : 0;
                    shape[(prec * 3 + k) * 2 + 1] = P->use_rows1d ? t.wg : 0;
                }
                if (ms) ms[prec * 3 + k] = P->use_rows1d ? t.ms : 0.0;
            }
        return 0;
    });
}

int tmi_row_tuning_set(void *handle, int prec, int kind, int per_lane, int lanes)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!P->use_rows1d) throw std::invalid_argument("the problem does not use the 1D row kernels");
        if (prec < 0 || prec > 1 || kind < 0 || kind > 2) throw std::invalid_argument("prec must be 0..1, kind 0..2");
        if (per_lane != 4 && per_lane != 8 && per_lane != 16) throw std::invalid_argument("outputs per lane: 4, 8 or 16");
        if (lanes != 64 && lanes != 128 && lanes != 256) throw std::invalid_argument("lanes per workgroup: 64, 128 or 256");
        TMI_HIP_CHECK(hipDeviceSynchronize());
        P->tune[prec][kind].B = per_lane;
        P->tune[prec][kind].wg = lanes;
        P->tune[prec][kind].ms = 0.0;
        return 0;
    });
}

int tmi_batch_create(const tmi_params *fluid, int npores, const int *nodes, const double *rho_b, const double *vext,
                     const unsigned char *zero_tf, void **handle)
{
    return guarded([&] {
        if (!fluid || !handle || !nodes || !rho_b) throw std::invalid_argument("null argument");
        BatchLayout bl;
        bl.npores = npores;
        bl.nodes = nodes;
        bl.rho_b = rho_b;
        *handle = create_problem(*fluid, vext, zero_tf, &bl);
        return 0;
    });
}

int tmi_batch_info(void *handle, int *npores, int *pore_off)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (npores) *npores = P->npores;
        if (pore_off && P->npores > 0) std::memcpy(pore_off, P->pore_off.data(), sizeof(int) * P->pore_off.size());
        return 0;
    });
}

int tmi_batch_newton_solve(void *handle, double *x_d, const tmi_newton_opts *opts, tmi_newton_stats *stats,
                           int *pore_iters, int *pore_converged, int *pore_gmres, double *pore_resid, void *stream)
{
    return guarded([&] {
        Problem *P = as_problem(handle);
        if (!opts || !stats || !x_d) throw std::invalid_argument("null argument");
        if (P->npores < 1) throw std::invalid_argument("not a batched problem");
        if (opts->convergence != TMI_CONV_RESID) throw std::domain_error("batches of pores converge on ||F_p||_2");
        return batch_newton_solve(*P, x_d, *opts, *stats, pore_iters, pore_converged, pore_gmres, pore_resid,
                                  (hipStream_t)stream);
    });
}

}  // extern "C"
