// Host-side setup of libtramonto_mi355x: quadrature stencils, LJ 12-6 mean-field
// attraction, and the bulk constants the fill consumes.  Runs once per problem.
//
// Stencil construction follows calc_stencils (src/dft_stencil.c:40-287): each
// grid element inside the bounding box of the stencil is classified against
// the stencil radius (calc_in_out_on, :753-812), integrated by Gauss quadrature
// when fully inside (:329-397) or by a midpoint rule when it straddles the
// radius (:399-496), its weight is split bilinearly onto the element's nodes
// (:630-690), and the table is renormalised to the analytic volume (:653-683).
// The arithmetic is kept in the reference's operation order so that the
// weights agree to the last bit with a faithful restatement.
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "tmi_internal.hpp"

namespace tmi {
namespace {

// Gauss-Legendre rules on [0,1] in set_gauss_quad's point order
// (src/dft_quadrature.c:40-159).  The 1/3/6/12-point rules use the
// reference's truncated decimals; 20/40-point rules are computed.
struct GaussRule {
    int n = 0;
    double x[40], w[40];
};

void legendre_roots(int n, double *xs, double *ws)
{
    // positive roots ascending (Newton iteration on P_n)
    const int h = n / 2;
    for (int k = 0; k < h; ++k) {
        double x = std::cos(kPi * (h - k - 0.25) / (n + 0.5));
        double dpn = 1.0;
        for (int it = 0; it < 64; ++it) {
            double pm = 1.0, pn = x;
            for (int j = 2; j <= n; ++j) {
                double pj = ((2.0 * j - 1.0) * x * pn - (j - 1.0) * pm) / j;
                pm = pn;
                pn = pj;
            }
            dpn = n * (x * pn - pm) / (x * x - 1.0);
            double step = pn / dpn;
            x -= step;
            if (std::fabs(step) < 1e-17) break;
        }
        double pm = 1.0, pn = x;
        for (int j = 2; j <= n; ++j) {
            double pj = ((2.0 * j - 1.0) * x * pn - (j - 1.0) * pm) / j;
            pm = pn;
            pn = pj;
        }
        dpn = n * (x * pn - pm) / (x * x - 1.0);
        xs[k] = x;
        ws[k] = 2.0 / ((1.0 - x * x) * dpn * dpn);
    }
}

GaussRule make_gauss_rule(int n);

// rules are pure functions of n: build each once
const GaussRule &gauss_rule(int n)
{
    static const GaussRule r1 = make_gauss_rule(1), r3 = make_gauss_rule(3), r6 = make_gauss_rule(6),
                           r12 = make_gauss_rule(12), r20 = make_gauss_rule(20), r40 = make_gauss_rule(40);
    switch (n) {
    case 1: return r1;
    case 3: return r3;
    case 6: return r6;
    case 12: return r12;
    case 20: return r20;
    case 40: return r40;
    default: throw std::runtime_error("unsupported Gauss rule");
    }
}

GaussRule make_gauss_rule(int n)
{
    GaussRule r;
    r.n = n;
    double xs[20], ws[20];
    int h = n / 2;
    switch (n) {
    case 1:
        r.x[0] = (0.0 + 1.0) / 2.0;
        r.w[0] = 2.0 / 2.0;
        return r;
    case 3:
        r.x[0] = (0.7745966692 + 1.0) / 2.0;
        r.w[0] = 0.555555555555556 / 2.0;
        r.x[1] = (0.0 + 1.0) / 2.0;
        r.w[1] = 0.888888888888888 / 2.0;
        r.x[2] = (-0.7745966692 + 1.0) / 2.0;
        r.w[2] = 0.555555555555556 / 2.0;
        return r;
    case 6: {
        const double a[3] = {0.238619186083197, 0.661209386466265, 0.932469514203152};
        const double b[3] = {0.467913934572691, 0.360761573048139, 0.171324492379170};
        for (int i = 0; i < 3; ++i) xs[i] = a[i], ws[i] = b[i];
        break;
    }
    case 12: {
        const double a[6] = {0.125233408511469, 0.367831498998180, 0.587317954286617,
                             0.769902674194305, 0.904117256370475, 0.981560634246719};
        const double b[6] = {0.249147045813403, 0.233492536538355, 0.203167426723066,
                             0.160078328543346, 0.106939325995318, 0.047175336386512};
        for (int i = 0; i < 6; ++i) xs[i] = a[i], ws[i] = b[i];
        break;
    }
    case 20:
    case 40:
        legendre_roots(n, xs, ws);
        break;
    default:
        throw std::runtime_error("unsupported Gauss rule");
    }
    for (int i = 0; i < h; ++i) {
        r.x[i] = (xs[i] + 1.0) / 2.0;
        r.w[i] = ws[i] / 2.0;
        r.x[h + i] = (-xs[i] + 1.0) / 2.0;
        r.w[h + i] = ws[i] / 2.0;
    }
    return r;
}

// LJ 12-6 pieces (src/dft_pairPot_LJ12_6.c:38-48, :128-228)
double lj_cs(double r, double sigma, double eps, double rcut)
{
    if (r > rcut) return 0.0;
    double u = (std::pow(sigma / r, 12) - std::pow(sigma / rcut, 12)) - (std::pow(sigma / r, 6) - std::pow(sigma / rcut, 6));
    return 4.0 * eps * u;
}

struct Attraction {
    double sigma, eps, rcut, rmin;
    // u_att with the constant core u(r<r_min) = u(r_min), cut and shifted
    double cs(double r) const
    {
        double s2 = sigma * sigma, s6 = s2 * s2 * s2;
        double rc_inv = 1.0 / rcut, rc2 = rc_inv * rc_inv, rc6 = rc2 * rc2 * rc2, rc12 = rc6 * rc6;
        if (r > rcut) return 0.0;
        if (r < rmin) r = rmin;
        double ri = 1.0 / r, r2 = ri * ri, r6 = r2 * r2 * r2, r12 = r6 * r6;
        return 4.0 * eps * s6 * (s6 * (r12 - rc12) - (r6 - rc6));
    }
    double nocs(double r) const
    {
        double s2 = sigma * sigma, s6 = s2 * s2 * s2;
        if (r < rmin) r = rmin;
        double ri = 1.0 / r, r2 = ri * ri, r6 = r2 * r2 * r2, r12 = r6 * r6;
        return 4.0 * eps * s6 * (s6 * r12 - r6);
    }
    double integral(double r) const
    {
        double s2 = sigma * sigma, s6 = s2 * s2 * s2;
        double ri = 1.0 / r, r3 = ri * ri * ri, r9 = r3 * r3 * r3;
        return 16 * kPi * eps * s6 * (-s6 * r9 / 9.0 + r3 / 3.0);
    }
    // analytic stencil volume, src/dft_stencil_theta_uattr.c:41-63 (one core constant)
    double volume() const
    {
        return (4.0 / 3.0) * kPi * std::pow(rmin, 3.0) * nocs(rmin) - (4.0 / 3.0) * kPi * std::pow(rcut, 3.0) * nocs(rcut) +
               integral(rcut) - integral(rmin);
    }
};

enum Kind { kDelta = TMI_DELTA, kTheta = TMI_THETA, kUatt = TMI_UATT };

class StencilBuilder {
  public:
    StencilBuilder(const tmi_params &p, Kind kind, int ic, int jc, const Attraction *att)
        : p_(p), kind_(kind), ic_(ic), jc_(jc), att_(att), nd_(p.ndim)
    {
        radius_ = (kind == kUatt) ? p.cut_ff[ic][jc] : p.hs_diam[ic] / 2.0;
        if (kind == kDelta)
            volume_ = kPi * std::pow(p.hs_diam[ic], 2);
        else if (kind == kTheta)
            volume_ = kPi * std::pow(p.hs_diam[ic], 3) / 6.0;
        else
            volume_ = att->volume();
        el_vol_ = 1.0;
        for (int d = 0; d < nd_; ++d) el_vol_ *= p.dx[d];
    }

    HostStencil build()
    {
        int reach[3] = {0, 0, 0}, span[3] = {1, 1, 1};
        size_t cells = 1;
        for (int d = 0; d < nd_; ++d) {
            reach[d] = (int)((float)radius_ / (float)p_.dx[d] + 0.999);
            span[d] = 2 * reach[d] + 1;
            cells *= (size_t)span[d];
            if (p_.bc[d][0] == TMI_PERIODIC && span[d] > p_.nx[d])
                throw std::runtime_error("stencil wider than the periodic domain");
        }
        std::vector<int> slot(cells, -1);
        HostStencil s;
        const int lo[3] = {-reach[0], nd_ >= 2 ? -reach[1] : 0, nd_ == 3 ? -reach[2] : 0};
        const int hi[3] = {reach[0], nd_ >= 2 ? reach[1] : 1, nd_ == 3 ? reach[2] : 1};
        const int corners = 1 << nd_;
        for (int ex = lo[0]; ex < hi[0]; ++ex)
            for (int ey = lo[1]; ey < hi[1]; ++ey)
                for (int ez = lo[2]; ez < hi[2]; ++ez) {
                    const int e[3] = {ex, ey, ez};
                    double xl[3], xr[3], rc2 = 0.0, nodal[8] = {0, 0, 0, 0, 0, 0, 0, 0};
                    for (int d = 0; d < nd_; ++d) {
                        xl[d] = e[d] * p_.dx[d];
                        xr[d] = (e[d] + 1) * p_.dx[d];
                    }
                    for (int d = 0; d < nd_; ++d) {
                        double mid = xl[d] + 0.5 * p_.dx[d];
                        rc2 += mid * mid;
                    }
                    int cls = classify(xl, xr);
                    if (cls == 1 || (cls == -1 && nd_ == 3 && kind_ == kDelta)) continue;
                    if (cls == -1)
                        interior_element(xl, rc2, nodal);
                    else
                        boundary_element(xl, rc2, nodal);
                    for (int c = 0; c < corners; ++c) {
                        if (nodal[c] == 0.0) continue;
                        int o[3] = {ex + (c & 1), ey + ((c >> 1) & 1), ez + ((c >> 2) & 1)};
                        size_t key = 0, mul = 1;
                        for (int d = 0; d < nd_; ++d) {
                            key += (size_t)(o[d] + reach[d]) * mul;
                            mul *= (size_t)span[d];
                        }
                        if (slot[key] < 0) {
                            slot[key] = s.L++;
                            s.w.push_back(nodal[c]);
                            s.hw.resize(8 * (size_t)s.L, 0.0);
                            s.hw[8 * (size_t)slot[key] + c] = nodal[c];
                            for (int d = 0; d < 3; ++d) s.off.push_back(d < nd_ ? o[d] : 0);
                        } else {
                            s.w[slot[key]] += nodal[c];
                            s.hw[8 * (size_t)slot[key] + c] += nodal[c];
                        }
                    }
                }
        double total = 0.0;
        for (double v : s.w) total += v;
        for (double v : s.w)
            if (std::isnan(v)) throw std::runtime_error("NaN stencil weight");
        if (total != volume_) {
            double ratio = volume_ / total;
            for (double &v : s.w) v *= ratio;
            for (double &v : s.hw) v *= ratio;   // src/dft_stencil.c:700-706
        }
        return s;
    }

  private:
    // 1: outside, -1: inside, 0: straddles (corner test)
    int classify(const double *xl, const double *xr) const
    {
        bool in = false, out = false;
        for (int c = 0; c < (1 << nd_); ++c) {
            double r;
            if (nd_ == 1) {
                r = std::fabs(c ? xr[0] : xl[0]);
            } else {
                double s = 0.0;
                for (int d = 0; d < nd_; ++d) {
                    double v = ((c >> d) & 1) ? xr[d] : xl[d];
                    s += v * v;
                }
                r = std::sqrt(s);
            }
            (r > radius_ ? out : in) = true;
        }
        if (!in && out) return 1;
        if (in && !out) return -1;
        return 0;
    }

    double weight_at(double rsq, const GaussRule &gu) const
    {
        const double R = radius_;
        if (kind_ == kDelta) {
            if (nd_ == 1) return 2.0 * kPi * R;
            if (nd_ == 2) return 2.0 / std::sqrt(1.0 - rsq);
            return 1.0;
        }
        if (kind_ == kTheta) {
            if (nd_ == 1) return (1.0 - rsq) * kPi * R * R;
            if (nd_ == 2) return std::sqrt(1.0 - rsq) * 2.0 * R;
            return 1.0;
        }
        const double rc = p_.cut_ff[ic_][jc_];
        const double zmax = std::sqrt(1 - rsq);
        double acc = 0.0;
        if (nd_ == 1) {
            for (int q = 0; q < gu.n; ++q) {
                double z = zmax * gu.x[q];
                acc += gu.w[q] * z * att_->cs(std::sqrt(rsq + z * z) * rc);
            }
            return 2.0 * kPi * acc * rc * rc * zmax;
        }
        if (nd_ == 2) {
            for (int q = 0; q < gu.n; ++q) {
                double z = zmax * gu.x[q];
                acc += gu.w[q] * att_->cs(std::sqrt(rsq + z * z) * rc);
            }
            return 2.0 * acc * rc * zmax;
        }
        return att_->cs(std::sqrt(rsq) * rc);
    }

    int integrand_points(double rc2) const
    {
        if (kind_ != kUatt) return 0;
        if (rc2 <= 4.000000001) return 40;
        if (rc2 <= 16.00000001) return 20;
        return 12;
    }

    void spread(double weight, const double *qx, const double *qy, const double *qz, double *nodal) const
    {
        if (nd_ == 3) {
            double a = *qx, b = *qy, c = *qz;
            nodal[0] += weight * (1.0 - a) * (1.0 - b) * (1.0 - c);
            nodal[1] += weight * a * (1.0 - b) * (1.0 - c);
            nodal[2] += weight * (1.0 - a) * b * (1.0 - c);
            nodal[3] += weight * a * b * (1.0 - c);
            nodal[4] += weight * (1.0 - a) * (1.0 - b) * c;
            nodal[5] += weight * a * (1.0 - b) * c;
            nodal[6] += weight * (1.0 - a) * b * c;
            nodal[7] += weight * a * b * c;
        } else if (nd_ == 2) {
            double a = *qx, b = *qy;
            nodal[0] += weight * (1.0 - a) * (1.0 - b);
            nodal[1] += weight * a * (1.0 - b);
            nodal[2] += weight * (1.0 - a) * b;
            nodal[3] += weight * a * b;
        } else {
            nodal[0] += weight * (1.0 - *qx);
            nodal[1] += weight * *qx;
        }
    }

    void interior_element(const double *xl, double rc2, double *nodal) const
    {
        const int ng = (kind_ == kUatt) ? 1 : (rc2 <= 4.000000001 ? 6 : 3);
        const GaussRule &g = gauss_rule(ng);
        const int nu = integrand_points(rc2);
        const GaussRule &gu = gauss_rule(nu > 0 ? nu : 1);
        const double inv_r2 = 1.0 / (radius_ * radius_);
        double pt[3];
        for (int i = 0; i < g.n; ++i) {
            pt[0] = xl[0] + g.x[i] * p_.dx[0];
            if (nd_ == 1) {
                double rsq = pt[0] * pt[0] * inv_r2;
                double wt = weight_at(rsq, gu);
                wt *= g.w[i] * el_vol_;
                spread(wt, &g.x[i], nullptr, nullptr, nodal);
                continue;
            }
            for (int j = 0; j < g.n; ++j) {
                pt[1] = xl[1] + g.x[j] * p_.dx[1];
                if (nd_ == 2) {
                    double rsq = (pt[0] * pt[0] + pt[1] * pt[1]) * inv_r2;
                    double wt = weight_at(rsq, gu);
                    wt *= g.w[i] * g.w[j] * el_vol_;
                    spread(wt, &g.x[i], &g.x[j], nullptr, nodal);
                    continue;
                }
                for (int k = 0; k < g.n; ++k) {
                    pt[2] = xl[2] + g.x[k] * p_.dx[2];
                    double rsq = (pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]) * inv_r2;
                    double wt = weight_at(rsq, gu);
                    wt *= g.w[i] * g.w[j] * g.w[k] * el_vol_;
                    spread(wt, &g.x[i], &g.x[j], &g.x[k], nodal);
                }
            }
        }
    }

    void boundary_element(const double *xl, double rc2, double *nodal) const
    {
        int npt;
        if (kind_ == kDelta)
            npt = (nd_ == 3) ? 40 : 150;
        else if (kind_ == kTheta)
            npt = (nd_ == 3) ? 20 : 40;
        else
            npt = (nd_ == 3) ? 1 : 20;
        const int nu = integrand_points(rc2);
        const GaussRule &gu = gauss_rule(nu > 0 ? nu : 1);
        const double inv_npt = 1.0 / (double)npt;
        const double inv_r2 = 1.0 / (radius_ * radius_);
        const bool delta = (kind_ == kDelta);
        double shell = 0.0;
        if (nd_ == 3 && delta) {
            double ss = 0.0;
            for (int d = 0; d < nd_; ++d) ss += p_.dx[d] * p_.dx[d];
            shell = 0.5 * inv_npt * std::sqrt(ss);
        }
        double q[3], pt[3];
        for (int i = 0; i < npt; ++i) {
            q[0] = ((double)i + 0.5) * inv_npt;
            pt[0] = xl[0] + q[0] * p_.dx[0];
            if (nd_ == 1) {
                double rsq = pt[0] * pt[0] * inv_r2;
                if (rsq <= 1.000001) {
                    double wt = weight_at(rsq, gu);
                    if (std::fabs(rsq - 1.0) <= 1.e-8) wt *= 0.5;
                    wt *= inv_npt * el_vol_;
                    spread(wt, &q[0], nullptr, nullptr, nodal);
                }
                continue;
            }
            for (int j = 0; j < npt; ++j) {
                q[1] = ((double)j + 0.5) * inv_npt;
                pt[1] = xl[1] + q[1] * p_.dx[1];
                if (nd_ == 2) {
                    double rsq = (pt[0] * pt[0] + pt[1] * pt[1]) * inv_r2;
                    if (rsq <= 1.0000001) {
                        double wt = weight_at(rsq, gu);
                        if (std::fabs(rsq - 1.0) <= 1.e-8) wt *= 0.5;
                        wt *= inv_npt * inv_npt * el_vol_;
                        spread(wt, &q[0], &q[1], nullptr, nodal);
                    }
                    continue;
                }
                for (int k = 0; k < npt; ++k) {
                    q[2] = ((double)k + 0.5) * inv_npt;
                    pt[2] = xl[2] + q[2] * p_.dx[2];
                    double rsq = (pt[0] * pt[0] + pt[1] * pt[1] + pt[2] * pt[2]) * inv_r2;
                    bool take = delta ? (std::fabs(rsq - 1.0) < 2 * shell) : (rsq <= 1.00000001);
                    if (!take) continue;
                    double wt = weight_at(rsq, gu);
                    if (!delta && std::fabs(rsq - 1.0) <= 1.e-8) wt *= 0.5;
                    wt *= (inv_npt * inv_npt * inv_npt) * el_vol_;
                    spread(wt, &q[0], &q[1], &q[2], nodal);
                }
            }
        }
    }

    const tmi_params &p_;
    Kind kind_;
    int ic_, jc_;
    const Attraction *att_;
    int nd_;
    double radius_ = 0.0, volume_ = 0.0, el_vol_ = 1.0;
};

Attraction make_attraction(const tmi_params &p, const BulkConstants &bc, int i, int j)
{
    Attraction a;
    a.sigma = p.sigma_ff[i][j];
    a.eps = p.eps_ff[i][j];
    a.rcut = p.cut_ff[i][j];
    a.rmin = bc.rmin[i][j];
    return a;
}

}  // namespace

// pairPot_find_rmin, src/dft_pairPot.c:98-148: coarse scan at the grid
// spacing, then successive 100x refinements around the minimum.
double lj_potential_minimum(double sigma, double eps, double rcut, double dx0)
{
    bool found = false;
    double r = 2.0 * dx0, step = dx0, best = lj_cs(r, sigma, eps, rcut), rbest = sigma, rlast = 0.0;
    while (r < rcut) {
        r += step;
        double u = lj_cs(r, sigma, eps, rcut);
        if (u < best) {
            best = u;
            rbest = r;
            found = true;
        }
        rlast = r;
    }
    if (!found || std::fabs(rbest - rlast) < 1.e-6) return sigma;
    double err = 1.0;
    while (err > 1.e-8) {
        double prev = best, step_old = step;
        double start = rbest - step_old;
        r = start;
        best = lj_cs(start, sigma, eps, rcut);
        step = step_old / 100;
        while (r < start + 2.0 * step_old) {
            r += step;
            double u = lj_cs(r, sigma, eps, rcut);
            if (u < best) {
                best = u;
                rbest = r;
            }
        }
        err = 100 * std::fabs(best - prev) / std::fabs(best);
    }
    return rbest;
}

// calc_InvR_params, rhobar_icomp, dphi_drb_bulk, chempot_FMT_hs
// (src/dft_thermo_hs.c:58-75, :292-334, :136-154)
void compute_bulk_constants(const tmi_params &p, BulkConstants &bc)
{
    const int nc = p.ncomp;
    bc.inv_4pi = 1.0 / (4.0 * kPi);
    for (int i = 0; i < nc; ++i) {
        bc.inv_rad[i] = 2.0 / p.hs_diam[i];
        bc.inv_4pir[i] = bc.inv_4pi * bc.inv_rad[i];
        bc.inv_4pirsq[i] = bc.inv_4pir[i] * bc.inv_rad[i];
    }
    for (int q = 0; q < 10; ++q) bc.rhobar_b[q] = bc.dphi_b[q] = 0.0;
    for (int i = 0; i < nc; ++i) {
        double d = p.hs_diam[i];
        double vol = kPi * d * d * d / 6.0, area = kPi * d * d, rho = p.rho_b[i];
        bc.rhobar_b[0] += 1.0 * vol * rho;
        bc.rhobar_b[1] += 1.0 * area * rho;
        bc.rhobar_b[2] += 1.0 * area * rho * bc.inv_4pir[i];
        bc.rhobar_b[3] += 1.0 * area * rho * bc.inv_4pirsq[i];
    }
    // first derivatives at the bulk state (vector densities vanish)
    const double n0 = bc.rhobar_b[3], n1 = bc.rhobar_b[2], n2 = bc.rhobar_b[1], n3 = bc.rhobar_b[0];
    const double a = 1.0 / (1.0 - n3), a2 = a * a, a3 = a2 * a;
    bc.dphi_b[0] = std::log(a);
    bc.dphi_b[1] = n2 * a;
    if (p.type_func == TMI_FMT1 || p.type_func == TMI_FMT2) {   // FMT2 equals FMT1 at nv2 = 0
        bc.dphi_b[2] = n1 * a + (n2 * n2) * a2 / (8.0 * kPi);
        bc.dphi_b[3] = n0 * a + (n1 * n2) * a2 + (n2 * n2 * n2) * a3 / (12.0 * kPi);
    } else if (n3 > 1.e-10) {   // FMT3, and FMT4, which equals it at nv2 = 0
        const double L = std::log(1. - n3), g = n3 + (1. - n3) * (1. - n3) * L;
        bc.dphi_b[2] = n1 * a + (n2 * n2) * a2 * g / (12.0 * kPi * n3 * n3);
        bc.dphi_b[3] = n0 * a + (n1 * n2) * a2 +
                       (((n2 * n2 * n2) * a2) / (36. * kPi * n3 * n3)) * (+2. * g * a + (-2. * (1. - n3) * L + n3) - 2. * (g / n3));
    }
    for (int i = 0; i < nc; ++i) {
        double R = p.hs_diam[i] / 2.;
        double ss[4] = {1., R, (4. * kPi) * std::pow(R, 2), (4. * kPi / 3.) * std::pow(R, 3)};
        bc.betamu_hs_ex[i] = 0.0;
        for (int q = 0; q < 4; ++q) bc.betamu_hs_ex[i] += bc.dphi_b[q] * ss[q] * 1.0;
    }
}

void build_all_stencils(const tmi_params &p, const BulkConstants &bc, HostStencil (*st)[kNC * kNC])
{
    const int nc = p.ncomp;
    for (int i = 0; i < nc; ++i) {
        st[TMI_DELTA][i] = StencilBuilder(p, kDelta, i, 0, nullptr).build();
        st[TMI_THETA][i] = StencilBuilder(p, kTheta, i, 0, nullptr).build();
    }
    if (p.type_attr >= 0) {
        for (int i = 0; i < nc; ++i)
            for (int j = 0; j < nc; ++j) {
                Attraction a = make_attraction(p, bc, i, j);
                st[TMI_UATT][i + nc * j] = StencilBuilder(p, kUatt, i, j, &a).build();
            }
    }
}

// calc_Avdw_att / chempot_att (src/dft_thermo_att.c:44-61, :86-97)
void compute_attraction_constants(const tmi_params &p, BulkConstants &bc, HostStencil (*st)[kNC * kNC])
{
    const int nc = p.ncomp;
    for (int i = 0; i < nc; ++i)
        for (int j = 0; j < nc; ++j) {
            double s = 0.0;
            for (double w : st[TMI_UATT][i + nc * j].w) s += w;
            bc.avdw[i][j] = s;
        }
    for (int i = 0; i < nc; ++i) {
        bc.betamu_att[i] = 0.0;
        for (int j = 0; j < nc; ++j) bc.betamu_att[i] += p.rho_b[j] * bc.avdw[i][j];
    }
}

}  // namespace tmi
