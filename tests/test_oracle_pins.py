"""Independent checks that pin each oracle function (CPU only).

* Gauss rules: polynomial exactness.
* Stencils: 1D closed forms of the projected delta/theta kernels, the
  attraction-stencil volume against direct numerical integration of
  4 pi r^2 u(r), point symmetry, r_min against 2^(1/6) sigma.
* FMT: first derivatives = finite differences of Phi; the contracted second
  derivatives = kernel vector . finite differences of the first derivatives;
  bulk chemical potential and pressure = Percus-Yevick (Rosenfeld) and
  Carnahan-Starling / BMCSL (White Bear) closed forms.
* Residual: a uniform bulk state gives F = 0; on a periodic grid the
  Euler-Lagrange rows equal the gradient of an independently assembled
  discrete grand potential (numpy), which fixes every sign, prefactor and
  the adjoint stencil direction.
* Jacobian: central differences of the residual, entry by entry.
"""
import math

import ctypes as ct

import numpy as np
import pytest
from scipy import integrate

from oracle.oracle import Oracle
from tramonto_amd.problem import FMT1, FMT3, IN_BULK, PERIODIC, ProblemSpec, small_case

PI = math.pi


def one_comp_1d(nx=81, dx=0.05, bc=(IN_BULK, IN_BULK), func=FMT1, attr=-1, rho_b=0.6, cut=2.5, eps=1.0):
    return ProblemSpec(name="t", ndim=1, ncomp=1, nx=[nx], dx=[dx], bc=[list(bc)], type_func=func, type_attr=attr,
                       hs_diam=[1.0], sigma_ff=np.ones((1, 1)), eps_ff=np.full((1, 1), eps),
                       cut_ff=np.full((1, 1), cut), rho_b=[rho_b], rho0=np.full((1, nx), rho_b))


# ------------------------------------------------------------------ stencils
def hat_integral(f, xk, dx, lo, hi):
    """int f(x) phi_k(x) dx over [lo, hi] for the hat function at xk."""
    a = integrate.quad(lambda x: f(x) * (1 - abs(x - xk) / dx), max(lo, xk - dx), min(hi, xk), epsabs=1e-15,
                       epsrel=1e-15)[0]
    b = integrate.quad(lambda x: f(x) * (1 - abs(x - xk) / dx), max(lo, xk), min(hi, xk + dx), epsabs=1e-15,
                       epsrel=1e-15)[0]
    return a + b


@pytest.mark.parametrize("dx", [0.05, 0.1, 0.025])
def test_1d_delta_and_theta_closed_form(dx):
    o = Oracle.from_spec(one_comp_1d(dx=dx))
    R = 0.5
    off, w = o.stencil(0)
    exact = np.array([hat_integral(lambda x: 2 * PI * R, k * dx, dx, -R, R) for k in off[:, 0]])
    assert np.allclose(w, exact, rtol=1e-13, atol=0)
    off, w = o.stencil(1)
    exact = np.array([hat_integral(lambda x: PI * (R * R - x * x), k * dx, dx, -R, R) for k in off[:, 0]])
    assert np.allclose(w, exact, rtol=1e-12, atol=1e-15)


def test_stencil_sums_and_symmetry():
    for case in ["2d_periodic", "3d_periodic_small", "3d_mix_bulk"]:
        sp = small_case(case)
        o = Oracle.from_spec(sp)
        for i in range(sp.ncomp):
            d = sp.hs_diam[i]
            off, w = o.stencil(0, i)
            assert abs(w.sum() - PI * d * d) < 1e-13 * PI * d * d
            key = {tuple(r): v for r, v in zip(off.tolist(), w)}
            for r, v in key.items():
                assert abs(key[tuple(-x for x in r)] - v) <= 1e-12 * abs(v)
            off, w = o.stencil(1, i)
            assert abs(w.sum() - PI * d ** 3 / 6) < 1e-13 * PI * d ** 3 / 6


def test_attraction_volume_and_rmin():
    sp = one_comp_1d(dx=0.025, attr=0, cut=3.0)
    o = Oracle.from_spec(sp)
    b = o.bulk()
    rmin = b[44]
    assert abs(rmin - 2 ** (1 / 6)) < 1e-3
    rc = 3.0

    def u(r):
        rr = max(r, rmin)
        lj = lambda s: 4 * (s ** -12 - s ** -6)
        return lj(rr) - lj(rc) if r <= rc else 0.0
    vol = integrate.quad(lambda r: 4 * PI * r * r * u(r), 0, rc, points=[rmin], epsabs=1e-13, limit=200)[0]
    assert abs(b[28] - vol) < 1e-10 * abs(vol)         # Avdw = renormalised stencil sum = analytic volume
    off, w = o.stencil(2)
    # the renormalisation only corrects quadrature error
    # (1D weight of node k: projected potential 2 pi int_|x|^rc u(r) r dr)
    proj = lambda x: 2 * PI * integrate.quad(lambda r: u(r) * r, abs(x), rc, points=[rmin] if abs(x) < rmin else None,
                                             limit=200)[0]
    mid = np.array([proj(k * 0.025) * 0.025 for k in off[:, 0]])
    assert np.abs(w - mid).max() < 2e-3 * np.abs(mid).max()


@pytest.mark.parametrize("n", [1, 3, 6, 12, 20, 40])
def test_gauss_rules_exact(n):
    """each n-point rule of the stencils (set_gauss_quad, src/dft_quadrature.c:40-159) integrates x^k on
    [0, 1] exactly for k < 2n and not for k = 2n: Gauss-Legendre of that order.  The 3-point rule keeps
    the reference's 10-digit abscissa 0.7745966692, hence its 1e-10."""
    from oracle.oracle import lib
    gp, gw = np.zeros(n), np.zeros(n)
    lib().tmo_gauss_rule(n, gp.ctypes.data_as(ct.c_void_p), gw.ctypes.data_as(ct.c_void_p))
    tol = 1e-10 if n == 3 else 2e-14
    assert np.all((gp > 0) & (gp < 1)) and np.all(gw > 0)
    for k in range(2 * n):
        assert abs(np.dot(gw, gp ** k) - 1.0 / (k + 1)) < tol, (n, k)
    if n <= 6:   # (beyond, the degree-2n error is below round-off)
        assert abs(np.dot(gw, gp ** (2 * n)) - 1.0 / (2 * n + 1)) > 1e-8


# ------------------------------------------------------------------ FMT physics
def rand_n(rng, nd, scale=1.0):
    """A state well inside the physical region (n3 < 1, |nv| < n2)."""
    eta = rng.uniform(0.1, 0.45)
    R = 0.5
    n3 = eta
    n2 = 3 * eta / R
    n1 = n2 / (4 * PI * R)
    n0 = n2 / (4 * PI * R * R)
    v2 = rng.uniform(-0.3, 0.3, nd) * n2 * scale
    v1 = v2 / (4 * PI * R) * rng.uniform(0.8, 1.2, nd)
    return np.concatenate([[n0, n1, n2, n3], v1, v2])


@pytest.mark.parametrize("func", [FMT1, 1, FMT3, 3])
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_fmt_first_derivatives_fd(func, nd):
    sp = one_comp_1d(func=func)
    sp.ndim = nd
    sp.nx = [12] * nd
    sp.dx = [0.1] * nd
    sp.bc = [[PERIODIC, PERIODIC]] * nd
    sp.rho0 = np.full((1, 12 ** nd), 0.6)
    o = Oracle.from_spec(sp)
    rng = np.random.default_rng(3)
    for _ in range(5):
        n = rand_n(rng, nd)
        g = o.dphi(n)
        for a in range(len(n)):
            h = 1e-6 * max(1.0, abs(n[a]))
            e = np.zeros_like(n)
            e[a] = h
            fd = (o.phi(n + e) - o.phi(n - e)) / (2 * h)
            assert abs(fd - g[a]) < 1e-7 * max(1.0, abs(g[a])), (a, fd, g[a])


@pytest.mark.parametrize("func", [FMT1, FMT3])
def test_fmt_second_derivatives_fd(func):
    nd = 3
    sp = one_comp_1d(func=func)
    sp.ndim, sp.nx, sp.dx, sp.bc = nd, [12] * 3, [0.1] * 3, [[PERIODIC, PERIODIC]] * 3
    sp.rho0 = np.full((1, 12 ** 3), 0.6)
    o = Oracle.from_spec(sp)
    rng = np.random.default_rng(4)
    R = 0.5
    for _ in range(4):
        n = rand_n(rng, nd)
        H = np.zeros((len(n), len(n)))
        for b in range(len(n)):
            h = 1e-6 * max(1.0, abs(n[b]))
            e = np.zeros_like(n)
            e[b] = h
            H[:, b] = (o.dphi(n + e) - o.dphi(n - e)) / (2 * h)
        assert np.allclose(H, H.T, atol=1e-6 * np.abs(H).max())
        # theta kernel: row 3 of H
        t = o.d2_theta(n)
        assert np.allclose(t, H[3], rtol=1e-6, atol=1e-6 * np.abs(H[3]).max())
        # delta kernel vector c = (1/4piR^2, 1/4piR, 1, 0, -v/4piR, -v) with v = off dx / R
        off = rng.integers(-5, 6, nd).astype(np.int32)
        v = off * 0.1 / R
        c = np.concatenate([[1 / (4 * PI * R * R), 1 / (4 * PI * R), 1.0, 0.0], -v / (4 * PI * R), -v])
        d = o.d2_delta(n, off, np.ones(nd))
        ref = c @ H
        assert np.allclose(d, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
        if func == FMT1:
            # mirror image: vector densities and kernel-vector components change sign
            sg = np.array([-1.0, 1.0, 1.0])
            s_full = np.concatenate([[1, 1, 1, 1], sg, sg])
            nm = n * s_full
            Hm = np.zeros_like(H)
            for b in range(len(n)):
                h = 1e-6 * max(1.0, abs(nm[b]))
                e = np.zeros_like(n)
                e[b] = h
                Hm[:, b] = (o.dphi(nm + e) - o.dphi(nm - e)) / (2 * h)
            ref = (c @ Hm) * s_full
            d = o.d2_delta(n, off, sg)
            assert np.allclose(d, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())


def _hessian_fd(o, n):
    H = np.zeros((len(n), len(n)))
    for b in range(len(n)):
        h = 1e-6 * max(1.0, abs(n[b]))
        e = np.zeros_like(n)
        e[b] = h
        H[:, b] = (o.dphi(n + e) - o.dphi(n - e)) / (2 * h)
    return H


@pytest.mark.parametrize("func", [1, 3])
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_fmt2_fmt4_second_derivatives(func, nd):
    """FMT2 and FMT4 (src/dft_physics_FMT2.c:106-250, FMT4.c:101-276, restated by the oracle): the theta
    Hessian row and every delta-Hessian column except nv2 equal finite differences of the first
    derivatives.  The nv2 columns of the delta Hessian do not (FMT2 in 2D and 3D, FMT4 in every
    dimension): the reference's assembled Jacobian is inexact there, which is why the device, whose
    product is the exact derivative, is compared with it on the other columns only."""
    sp = one_comp_1d(func=func)
    sp.ndim, sp.nx, sp.dx, sp.bc = nd, [12] * nd, [0.1] * nd, [[PERIODIC, PERIODIC]] * nd
    sp.rho0 = np.full((1, 12 ** nd), 0.6)
    o = Oracle.from_spec(sp)
    rng = np.random.default_rng(4)
    R = 0.5
    worst_nv2 = 0.0
    for _ in range(4):
        n = rand_n(rng, nd)
        H = _hessian_fd(o, n)
        t = o.d2_theta(n)
        assert np.allclose(t, H[3], rtol=1e-6, atol=1e-6 * np.abs(H[3]).max())
        off = rng.integers(-5, 6, nd).astype(np.int32)
        v = off * 0.1 / R
        c = np.concatenate([[1 / (4 * PI * R * R), 1 / (4 * PI * R), 1.0, 0.0], -v / (4 * PI * R), -v])
        d = o.d2_delta(n, off, np.ones(nd))
        ref = c @ H
        scale = np.abs(ref).max()
        assert np.allclose(d[: 4 + nd], ref[: 4 + nd], rtol=1e-6, atol=1e-6 * scale)
        worst_nv2 = max(worst_nv2, float(np.abs(d[4 + nd:] - ref[4 + nd:]).max() / scale))
    if func == 1 and nd == 1:
        assert worst_nv2 < 1e-6
    else:
        assert worst_nv2 > 1e-3


@pytest.mark.parametrize("rho", [0.1, 0.5, 0.8])
def test_bulk_fmt2_and_fmt4_reduce_to_rosenfeld_and_white_bear(rho):
    """at nv2 = 0, FMT2 is FMT1 (Percus-Yevick) and FMT4 is White Bear (Carnahan-Starling)"""
    for func, other in ((1, FMT1), (3, FMT3)):
        b = Oracle.from_spec(one_comp_1d(rho_b=rho, func=func)).bulk()
        bo = Oracle.from_spec(one_comp_1d(rho_b=rho, func=other)).bulk()
        assert np.allclose(b, bo, rtol=1e-14, atol=1e-14)
    eta = PI / 6 * rho
    mu_cs = eta * (8 - 9 * eta + 3 * eta * eta) / (1 - eta) ** 3
    assert abs(Oracle.from_spec(one_comp_1d(rho_b=rho, func=3)).bulk()[20] - mu_cs) < 1e-11 * max(1, abs(mu_cs))


def py_mu_p(etas_by_comp, d):
    """Percus-Yevick (compressibility) mixture: beta mu_ex_i, beta p."""
    xi = [sum(PI / 6 * r * di ** k for r, di in zip(etas_by_comp, d)) for k in range(4)]
    one = 1 - xi[3]
    p = 6 / PI * (xi[0] / one + 3 * xi[1] * xi[2] / one ** 2 + 3 * xi[2] ** 3 / one ** 3)
    mu = [-math.log(one) + PI / 6 * p * di ** 3 + 3 * xi[2] * di / one + 3 * xi[1] * di ** 2 / one
          + 4.5 * (xi[2] * di) ** 2 / one ** 2 for di in d]
    return mu, p


def test_oracle_reports_bad_problems():
    sp = one_comp_1d(nx=5, dx=0.1, bc=(PERIODIC, PERIODIC))
    with pytest.raises(ValueError, match="periodic"):
        Oracle.from_spec(sp)


@pytest.mark.parametrize("rho", [0.1, 0.5, 0.8])
def test_bulk_rosenfeld_is_percus_yevick(rho):
    o = Oracle.from_spec(one_comp_1d(rho_b=rho))
    b = o.bulk()
    mu, p = py_mu_p([rho], [1.0])
    assert abs(b[20] - mu[0]) < 1e-12 * max(1, abs(mu[0]))


def test_bulk_rosenfeld_mixture_is_percus_yevick():
    sp = small_case("1d_pore_mix")
    sp.type_func = FMT1
    o = Oracle.from_spec(sp)
    b = o.bulk()
    mu, p = py_mu_p(sp.rho_b, sp.hs_diam)
    for i in range(2):
        assert abs(b[20 + i] - mu[i]) < 1e-12 * max(1, abs(mu[i]))


@pytest.mark.parametrize("rho", [0.1, 0.5, 0.8])
def test_bulk_white_bear_is_carnahan_starling(rho):
    o = Oracle.from_spec(one_comp_1d(rho_b=rho, func=FMT3))
    b = o.bulk()
    eta = PI / 6 * rho
    mu_cs = eta * (8 - 9 * eta + 3 * eta * eta) / (1 - eta) ** 3
    assert abs(b[20] - mu_cs) < 1e-11 * max(1, abs(mu_cs))


# ------------------------------------------------------------------ residual
@pytest.mark.parametrize("case", ["2d_periodic", "3d_periodic_small", "3d_periodic_wb", "1d_reflect_wca"])
def test_uniform_bulk_state_has_zero_residual(case):
    sp = small_case(case)
    sp.rho0 = np.tile(np.array(sp.rho_b)[:, None], (1, sp.nnodes))
    o = Oracle.from_spec(sp)
    x = o.init_fields(sp.initial_x())
    F = o.residual(x)
    assert np.abs(F).max() < 1e-12 * 10


def test_uniform_mixture_with_bulk_boundaries_has_zero_residual():
    sp = small_case("1d_pore_mix")
    sp.bc = [[IN_BULK, IN_BULK]]
    sp.nwall = 0
    sp.vext = None
    sp.zero_tf = None
    sp.rho0 = np.tile(np.array(sp.rho_b)[:, None], (1, sp.nnodes))
    o = Oracle.from_spec(sp)
    x = o.init_fields(sp.initial_x())
    assert np.abs(o.residual(x)).max() < 1e-12 * 10


def periodic_conv(field, off, w, shape):
    out = np.zeros(shape)
    f = field.reshape(shape[::-1]).transpose() if len(shape) > 1 else field
    for (k, wk) in zip(off, w):
        out += wk * np.roll(f, tuple(-k[: len(shape)]), axis=tuple(range(len(shape))))
    return out


@pytest.mark.parametrize("case", ["2d_periodic", "3d_periodic_small", "3d_periodic_wb"])
def test_el_rows_are_grand_potential_gradient(case):
    """F_EL(i) = d Omega / d rho_i with Omega = sum_j Phi(n(rho)_j)
    + sum_i rho_i (ln rho_i - 1 - ln rho_b - mu_hs - mu_att) + 1/2 sum rho u rho,
    n and the mean field from numpy convolutions with the oracle's stencils."""
    sp = small_case(case)
    o = Oracle.from_spec(sp)
    shape = tuple(sp.nx[: sp.ndim])
    nd = sp.ndim
    rng = np.random.default_rng(8)
    rho = sp.rho0[0] * (1 + 0.05 * rng.uniform(-1, 1, sp.nnodes))
    off_d, w_d = o.stencil(0)
    off_t, w_t = o.stencil(1)
    off_u, w_u = o.stencil(2)
    R = 0.5
    b = o.bulk()
    mu = -math.log(sp.rho_b[0]) - b[20] - b[24]

    def to_grid(a):
        return a.reshape(shape[::-1]).transpose()

    def omega(r):
        g = to_grid(r)
        n3 = periodic_conv(r, off_t, w_t, shape)
        n2 = periodic_conv(r, off_d, w_d, shape)
        nv2 = [periodic_conv(r, off_d, w_d * off_d[:, d] * sp.dx[d] / R, shape) for d in range(nd)]
        n = [n2 / (4 * PI * R * R), n2 / (4 * PI * R), n2, n3] + [v / (4 * PI * R) for v in nv2] + nv2
        phi = np.vectorize(lambda *a: o.phi(np.array(a)))(*n)
        mf = periodic_conv(r, off_u, w_u, shape)
        return phi.sum() + (g * (np.log(g) - 1 + mu)).sum() + 0.5 * (g * mf).sum()

    x = sp.initial_x()
    x[0] = rho
    x = o.init_fields(x)
    F = o.residual(x)[0]
    def shifted(node, t):
        r = rho.copy()
        r[node] += t
        return omega(r)

    for node in rng.integers(0, sp.nnodes, 4):
        h = 1e-4   # fourth-order central difference
        fd = (-shifted(node, 2 * h) + 8 * shifted(node, h) - 8 * shifted(node, -h) + shifted(node, -2 * h)) / (12 * h)
        assert abs(fd - F[node]) < 1e-7 * max(1.0, abs(F[node])), (node, fd, F[node])


# ------------------------------------------------------------------ Jacobian
@pytest.mark.parametrize("case", ["1d_hw", "1d_reflect_wca", "1d_pore_mix"])
def test_jacobian_entries_central_differences(case):
    """do_numerical_jacobian-style check (src/dft_newton.c:405-560)."""
    sp = small_case(case)
    o = Oracle.from_spec(sp)
    x = o.init_fields(sp.initial_x())
    rng = np.random.default_rng(9)
    x[sp.ncomp:] *= 1 + 0.01 * rng.uniform(-1, 1, x[sp.ncomp:].shape)
    J = o.jacobian_dense(x)
    n = J.shape[0]
    cols = rng.choice(n, size=min(n, 60), replace=False)
    for j in cols:
        xp = x.copy().reshape(-1)
        xm = x.copy().reshape(-1)
        h = 1e-7 * max(1.0, abs(xp[j]))
        xp[j] += h
        xm[j] -= h
        fd = (o.residual(xp.reshape(x.shape)) - o.residual(xm.reshape(x.shape))).reshape(-1) / (2 * h)
        scale = max(1.0, np.abs(J[:, j]).max())
        assert np.abs(fd - J[:, j]).max() < 1e-5 * scale, j


def test_schur_block_a11_is_minus_identity():
    sp = small_case("1d_pore_mix")
    o = Oracle.from_spec(sp)
    x = o.init_fields(sp.initial_x())
    J = o.jacobian_dense(x)
    n2 = sp.ncomp * sp.nnodes
    assert np.array_equal(J[n2:, n2:], -np.eye(J.shape[0] - n2))


def test_residual_nodes_equals_full_residual_columns():
    """tmo_residual_nodes (the bounded-sample fill) gives the listed columns of tmo_residual."""
    from oracle.oracle import Oracle
    from tramonto_amd.problem import small_case
    for case in ("1d_hw", "2d_rough_wall", "3d_solute_small"):
        sp = small_case(case)
        o = Oracle.from_spec(sp)
        x = o.init_fields(sp.initial_x())
        x[sp.ncomp:] *= 1.01
        F = o.residual(x).reshape(o.nunk, -1)
        nodes = np.random.default_rng(3).choice(sp.nnodes, 37, replace=False)
        assert np.array_equal(o.residual_nodes(x, nodes), F[:, nodes])
