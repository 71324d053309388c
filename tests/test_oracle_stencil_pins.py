"""The 2D and 3D stencil weights against their exact integrals (CPU).

The reference's weight of stencil point k is the integral of the kernel (the sphere's surface for
delta, the ball for theta, both projected onto the plane in 2D) against the hat function of node
k, approximated by Gauss points on elements inside the sphere and mid-point sub-cells on the
elements it crosses, then renormalised (src/dft_stencil.c:329-496, 653-683; kernels
src/dft_stencil_delta_R.c:70-78, src/dft_stencil_theta_R.c:66-74).  tests/stencil_exact.py
evaluates the same integrals to ~1e-13 by closed-form circle averages of the hat and kink-split
Gauss sums; these tests pin the oracle's (and, by the bitwise test in test_capi.py, the library's)
weights against them:

* which points carry weight: exactly the nodes whose hat support meets the sphere (delta, 3D) or
  the ball; at sigma = 1, dx = 0.1 sigma that is 844 delta and 1015 theta points in 3D;
* points whose whole hat support lies inside the sphere: weight / exact weight is one constant
  (the renormalisation factor) to 1e-6, the interior Gauss rule's accuracy;
* every weight within 1.5x the reference rule's own error at these geometries (the mid-point
  sub-cells on crossing elements, and the 3D delta shell band);
* the 3D weights summed over y and z offsets reproduce the 1D closed forms (2 pi R, pi (R^2 - x^2))
  to the 3D rule's error, and a 3D slab uniform in y and z gives the 1D residual to 1e-4;
* the same planar wall along x, y or z gives permuted-identical residuals.

Each tolerance is shown to have teeth: halving the crossing-element sub-points, lowering the
interior Gauss order from 6 to 3, or narrowing the 3D delta shell band (tmo_set_stencil_rule, an
oracle hook for these tests) breaks it.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle, lib
from stencil_exact import exact_weights, support_points
from tramonto_amd.problem import IN_BULK, IN_WALL, PERIODIC, FMT1, ProblemSpec

R = 0.5

# max |w - w_exact| / max w_exact of the reference rule at sigma = 1 (measured: 2D delta 4.81e-3 /
# 5.16e-3, 2D theta 8.48e-5 / 8.31e-5 at dx 0.1 / 0.05; 3D delta 1.08e-2, 3D theta 2.69e-3 at dx 0.1),
# times 1.5
TOL = {(2, "delta", 0.1): 7.5e-3, (2, "delta", 0.05): 7.5e-3, (2, "theta", 0.1): 1.3e-4,
       (2, "theta", 0.05): 1.3e-4, (3, "delta", 0.1): 1.6e-2, (3, "theta", 0.1): 4.0e-3}


def stencil(nd, dx, t, rule=(0, 0, 0.0)):
    sp = ProblemSpec(name="s", ndim=nd, ncomp=1, nx=[32] * nd, dx=[dx] * nd, bc=[[PERIODIC, PERIODIC]] * nd,
                     hs_diam=[2 * R], sigma_ff=np.ones((1, 1)), rho_b=[0.5])
    lib().tmo_set_stencil_rule(*rule)
    try:
        off, w = Oracle.from_spec(sp).stencil(t, 0, 0)
    finally:
        lib().tmo_set_stencil_rule(0, 0, 0.0)
    return {tuple(int(v) for v in o[:nd]): float(x) for o, x in zip(np.asarray(off), np.asarray(w))}


_EXACT = {}


def exact(nd, kind, dx):
    key = (nd, kind, dx)
    if key not in _EXACT:
        reach = int((R / dx) + 1)
        pts = support_points(kind, nd, [dx] * 3, R, reach)
        _EXACT[key] = dict(zip(pts, exact_weights(kind, nd, np.array(pts), [dx] * 3, R)))
    return _EXACT[key]


def max_err(w, we):
    keys = set(w) | set(we)
    return max(abs(w.get(k, 0.0) - we.get(k, 0.0)) for k in keys) / max(we.values())


def interior_spread(w, we, nd, dx):
    """spread of w / w_exact over points whose hat support lies inside the sphere"""
    r = [w[k] / we[k] for k in we if np.sqrt(np.sum(((np.abs(np.array(k)) + 1) * dx) ** 2)) < R]
    r = np.array(r)
    return (r.max() - r.min()) / r.mean(), len(r)


CASES = sorted(TOL)


@pytest.mark.parametrize("nd,kind,dx", CASES)
def test_exact_integrals_are_exact(nd, kind, dx):
    """the checker itself: exact weights sum to the analytic total (the hats sum to 1), and in 3D a
    point whose hat support lies inside the ball weighs exactly dx^3"""
    we = exact(nd, kind, dx)
    total = np.pi * (2 * R) ** 2 if kind == "delta" else np.pi * (2 * R) ** 3 / 6
    assert abs(sum(we.values()) - total) < 1e-12 * total
    if nd == 3 and kind == "theta":
        inner = [v for k, v in we.items() if np.sqrt(np.sum(((np.abs(np.array(k)) + 1) * dx) ** 2)) < R]
        assert len(inner) > 100 and np.abs(np.array(inner) / dx ** 3 - 1.0).max() < 1e-13


@pytest.mark.parametrize("nd,kind,dx", CASES)
def test_points_are_exactly_the_hat_supports_meeting_the_kernel(nd, kind, dx):
    w = stencil(nd, dx, 0 if kind == "delta" else 1)
    we = exact(nd, kind, dx)
    assert set(w) == set(we)
    if (nd, dx) == (3, 0.1):
        assert len(w) == (844 if kind == "delta" else 1015)


@pytest.mark.parametrize("nd,kind,dx", CASES)
def test_weights_match_exact_integrals(nd, kind, dx):
    w = stencil(nd, dx, 0 if kind == "delta" else 1)
    we = exact(nd, kind, dx)
    assert max_err(w, we) < TOL[(nd, kind, dx)], max_err(w, we)


@pytest.mark.parametrize("nd,kind,dx", [c for c in CASES if c[0] == 2])
def test_interior_gauss_rule(nd, kind, dx):
    w = stencil(nd, dx, 0 if kind == "delta" else 1)
    spread, n = interior_spread(w, exact(nd, kind, dx), nd, dx)
    assert n > 30 and spread < 1e-6, spread


@pytest.mark.parametrize("mutation", ["gauss3", "npt_half", "band1"])
@pytest.mark.parametrize("nd,kind,dx", [(2, "delta", 0.1), (2, "theta", 0.1), (2, "delta", 0.05), (2, "theta", 0.05),
                                        (3, "delta", 0.1), (3, "theta", 0.1)])
def test_pins_have_teeth(nd, kind, dx, mutation):
    """a plausible mistake in the rule fails one of the pins above"""
    t = 0 if kind == "delta" else 1
    npt = {(2, 0): 150, (2, 1): 40, (3, 0): 40, (3, 1): 20}[(nd, t)]
    rule = {"gauss3": (3, 0, 0.0), "npt_half": (0, npt // 2, 0.0), "band1": (0, 0, 1.0)}[mutation]
    if mutation == "band1" and (nd, kind) != (3, "delta"):
        pytest.skip("the shell band exists for 3D delta stencils only")
    if mutation == "gauss3" and nd == 3:
        pytest.skip("3D: the delta rule has no interior elements and the theta kernel is constant, so every "
                    "Gauss order integrates the trilinear hat exactly")
    w = stencil(nd, dx, t, rule)
    we = exact(nd, kind, dx)
    caught = max_err(w, we) >= TOL[(nd, kind, dx)]
    if nd == 2:
        caught = caught or interior_spread(w, we, nd, dx)[0] >= 1e-6
    assert caught, (mutation, max_err(w, we))


def one_d_weights(kind, dx):
    """the oracle's 1D weights, pinned by their closed forms in test_oracle_pins.py"""
    sp = ProblemSpec(name="s", ndim=1, ncomp=1, nx=[64], dx=[dx], bc=[[PERIODIC, PERIODIC]],
                     hs_diam=[2 * R], sigma_ff=np.ones((1, 1)), rho_b=[0.5])
    off, w = Oracle.from_spec(sp).stencil(0 if kind == "delta" else 1, 0, 0)
    return {int(o[0]): float(x) for o, x in zip(np.asarray(off), np.asarray(w))}


@pytest.mark.parametrize("kind", ["delta", "theta"])
def test_3d_weights_project_to_the_1d_closed_forms(kind):
    """sum over (ky, kz) of the 3D weight of (kx, ky, kz) = the 1D weight of kx: exactly for the exact
    integrals (the y and z hats sum to 1), within the 3D rule's error for the reference rule"""
    dx = 0.1
    w1 = one_d_weights(kind, dx)
    we = exact(3, kind, dx)
    w3 = stencil(3, dx, 0 if kind == "delta" else 1)
    pe, pr = {}, {}
    for k, v in we.items():
        pe[k[0]] = pe.get(k[0], 0.0) + v
    for k, v in w3.items():
        pr[k[0]] = pr.get(k[0], 0.0) + v
    scale = max(w1.values())
    assert max(abs(pe[k] - w1[k]) for k in w1) < 1e-4 * scale     # 1D rule error only
    assert max(abs(pr[k] - w1[k]) for k in w1) < (2e-2 if kind == "delta" else 5e-3) * scale


def slab_spec(nd, dx, axis=0):
    """hard spheres at a hard planar wall (IN_WALL, V_ext = VEXT_MAX and zero density within sigma/2)
    normal to `axis`, periodic along the other axes (one stencil width); the same seeded profile along
    the normal"""
    from tramonto_amd.problem import SplitMix64
    n_norm = int(round(4.0 / dx)) + 1
    n_per = 2 * int(R / dx + 1) + 1
    nx = [n_per] * nd
    nx[axis] = n_norm
    bc = [[PERIODIC, PERIODIC] for _ in range(nd)]
    bc[axis] = [IN_WALL, IN_BULK]
    z = np.arange(n_norm) * dx
    excl = z < 0.5 - 1e-12
    prof = 0.6 * (1 + 0.1 * SplitMix64(3).uniform(n_norm, -1, 1))
    prof[excl] = 0.0
    shape = [1] * nd
    shape[nd - 1 - axis] = n_norm                   # arrays [z][y][x]: node index i + j nx + k nx ny
    rho = np.broadcast_to(prof.reshape(shape), nx[::-1]).reshape(-1)
    zero = np.broadcast_to(excl.reshape(shape), nx[::-1]).reshape(-1).astype(np.uint8)
    return ProblemSpec(name="slab", ndim=nd, ncomp=1, nx=nx, dx=[dx] * nd, bc=bc, type_func=FMT1, type_attr=-1,
                       nwall=1, vext_max=10.0, hs_diam=[2 * R], sigma_ff=np.ones((1, 1)), rho_b=[0.6],
                       vext=np.where(zero == 1, 10.0, 0.0)[:, None], zero_tf=zero[:, None], rho0=rho[None, :])


def swap_axes(a, axis):
    """the fields [u][z][y][x] (one component, 3D) with axis 0 (x) exchanged with `axis`: arrays
    transposed, vector components of rho-bar (rows 5-7 nv2, 8-10 nv1) exchanged; an involution"""
    if axis == 0:
        return a
    perm = {1: (0, 1, 3, 2), 2: (0, 3, 2, 1)}[axis]
    comp = {1: (1, 0, 2), 2: (2, 1, 0)}[axis]
    order = list(range(5)) + [5 + p for p in comp] + [8 + p for p in comp]
    return np.transpose(a, perm)[order]


def test_planar_wall_along_each_axis_gives_permuted_residuals():
    """F of the same wall normal to x, y or z, at the same (permuted) state, is the same field with the
    axes permuted: the 3D stencils are symmetric under axis exchange on a cubic grid"""
    dx = 0.1
    s0 = slab_spec(3, dx, 0)
    o0 = Oracle.from_spec(s0)
    x0 = o0.init_fields(s0.initial_x())
    rng = np.random.default_rng(5)
    x0[1:] *= 1.0 + 0.02 * rng.uniform(-1.0, 1.0, x0[1:].shape)
    F0 = o0.residual(x0).reshape(s0.nunk, *s0.nx[::-1])
    scale = np.abs(F0).max()
    for axis in (1, 2):
        sp = slab_spec(3, dx, axis)
        o = Oracle.from_spec(sp)
        xa = np.ascontiguousarray(swap_axes(x0.reshape(s0.nunk, *s0.nx[::-1]), axis)).reshape(sp.nunk, -1)
        assert np.array_equal(xa[0], sp.rho0[0])
        Fa = o.residual(xa).reshape(sp.nunk, *sp.nx[::-1])
        assert np.abs(swap_axes(Fa, axis) - F0).max() < 1e-13 * scale, axis


def test_slab_uniform_in_y_and_z_approaches_the_1d_residual():
    """a 3D wall slab uniform in y and z, at the 1D problem's state: its residual rows equal the 1D
    ones to the 3D rule's error (1e-4 of max |F| at dx 0.1 and 0.05 sigma: the rule's sub-point
    counts are per element, so its relative error does not shrink with dx); the y and z vector rows
    vanish by symmetry"""
    diffs = []
    for dx in (0.1, 0.05):
        s1, s3 = slab_spec(1, dx), slab_spec(3, dx)
        o1, o3 = Oracle.from_spec(s1), Oracle.from_spec(s3)
        x1 = o1.init_fields(s1.initial_x())
        x1[1:] *= 1.01
        F1 = o1.residual(x1)
        n = s1.nnodes
        reps = s3.nnodes // n
        x3 = np.zeros((s3.nunk, s3.nnodes))
        x3[:5] = np.tile(x1[:5], (1, reps))            # rho, n3, n2, n1, n0: uniform in y and z
        x3[5] = np.tile(x1[5], reps)                   # nv2_x
        x3[8] = np.tile(x1[6], reps)                   # nv1_x
        F3 = o3.residual(x3).reshape(s3.nunk, reps, n)
        assert np.abs(F3[[6, 7, 9, 10]]).max() < 1e-12 * np.abs(F1).max()
        G = np.concatenate([F3[:6], F3[8:9]])          # the 1D unknowns' rows
        diffs.append(np.abs(G - F1[:, None, :]).max() / np.abs(F1).max())
    assert max(diffs) < 2e-4, diffs
