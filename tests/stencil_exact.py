"""Exact stencil weights, computed independently of the reference's quadrature rule.

The reference builds the weight of stencil point k as the integral of the weight kernel against
the hat function phi_k of node k (the tensor product of the 1D hats h(t) = max(0, 1 - |t|) at
the node), approximated element by element with Gauss points inside the sphere and mid-point
sub-cells on the elements the sphere crosses, then renormalised to the analytic total
(src/dft_stencil.c:329-496, 653-683).  Here the same integrals are evaluated to ~1e-13 by
another route:

* A(rho): the integral of the 2D hat h_x(rho cos t) h_y(rho sin t) over the circle of radius rho,
  in closed form: on each arc between the angles where a hat kink or support edge is crossed,
  the integrand is (a + b cos t)(c + d sin t).
* 3D delta (the sphere |r| = R):  S = R^2 int_0^pi h_z(R cos s) A(R sin s) sin s ds.
* 3D theta (the ball):            V = int_-R^R h_z(z) int_0^rho(z) r A(r) dr dz,  rho(z) = sqrt(R^2 - z^2).
* 2D delta (the sphere projected on the plane): 2 R^2 int_0^{pi/2} A(R sin s) sin s ds.
* 2D theta (the ball projected): int_0^R 2 sqrt(R^2 - r^2) r A(r) dr.

Every 1D integral is split at the abscissae where the integrand has a kink (where a circle
or a sphere section meets a hat kink or support edge) and summed by Gauss-Legendre on each
piece, so the pieces are smooth.  Since the hats of all nodes sum to 1, the exact weights sum
to the analytic totals pi d^2 and pi d^3/6 without renormalisation.
"""
from __future__ import annotations

import numpy as np

GL_N = 24
_GX, _GW = np.polynomial.legendre.leggauss(GL_N)


_U = 0.5 * (_GX + 1.0)
_UW = 0.5 * _GW


def _gauss_pieces(breaks):
    """Nodes and weights on the pieces of sorted breakpoints (array [B]): Gauss-Legendre in u on each
    piece with x = a + (b - a)(3u^2 - 2u^3).  The map's derivative 6u(1 - u) vanishes at both ends,
    so an end point where the integrand behaves like sqrt(x - a) (a circle tangent to a hat kink
    line) becomes smooth in u and the rule keeps its spectral accuracy."""
    a, b = breaks[:-1], breaks[1:]
    keep = b - a > 1e-15
    a, b = a[keep], b[keep]
    g = 3 * _U ** 2 - 2 * _U ** 3
    dg = 6 * _U * (1 - _U)
    x = a[:, None] + (b - a)[:, None] * g[None, :]
    w = (b - a)[:, None] * (_UW * dg)[None, :]
    return x.ravel(), w.ravel()


def circle_hat_average(px, py, dx, dy, rho):
    """A(rho) = int_0^{2pi} h((rho cos t - px)/dx) h((rho sin t - py)/dy) dt for an array rho, exactly."""
    r = np.atleast_1d(np.asarray(rho, dtype=np.float64))[:, None]
    cx = np.array([px - dx, px, px + dx])[None, :]
    cy = np.array([py - dy, py, py + dy])[None, :]
    rs = np.maximum(r, 1e-300)
    with np.errstate(invalid="ignore"):
        ax = np.where(np.abs(cx) < r, np.arccos(np.clip(cx / rs, -1.0, 1.0)), 0.0)
        ay = np.where(np.abs(cy) < r, np.arcsin(np.clip(cy / rs, -1.0, 1.0)), 0.0)
    two = 2 * np.pi
    br = np.concatenate([np.zeros_like(r), np.full_like(r, two), ax, two - ax, ay % two, (np.pi - ay) % two], axis=1)
    br = np.sort(br, axis=1)   # candidates that do not apply sit at 0, pi or 2 pi: empty or harmless splits
    t0, t1 = br[:, :-1], br[:, 1:]
    tm = 0.5 * (t0 + t1)
    ux = r * np.cos(tm) - px
    uy = r * np.sin(tm) - py
    live = (np.abs(ux) < dx) & (np.abs(uy) < dy) & (t1 > t0)
    sx = np.sign(ux)
    sy = np.sign(uy)
    # h_x = 1 - sx (r cos t - px)/dx = alpha + beta cos t;  h_y = gamma + delta sin t
    alpha = 1.0 + sx * px / dx
    beta = -sx * r / dx
    gamma = 1.0 + sy * py / dy
    delta = -sy * r / dy
    val = (alpha * gamma * (t1 - t0) + alpha * delta * (np.cos(t0) - np.cos(t1)) +
           beta * gamma * (np.sin(t1) - np.sin(t0)) + 0.5 * beta * delta * (np.sin(t1) ** 2 - np.sin(t0) ** 2))
    return np.sum(np.where(live, val, 0.0), axis=1)


def _radial_breaks(px, py, dx, dy):
    """radii where the structure of A(rho) changes: the circle touches a kink line or corner"""
    cx = np.array([px - dx, px, px + dx])
    cy = np.array([py - dy, py, py + dy])
    b = list(np.abs(cx)) + list(np.abs(cy))
    b += [float(np.hypot(a, c)) for a in cx for c in cy]
    return np.array(b)


def hat_1d(t):
    return np.maximum(0.0, 1.0 - np.abs(t))


def sphere_weight_3d(k, dx, R):
    """int over the sphere |r| = R of the trilinear hat of node k (offset k, spacing dx)."""
    px, py, pz = (k[d] * dx[d] for d in range(3))
    # s in [0, pi]: z = R cos s, rho = R sin s
    br = [0.0, np.pi]
    for c in (pz - dx[2], pz, pz + dx[2]):
        if abs(c) < R:
            br.append(np.arccos(c / R))
    for b in _radial_breaks(px, py, dx[0], dx[1]):
        if b < R:
            a = np.arcsin(b / R)
            br += [a, np.pi - a]
    s, w = _gauss_pieces(np.unique(np.array(br)))
    A = circle_hat_average(px, py, dx[0], dx[1], R * np.sin(s))
    return float(R * R * np.sum(w * hat_1d((R * np.cos(s) - pz) / dx[2]) * A * np.sin(s)))


def _disk_integral(px, py, dx, dy, rho_max):
    """B(rho_max) = int_0^rho_max r A(r) dr for an array of rho_max"""
    rb = _radial_breaks(px, py, dx, dy)
    xs, ws, owner = [], [], []
    for i, rm in enumerate(rho_max):
        br = np.unique(np.concatenate([[0.0, rm], rb[rb < rm]]))
        r, w = _gauss_pieces(br)
        xs.append(r)
        ws.append(w)
        owner.append(np.full(len(r), i))
    r, w, owner = np.concatenate(xs), np.concatenate(ws), np.concatenate(owner)
    return np.bincount(owner, weights=w * r * circle_hat_average(px, py, dx, dy, r), minlength=len(rho_max))


def ball_weight_3d(k, dx, R):
    """int over the ball |r| < R of the trilinear hat of node k."""
    px, py, pz = (k[d] * dx[d] for d in range(3))
    br = [-R, R]
    for c in (pz - dx[2], pz, pz + dx[2]):
        if abs(c) < R:
            br.append(c)
    for b in _radial_breaks(px, py, dx[0], dx[1]):
        if b < R:
            z = np.sqrt(R * R - b * b)
            br += [z, -z]
    z, w = _gauss_pieces(np.unique(np.array(br)))
    hz = hat_1d((z - pz) / dx[2])
    live = hz > 0
    B = np.zeros_like(z)
    B[live] = _disk_integral(px, py, dx[0], dx[1], np.sqrt(np.maximum(R * R - z[live] ** 2, 0.0)))
    return float(np.sum(w * hz * B))


def sphere_weight_2d(k, dx, R):
    """the sphere's surface projected on the plane: int 2R/sqrt(R^2 - r^2) phi_k over the disk"""
    px, py = k[0] * dx[0], k[1] * dx[1]
    br = [0.0, 0.5 * np.pi]
    for b in _radial_breaks(px, py, dx[0], dx[1]):
        if b < R:
            br.append(np.arcsin(b / R))
    s, w = _gauss_pieces(np.unique(np.array(br)))
    return float(2 * R * R * np.sum(w * circle_hat_average(px, py, dx[0], dx[1], R * np.sin(s)) * np.sin(s)))


def ball_weight_2d(k, dx, R):
    """the ball projected on the plane: int 2 sqrt(R^2 - r^2) phi_k over the disk"""
    px, py = k[0] * dx[0], k[1] * dx[1]
    rb = _radial_breaks(px, py, dx[0], dx[1])
    # r = R sin s removes the square-root end point: dr = R cos s ds, sqrt(R^2 - r^2) = R cos s
    br = [0.0, 0.5 * np.pi] + [float(np.arcsin(b / R)) for b in rb if b < R]
    s, w = _gauss_pieces(np.unique(np.array(br)))
    r = R * np.sin(s)
    return float(np.sum(w * 2 * R * np.cos(s) * r * circle_hat_average(px, py, dx[0], dx[1], r) * R * np.cos(s)))


def exact_weights(kind, ndim, offsets, dx, R):
    """exact weights of the stencil points `offsets` ([L][3] ints): kind 'delta' or 'theta'"""
    fn = {("delta", 3): sphere_weight_3d, ("theta", 3): ball_weight_3d,
          ("delta", 2): sphere_weight_2d, ("theta", 2): ball_weight_2d}[(kind, ndim)]
    cache = {}
    out = np.zeros(len(offsets))
    for q, o in enumerate(offsets):
        a = [abs(int(v)) for v in o[:ndim]]
        key = tuple(sorted(a)) if all(d == dx[0] for d in dx[:ndim]) else tuple(a)
        if key not in cache:
            cache[key] = fn(np.array(a, dtype=float), dx, R)
        out[q] = cache[key]
    return out


def support_points(kind, ndim, dx, R, reach):
    """offsets (|o_d| <= reach) whose hat support (the open cube of the 2^ndim elements around the
    node) meets the sphere (delta) or the ball (theta), i.e. whose exact weight is nonzero"""
    pts = []
    grid = np.array(np.meshgrid(*([np.arange(-reach, reach + 1)] * ndim), indexing="ij")).reshape(ndim, -1).T
    for o in grid:
        lo = (np.abs(o) - 1) * np.asarray(dx[:ndim])
        lo = np.maximum(lo, 0.0)
        hi = (np.abs(o) + 1) * np.asarray(dx[:ndim])
        dmin = float(np.sqrt(np.sum(lo ** 2)))
        dmax = float(np.sqrt(np.sum(hi ** 2)))
        if kind == "theta" or ndim == 2:   # the projected 2D kernels cover the whole disk
            hit = dmin < R
        else:
            hit = dmin < R < dmax
        if hit:
            pts.append(tuple(int(v) for v in o))
    return pts
