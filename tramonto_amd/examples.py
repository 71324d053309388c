"""State points described with Tramonto input parameters (1D planar walls).

`spec_from_tramonto` turns the parameters of a Tramonto 1D example (as
recorded in tests/golden/examples_1d.json) into a ProblemSpec:

* node positions x_i = -Size_x/2 + i Esize_x (node_to_position,
  src/dft_mesh_lib.c:579-620);
* planar walls of half-width WallParam centred at WallPos;
* the cut-and-shifted LJ 9-3 wall potential by distance to the surface
  (Vext_LJ9_3_CS, src/dft_vext_1D.c:40-56; setup_vext_XRSurf,
  src/dft_vext.c:364-434): VEXT_MAX inside the wall and within 0.001 of it,
  otherwise min(V, VEXT_MAX); Zero_density_TF where V_ext >= VEXT_MAX
  (src/dft_vext.c:155-166);
* Lorentz-Berthelot wall-fluid mixing when Mix_type = 0
  (src/dft_potential_params.c);
* the rho_b exp(-V_ext) initial guess capped at 0.99
  (setup_exp_density, src/dft_guess_EL_DENSITY.c:171-201).

`adsorption` is calc_adsorption's integral (src/dft_out_ads.c:33-80,
integrateInSpace src/dft_out_utils.c:38-62) with the Nel_hit2 node weights
of src/dft_out_utils.c:225-270.
"""
from __future__ import annotations

import math
from typing import Dict

import numpy as np

from .problem import FMT1, IN_BULK, IN_WALL, PERIODIC, REFLECT, ProblemSpec


def lj93_cs(z, eps_wf, rho_w, sigma, cut):
    """Vext_LJ9_3_CS, src/dft_vext_1D.c:40-56."""
    z = np.asarray(z, dtype=np.float64)
    with np.errstate(divide="ignore", over="ignore"):
        v = eps_wf * rho_w * (2.0 * math.pi / 3.0) * sigma ** 3 * (
            (2.0 / 15.0) * ((sigma / z) ** 9 - (sigma / cut) ** 9) - ((sigma / z) ** 3 - (sigma / cut) ** 3))
    return np.where(z <= cut, v, 0.0)


def wall_fluid_params(g: Dict):
    C = g["ncomp"]
    if g["mix_type"] == 0:
        sig_ff = np.array(g["sigma_ff"]).reshape(C, C)
        eps_ff = np.array(g["eps_ff"]).reshape(C, C)
        cut_ff = np.array(g["cut_ff"]).reshape(C, C)
        sw = [0.5 * (sig_ff[i, i] + g["sigma_ww"]) for i in range(C)]
        ew = [math.sqrt(eps_ff[i, i] * g["eps_ww"]) for i in range(C)]
        cw = [0.5 * (cut_ff[i, i] + g["cut_ww"]) for i in range(C)]
        return sw, ew, cw
    return list(g["sigma_wf"]), list(g["eps_wf"]), list(g["cut_wf"])


def spec_from_tramonto(g: Dict, name: str = "example") -> ProblemSpec:
    if g["ndim"] != 1 or g["surf_type"] != 0 or g["ipot_wf_n"] != 2 or g["lhard_surf"] != 0:
        raise ValueError("only 1D planar LJ 9-3 walls without hard-surface weights are described here")
    C = g["ncomp"]
    size, dx = g["size_x"][0], g["esize_x"][0]
    periodic = g["type_bc_x"][0] == PERIODIC
    nx = int(math.floor(size / dx + 0.5)) + (0 if periodic else 1)
    x = -0.5 * size + dx * np.arange(nx)
    vmax = g["vext_max"]
    sw, ew, cw = wall_fluid_params(g)
    h = g["wall_param"]
    vext = np.zeros((nx, C))
    inside = np.zeros(nx, dtype=bool)
    for w in g["walls"]:
        c = w["pos"][0]
        d = np.abs(x - c) - h
        inside |= d < 0.0
        for i in range(C):
            vv = np.full(nx, vmax)
            ok = d > 0.001
            vv[ok] = lj93_cs(d[ok], ew[i], g["rho_w"], sw[i], cw[i])
            vext[:, i] += vv
    vext = np.minimum(vext, vmax)
    vext[inside] = vmax
    zero = (vext >= vmax).astype(np.uint8)
    sig = np.array(g["sigma_ff"]).reshape(C, C)
    eps = np.array(g["eps_ff"]).reshape(C, C)
    cut = np.array(g["cut_ff"]).reshape(C, C)
    if g["mix_type"] == 0:
        for i in range(C):
            for j in range(C):
                sig[i, j] = 0.5 * (sig[i, i] + sig[j, j])
                eps[i, j] = math.sqrt(eps[i, i] * eps[j, j])
                cut[i, j] = 0.5 * (cut[i, i] + cut[j, j])
    rho_b = list(g["rho_b"])
    rho0 = np.zeros((C, nx))
    for i in range(C):
        if g["iguess"] == 1:
            r = rho_b[i] * np.exp(-vext[:, i])
        else:
            r = np.full(nx, rho_b[i])
        r = np.where(r >= 1.0, 0.99, r)
        r[zero[:, i] == 1] = 0.0
        rho0[i] = r
    bc = [[int(t) for t in g["type_bc_x"]]]
    return ProblemSpec(name=name, ndim=1, ncomp=C, nx=[nx], dx=[dx], bc=bc, type_func=g["type_func"],
                       type_attr=g["type_attr"], nwall=g["nwall"], vext_max=vmax,
                       hs_diam=[sig[i, i] for i in range(C)], sigma_ff=sig, eps_ff=eps, cut_ff=cut, rho_b=rho_b,
                       vext=vext, zero_tf=zero, rho0=rho0)


def node_weights_1d(spec: ProblemSpec, wall_elem: np.ndarray) -> np.ndarray:
    """Nel_hit2 for a 1D grid: 2 minus elements outside an IN_WALL edge and
    wall elements; halved (integer division) at IN_BULK / REFLECT edges."""
    nx = spec.nx[0]
    left, right = spec.bc[0]
    w = np.zeros(nx, dtype=np.int64)
    for i in range(nx):
        n = 2
        for e, side in ((i - 1, left), (i, right)):
            if e < 0 or e >= nx - 1:
                if side == IN_WALL:
                    n -= 1
                continue
            if wall_elem[e]:
                n -= 1
        if i == 0 and left in (REFLECT, IN_BULK):
            n //= 2
        if i == nx - 1 and right in (REFLECT, IN_BULK):
            n //= 2
        w[i] = n
    return w


def integration_counts(spec: ProblemSpec, g: Dict) -> np.ndarray:
    """The [ncomp][N] element counts Nel_hit = Nel_hit2 of a 1D example without hard surfaces
    (setup_integrals, src/dft_out_utils.c:182-270): an element is a wall element when its centre
    lies within the wall's half-width of a planar wall."""
    nx, dx = spec.nx[0], spec.dx[0]
    x = -0.5 * g["size_x"][0] + dx * np.arange(nx)
    mid = x[:-1] + 0.5 * dx
    wall_elem = np.zeros(nx - 1, dtype=bool)
    for w in g["walls"]:
        wall_elem |= np.abs(mid - w["pos"][0]) < g["wall_param"]
    wt = node_weights_1d(spec, wall_elem)
    return np.tile(wt.astype(np.int32), (spec.ncomp, 1))


def adsorption(spec: ProblemSpec, g: Dict, rho: np.ndarray) -> np.ndarray:
    """Integral of each component's density over the fluid (calc_adsorption)."""
    wt = integration_counts(spec, g)[0]
    return np.array([(rho[c] * wt * spec.dx[0] / 2.0).sum() for c in range(spec.ncomp)])
