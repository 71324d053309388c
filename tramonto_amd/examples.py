"""State points described with Tramonto input parameters (1D planar walls).

`read_tramonto_input` reads those parameters from a Tramonto input directory
(dft_input.dat, dft_surfaces.dat); `spec_from_tramonto` turns them (or the
records of tests/golden/examples_1d.json) into a ProblemSpec:

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
import os
import re
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


# ---------------------------------------------------------------- input files
_NUM = re.compile(r"^[-+]?(\d+\.?\d*|\.\d+)([eE][-+]?\d+)?$")


def _at_lines(path):
    out = []
    with open(path) as f:
        for line in f:
            if not line.startswith("@"):
                continue
            toks = line[1:].split()
            vals = []
            for t in toks:
                if _NUM.match(t):
                    vals.append(float(t))
                else:
                    break
            label = " ".join(toks[len(vals):])
            out.append((vals, label))
    return out


def _find(lines, prefix, start=0):
    for k in range(start, len(lines)):
        if lines[k][1].startswith(prefix):
            return lines[k][0], k
    raise KeyError(prefix)


def read_tramonto_input(d: str) -> Dict:
    """The parameters of a Tramonto input directory (dft_input.dat, dft_surfaces.dat) that the
    nonlocal-density fill and its Newton solve consume, read from the positional `@` lines of
    dft_input.dat (read_input_file, src/dft_input.c:47) and the wall rows of dft_surfaces.dat."""
    L = _at_lines(os.path.join(d, "dft_input.dat"))
    g = {}
    v, _ = _find(L, "Length_ref")
    g["vext_max"] = v[4]
    g["ndim"] = int(_find(L, "Ndim")[0][0])
    g["size_x"] = _find(L, "Size_x")[0][: g["ndim"]]
    g["esize_x"] = _find(L, "Esize_x")[0][: g["ndim"]]
    g["type_bc_x"] = [int(t) for t in _find(L, "Type_bc(x0")[0][:2]]
    v, _ = _find(L, "Type_func")
    g["type_func"] = int(v[0])
    v, _ = _find(L, "Type_attr")
    g["type_attr"], g["type_pairpot"] = int(v[0]), int(v[1])
    v, _ = _find(L, "Nwall_types")
    g["nwall_type"], g["nwall"] = int(v[0]), int(v[1])
    g["surf_type"] = int(_find(L, "Surf_type")[0][0])
    g["orientation"] = int(_find(L, "Orientation[iwall_type]")[0][0])
    g["wall_param"] = _find(L, "WallParam[iwall_type]")[0][0]
    g["ipot_wf_n"] = int(_find(L, "Type of neutral surface-fluid")[0][0])
    g["lhard_surf"] = int(_find(L, "Lhard_surf")[0][0])
    g["type_vext"] = int(_find(L, "Type_vext")[0][0])
    g["vext_potential_id"] = int(_find(L, "Vext_PotentialID")[0][0])
    v, _ = _find(L, "Ncomp Mix_type")
    g["ncomp"], g["mix_type"] = int(v[0]), int(v[1])
    C = g["ncomp"]
    for key, lab in (("sigma_ff", "Sigma_ff"), ("eps_ff", "Eps_ff"), ("cut_ff", "Cut_ff")):
        g[key] = _find(L, lab)[0][: C * C]
    for key, lab in (("rho_w", "Rho_w"), ("sigma_ww", "Sigma_ww"), ("eps_ww", "Eps_ww"), ("cut_ww", "Cut_ww")):
        g[key] = _find(L, lab)[0][0]
    for key, lab in (("sigma_wf", "Sigma_wf"), ("eps_wf", "Eps_wf"), ("cut_wf", "Cut_wf")):
        g[key] = _find(L, lab)[0][:C]
    g["rho_b"] = _find(L, "Rho_b_0")[0][:C]
    v, _ = _find(L, "Iguess")
    g["iguess"] = int(v[0])
    v, _ = _find(L, "NL_Solver")
    g["nl_solver"], g["max_nl_iter"], g["att_in_a22"] = int(v[0]), int(v[1]), int(v[3])
    v, _ = _find(L, "NL_rel_tol NL_abs_tol")
    g["nl_rel_tol"], g["nl_abs_tol"], g["nl_update_scaling"] = v[0], v[1], v[2]
    with open(os.path.join(d, "dft_surfaces.dat")) as f:
        rows = [r.split() for r in f if r.strip()]
    g["walls"] = [{"type": int(r[0]), "pos": [float(t) for t in r[2:2 + g["ndim"]]]} for r in rows]
    return g
