"""State points described with Tramonto input parameters (1D planar walls).

`read_tramonto_input` reads those parameters from a Tramonto input directory
(dft_input.dat, dft_surfaces.dat); `spec_from_tramonto` turns them (or the
records of tests/golden/examples_1d.json) into a ProblemSpec:

* node positions x_i = -Size_x/2 + i Esize_x (node_to_position,
  src/dft_mesh_lib.c:579-620);
* planar walls of half-width WallParam centred at WallPos;
* either the cut-and-shifted LJ 9-3 wall potential by distance to the surface
  (Ipot_wf_n = VEXT_DIST_TO_SURF with Lhard_surf = 0: Vext_LJ9_3_CS,
  src/dft_vext_1D.c:40-56; setup_vext_XRSurf, src/dft_vext.c:364-434):
  VEXT_MAX inside the wall and within 0.001 of it, otherwise min(V, VEXT_MAX);
  Zero_density_TF where V_ext >= VEXT_MAX (src/dft_vext.c:155-166);
* or hard walls (Ipot_wf_n = VEXT_HARD with Lhard_surf = 1):
  V_ext = 0 and Zero_density_TF from the element lists of the mesh
  (src/dft_mesh.c:1340-1470): component i's list flags the elements whose
  centre lies within WallParam + Sigma_ff[i][i]/2 of a wall
  (src/dft_surfaces.c:357-395, src/dft_surfaces_planar.c:40-68), and a node
  all of whose elements are flagged has zero density;
* Lorentz-Berthelot wall-fluid mixing when Mix_type = 0
  (src/dft_potential_params.c);
* the rho_b exp(-V_ext) initial guess capped at 0.99
  (setup_exp_density, src/dft_guess_EL_DENSITY.c:171-201).

`integration_weights` returns the Nel_hit / Nel_hit2 node weights of
setup_integrals (src/dft_out_utils.c:182-270) and the scale factor of
integrateInSpace (src/dft_out_utils.c:38-62); `adsorption` is
calc_adsorption's integral (src/dft_out_ads.c:33-80) with Nel_hit2;
`force_weights` gives the surface sums of sum_rho_wall
(src/dft_out_force.c:223-280) as node weights per linked wall.
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


VEXT_HARD = 1            # Ipot_wf_n, src/dft_globals_const.h:485-489
VEXT_DIST_TO_SURF = 2


def _grid(g: Dict):
    size, dx = g["size_x"][0], g["esize_x"][0]
    periodic = g["type_bc_x"][0] == PERIODIC
    nx = int(math.floor(size / dx + 0.5)) + (0 if periodic else 1)
    return nx, dx, -0.5 * size + dx * np.arange(nx)


def hard_lists(g: Dict) -> bool:
    """Lhard_surf with walls: Ncomp + 1 element lists (src/dft_mesh.c:330-345)"""
    return bool(g["lhard_surf"]) and g["nwall"] > 0


def wall_elements(g: Dict, adjust: float = 0.0) -> np.ndarray:
    """The owning wall of each element of a 1D grid, -1 for fluid: the element's centre lies within
    WallParam + adjust of the wall's centre (the planar in-surface test with its 1e-11 offset,
    src/dft_surfaces_planar.c:40-68, at the element centre, src/dft_surfaces.c:357-395)."""
    nx, dx, x = _grid(g)
    mid = x[:-1] + 0.5 * dx
    own = np.full(nx - 1, -1, dtype=np.int64)
    for iw, w in enumerate(g["walls"]):
        hit = (np.abs(w["pos"][0] - mid) <= g["wall_param"] + adjust - 1e-11) & (own < 0)
        own[hit] = iw
    return own


def _node_elements(nx: int, bc, own: np.ndarray, i: int):
    """The owners of the two elements touching node i (left, right).  Past an IN_WALL or REFLECT
    edge the element takes the state of the edge element (Touch_domain_boundary, reflection), past
    an IN_BULK edge it is fluid."""
    out = []
    for e, side in ((i - 1, bc[0]), (i, bc[1])):
        if 0 <= e < nx - 1:
            out.append(int(own[e]))
        elif side in (IN_WALL, REFLECT):
            out.append(int(own[0 if e < 0 else nx - 2]))
        else:
            out.append(-1)
    return out


def _list_adjust(g: Dict, icomp: int) -> float:
    """dist_adjustments of component icomp's list (src/dft_surfaces.c:381-383)"""
    return 0.5 * float(_pair_matrix(g, "sigma_ff")[icomp, icomp])


def hard_surface_mask(g: Dict, own: np.ndarray) -> np.ndarray:
    """Per node of a 1D grid: 0 unless the node is a wall-fluid boundary node of the element list
    `own` (Nodes_2_boundary_wall != -1, src/dft_mesh.c:1452-1470); then bit 0 is set when the
    element to its right (the element that has it as local node 0) is fluid and bit 1 when the
    element to its left is (HW_boundary_weight, src/dft_utils.c:41-62, which counts elements inside
    the domain or their reflections only)."""
    nx, _, _ = _grid(g)
    bc = [int(t) for t in g["type_bc_x"]]
    m = np.zeros(nx, dtype=np.uint8)
    for n in range(nx):
        left, right = _node_elements(nx, bc, own, n)
        if (left >= 0) == (right >= 0):
            continue
        for bit, e, side in ((0, n, bc[1]), (1, n - 1, bc[0])):
            if 0 <= e < nx - 1:
                fluid = own[e] < 0
            elif side == REFLECT:
                fluid = own[0 if e < 0 else nx - 2] < 0
            else:
                fluid = False
            if fluid:
                m[n] |= 1 << bit
    return m


def _pair_matrix(g: Dict, key: str) -> np.ndarray:
    """a C x C parameter matrix; inputs with Lorentz-Berthelot mixing may list the diagonal only"""
    C = g["ncomp"]
    v = np.array(g[key], dtype=np.float64)
    return np.diag(v) if v.size == C and C > 1 else v.reshape(C, C)


def spec_from_tramonto(g: Dict, name: str = "example") -> ProblemSpec:
    hard = g["ipot_wf_n"] == VEXT_HARD and g["lhard_surf"] == 1
    lj = g["ipot_wf_n"] == VEXT_DIST_TO_SURF and g["lhard_surf"] == 0
    if g["ndim"] != 1 or g["surf_type"] != 0 or not (hard or lj):
        raise ValueError("only 1D planar walls are described here: LJ 9-3 walls without hard-surface "
                         "weights, or hard walls with them")
    C = g["ncomp"]
    nx, dx, x = _grid(g)
    vmax = g["vext_max"]
    vext = np.zeros((nx, C))
    if lj:
        sw, ew, cw = wall_fluid_params(g)
        h = g["wall_param"]
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
    else:
        bc = [int(t) for t in g["type_bc_x"]]
        zero = np.zeros((nx, C), dtype=np.uint8)
        hw_mask = np.zeros((C, nx), dtype=np.uint8)
        for i in range(C):
            own = wall_elements(g, _list_adjust(g, i))
            for n in range(nx):
                zero[n, i] = all(o >= 0 for o in _node_elements(nx, bc, own, n))
            hw_mask[i] = hard_surface_mask(g, own)
    sig, eps, cut = (_pair_matrix(g, k) for k in ("sigma_ff", "eps_ff", "cut_ff"))
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
                       type_attr=g["type_attr"], nwall=g["nwall"], lhard_surf=int(hard),
                       hw_mask=hw_mask if hard else None, vext_max=vmax,
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


def integration_weights(spec: ProblemSpec, g: Dict):
    """The [ncomp][N] element counts (Nel_hit, Nel_hit2) of a 1D example and integrateInSpace's
    scale factor (setup_integrals, src/dft_out_utils.c:182-270; src/dft_out_utils.c:38-62).
    Nel_hit counts the elements outside the walls themselves; Nel_hit2 those outside each
    component's list (the walls widened by Sigma_ff[i][i]/2 with hard surfaces, the walls
    themselves otherwise).  The scale is Fac_vol / (Fac_area Area) per unit area (Lper_area) and
    Fac_vol otherwise, with Fac_vol = 2 for one REFLECT edge when Lcount_reflect
    (src/dft_out_main.c:244-262)."""
    C = spec.ncomp
    hit = node_weights_1d(spec, wall_elements(g) >= 0).astype(np.int32)
    nel_hit = np.tile(hit, (C, 1))
    if hard_lists(g):
        nel_hit2 = np.stack([node_weights_1d(spec, wall_elements(g, _list_adjust(g, i)) >= 0).astype(np.int32)
                             for i in range(C)])
    else:
        nel_hit2 = nel_hit.copy()
    fac_vol, fac_area = _reflect_factors(g)
    area = surface_area(g)
    scale = fac_vol / (fac_area * area) if (g.get("lper_area", 0) and area > 0.0) else fac_vol
    return nel_hit, nel_hit2, scale


def _reflect_factors(g: Dict):
    """(Fac_vol, Fac_area) of a 1D domain (src/dft_out_main.c:244-262)"""
    fac_vol, fac_area = 1.0, 1.0
    left, right = g["type_bc_x"]
    if not (left == REFLECT and right == REFLECT) and g.get("lcount_reflect", 0):
        if left == REFLECT or right == REFLECT:
            fac_vol *= 2.0
            edge = -0.5 * g["size_x"][0] if left == REFLECT else 0.5 * g["size_x"][0]
            if g["walls"] and g["walls"][0]["pos"][0] == edge:
                fac_area *= 2.0
    return fac_vol, fac_area


def integration_counts(spec: ProblemSpec, g: Dict) -> np.ndarray:
    """Nel_hit2 (the profile terms' weights) of integration_weights."""
    return integration_weights(spec, g)[1]


def adsorption(spec: ProblemSpec, g: Dict, rho: np.ndarray) -> np.ndarray:
    """Integral of each component's density over the fluid (calc_adsorption)."""
    _, h2, scale = integration_weights(spec, g)
    return np.array([(rho[c] * h2[c] * spec.dx[0] / 2.0).sum() * scale for c in range(spec.ncomp)])


def _boundary_nodes(g: Dict, own: np.ndarray):
    """(node, wall, side) of the wall-fluid boundary nodes of one element list in 1D: one element
    of the node is a wall element and the other a fluid element (Nodes_2_boundary_wall,
    src/dft_mesh.c:1452-1470).  side is -1 when the wall element lies at lower x."""
    nx, _, _ = _grid(g)
    bc = [int(t) for t in g["type_bc_x"]]
    out = []
    for n in range(nx):
        left, right = _node_elements(nx, bc, own, n)
        if (left >= 0) == (right >= 0):
            continue
        if left >= 0 and right >= 0 and left != right:
            raise ValueError("a node between two walls (a wall-wall boundary) is not described here")
        out.append((n, left if left >= 0 else right, -1 if left >= 0 else 1))
    return out


def surface_area(g: Dict) -> float:
    """Area of the integrals per unit area (src/dft_out_utils.c:160-175): the surface elements
    of the walls linked to wall 0 in the walls' own list (S_area_tot, src/dft_mesh.c:2352-2390),
    1 per boundary node in 1D; 1 without walls."""
    if g["nwall"] == 0:
        return 1.0
    links = [w.get("link", i) for i, w in enumerate(g["walls"])]
    nlink = len(set(links))
    b = _boundary_nodes(g, wall_elements(g))
    if nlink == g["nwall"]:
        return float(sum(1 for _, w, _ in b if w == 0))
    return float(sum(1 for _, w, _ in b if links[w] == links[0]))


def force_weights(spec: ProblemSpec, g: Dict) -> np.ndarray:
    """[Nlink][ncomp][N] node weights of the solvation force on each linked wall in x: the reference
    sums, over the boundary nodes of each component's list, -sign(surface normal) Area_surf_el
    times the density (sum_rho_wall, src/dft_out_force.c:223-280; the 1D surface elements and
    normals of src/dft_mesh.c:1796-1810: normal -1 when the wall element lies at higher x and +1
    when it lies at lower x, so the weight is +1 for a wall above the node, -1 below it), and
    divides by the area per unit area (Lper_area), or multiplies by Fac_area otherwise
    (src/dft_out_force.c:153-167).  The result, sum(W * rho), is the reference's ptilde; the
    force adds the bulk pressure to it for walls at an IN_WALL edge."""
    if not hard_lists(g):
        raise ValueError("the contact sums of the force need hard surfaces (Lhard_surf)")
    C, N = spec.ncomp, spec.nnodes
    links = [w.get("link", i) for i, w in enumerate(g["walls"])]
    nlink = max(links) + 1
    W = np.zeros((nlink, C, N))
    nlists = C + 1
    for c in range(C):
        lst = 0 if nlists <= 2 else c
        for n, w, side in _boundary_nodes(g, wall_elements(g, _list_adjust(g, lst))):
            W[links[w], c, n] += float(side)
    area = surface_area(g)
    if g.get("lper_area", 0) and area > 0.0:
        W /= area
    else:
        W *= _reflect_factors(g)[1]
    return W


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
    v, _ = _find(L, "Lper_area")
    g["lper_area"], g["lcount_reflect"] = int(v[0]), int(v[1])
    v, _ = _find(L, "Iguess")
    g["iguess"] = int(v[0])
    v, _ = _find(L, "NL_Solver")
    g["nl_solver"], g["max_nl_iter"], g["att_in_a22"] = int(v[0]), int(v[1]), int(v[3])
    v, _ = _find(L, "NL_rel_tol NL_abs_tol")
    g["nl_rel_tol"], g["nl_abs_tol"], g["nl_update_scaling"] = v[0], v[1], v[2]
    with open(os.path.join(d, "dft_surfaces.dat")) as f:
        rows = [r.split() for r in f if r.strip()]
    g["walls"] = [{"type": int(r[0]), "link": int(r[1]), "pos": [float(t) for t in r[2:2 + g["ndim"]]]}
                  for r in rows]
    return g
