"""The oracle on the reference's hard-wall Examples (Lhard_surf = 1, no attraction).

tests/golden/examples_hs_1d.json (written by make_fixtures.py) holds the inputs
and README key outputs of HS1-HS5_FMT1_1D, MIX1_FMT1_1D and HS8_FMT4_1D: hard
spheres (Rosenfeld FMT; FMT4 for HS8) between hard planar walls, where the rho-bar stencils take the
hard-surface weights at the contact nodes (HW_boundary_weight,
src/dft_utils.c:41-62, in src/dft_fill_CLSmf.c:140-160).  The oracle's Newton
solve must reproduce the README adsorption, energy and wall force to the digits
printed there, and niter within the reference harness' +-4
(Examples/check_results/main.cpp:120-139).  Without the boundary-node weights
the same solves miss the README values by far more than the harness' 1e-3.
"""
import dataclasses
import json
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from tramonto_amd.examples import force_weights, integration_weights, spec_from_tramonto

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "examples_hs_1d.json")) as f:
    EXAMPLES = json.load(f)


def printed_tol(text):
    """half a unit in the last digit the README prints"""
    dec = len(text.split(".")[1]) if "." in text else 0
    return 0.5 * 10.0 ** (-dec) + 1e-9


def solve(spec, g):
    o = Oracle.from_spec(spec)
    x, st = o.newton(o.init_fields(spec.initial_x()), resid_tol=1e-10, update_scaling=g["nl_update_scaling"])
    assert st["converged"], st
    return o, x, st


def outputs(spec, g, o, x):
    h1, h2, scale = integration_weights(spec, g)
    obs = o.observables(x, h1, h2, scale)
    W = force_weights(spec, g)
    force = [float((W[k] * x[: spec.ncomp]).sum()) for k in range(W.shape[0])]
    return obs, force


@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_hard_wall_example_readme_outputs(name):
    g = EXAMPLES[name]
    ref = g["key_outputs"]
    spec = spec_from_tramonto(g, name)
    assert spec.lhard_surf == 1 and spec.hw_mask is not None
    o, x, st = solve(spec, g)
    assert abs(st["iters"] - ref["niter"]) <= 4
    obs, force = outputs(spec, g, o, x)
    assert abs(obs["omega_s"] - ref["energy"]) < printed_tol(ref["energy_text"]), (obs["omega_s"], ref)
    for c in range(spec.ncomp):   # to the printed digits (HS8's README prints 5 decimals), at least 2e-6
        tol = max(2e-6, printed_tol(repr(ref["adsorption"][c])))
        assert abs(obs["ads"][c] - ref["adsorption"][c]) < tol, (c, obs["ads"], ref)
    # the README's "force" is one of the walls' values (ptilde: contact sums per unit area): the
    # reference harness matches each key against every value of the output line
    # (Examples/check_results/main.cpp, matchParameters)
    tol = printed_tol(ref["force_text"])
    assert min(abs(f - ref["force"]) for f in force) < tol, (force, ref)


@pytest.mark.parametrize("name", ["HS1_FMT1_1D", "HS5_FMT1_1D"])
def test_boundary_node_weights_are_needed(name):
    """the same solve with full stencil weights at the contact nodes misses the README values"""
    g = EXAMPLES[name]
    spec = spec_from_tramonto(g, name)
    o, x, _ = solve(dataclasses.replace(spec, hw_mask=None), g)
    obs, force = outputs(spec, g, o, x)
    assert abs(obs["omega_s"] - g["key_outputs"]["energy"]) > 0.1
    assert min(abs(f - g["key_outputs"]["force"]) for f in force) > 0.1


def test_contact_theorem_single_wall():
    """One hard wall in contact with bulk: the contact density (the wall force per unit area)
    approaches the bulk pressure of the functional, the Percus-Yevick compressibility pressure for
    Rosenfeld FMT, as the mesh is refined (contact theorem); HS5 (dx = 0.1) is within 5 %."""
    g = EXAMPLES["HS5_FMT1_1D"]
    spec = spec_from_tramonto(g, "HS5")
    o, x, _ = solve(spec, g)
    _, force = outputs(spec, g, o, x)
    rho = g["rho_b"][0]
    eta = np.pi / 6.0 * rho
    p_py = rho * (1 + eta + eta ** 2) / (1 - eta) ** 3
    assert abs(abs(force[0]) - p_py) < 0.05 * p_py, (force, p_py)


def test_reflective_example_equals_half_slit():
    """HS3 (one wall, reflecting edge at the slit's middle) and HS4 (its mirror image) are half
    of the symmetric slit: same profile mirrored, same integrals, opposite forces."""
    s3, s4 = (spec_from_tramonto(EXAMPLES[n], n) for n in ("HS3_FMT1_1D", "HS4_FMT1_1D"))
    o3, x3, _ = solve(s3, EXAMPLES["HS3_FMT1_1D"])
    o4, x4, _ = solve(s4, EXAMPLES["HS4_FMT1_1D"])
    assert np.abs(x3[0] - x4[0][::-1]).max() < 1e-9
    a3, f3 = outputs(s3, EXAMPLES["HS3_FMT1_1D"], o3, x3)
    a4, f4 = outputs(s4, EXAMPLES["HS4_FMT1_1D"], o4, x4)
    assert abs(a3["omega_s"] - a4["omega_s"]) < 1e-9 and abs(f3[0] + f4[0]) < 1e-9


def test_one_size_rows_give_the_newton_steps_of_two_identical_components():
    """With one component the reference fills n1, n0 and nv1 as multiples of n2 and nv2 and
    inverts A11 = [-I 0; B -I] exactly (src/dft_fill_FMT.c:215-236, 305-320;
    src/dft_HardSphereA11_Epetra_Operator.cpp:127-172); with two components (MIX1) it fills their
    stencil sums.  HS5 and MIX1 are the same fluid (two identical species at half the density), so
    every Newton iterate of HS5 equals the sum of MIX1's two densities: a wrong factor in B, in the
    Schur right-hand side or in the back-substitution changes the steps."""
    g5, g1 = EXAMPLES["HS5_FMT1_1D"], EXAMPLES["MIX1_FMT1_1D"]
    s5, s1 = spec_from_tramonto(g5, "HS5"), spec_from_tramonto(g1, "MIX1")
    o5, o1 = Oracle.from_spec(s5), Oracle.from_spec(s1)
    x5, x1 = o5.init_fields(s5.initial_x()), o1.init_fields(s1.initial_x())
    r5 = o5.residual(x5).reshape(o5.nunk, -1)
    # n1, n0 (unknowns 3, 4): the algebraic rows hold for the rho-bars built by stencil sums
    assert np.abs(r5[o5.iu_rb + 2: o5.iu_rb + 4]).max() < 1e-14
    for k in (1, 2, 3):
        y5, _ = o5.newton(x5.copy(), max_iter=k, resid_tol=0.0, update_scaling=g5["nl_update_scaling"])
        y1, _ = o1.newton(x1.copy(), max_iter=k, resid_tol=0.0, update_scaling=g1["nl_update_scaling"])
        assert np.abs(y5[0] - (y1[0] + y1[1])).max() < 1e-10 * np.abs(y5[0]).max(), k


@pytest.mark.parametrize("attraction", [False, True])
def test_reflected_boundary_node_weights_match_full_line(attraction):
    """HW_boundary_weight through a REFLECT edge (src/dft_utils.c:41-62 with the reflect flags of
    offset_to_node_box, src/dft_mesh_lib.c:203-297 and 896-937): a hard slab centred on a REFLECT
    plane is half of the full line holding the whole slab.  In the half line the stencil points of
    the rows near the plane reach the slab's surface node through the reflection, where the fluid
    element is on the mirror side; the full line reaches the mirror node directly.  Residual rows
    (rho-bar, mean field, Euler-Lagrange) must agree.  Reading the unflipped element bit for the
    reflected points changes the rho-bar rows near the plane by ~1e-2 and fails this."""
    from tramonto_amd.problem import reflect_slab_1d
    full, half = reflect_slab_1d(False, attraction=attraction), reflect_slab_1d(True, attraction=attraction)
    of, oh = Oracle.from_spec(full), Oracle.from_spec(half)
    xf = of.init_fields(full.initial_x())
    M = half.nnodes - 1
    xh = np.ascontiguousarray(xf[:, M:])
    # exact mirror symmetry of the state (odd vector rows), so both lines hold the same fluid
    nd, C = 1, 1
    for u in range(xf.shape[0]):
        odd = C + 4 <= u < C + 4 + 2 * nd
        xh[u] = 0.5 * (xf[u, M:] + (-1 if odd else 1) * xf[u, M::-1])
        xf[u, M:] = xh[u]
        xf[u, :M + 1] = ((-1 if odd else 1) * xh[u])[::-1]
    Ff = of.residual(xf).reshape(xf.shape[0], -1)[:, M:]
    Fh = oh.residual(xh).reshape(xh.shape[0], -1)
    assert np.abs(Fh - Ff).max() <= 1e-13 * np.abs(Ff).max(), np.abs(Fh - Ff).max()
    # the hard-surface weights are in play: full stencil weights at the surface node change the rows
    o0 = Oracle.from_spec(dataclasses.replace(half, hw_mask=None))
    assert np.abs(o0.residual(xh).reshape(xh.shape[0], -1) - Fh).max() > 1e-3
