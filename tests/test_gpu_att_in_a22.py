"""The reference's ATTInA22Block = TRUE layout (att_in_a22 = 1; run on an MI355X).

With the attraction in the A22 block the reference keeps no mean-field
unknowns: x = [rho_c | rho-bar], and the Euler-Lagrange rows add u * rho
directly (src/dft_fill_EL.c:143-156 with the mean-field sum instead of the
unknown).  The library serves that layout by expanding x with the mean-field
rows u * rho before each device operation and returning the caller's rows, so:
* F in this layout equals the oracle's residual (which keeps mean-field
  unknowns) at x extended by u * rho, without its mean-field rows;
* J in this layout equals J_int[e, e] + J_int[e, mf] J_int[mf, e] (the
  mean-field rows eliminated), checked entry by entry on 1D cases;
* the Schur solve returns dx with J dx = -F, and the Newton solve the profile
  of the mean-field layout; the README outputs of the Examples hold.
"""
import dataclasses
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd.examples import integration_counts, spec_from_tramonto  # noqa: E402
from tramonto_amd.problem import small_case  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "examples_1d.json")) as f:
    EXAMPLES = json.load(f)

CASES = ["1d_reflect_wca", "1d_pore_mix", "2d_rough_wall", "2d_periodic", "3d_periodic_small", "3d_solute_small"]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def pair(case):
    """the same state point in the mean-field layout (for the oracle) and in the att_in_a22 layout"""
    spec = small_case(case)
    assert spec.type_attr >= 0, case
    return spec, dataclasses.replace(spec, att_in_a22=1)


def state(spec, o, seed):
    x = o.init_fields(spec.initial_x())
    rng = np.random.default_rng(seed)
    x[: spec.ncomp] *= 1.0 + 0.05 * rng.uniform(-1.0, 1.0, x[: spec.ncomp].shape)
    return o.init_fields(x)   # rho-bar and mean-field rows of the perturbed densities


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_residual_matches_oracle_without_mean_field_rows(case):
    from tramonto_amd.device import DeviceProblem
    spec, spec_a = pair(case)
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec_a)
    ne = spec_a.nunk
    assert dp.info.nunk == ne == spec.nunk - spec.ncomp and dp.info.n_mf == 0
    x = state(spec, o, 3)
    Fo = o.residual(x)[:ne]
    Fg, s = dp.fill_residual(dev(x[:ne]))
    Fg = Fg.cpu().numpy()
    assert np.abs(Fg - Fo).max() < 1e-12 * np.abs(Fo).max()
    assert abs(s - float((Fo * Fo).sum())) <= 1e-10 * float((Fo * Fo).sum())


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["1d_reflect_wca", "1d_pore_mix"])
def test_jacobian_eliminates_the_mean_field_rows(case):
    from tramonto_amd.device import DeviceProblem
    spec, spec_a = pair(case)
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec_a)
    x = state(spec, o, 4)
    N, C, ne = spec.nnodes, spec.ncomp, spec_a.nunk
    J = o.jacobian_dense(x)
    e = np.arange(ne * N)
    mf = np.arange(ne * N, spec.nunk * N)
    Je = J[np.ix_(e, e)] + J[np.ix_(e, mf)] @ J[np.ix_(mf, e)]
    rng = np.random.default_rng(5)
    for _ in range(3):
        v = rng.uniform(-1.0, 1.0, ne * N)
        Jv = dp.apply_jacobian(dev(x[:ne]), dev(v.reshape(ne, N))).cpu().numpy().ravel()
        ref = Je @ v
        assert np.abs(Jv - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["2d_rough_wall", "3d_periodic_small", "3d_solute_small"])
def test_jvp_and_schur_solve_in_the_att_in_a22_layout(case):
    from tramonto_amd.device import DeviceProblem
    spec, spec_a = pair(case)
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec_a)
    ne = spec_a.nunk
    x = dev(state(spec, o, 6)[:ne])
    v = torch.rand_like(x) - 0.5
    Jv = dp.apply_jacobian(x, v)
    h = 1e-6
    Fp, _ = dp.fill_residual(x + h * x * v)
    Fm, _ = dp.fill_residual(x - h * x * v)
    fd = (Fp - Fm) / (2 * h)
    Jxv = dp.apply_jacobian(x, x * v)
    assert float((Jxv - fd).abs().max() / fd.abs().max()) < 1e-6
    F, _ = dp.fill_residual(x)
    dx, its = dp.schur_solve(x, F, rtol=1e-10, restart=200, maxit=2000)
    r = dp.apply_jacobian(x, dx) + F
    assert float(r.norm() / F.norm()) < 1e-8, (float(r.norm() / F.norm()), its)
    del Jv


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_examples_in_their_own_layout(name):
    """The Examples' inputs set ATTInA22Block = 1: solved in that layout they give the README key
    outputs and the profile of the mean-field layout."""
    from tramonto_amd.device import DeviceProblem
    g = EXAMPLES[name]
    assert g["att_in_a22"] == 1
    spec = spec_from_tramonto(g, name)
    spec_a = dataclasses.replace(spec, att_in_a22=1)
    kw = dict(max_iter=g["max_nl_iter"], nl_rel_tol=g["nl_rel_tol"], nl_abs_tol=g["nl_abs_tol"],
              update_scaling=g["nl_update_scaling"], resid_tol=1e-10, gmres_tol=1e-9, gmres_restart=100,
              gmres_max_iter=2000)
    out = {}
    for s in (spec, spec_a):
        dp = DeviceProblem(s)
        x = dp.initial_state()
        st = dp.newton(x, **kw)
        assert st["converged"], st
        out[s.att_in_a22] = (dp, x)
    dp, xa = out[1]
    assert xa.shape[0] == spec_a.nunk
    assert float((xa[: spec.ncomp] - out[0][1][: spec.ncomp]).abs().max()) < 1e-10
    h = integration_counts(spec, g)
    obs = dp.observables(xa, h)
    assert abs(obs["omega_s"] - g["key_outputs"]["energy"]) < 2e-6, (obs["omega_s"], g["key_outputs"])
    for c in range(spec.ncomp):
        assert abs(obs["ads"][c] - g["key_outputs"]["adsorption"][c]) < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["1d_reflect_wca", "1d_pore_mix"])
def test_transpose_product_eliminates_the_mean_field_rows(case):
    from tramonto_amd.device import DeviceProblem
    spec, spec_a = pair(case)
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec_a)
    x = state(spec, o, 8)
    N, ne = spec.nnodes, spec_a.nunk
    J = o.jacobian_dense(x)
    e = np.arange(ne * N)
    mf = np.arange(ne * N, spec.nunk * N)
    Je = J[np.ix_(e, e)] + J[np.ix_(e, mf)] @ J[np.ix_(mf, e)]
    w = np.random.default_rng(9).uniform(-1.0, 1.0, ne * N)
    got = dp.apply_jacobian_transpose(dev(x[:ne]), dev(w.reshape(ne, N))).cpu().numpy().ravel()
    ref = Je.T @ w
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["2d_rough_wall", "3d_periodic_small"])
def test_adjoint_identity_in_the_att_in_a22_layout(case):
    from tramonto_amd.device import DeviceProblem
    spec, spec_a = pair(case)
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec_a)
    ne = spec_a.nunk
    x = dev(state(spec, o, 10)[:ne])
    g = torch.Generator(device="cpu").manual_seed(3)
    v = (torch.rand(x.shape, generator=g, dtype=torch.float64) - 0.5).cuda()
    w = (torch.rand(x.shape, generator=g, dtype=torch.float64) - 0.5).cuda()
    a = float((w * dp.apply_jacobian(x, v)).sum())
    b = float((dp.apply_jacobian_transpose(x, w) * v).sum())
    assert abs(a - b) <= 1e-12 * max(abs(a), abs(b)), (a, b)
