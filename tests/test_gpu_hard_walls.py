"""Hard surfaces on the device (Lhard_surf = 1, no attraction; run on an MI355X).

The reference's hard-wall Examples (tests/golden/examples_hs_1d.json): the fill
with the boundary-node weights of HW_boundary_weight (src/dft_utils.c:41-62 in
src/dft_fill_CLSmf.c:140-160) against the oracle to 1e-12 at perturbed states
(F, J v against the assembled J entry by entry, J^T w), and the device Newton
solve with the device observables and wall forces (tmi_compute_force)
reproducing the README adsorption, energy and force to the printed digits.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd.examples import force_weights, integration_weights, spec_from_tramonto  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "examples_hs_1d.json")) as f:
    EXAMPLES = json.load(f)


def printed_tol(text):
    dec = len(text.split(".")[1]) if "." in text else 0
    return 0.5 * 10.0 ** (-dec) + 1e-9


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def perturbed(spec, o, seed):
    x = o.init_fields(spec.initial_x())
    rng = np.random.default_rng(seed)
    C = spec.ncomp
    x[:C] *= 1.0 + 0.1 * rng.uniform(-1.0, 1.0, x[:C].shape)
    x[C:] *= 1.0 + 0.05 * rng.uniform(-1.0, 1.0, x[C:].shape)   # rho-bar rows off their definitions
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_residual_and_products_match_oracle(name):
    from tramonto_amd.device import DeviceProblem
    g = EXAMPLES[name]
    spec = spec_from_tramonto(g, name)
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec)
    x = perturbed(spec, o, 5)
    F, _ = dp.fill_residual(dev(x))
    ref = o.residual(x)
    assert np.abs(F.cpu().numpy().ravel() - ref.ravel()).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    J = o.jacobian_dense(x)
    rng = np.random.default_rng(7)
    v = rng.uniform(-1.0, 1.0, x.shape)
    Jv = dp.apply_jacobian(dev(x), dev(v)).cpu().numpy().ravel()
    r = J @ v.ravel()
    assert np.abs(Jv - r).max() <= 1e-12 * np.abs(r).max()
    w = rng.uniform(-1.0, 1.0, x.shape)
    JTw = dp.apply_jacobian_transpose(dev(x), dev(w)).cpu().numpy().ravel()
    r = J.T @ w.ravel()
    assert np.abs(JTw - r).max() <= 1e-12 * np.abs(r).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_device_newton_reproduces_hard_wall_example(name):
    from tramonto_amd.device import DeviceProblem
    g = EXAMPLES[name]
    ref = g["key_outputs"]
    spec = spec_from_tramonto(g, name)
    dp = DeviceProblem(spec)
    x = dp.initial_state()
    st = dp.newton(x, max_iter=g["max_nl_iter"], nl_rel_tol=g["nl_rel_tol"], nl_abs_tol=g["nl_abs_tol"],
                   update_scaling=g["nl_update_scaling"], resid_tol=1e-10, gmres_tol=1e-9, gmres_restart=100,
                   gmres_max_iter=2000)
    assert st["converged"], st
    o = Oracle.from_spec(spec)
    xo, sto = o.newton(o.init_fields(spec.initial_x()), resid_tol=1e-10, update_scaling=g["nl_update_scaling"])
    xg = x.cpu().numpy()
    assert np.abs(xg[: spec.ncomp] - xo[: spec.ncomp]).max() < 1e-8
    h1, h2, scale = integration_weights(spec, g)
    obs = dp.observables(x, h1, h2, scale)
    assert abs(obs["omega_s"] - ref["energy"]) < printed_tol(ref["energy_text"]), (obs, ref)
    for c in range(spec.ncomp):
        assert abs(obs["ads"][c] - ref["adsorption"][c]) < 2e-6, (obs, ref)
    W = force_weights(spec, g)
    force = dp.force(x, W)
    assert min(abs(f - ref["force"]) for f in force) < printed_tol(ref["force_text"]), (force, ref)
    fo = [(W[k] * xg[: spec.ncomp]).sum() for k in range(W.shape[0])]
    assert np.abs(np.array(fo) - force).max() < 1e-12 * max(1.0, np.abs(fo).max())
