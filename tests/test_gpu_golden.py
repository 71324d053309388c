"""The device path against the reference Examples (run on an MI355X).

1. The device Newton-Krylov solve reproduces each 1D Example's README
   adsorption (absolute 1e-3, the reference harness' rule) and the oracle's
   converged profile to 1e-8; the device observables (tmi_compute_observables)
   of that profile reproduce the README energy and adsorption to 2e-6 and the
   oracle's observables of the same profile to 1e-12.
2. The reference's own Newton call sequence (src/dft_newton.c:131-196) driven
   through the LinProbMgr C ABI and fill_resid_and_matrix_control, with
   host double** arrays exactly as the reference passes them, converges to
   the same profile.
"""
import ctypes as ct
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd import _lib  # noqa: E402
from tramonto_amd.examples import adsorption, integration_counts, spec_from_tramonto  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "examples_1d.json")) as f:
    EXAMPLES = json.load(f)


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(EXAMPLES))
def test_device_newton_reproduces_example(name):
    from tramonto_amd.device import DeviceProblem
    g = EXAMPLES[name]
    spec = spec_from_tramonto(g, name)
    dp = DeviceProblem(spec)
    x = dp.initial_state()
    st = dp.newton(x, max_iter=g["max_nl_iter"], nl_rel_tol=g["nl_rel_tol"], nl_abs_tol=g["nl_abs_tol"],
                   update_scaling=g["nl_update_scaling"], resid_tol=1e-10, gmres_tol=1e-9, gmres_restart=100,
                   gmres_max_iter=2000)
    assert st["converged"], st
    xg = x.cpu().numpy()
    ads = adsorption(spec, g, xg[: spec.ncomp])
    for c in range(spec.ncomp):
        assert abs(ads[c] - g["key_outputs"]["adsorption"][c]) < 1e-3, (ads, g["key_outputs"])
    o = Oracle.from_spec(spec)
    xo, sto = o.newton(o.init_fields(spec.initial_x()), resid_tol=1e-10, update_scaling=g["nl_update_scaling"])
    assert sto["converged"]
    assert np.abs(xg[: spec.ncomp] - xo[: spec.ncomp]).max() < 1e-8
    h = integration_counts(spec, g)
    obs = dp.observables(x, h)
    assert abs(obs["omega_s"] - g["key_outputs"]["energy"]) < 2e-6, (obs, g["key_outputs"])
    for c in range(spec.ncomp):
        assert abs(obs["ads"][c] - g["key_outputs"]["adsorption"][c]) < 2e-6, (obs, g["key_outputs"])
    ref = o.observables(xg, h, h)
    for k in ("omega", "omega_s", "ideal", "mu", "vext", "hs", "att"):
        assert abs(obs[k] - ref[k]) <= 1e-12 * max(1.0, abs(ref[k])), (k, obs[k], ref[k])


def rows(a):
    """double** view of a C-contiguous [nunk][N] array."""
    P = ct.POINTER(ct.c_double)
    arr = (P * a.shape[0])()
    for i in range(a.shape[0]):
        arr[i] = a[i].ctypes.data_as(P)
    return arr


def update_solution(x, dx, ncomp, scaling):
    """The caller-side update of src/dft_newton.c:266-380 (positivity fraction, clamps)."""
    rho, drho = x[:ncomp], dx[:ncomp]
    neg = rho + drho < 0.0
    frac = 1.0
    if neg.any():
        f = np.maximum(np.minimum(1.0, rho[neg] / (-drho[neg])), scaling)
        frac = min(1.0, f.min())
    new = x + frac * dx
    low = (new[:ncomp] < 1e-99)
    x[:ncomp] = np.where(low, 0.1 * x[:ncomp], new[:ncomp])
    n3 = x[ncomp]
    big = new[ncomp] > 1.0
    x[ncomp] = np.where(big, n3 + 0.5 * (1.0 - n3), new[ncomp])
    x[ncomp + 1:] = new[ncomp + 1:]


@pytest.mark.gpu
@pytest.mark.parametrize("att_in_a22", [0, 1])
def test_reference_abi_newton_sequence(att_in_a22):
    """initializeproblemvalues -> fill(x, iter, FALSE) -> finalizeproblemvalues
    -> setupsolver -> solve -> getlhs -> x += dx, as dft_newton.c does; in the
    mean-field layout and in the example's own ATTInA22Block = 1 layout."""
    import dataclasses
    from tramonto_amd.device import DeviceProblem
    L = _lib.lib()
    g = EXAMPLES["LJ1_FMT1_WCA_1D"]
    spec0 = spec_from_tramonto(g, "LJ1")
    spec = dataclasses.replace(spec0, att_in_a22=att_in_a22)
    dp = DeviceProblem(spec)
    nunk, N = spec.nunk, spec.nnodes
    mgr = L.dft_hardsphere_lin_prob_mgr_create(nunk, None, 0)
    ind = np.arange(spec.ncomp, nunk, dtype=np.int32)
    dens = np.arange(0, spec.ncomp, dtype=np.int32)
    gids = np.arange(N, dtype=np.int32)
    vp = lambda a: a.ctypes.data_as(ct.c_void_p)
    assert L.dft_hardsphere_lin_prob_mgr_setindnonlocalequationids(mgr, len(ind), vp(ind)) == 0
    assert L.dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids(mgr, 0, None) == 0
    assert L.dft_hardsphere_lin_prob_mgr_setdensityequationids(mgr, len(dens), vp(dens)) == 0
    assert L.dft_linprobmgr_setnodalrowmap(mgr, N, vp(gids)) == 0
    assert L.dft_linprobmgr_setnodalcolmap(mgr, N, vp(gids)) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(mgr) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(mgr) == 1
    assert L.tmi_set_active_problem(dp.handle, mgr) == 0
    x = np.ascontiguousarray(dp.initial_state().cpu().numpy())
    dx = np.zeros_like(x)
    xr, dxr = rows(x), rows(dx)
    converged = False
    for it in range(1, 30):
        assert L.dft_linprobmgr_initializeproblemvalues(mgr) == 0
        l2 = L.fill_resid_and_matrix_control(ct.cast(xr, ct.c_void_p), it, 0)
        assert l2 >= 0.0
        assert L.dft_linprobmgr_finalizeproblemvalues(mgr) == 0
        if np.sqrt(l2) < 1e-10:
            converged = True
            break
        assert L.dft_linprobmgr_setupsolver(mgr) == 0
        assert L.dft_linprobmgr_solve(mgr) == 0
        assert L.dft_linprobmgr_getlhs(mgr, ct.cast(dxr, ct.c_void_p)) == 0
        update_solution(x, dx, spec.ncomp, g["nl_update_scaling"])
    assert converged
    o = Oracle.from_spec(spec0)
    xo, _ = o.newton(o.init_fields(spec0.initial_x()), resid_tol=1e-10, update_scaling=1.0)
    assert np.abs(x[: spec.ncomp] - xo[: spec.ncomp]).max() < 1e-8
    # applymatrix: J dx against the device product
    v = np.ascontiguousarray(np.random.default_rng(1).uniform(-1, 1, x.shape))
    b = np.zeros_like(v)
    assert L.dft_linprobmgr_applymatrix(mgr, ct.cast(rows(v), ct.c_void_p), ct.cast(rows(b), ct.c_void_p)) == 0
    Jv = dp.apply_jacobian(torch.from_numpy(x).cuda(), torch.from_numpy(v).cuda()).cpu().numpy()
    assert np.abs(b - Jv).max() <= 1e-13 * np.abs(Jv).max()
    L.dft_linprobmgr_destruct(mgr)
    L.tmi_set_active_problem(None, None)
