"""The adjoint harness: torch.autograd.Function over the C ABI (run on an MI355X).

* Residual.backward is the device transpose product: torch.autograd.gradcheck
  compares it with central differences of the device residual.
* The transpose solve J^T lam = g meets its tolerance (checked with an
  independent transpose product).
* dGamma/d eps_w and dGamma/d rho_b (hence dGamma/d mu) of the rough-wall
  adsorption from one backward pass agree with central differences of fully
  re-solved equilibrium profiles.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from tramonto_amd.problem import rough_wall_2d, small_case  # noqa: E402


def rough_model(nx=40, ny=32, dx=0.1, in_place=False):
    from tramonto_amd.adjoint import ParametricModel

    def make(theta):
        return rough_wall_2d(seed=7, nx=nx, ny=ny, dx=dx, cut=1.5, wall_cut=2.0, eps_w=float(theta[0]),
                             rho_b=float(theta[1]))
    return ParametricModel(make, in_place=in_place)


@pytest.mark.gpu
def test_residual_function_gradcheck():
    from tramonto_amd.adjoint import residual
    from tramonto_amd.device import DeviceProblem
    spec = small_case("1d_reflect_wca")
    spec.nx = [41]
    spec.rho0 = spec.rho0[:, :41]
    dp = DeviceProblem(spec)
    x = dp.initial_state().requires_grad_(True)
    assert torch.autograd.gradcheck(lambda t: residual(t, dp), (x,), eps=1e-6, atol=1e-7, rtol=1e-6)


@pytest.mark.gpu
def test_transpose_solve_residual():
    from tramonto_amd.adjoint import transpose_solve
    model = rough_model()
    theta = np.array([1.5, 0.5])
    dp = model.problem(theta)
    x = dp.initial_state()
    st = dp.newton(x, resid_tol=1e-10, gmres_tol=1e-9, max_iter=60)
    assert st["converged"], st
    g = torch.rand_like(x)
    lam, info = transpose_solve(dp, x, g)
    assert info["rel_resid"] < 1e-9, info
    # independent check: <lam, J v> = <g, v> for random v
    v = torch.rand_like(x)
    Jv = dp.apply_jacobian(x, v)
    a, b = float((lam * Jv).sum()), float((g * v).sum())
    assert abs(a - b) < 1e-8 * abs(b), (a, b)


@pytest.mark.gpu
def test_in_place_parameter_updates_match_new_problems(monkeypatch):
    """set_rho_b / set_vext on a problem give the residual of a problem built with the new values
    (bit for bit: both problems get the same stencil-sum form, which set-up otherwise picks by
    timing and which may then differ between two problems in the last bits)."""
    monkeypatch.setenv("TMI_FMT_FFT", "1")
    monkeypatch.setenv("TMI_FFT_XFUSE", "1")
    model = rough_model()
    dp = model.problem(np.array([1.5, 0.5]))
    x = dp.initial_state()
    new = model.make_spec(np.array([1.2, 0.45]))
    dp.set_rho_b(new.rho_b)
    dp.set_vext(new.vext)
    Fa, _ = dp.fill_residual(x)
    Fb, _ = model.problem(np.array([1.2, 0.45])).fill_residual(x)
    assert float((Fa - Fb).abs().max()) == 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("in_place", [False, True])
def test_adsorption_sensitivities_match_finite_differences(in_place):
    from tramonto_amd.adjoint import dbetamu_drho_b, equilibrium_profile, excess_adsorption
    model = rough_model(in_place=in_place)
    theta0 = np.array([1.5, 0.5])
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    x = equilibrium_profile(theta, model)
    spec = model.make_spec(theta0)
    area = spec.nx[1] * spec.dx[1]
    gam = excess_adsorption(x, spec, theta[1:2], area)
    gam.backward()
    grad = theta.grad.numpy().copy()
    assert model.last_adjoint["rel_resid"] < 1e-9, model.last_adjoint

    def gamma_at(th):
        dp = model.problem(th)
        xs = dp.initial_state()
        st = dp.newton(xs, resid_tol=1e-12, gmres_tol=1e-11, max_iter=80)
        assert st["converged"], st
        sp = model.make_spec(th)
        return float(excess_adsorption(xs, sp, torch.tensor(th[1:2]), area))

    for k in range(2):
        h = 1e-4 * theta0[k]
        tp, tm = theta0.copy(), theta0.copy()
        tp[k] += h
        tm[k] -= h
        fd = (gamma_at(tp) - gamma_at(tm)) / (2 * h)
        assert abs(grad[k] - fd) < 1e-5 * max(abs(fd), 1e-3), (k, grad[k], fd)
    # dGamma/d(beta mu) by the chain rule through the bulk equation of state
    dmu = dbetamu_drho_b(spec)[0, 0]
    assert dmu > 0
    assert np.isfinite(grad[1] / dmu)


@pytest.mark.gpu
def test_full_size_rough_wall_jvp_vjp():
    """256 x 256 rough wall at dx = 0.05 (too large for the assembled oracle
    Jacobian): J v against central differences (h = 1e-6, direction relative
    to x) and <w, J v> = <J^T w, v>."""
    from tramonto_amd.device import DeviceProblem
    dp = DeviceProblem(rough_wall_2d())
    x = dp.initial_state()
    g = torch.Generator(device="cuda").manual_seed(3)
    v = (torch.rand(x.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1) * x.abs() * 0.1
    h = 1e-6
    Fp, _ = dp.fill_residual(x + h * v)
    Fm, _ = dp.fill_residual(x - h * v)
    fd = (Fp - Fm) / (2 * h)
    Jv = dp.apply_jacobian(x, v)
    assert float((Jv - fd).abs().max() / fd.abs().max()) < 1e-6
    w = torch.rand(x.shape, generator=g, device="cuda", dtype=torch.float64) * 2 - 1
    JTw = dp.apply_jacobian_transpose(x, w)
    a, b = float((w * Jv).sum()), float((JTw * v).sum())
    assert abs(a - b) < 1e-11 * float((w * Jv).abs().sum())
