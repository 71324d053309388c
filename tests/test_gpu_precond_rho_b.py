"""A new bulk density (tmi_problem_set_rho_b) marks the bulk-symbol Schur preconditioner stale;
the next GMRES (or preconditioner query) rebuilds it from the new rho_b.  On the walled rough-wall
grid the symbol comes from a periodic bulk problem on the padded FFT grid, so a problem created
at rho_b = 0.5 and moved to 0.55 (the liquid side: kT = 1 is below the mean-field critical
temperature) must solve exactly as one created at 0.55: same Newton steps,
same GMRES iterations, same profile."""
import pytest

from tramonto_amd.problem import rough_wall_2d

pytestmark = pytest.mark.gpu


def test_new_bulk_density_rebuilds_the_preconditioner(monkeypatch):
    from tramonto_amd.device import DeviceProblem
    monkeypatch.setenv("TMI_FMT_FFT", "1")   # one FMT form for both problems: identical arithmetic
    monkeypatch.setenv("TMI_FFT_XFUSE", "0")
    moved = DeviceProblem(rough_wall_2d(eps_w=1.5, rho_b=0.5))
    fresh = DeviceProblem(rough_wall_2d(eps_w=1.5, rho_b=0.55))
    assert moved.info.precond == 1 and fresh.info.precond == 1
    moved.set_rho_b([0.55])
    x_fresh = fresh.initial_state()
    x_moved = x_fresh.clone()
    kw = dict(max_iter=30, resid_tol=1e-8, gmres_tol=1e-7)
    s_fresh = fresh.newton(x_fresh, **kw)
    s_moved = moved.newton(x_moved, **kw)
    assert s_fresh["converged"] and s_moved["converged"], (s_fresh, s_moved)
    assert s_moved["iters"] == s_fresh["iters"]
    assert s_moved["gmres_iters"] == s_fresh["gmres_iters"]
    assert float((x_moved - x_fresh).abs().max()) < 1e-12 * float(x_fresh.abs().max())
    assert moved.info.precond == 1
