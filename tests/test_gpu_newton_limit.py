"""Newton's step limit where a density reaches zero (update_solution, src/dft_newton.c:286-358):
frac is the smallest x / (-dx) over the densities, and the density that set it reaches
x + frac dx = 0 in exact arithmetic, which the reference's test x + frac dx < 1e-99 sends to
0.1 x.  In floating point that sum is a residue of either sign; a positive one left the node at
rho ~ 1e-21, whose 1/rho term wrecked the next Jacobian, and Newton stalled or diverged
depending on the last bit of the step.  The rough-wall state (256 x 256, eps_w = 1.5,
rho_b = 0.5) takes such a limited first step (frac ~ 0.585) under every FMT form; each form
must leave no density in (0, 1e-12) after it, converge, and agree with the others."""
import numpy as np
import pytest
import torch

from tramonto_amd.problem import rough_wall_2d

pytestmark = pytest.mark.gpu

FORMS = {"direct": {"TMI_FMT_FFT": "0"},
         "fft": {"TMI_FMT_FFT": "1", "TMI_FFT_XFUSE": "0"},
         "fft_fused_x": {"TMI_FMT_FFT": "1"}}


def make(form, monkeypatch):
    from tramonto_amd.device import DeviceProblem
    for k in ("TMI_FMT_FFT", "TMI_FFT_XFUSE"):
        monkeypatch.delenv(k, raising=False)
    for k, v in FORMS[form].items():
        monkeypatch.setenv(k, v)
    dp = DeviceProblem(rough_wall_2d(eps_w=1.5, rho_b=0.5))
    for k in FORMS[form]:
        monkeypatch.delenv(k, raising=False)
    return dp


def test_limited_step_leaves_no_residue_density(monkeypatch):
    sols = {}
    for form in FORMS:
        dp = make(form, monkeypatch)
        x = dp.initial_state()
        st = dp.newton(x, max_iter=1, resid_tol=1e-8, gmres_tol=1e-7)
        assert st["iters"] == 1
        rho = x[0].cpu().numpy()
        tiny = (rho > 0.0) & (rho < 1e-12)
        assert not tiny.any(), (form, np.sort(rho[rho > 0.0])[:3])
        st = dp.newton(x, max_iter=20, resid_tol=1e-8, gmres_tol=1e-7)
        assert st["converged"] and st["iters"] <= 12, (form, st)
        sols[form] = x.clone()
    ref = sols["direct"]
    for form, x in sols.items():
        assert float((x - ref).abs().max()) < 1e-7 * float(ref.abs().max()), form
