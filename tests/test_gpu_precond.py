"""The bulk-circulant Schur preconditioner (run on an MI355X).

GMRES is right-preconditioned, so the preconditioner changes the Krylov
iterates, not the solution: every solve must reach the same Newton update and
the same converged profile as the reference's A22-diagonal preconditioner
(and the oracle) to the solver tolerance.  At the uniform bulk state the
preconditioner is the exact inverse of the Schur operator, so GMRES must
converge in one iteration there.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd.problem import small_case  # noqa: E402

PERIODIC = ["2d_periodic", "3d_periodic_small", "3d_periodic_wb", "3d_mix_periodic", "3d_solute_small"]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


@pytest.fixture(scope="module")
def problems():
    from tramonto_amd.device import DeviceProblem
    cache = {}

    def get(case):
        if case not in cache:
            spec = small_case(case)
            cache[case] = (spec, Oracle.from_spec(spec), DeviceProblem(spec))
        return cache[case]
    return get


def generic_state(spec, o, seed):
    x = o.init_fields(spec.initial_x())
    rng = np.random.default_rng(seed)
    x[spec.ncomp:] *= 1.0 + 0.01 * rng.uniform(-1.0, 1.0, x[spec.ncomp:].shape)
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("case", PERIODIC)
def test_bulk_preconditioner_is_default_on_periodic_grids(case, problems):
    from tramonto_amd.device import PRECOND_BULK
    spec, o, dp = problems(case)
    assert dp.info.precond == PRECOND_BULK


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["3d_periodic_small", "3d_mix_periodic", "2d_periodic"])
def test_bulk_preconditioner_exact_at_bulk(case, problems):
    """At x = uniform bulk (no zero rows) the preconditioned Schur operator is the identity."""
    from tramonto_amd.device import PRECOND_BULK, PRECOND_DIAG
    spec, o, dp = problems(case)
    x = np.zeros((spec.nunk, spec.nnodes))
    for c in range(spec.ncomp):
        x[c] = spec.rho_b[c]
    xg = dp.init_fields(dev(x))
    rng = np.random.default_rng(5)
    F = dev(rng.uniform(-1.0, 1.0, (spec.nunk, spec.nnodes)))
    try:
        dp.set_preconditioner(PRECOND_DIAG)
        _, its_diag = dp.schur_solve(xg, F, rtol=1e-10)
        dp.set_preconditioner(PRECOND_BULK)
        _, its_bulk = dp.schur_solve(xg, F, rtol=1e-10)
    finally:
        dp.set_preconditioner(PRECOND_BULK)
    assert its_bulk <= 2 < its_diag, (its_bulk, its_diag)


def near_bulk_state(spec, dp, seed):
    """uniform bulk densities perturbed by 5 %, with their rho-bar and mean-field rows"""
    x = np.zeros((spec.nunk, spec.nnodes))
    rng = np.random.default_rng(seed)
    for c in range(spec.ncomp):
        x[c] = spec.rho_b[c] * (1.0 + 0.05 * rng.uniform(-1.0, 1.0, spec.nnodes))
    return dp.init_fields(dev(x))


@pytest.mark.gpu
@pytest.mark.parametrize("case", PERIODIC)
@pytest.mark.parametrize("state", ["seeded", "near_bulk"])
def test_bulk_preconditioner_same_update(case, state, problems):
    """The same Newton update with either preconditioner, at the seeded state (densities from the
    clip bounds 0.05 to 0.9, far from the uniform fluid) and near the bulk.  Near the bulk, where
    the symbol is nearly the operator's inverse, it also takes fewer GMRES iterations; far from it
    the A22 diagonal (the local densities) can win, so there only the update is compared."""
    from tramonto_amd.device import PRECOND_BULK, PRECOND_DIAG
    spec, o, dp = problems(case)
    if state == "seeded":
        x = dev(generic_state(spec, o, 41))
    else:
        x = near_bulk_state(spec, dp, 43)
    F, _ = dp.fill_residual(x)
    out = {}
    try:
        for kind in (PRECOND_DIAG, PRECOND_BULK):
            dp.set_preconditioner(kind)
            out[kind] = dp.schur_solve(x, F, rtol=1e-12, restart=200, maxit=2000)
    finally:
        dp.set_preconditioner(PRECOND_BULK)
    a, b = out[PRECOND_BULK][0], out[PRECOND_DIAG][0]
    assert float((a - b).abs().max() / b.abs().max()) < 1e-7
    print(f"{case} {state}: GMRES iterations bulk {out[PRECOND_BULK][1]}, diagonal {out[PRECOND_DIAG][1]}")
    if state == "near_bulk" and case != "3d_solute_small":   # the solute's zero rows are no bulk
        assert out[PRECOND_BULK][1] < out[PRECOND_DIAG][1], (out[PRECOND_BULK][1], out[PRECOND_DIAG][1])


@pytest.fixture(scope="module")
def oracle_profiles(problems):
    cache = {}

    def get(case):
        if case not in cache:
            spec, o, _ = problems(case)
            x0 = o.init_fields(spec.initial_x())
            xo, st = o.newton(x0, resid_tol=1e-10)
            assert st["converged"], st
            cache[case] = (x0, xo)
        return cache[case]
    return get


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", ["3d_periodic_small", "3d_mix_periodic", "3d_solute_small"])
def test_bulk_preconditioned_newton_matches_oracle(case, precision, problems, oracle_profiles):
    spec, o, dp = problems(case)
    x0, xo = oracle_profiles(case)
    xg = dev(x0)
    st = dp.newton(xg, resid_tol=1e-10, gmres_tol=1e-8, max_iter=60, precision=precision)
    assert st["converged"], st
    assert np.abs(xg.cpu().numpy()[: spec.ncomp] - xo[: spec.ncomp]).max() < 1e-8


@pytest.mark.gpu
def test_bulk_preconditioner_follows_rho_b(problems):
    """After set_rho_b the symbol is rebuilt: still exact at the new bulk state."""
    spec, o, dp = problems("3d_periodic_small")
    old = list(spec.rho_b)
    try:
        dp.set_rho_b([0.45])
        x = np.zeros((spec.nunk, spec.nnodes))
        x[0] = 0.45
        xg = dp.init_fields(dev(x))
        F = dev(np.random.default_rng(6).uniform(-1.0, 1.0, (spec.nunk, spec.nnodes)))
        _, its = dp.schur_solve(xg, F, rtol=1e-10)
        assert its <= 2, its
    finally:
        dp.set_rho_b(old)


WALLED = ["2d_wall_reflect", "2d_rough_wall", "3d_mix_bulk"]


@pytest.mark.gpu
@pytest.mark.parametrize("case", WALLED)
def test_walled_grid_preconditioner_same_update(case, problems):
    """Walls, reflecting or bulk boundaries: the FFT grid pads those axes and the symbol comes from
    a periodic bulk fluid on that grid.  It is the default there too; the Schur update equals the
    diagonal-preconditioned one to the solver tolerance at a seeded state, in fewer GMRES iterations
    on the rough wall (the 2D workload's geometry)."""
    from tramonto_amd.device import PRECOND_BULK, PRECOND_DIAG
    spec, o, dp = problems(case)
    assert dp.info.precond == PRECOND_BULK
    x = dev(generic_state(spec, o, 47))
    F, _ = dp.fill_residual(x)
    out = {}
    try:
        for kind in (PRECOND_DIAG, PRECOND_BULK):
            dp.set_preconditioner(kind)
            out[kind] = dp.schur_solve(x, F, rtol=1e-12, restart=200, maxit=3000)
    finally:
        dp.set_preconditioner(PRECOND_BULK)
    a, b = out[PRECOND_BULK][0], out[PRECOND_DIAG][0]
    assert float((a - b).abs().max() / b.abs().max()) < 1e-7
    print(f"{case}: GMRES iterations bulk symbol {out[PRECOND_BULK][1]}, diagonal {out[PRECOND_DIAG][1]}")
    if case == "2d_rough_wall":
        assert out[PRECOND_BULK][1] < out[PRECOND_DIAG][1], (out[PRECOND_BULK][1], out[PRECOND_DIAG][1])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", WALLED)
def test_walled_grid_preconditioned_newton_matches_oracle(case, precision, problems, oracle_profiles):
    spec, o, dp = problems(case)
    x0, xo = oracle_profiles(case)
    xg = dev(x0)
    st = dp.newton(xg, resid_tol=1e-10, gmres_tol=1e-8, max_iter=60, precision=precision)
    assert st["converged"], st
    assert np.abs(xg.cpu().numpy()[: spec.ncomp] - xo[: spec.ncomp]).max() < 1e-8


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["3d_periodic_small", "3d_mix_periodic", "2d_periodic", "3d_solute_small"])
def test_fused_preconditioner_matches_separate_kernels(case, monkeypatch):
    """The preconditioner applied through fused-end passes and the fused x pass (real symbol tables)
    and through the separate gather, complex product and extract kernels: the same Krylov iterates,
    so the same iteration count and the same update to round-off."""
    from tramonto_amd.device import DeviceProblem
    spec = small_case(case)
    o = Oracle.from_spec(spec)
    out = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("TMI_PRECOND_FUSED", fused)
        dp = DeviceProblem(spec)
        x = dev(generic_state(spec, o, 43))
        F, _ = dp.fill_residual(x)
        out[fused] = dp.schur_solve(x, F, rtol=1e-10, restart=200, maxit=2000)
    a, b = out["1"][0], out["0"][0]
    assert out["1"][1] == out["0"][1], (out["1"][1], out["0"][1])
    assert float((a - b).abs().max() / b.abs().max()) < 1e-10
