"""The reference's LinProbMgr ABI as NOX and the C Newton driver call it (run on an MI355X).

The call sequences are those of src/NOXLOCA_Tramonto_Group.cpp:176-313:

* computeF: setrhs(0) -> importr2c -> fill_resid_and_matrix_control(x, 0, 1) -> getrhs = -F;
* computeJacobian: initializeproblemvalues -> fill(x, 1, 0) -> finalizeproblemvalues;
* applyJacobian: importr2c -> applymatrix;
* applyJacobianInverse: setrhs(b) -> setupsolver -> solve -> getlhs.

Each result is checked against the oracle: F to 1e-12, the assembled Jacobian entries
(getmatrixvalue, writeMatrix) to 1e-12, and the solve by J dx = b.  The solver options
come from a ParameterList as the reference builds it (dft_parameterlist_set_all_aztec,
src/dft_mesh.c:3187-3188).
"""
import ctypes as ct
import os

import numpy as np
import pytest
import scipy.sparse as sps

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd import _lib  # noqa: E402
from tramonto_amd.problem import small_case  # noqa: E402

RTOL = 1e-12
vp = lambda a: a.ctypes.data_as(ct.c_void_p)


def rows(a):
    P = ct.POINTER(ct.c_double)
    arr = (P * a.shape[0])()
    for i in range(a.shape[0]):
        arr[i] = a[i].ctypes.data_as(P)
    return ct.cast(arr, ct.c_void_p)


def aztec_list(L, max_iter=500, tol=1e-11, kspace=500):
    """The ParameterList of src/dft_mesh.c:3130-3188 (options/params arrays, AztecOO indices)."""
    opts = np.zeros(47, dtype=np.int32)
    pars = np.zeros(30)
    opts[6], opts[10], pars[0] = max_iter, kspace, tol   # AZ_max_iter, AZ_kspace, AZ_tol
    plist = L.dft_parameterlist_create()
    L.dft_parameterlist_set_all_aztec(plist, vp(opts), vp(pars))
    return plist


def state(spec, o, seed):
    x = o.init_fields(spec.initial_x())
    rng = np.random.default_rng(seed)
    x[spec.ncomp:] *= 1.0 + 0.01 * rng.uniform(-1.0, 1.0, x[spec.ncomp:].shape)
    return np.ascontiguousarray(x)


def hs_ids(spec):
    C = spec.ncomp
    return np.arange(C, spec.nunk, dtype=np.int32), np.zeros(0, dtype=np.int32), np.arange(C, dtype=np.int32)


def make_manager(L, spec, plist, kind="hs"):
    create = {"hs": L.dft_hardsphere_lin_prob_mgr_create, "debug": L.dft_hardsphere_lin_prob_mgr_create_debug,
              "basic": L.dft_basic_lin_prob_mgr_create}[kind]
    mgr = create(spec.nunk, plist, 0)
    N = spec.nnodes
    gids = np.arange(N, dtype=np.int32)
    if kind != "basic":
        ind, dep, dens = hs_ids(spec)
        assert L.dft_hardsphere_lin_prob_mgr_setindnonlocalequationids(mgr, len(ind), vp(ind)) == 0
        assert L.dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids(mgr, 0, None) == 0
        assert L.dft_hardsphere_lin_prob_mgr_setdensityequationids(mgr, len(dens), vp(dens)) == 0
        assert L.dft_hardsphere_lin_prob_mgr_seta22blockisdiagonal(mgr, 0 if spec.type_attr >= 0 else 1) == 0
    assert L.dft_linprobmgr_setnodalrowmap(mgr, N, vp(gids)) == 0
    assert L.dft_linprobmgr_setnodalcolmap(mgr, N, vp(gids)) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(mgr) == 0
    return mgr


@pytest.fixture(scope="module")
def setup():
    from tramonto_amd.device import DeviceProblem
    spec = small_case("1d_reflect_wca")
    o = Oracle.from_spec(spec)
    dp = DeviceProblem(spec)
    yield spec, o, dp
    _lib.lib().tmi_set_active_problem(None, None)


@pytest.mark.gpu
def test_nox_compute_f_jacobian_and_inverse(setup):
    spec, o, dp = setup
    L = _lib.lib()
    plist = aztec_list(L)
    mgr = make_manager(L, spec, plist)
    assert L.tmi_set_active_problem(dp.handle, mgr) == 0
    x = state(spec, o, 3)
    xbox = np.zeros_like(x)
    f = np.zeros_like(x)
    Fo = o.residual(x)
    # computeF
    assert L.dft_linprobmgr_setrhs(mgr, rows(f)) == 0
    assert L.dft_linprobmgr_importr2c(mgr, rows(x), rows(xbox)) == 0
    assert np.array_equal(xbox, x)
    l2 = L.fill_resid_and_matrix_control(rows(xbox), 0, 1)
    assert L.dft_linprobmgr_getrhs(mgr, rows(f)) == 0
    assert np.abs(-f - Fo).max() <= RTOL * np.abs(Fo).max()
    assert abs(l2 - (Fo ** 2).sum()) <= 1e-10 * (Fo ** 2).sum()
    # computeJacobian: the rhs is reset by initializeproblemvalues and refilled
    assert L.dft_linprobmgr_initializeproblemvalues(mgr) == 0
    L.fill_resid_and_matrix_control(rows(xbox), 1, 0)
    assert L.dft_linprobmgr_finalizeproblemvalues(mgr) == 0
    assert L.dft_linprobmgr_getrhs(mgr, rows(f)) == 0
    assert np.abs(-f - Fo).max() <= RTOL * np.abs(Fo).max()
    # applyJacobian against the assembled Jacobian
    n = spec.nunk * spec.nnodes
    r, c, vals = o.jacobian_coo(x)
    J = sps.coo_matrix((vals, (r, c)), shape=(n, n)).tocsr()
    v = np.ascontiguousarray(np.random.default_rng(5).uniform(-1, 1, x.shape))
    vbox, Jv = np.zeros_like(v), np.zeros_like(v)
    assert L.dft_linprobmgr_importr2c(mgr, rows(v), rows(vbox)) == 0
    assert L.dft_linprobmgr_applymatrix(mgr, rows(vbox), rows(Jv)) == 0
    Jvo = (J @ v.reshape(-1)).reshape(x.shape)
    assert np.abs(Jv - Jvo).max() <= RTOL * np.abs(Jvo).max()
    # applyJacobianInverse
    b = np.ascontiguousarray(np.random.default_rng(6).uniform(-1, 1, x.shape))
    dx = np.zeros_like(b)
    assert L.dft_linprobmgr_setrhs(mgr, rows(b)) == 0
    assert L.dft_linprobmgr_setupsolver(mgr) == 0
    assert L.dft_linprobmgr_solve(mgr) == 0
    assert L.dft_linprobmgr_getlhs(mgr, rows(dx)) == 0
    res = J @ dx.reshape(-1) - b.reshape(-1)
    assert np.linalg.norm(res) <= 1e-8 * np.linalg.norm(b), np.linalg.norm(res) / np.linalg.norm(b)
    it = ct.c_int()
    assert L.tmi_linprobmgr_last_solve(mgr, ct.byref(it), None) == 0 and it.value > 1
    L.dft_linprobmgr_destruct(mgr)
    L.dft_parameterlist_destruct(plist)


@pytest.mark.gpu
def test_getmatrixvalue_and_writematrix_match_assembled_entries(setup, tmp_path):
    spec, o, dp = setup
    L = _lib.lib()
    mgr = make_manager(L, spec, None)
    assert L.tmi_set_active_problem(dp.handle, mgr) == 0
    x = state(spec, o, 4)
    assert L.dft_linprobmgr_initializeproblemvalues(mgr) == 0
    L.fill_resid_and_matrix_control(rows(x), 1, 0)
    N, nunk = spec.nnodes, spec.nunk
    n = nunk * N
    r, c, vals = o.jacobian_coo(x)
    J = sps.coo_matrix((vals, (r, c)), shape=(n, n)).toarray()
    scale = np.abs(J).max()
    rng = np.random.default_rng(9)
    nz = np.flatnonzero(J)
    picks = list(rng.choice(nz, 40, replace=False)) + list(rng.integers(0, n * n, 20))
    for k in picks:
        i, j = divmod(int(k), n)
        got = L.dft_linprobmgr_getmatrixvalue(mgr, i // N, i % N, j // N, j % N)
        assert abs(got - J[i, j]) <= RTOL * scale, (i, j, got, J[i, j])
    # out-of-range indices are rejected (0.0 and an error message), not read past the buffers
    assert L.dft_linprobmgr_getmatrixvalue(mgr, 0, N, 0, 0) == 0.0
    assert L.dft_linprobmgr_getmatrixvalue(mgr, 0, 0, nunk, 0) == 0.0
    fn = str(tmp_path / "J.mm")
    assert L.dft_linprobmgr_writeMatrix(mgr, fn.encode(), b"J", b"test") == 0
    lines = [ln for ln in open(fn) if not ln.startswith("%")]
    m, m2, nnz = map(int, lines[0].split())
    assert m == n and m2 == n and nnz == len(lines) - 1
    W = np.zeros((n, n))
    for ln in lines[1:]:
        a, b2, v = ln.split()
        W[int(a) - 1, int(b2) - 1] = float(v)
    assert np.abs(W - J).max() <= RTOL * scale
    L.dft_linprobmgr_destruct(mgr)


@pytest.mark.gpu
def test_residual_only_fill_keeps_the_linearisation_point(setup):
    """Only a full fill (flag 0) moves the Jacobian's point, as only it changes the reference's
    matrix: after a full fill at x0, resid-only fills (flags 1 and 4, as NOX computeF and the
    numerical Jacobian make them) at x1 leave applymatrix at J(x0)."""
    spec, o, dp = setup
    L = _lib.lib()
    mgr = make_manager(L, spec, None)
    assert L.tmi_set_active_problem(dp.handle, mgr) == 0
    x0, x1 = state(spec, o, 1), state(spec, o, 2)
    x1[: spec.ncomp] *= 1.05
    assert L.dft_linprobmgr_initializeproblemvalues(mgr) == 0
    L.fill_resid_and_matrix_control(rows(x0), 1, 0)
    for flag in (1, 4, 2, 3):
        f = np.zeros_like(x1)
        assert L.dft_linprobmgr_setrhs(mgr, rows(f)) == 0
        l2 = L.fill_resid_and_matrix_control(rows(x1), 0, flag)
        F1 = o.residual(x1)
        assert abs(l2 - (F1 ** 2).sum()) <= 1e-10 * (F1 ** 2).sum()
        assert L.dft_linprobmgr_getrhs(mgr, rows(f)) == 0
        if flag in (1, 4):
            assert np.abs(-f - F1).max() <= RTOL * np.abs(F1).max()
        else:
            assert not f.any()
    v = np.ascontiguousarray(np.random.default_rng(2).uniform(-1, 1, x0.shape))
    Jv = np.zeros_like(v)
    assert L.dft_linprobmgr_applymatrix(mgr, rows(v), rows(Jv)) == 0
    ref0 = dp.apply_jacobian(torch.from_numpy(x0).cuda(), torch.from_numpy(v).cuda()).cpu().numpy()
    ref1 = dp.apply_jacobian(torch.from_numpy(x1).cuda(), torch.from_numpy(v).cuda()).cpu().numpy()
    assert np.abs(Jv - ref0).max() <= 1e-13 * np.abs(ref0).max()
    assert np.abs(ref1 - ref0).max() > 1e-6 * np.abs(ref0).max()
    L.dft_linprobmgr_destruct(mgr)


@pytest.mark.gpu
def test_parameter_list_debug_and_basic_managers(setup, capfd):
    """Max_iter and Tol come from the ParameterList at solve time; the debug manager reports
    ||J dx - b||; the basic manager (no equation IDs) returns the same solution."""
    spec, o, dp = setup
    L = _lib.lib()
    x = state(spec, o, 7)
    b = np.ascontiguousarray(np.random.default_rng(8).uniform(-1, 1, x.shape))

    def solve(kind, plist):
        mgr = make_manager(L, spec, plist, kind)
        assert L.tmi_set_active_problem(dp.handle, mgr) == 0
        assert L.dft_linprobmgr_initializeproblemvalues(mgr) == 0
        L.fill_resid_and_matrix_control(rows(x), 1, 0)
        assert L.dft_linprobmgr_setrhs(mgr, rows(b)) == 0
        assert L.dft_linprobmgr_setupsolver(mgr) == 0
        assert L.dft_linprobmgr_solve(mgr) == 0
        dx = np.zeros_like(b)
        assert L.dft_linprobmgr_getlhs(mgr, rows(dx)) == 0
        it, dr = ct.c_int(), ct.c_double()
        assert L.tmi_linprobmgr_last_solve(mgr, ct.byref(it), ct.byref(dr)) == 0
        L.dft_linprobmgr_destruct(mgr)
        return dx, it.value, dr.value

    short = aztec_list(L, max_iter=2, tol=1e-14)
    _, its, _ = solve("hs", short)
    assert its == 2
    full = aztec_list(L, max_iter=500, tol=1e-11)
    dx_hs, its_hs, _ = solve("hs", full)
    assert its_hs > 2
    capfd.readouterr()
    dx_dbg, _, dres = solve("debug", full)
    out = capfd.readouterr().out
    assert "Global Residual for solution" in out
    assert 0.0 <= dres <= 1e-9 * np.linalg.norm(b)
    dx_basic, _, _ = solve("basic", full)
    assert np.abs(dx_dbg - dx_hs).max() <= 1e-13 * np.abs(dx_hs).max()
    assert np.abs(dx_basic - dx_hs).max() <= 1e-13 * np.abs(dx_hs).max()
    L.dft_parameterlist_destruct(short)
    L.dft_parameterlist_destruct(full)


@pytest.mark.gpu
def test_block_ids_are_checked_against_the_problem(setup):
    """linsolver_setup_HSTYPE's lists (src/dft_linsolver_setup.c:182-239) must match the problem's
    layout: a density row registered as nonlocal is refused at finalize and at registration."""
    spec, o, dp = setup
    L = _lib.lib()
    N = spec.nnodes
    gids = np.arange(N, dtype=np.int32)
    bad_ind = np.arange(0, spec.nunk - 1, dtype=np.int32)
    dens = np.array([spec.nunk - 1], dtype=np.int32)
    mgr = L.dft_hardsphere_lin_prob_mgr_create(spec.nunk, None, 0)
    assert L.dft_hardsphere_lin_prob_mgr_setindnonlocalequationids(mgr, len(bad_ind), vp(bad_ind)) == 0
    assert L.dft_hardsphere_lin_prob_mgr_setdensityequationids(mgr, 1, vp(dens)) == 0
    assert L.dft_linprobmgr_setnodalrowmap(mgr, N, vp(gids)) == 0
    assert L.dft_linprobmgr_setnodalcolmap(mgr, N, vp(gids)) == 0
    assert L.tmi_set_active_problem(dp.handle, mgr) == 0        # not finalised yet: nothing to check
    assert L.dft_linprobmgr_finalizeblockstructure(mgr) == -1
    assert b"density equation IDs" in L.tmi_last_error()
    L.dft_linprobmgr_destruct(mgr)
    good = make_manager(L, spec, None)
    short = np.arange(N - 1, dtype=np.int32)
    assert L.dft_linprobmgr_setnodalrowmap(good, N - 1, vp(short)) == 0
    assert L.tmi_set_active_problem(dp.handle, good) == -1      # row map shorter than the problem
    L.dft_linprobmgr_destruct(good)
    L.tmi_set_active_problem(None, None)
