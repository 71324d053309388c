"""Batched uneven 1D pores on the device against the oracle, pore by pore
(run on an MI355X).

Every pore of a batch is the problem the oracle (and the reference) solves on
its own; the batch must reproduce each pore's residual and Jacobian products
to 1e-12 and each pore's converged profile to 1e-8, with and without graph
replay, and in mixed precision.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sps

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd.batch import slit_pore_batch  # noqa: E402

RTOL = 1e-12


@pytest.fixture(scope="module")
def small():
    from tramonto_amd.batch import DeviceBatch
    b = slit_pore_batch(npores=5, seed=3, h_range=(3.0, 5.0))
    oracles = [Oracle.from_spec(p) for p in b.pores]
    return b, oracles, DeviceBatch(b)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def generic_state(b, oracles, seed):
    xs = [o.init_fields(p.initial_x()) for p, o in zip(b.pores, oracles)]
    x = np.concatenate(xs, axis=1)
    rng = np.random.default_rng(seed)
    x[b.ncomp:] *= 1.0 + 0.01 * rng.uniform(-1.0, 1.0, x[b.ncomp:].shape)
    return x


@pytest.mark.gpu
def test_batch_init_fields_per_pore(small):
    b, oracles, db = small
    xg = db.init_fields(dev(b.initial_x())).cpu().numpy()
    for xp, p, o in zip(b.split(xg), b.pores, oracles):
        xo = o.init_fields(p.initial_x())
        assert np.abs(xp - xo).max() < RTOL * np.abs(xo).max()


@pytest.mark.gpu
def test_batch_residual_per_pore(small):
    b, oracles, db = small
    x = generic_state(b, oracles, 21)
    Fg, _ = db.fill_residual(dev(x))
    for Fp, xp, o in zip(b.split(Fg.cpu().numpy()), b.split(x), oracles):
        Fo = o.residual(np.ascontiguousarray(xp))
        assert np.abs(Fp - Fo).max() < RTOL * np.abs(Fo).max()


@pytest.mark.gpu
def test_batch_jvp_per_pore(small):
    b, oracles, db = small
    x = generic_state(b, oracles, 22)
    rng = np.random.default_rng(4)
    v = rng.uniform(-1.0, 1.0, x.shape)
    Jv = db.apply_jacobian(dev(x), dev(v)).cpu().numpy()
    for p, o, xp, vp, Jp in zip(b.pores, oracles, b.split(x), b.split(v), b.split(Jv)):
        r, c, val = o.jacobian_coo(np.ascontiguousarray(xp))
        n = p.nunk * p.nnodes
        J = sps.coo_matrix((val, (r, c)), shape=(n, n)).tocsr()
        ref = (J @ np.ascontiguousarray(vp).reshape(-1)).reshape(p.nunk, p.nnodes)
        assert np.abs(Jp - ref).max() < RTOL * np.abs(ref).max()


@pytest.mark.gpu
def test_batch_transpose_is_unsupported(small):
    b, oracles, db = small
    x = db.initial_state()
    with pytest.raises(RuntimeError, match="not implemented"):
        db.apply_jacobian_transpose(x, x)


@pytest.fixture(scope="module")
def oracle_profiles(small):
    b, oracles, _ = small
    out = []
    for p, o in zip(b.pores, oracles):
        xo, st = o.newton(o.init_fields(p.initial_x()), resid_tol=1e-10)
        assert st["converged"], st
        out.append(xo)
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("graph", ["1", "0"])
def test_batch_newton_matches_oracle(graph, small, oracle_profiles, monkeypatch):
    b, oracles, db = small
    monkeypatch.setenv("TMI_BATCH_GRAPH", graph)
    x = db.initial_state()
    st = db.newton(x, resid_tol=1e-10, gmres_tol=1e-9, max_iter=60)
    assert st["converged"], st
    assert st["pore_converged"].all() and (st["pore_resid"] < 1e-10).all()
    assert (st["pore_iters"] > 0).all() and (st["pore_gmres"] > 0).all()
    for xp, xo, p in zip(b.split(x.cpu().numpy()), oracle_profiles, b.pores):
        assert np.abs(xp[: p.ncomp] - xo[: p.ncomp]).max() < 1e-8


@pytest.mark.gpu
def test_batch_newton_mixed_precision(small, oracle_profiles):
    from tramonto_amd.device import PREC_MIXED
    b, oracles, db = small
    x = db.initial_state()
    st = db.newton(x, resid_tol=1e-10, gmres_tol=1e-7, max_iter=60, precision=PREC_MIXED)
    assert st["converged"], st
    for xp, xo, p in zip(b.split(x.cpu().numpy()), oracle_profiles, b.pores):
        assert np.abs(xp[: p.ncomp] - xo[: p.ncomp]).max() < 1e-8


@pytest.mark.gpu
def test_batch_newton_matches_single_pore_solves(small):
    """The batch and one device problem per pore converge to the same profiles."""
    from tramonto_amd.device import DeviceProblem
    b, oracles, db = small
    x = db.initial_state()
    assert db.newton(x, resid_tol=1e-11, gmres_tol=1e-10, max_iter=60)["converged"]
    for xp, p in zip(b.split(x.cpu().numpy()), b.pores):
        dp = DeviceProblem(p)
        xs = dp.initial_state()
        assert dp.newton(xs, resid_tol=1e-11, gmres_tol=1e-10, max_iter=60)["converged"]
        assert np.abs(xp - xs.cpu().numpy()).max() < 1e-9 * max(1.0, np.abs(xp).max())


@pytest.mark.gpu
@pytest.mark.parametrize("per_lane", [4, 8, 16])
@pytest.mark.parametrize("lanes", [64, 128, 256])
def test_every_row_kernel_shape_matches_oracle(small, per_lane, lanes):
    """Every launch shape the 1D autotuner can choose (outputs per lane x lanes per workgroup, for the
    rho-bar, mean-field and Euler-Lagrange row kernels, FP64 and FP32), forced with
    tmi_row_tuning_set: residual and J v per pore against the oracle to 1e-12, the FP32 product
    against FP64 to 1e-5."""
    b, oracles, db = small
    db.set_row_tuning(per_lane, lanes)
    tun = db.row_tuning()
    assert all(tun[p][k][:2] == (per_lane, lanes) for p in tun for k in tun[p] if tun[p][k][0])
    try:
        x = generic_state(b, oracles, 30 + per_lane + lanes)
        Fg, _ = db.fill_residual(dev(x))
        for Fp, xp, o in zip(b.split(Fg.cpu().numpy()), b.split(x), oracles):
            Fo = o.residual(np.ascontiguousarray(xp))
            assert np.abs(Fp - Fo).max() < RTOL * np.abs(Fo).max()
        rng = np.random.default_rng(per_lane * lanes)
        v = rng.uniform(-1.0, 1.0, x.shape)
        Jv = db.apply_jacobian(dev(x), dev(v)).cpu().numpy()
        for p, o, xp, vp, Jp in zip(b.pores, oracles, b.split(x), b.split(v), b.split(Jv)):
            r, c, val = o.jacobian_coo(np.ascontiguousarray(xp))
            n = p.nunk * p.nnodes
            J = sps.coo_matrix((val, (r, c)), shape=(n, n)).tocsr()
            ref = (J @ np.ascontiguousarray(vp).reshape(-1)).reshape(p.nunk, p.nnodes)
            assert np.abs(Jp - ref).max() < RTOL * np.abs(ref).max()
        J32 = db.apply_jacobian(dev(x), dev(v), fp32=True).cpu().numpy()
        assert np.abs(J32 - Jv).max() < 1e-5 * np.abs(Jv).max()
    finally:
        db.set_row_tuning(8, 256)


# ---------------------------------------------------------------- 2D pieces
@pytest.fixture(scope="module")
def small2d():
    """four rough slit pores of 64 columns (periodic x, walls at both y ends) and uneven heights"""
    from tramonto_amd.batch import DeviceBatch, rough_slit_pore_batch_2d
    b = rough_slit_pore_batch_2d(npores=4, seed=5, nx=64, h_range=(3.0, 5.0), dx=0.1)
    oracles = [Oracle.from_spec(p) for p in b.pores]
    return b, oracles, DeviceBatch(b)


@pytest.mark.gpu
def test_batch_2d_residual_and_jvp_per_piece(small2d):
    """2D pieces stacked along y between wall rows: each piece's residual and Jacobian product equal
    the oracle's for that piece on its own (1e-12), at a perturbed state."""
    b, oracles, db = small2d
    assert db.info.fmt_fft == 0 and db.info.mf_fft == 0   # batches of 2D pieces keep the direct sums
    x = generic_state(b, oracles, 31)
    Fg, _ = db.fill_residual(dev(x))
    for Fp, xp, o in zip(b.split(Fg.cpu().numpy()), b.split(x), oracles):
        Fo = o.residual(np.ascontiguousarray(xp))
        assert np.abs(Fp - Fo).max() < RTOL * np.abs(Fo).max()
    rng = np.random.default_rng(6)
    v = rng.uniform(-1.0, 1.0, x.shape)
    Jv = db.apply_jacobian(dev(x), dev(v)).cpu().numpy()
    for p, o, xp, vp, Jp in zip(b.pores, oracles, b.split(x), b.split(v), b.split(Jv)):
        r, c, val = o.jacobian_coo(np.ascontiguousarray(xp))
        n = p.nunk * p.nnodes
        J = sps.coo_matrix((val, (r, c)), shape=(n, n)).tocsr()
        ref = (J @ np.ascontiguousarray(vp).reshape(-1)).reshape(p.nunk, p.nnodes)
        assert np.abs(Jp - ref).max() < RTOL * np.abs(ref).max()


@pytest.mark.gpu
@pytest.mark.parametrize("graph", ["1", "0"])
def test_batch_2d_newton_matches_single_solves(graph, small2d, monkeypatch):
    """The batched per-piece Newton-Krylov on 2D pieces: every piece converges, to the profile of
    the piece solved as one device problem (1e-8; that path's Newton profiles are pinned to the
    oracle's by the parity suite; the first piece is also checked against the oracle here)."""
    from tramonto_amd.device import DeviceProblem
    b, oracles, db = small2d
    monkeypatch.setenv("TMI_BATCH_GRAPH", graph)
    x = db.initial_state()
    st = db.newton(x, resid_tol=1e-10, gmres_tol=1e-9, gmres_restart=100, gmres_max_iter=2000, max_iter=60)
    print("2D batch:", st["pore_iters"], st["pore_gmres"], st["pore_resid"], flush=True)
    assert st["converged"], st
    assert st["pore_converged"].all() and (st["pore_resid"] < 1e-10).all()
    for k, (xp, p, o) in enumerate(zip(b.split(x.cpu().numpy()), b.pores, oracles)):
        if k == 0:
            xo, sto = o.newton(o.init_fields(p.initial_x()), resid_tol=1e-10)
            assert sto["converged"], sto
            assert np.abs(xp[: p.ncomp] - xo[: p.ncomp]).max() < 1e-8
        dp = DeviceProblem(p)
        xs = dp.initial_state()
        assert dp.newton(xs, resid_tol=1e-11, gmres_tol=1e-10, gmres_restart=100, gmres_max_iter=2000,
                         max_iter=60)["converged"]
        assert np.abs(xp[: p.ncomp] - xs.cpu().numpy()[: p.ncomp]).max() < 1e-8
