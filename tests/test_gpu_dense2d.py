"""The dense 2D convolution of the direct attraction rows against the per-point gather kernel it
replaces (TMI_STENCIL_KERNEL=gather): residual and Jacobian products in FP64 (1e-13 of the largest
row; the FP64 sums run on the matrix cores, k_mf_rows_2d_mfma and for one component the rho-bar and
Euler-Lagrange rows k_rb_rows_2d_mfma, k_el_rows_2d_mfma, or with TMI_MF2D_MFMA=0 TMI_RB2D_MFMA=0 in
the FMA kernel k_mf_rows_2d and the gather kernels) and in the FP32 sums of the mixed path (1e-5), on
walled/reflecting and periodic 2D grids, two components, and a batch of 2D pieces (gap rows between
the pieces)."""
import dataclasses

import numpy as np
import pytest
import torch

from oracle.oracle import Oracle
from tramonto_amd.problem import small_case

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def binary_2d():
    s = small_case("2d_wall_reflect")
    N = s.nnodes
    sig = np.array([[1.0, 0.9], [0.9, 0.8]])
    return dataclasses.replace(
        s, ncomp=2, hs_diam=[1.0, 0.8], sigma_ff=sig, eps_ff=np.array([[0.8, 0.7], [0.7, 0.6]]),
        cut_ff=1.5 * sig, rho_b=[0.3, 0.2], vext=np.repeat(s.vext, 2, axis=1), zero_tf=np.repeat(s.zero_tf, 2, axis=1),
        rho0=np.stack([0.55 * s.rho0[0], 0.35 * s.rho0[0]]).reshape(2, N))


def make_pair(spec_or_batch, monkeypatch, mfma="1"):
    from tramonto_amd.batch import BatchSpec, DeviceBatch
    from tramonto_amd.device import DeviceProblem
    monkeypatch.setenv("TMI_MF_FFT", "0")
    monkeypatch.setenv("TMI_FMT_FFT", "0")
    monkeypatch.setenv("TMI_MF2D_MFMA", mfma)
    monkeypatch.setenv("TMI_RB2D_MFMA", mfma)
    mk = DeviceBatch if isinstance(spec_or_batch, BatchSpec) else DeviceProblem
    monkeypatch.delenv("TMI_STENCIL_KERNEL", raising=False)
    dense = mk(spec_or_batch)
    monkeypatch.setenv("TMI_STENCIL_KERNEL", "gather")
    gather = mk(spec_or_batch)
    monkeypatch.delenv("TMI_STENCIL_KERNEL", raising=False)
    assert dense.info.mf_fft == 0 and gather.info.mf_fft == 0
    return dense, gather


def state(spec, seed):
    o = Oracle.from_spec(spec)
    x = o.init_fields(spec.initial_x())
    rng = np.random.default_rng(seed)
    x[spec.ncomp:] *= 1.0 + 0.01 * rng.uniform(-1.0, 1.0, x[spec.ncomp:].shape)
    return x


def check(dense, gather, x, seed):
    Fd, _ = dense.fill_residual(dev(x))
    Fg, _ = gather.fill_residual(dev(x))
    Fd, Fg = Fd.cpu().numpy(), Fg.cpu().numpy()
    assert np.abs(Fd - Fg).max() < 1e-13 * np.abs(Fg).max()
    v = np.random.default_rng(seed).uniform(-1.0, 1.0, x.shape)
    for fp32, tol in ((False, 1e-13), (True, 1e-5)):
        Jd = dense.apply_jacobian(dev(x), dev(v), fp32=fp32).cpu().numpy()
        Jg = gather.apply_jacobian(dev(x), dev(v), fp32=fp32).cpu().numpy()
        assert np.abs(Jd - Jg).max() < tol * np.abs(Jg).max(), fp32


@pytest.mark.parametrize("mfma", ["1", "0"])
@pytest.mark.parametrize("case", ["2d_wall_reflect", "2d_periodic", "binary"])
def test_dense_2d_attraction_matches_gather(case, mfma, monkeypatch):
    spec = binary_2d() if case == "binary" else small_case(case)
    dense, gather = make_pair(spec, monkeypatch, mfma)
    check(dense, gather, state(spec, 3), 4)


@pytest.mark.parametrize("mfma", ["1", "0"])
def test_dense_2d_attraction_in_a_batch_of_pieces(mfma, monkeypatch):
    from tramonto_amd.batch import rough_slit_pore_batch_2d
    b = rough_slit_pore_batch_2d(npores=3, seed=5, nx=64, h_range=(3.0, 5.0), dx=0.1)
    dense, gather = make_pair(b, monkeypatch, mfma)
    x = np.concatenate([state(p, 7 + k) for k, p in enumerate(b.pores)], axis=1)
    check(dense, gather, x, 8)
