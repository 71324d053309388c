"""Parity at the benchmark workloads and on the FFT kernels that carry them (run on an MI355X).

* The 64^3 periodic Rosenfeld + WCA workload (dx 0.1 sigma, cut-off 2.5 sigma, the seeded
  8-mode field): the device residual against the oracle's at full size to 1e-12, and the
  FFT-form Jacobian products against the direct stencil sums on the same problem.
* Periodic boxes whose axis lengths select each register FFT kernel (x lengths 64 and 128 for
  the fused x pass in registers; strided y/z lengths 32, 64 and 128 for the one-wave register
  pass, forced below its wave threshold): residual against the oracle, products against the
  direct form, both precisions.
* The 128^3 three-component White Bear + WCA solute at full size: the residual against the
  oracle's rows at ~4000 sampled nodes (a whole oracle fill takes half an hour on 16 cores);
  uniform bulk gives F = 0, J v equals central differences of F, and <w, J v> = <J^T w, v>.

Tolerances: 1e-12 relative to the vector's max-norm for FP64 parity, 1e-5 for the FP32
products against FP64, 1e-6 for central differences (h = 1e-6, direction relative to x).
"""
import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle.oracle import Oracle  # noqa: E402
from tramonto_amd.problem import periodic_3d, small_case, solute_3d  # noqa: E402

RTOL = 1e-12


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def host(t):
    return t.cpu().numpy()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def perturbed(spec, x, seed):
    rng = np.random.default_rng(seed)
    x = x.copy()
    x[spec.ncomp:] *= 1.0 + 0.01 * rng.uniform(-1.0, 1.0, x[spec.ncomp:].shape)
    return x


def make(spec, monkeypatch, **env):
    from tramonto_amd.device import DeviceProblem
    for k in ("TMI_FMT_FFT", "TMI_MF_FFT", "TMI_FFT_REGPASS", "TMI_FFT_XREG", "TMI_FFT_XFUSE"):
        monkeypatch.delenv(k, raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    dp = DeviceProblem(spec)
    for k in env:
        monkeypatch.delenv(k, raising=False)
    return dp


@pytest.fixture(scope="module")
def config2():
    """The 64^3 workload, its oracle and an oracle-initialised, perturbed state."""
    spec = periodic_3d()
    o = Oracle.from_spec(spec)
    x0 = o.init_fields(spec.initial_x())
    return spec, o, x0, perturbed(spec, x0, 101)


@pytest.mark.gpu
def test_config2_residual_matches_oracle_at_full_size(config2, monkeypatch):
    spec, o, x0, x = config2
    assert spec.nx == [64, 64, 64] and spec.dx[0] == 0.1 and float(spec.cut_ff[0, 0]) == 2.5
    dp = make(spec, monkeypatch)
    print(f"\n64^3: fmt_fft {dp.info.fmt_fft} mf_fft {dp.info.mf_fft} xreg {dp.info.fft_xreg} "
          f"regpass {dp.info.fft_regpass}", flush=True)
    assert dp.info.mf_fft == 1
    xi = host(dp.init_fields(dev(spec.initial_x())))
    assert rel(xi, x0) < RTOL
    Fo = o.residual(x).reshape(spec.nunk, -1)
    Fg, sumsq = dp.fill_residual(dev(x))
    Fg = host(Fg)
    assert rel(Fg, Fo) < RTOL
    for a, b in ((0, 1), (1, 1 + spec.nrb), (1 + spec.nrb, spec.nunk)):
        assert np.abs(Fg[a:b] - Fo[a:b]).max() < RTOL * max(np.abs(Fo[a:b]).max(), 1.0)
    assert abs(sumsq - float((Fo ** 2).sum())) <= 1e-10 * float((Fo ** 2).sum())


@pytest.mark.gpu
@pytest.mark.parametrize("fused", ["register", "lds"])
def test_config2_fft_products_match_direct_sums(config2, monkeypatch, fused):
    """FFT-form J v, J^T w and FP32 J v at 64^3 against the direct column kernels."""
    spec, o, x0, x = config2
    env = {"TMI_FMT_FFT": "1", "TMI_MF_FFT": "1"}
    if fused == "lds":
        env["TMI_FFT_XREG"] = "0"
    fft = make(spec, monkeypatch, **env)
    direct = make(spec, monkeypatch, TMI_FMT_FFT="0", TMI_MF_FFT="0")
    assert fft.info.fmt_fft == 2 and fft.info.mf_fft == 1 and direct.info.fmt_fft == 0 and direct.info.mf_fft == 0
    assert fft.info.fft_xreg == (3 if fused == "register" else 0)
    rng = np.random.default_rng(7)
    v = rng.uniform(-1.0, 1.0, x.shape)
    w = rng.uniform(-1.0, 1.0, x.shape)
    xd = dev(x)
    Jv_f, Jv_d = host(fft.apply_jacobian(xd, dev(v))), host(direct.apply_jacobian(xd, dev(v)))
    assert rel(Jv_f, Jv_d) < RTOL
    JTw_f, JTw_d = host(fft.apply_jacobian_transpose(xd, dev(w))), host(direct.apply_jacobian_transpose(xd, dev(w)))
    assert rel(JTw_f, JTw_d) < RTOL
    Jv32 = host(fft.apply_jacobian(xd, dev(v), fp32=True))
    assert rel(Jv32, Jv_d) < 1e-5
    Ff, _ = fft.fill_residual(xd)
    Fd, _ = direct.fill_residual(xd)
    assert rel(host(Ff), host(Fd)) < RTOL


KERNEL_CASES = {
    # case: (fft_xreg expected with the fused x pass forced, strided register pass expected)
    "3d_x64": 3,        # k_fft_xprod_reg<T, 8, 8, 3> (density) and <T, 8, 8, 1> (Euler-Lagrange)
    "3d_x128": 2,       # <T, 16, 8, 1> (Euler-Lagrange); the density product's 3 pairs run in LDS
    "3d_mix3_x128": 2,  # <T, 16, 8, 2>: three components' Euler-Lagrange rows
    "3d_yz64": 0,       # x = 12: LDS fused pass; y and z through k_fft_strided<T, +-1, mode, 8, 8>
    "3d_y128_z32": 0,   # y through k_fft_strided<T, +-1, mode, 16, 8>, z through <T, +-1, mode, 8, 4>
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(KERNEL_CASES))
def test_register_fft_kernels_match_oracle(case, monkeypatch):
    spec = small_case(case)
    o = Oracle.from_spec(spec)
    x = perturbed(spec, o.init_fields(spec.initial_x()), 103)
    fft = make(spec, monkeypatch, TMI_FMT_FFT="1", TMI_MF_FFT="1", TMI_FFT_REGPASS="2")
    print(f"\n{case}: fmt_fft {fft.info.fmt_fft} xreg {fft.info.fft_xreg} regpass {fft.info.fft_regpass}", flush=True)
    assert fft.info.fmt_fft == 2 and fft.info.mf_fft == 1 and fft.info.fft_regpass == 2
    assert fft.info.fft_xreg == KERNEL_CASES[case]
    Fo = o.residual(x)
    Fg, _ = fft.fill_residual(dev(x))
    assert rel(host(Fg).ravel(), Fo.ravel()) < RTOL
    direct = make(spec, monkeypatch, TMI_FMT_FFT="0", TMI_MF_FFT="0")
    rng = np.random.default_rng(11)
    v = rng.uniform(-1.0, 1.0, x.shape)
    w = rng.uniform(-1.0, 1.0, x.shape)
    xd = dev(x)
    Jv = host(fft.apply_jacobian(xd, dev(v)))
    assert rel(Jv, host(direct.apply_jacobian(xd, dev(v)))) < RTOL
    JTw = host(fft.apply_jacobian_transpose(xd, dev(w)))
    assert rel(JTw, host(direct.apply_jacobian_transpose(xd, dev(w)))) < RTOL
    assert rel(host(fft.apply_jacobian(xd, dev(v), fp32=True)), Jv) < 1e-5
    a, b = float((w * Jv).sum()), float((JTw * v).sum())
    assert abs(a - b) < 1e-12 * np.abs(w * Jv).sum()


@pytest.fixture(scope="module")
def config5():
    from tramonto_amd.device import DeviceProblem
    spec = solute_3d()
    assert spec.nx == [128, 128, 128] and spec.ncomp == 3 and spec.type_func == 2
    dp = DeviceProblem(spec)
    return spec, dp


def solute_sample(spec, seed, n_random=2048, n_surface=2048):
    """nodes of the 128^3 solute grid: random ones, and ones within a stencil reach of the excluded
    core of some species (zero-density neighbours, the potential's steep part)"""
    rng = np.random.default_rng(seed)
    n = spec.nx[0]
    g = (np.arange(n) - 0.5 * (n - 1)) * spec.dx[0]
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2).reshape(-1)
    shell = np.flatnonzero((r > 4.0 + 0.3) & (r < 4.0 + 1.8))
    nodes = np.concatenate([rng.choice(spec.nnodes, n_random, replace=False), rng.choice(shell, n_surface, replace=False)])
    return np.unique(nodes)


@pytest.mark.gpu
def test_config5_residual_matches_oracle_on_sampled_nodes(config5):
    """The 128^3 three-component solute at full size: the device residual against the oracle's rows
    at ~4000 sampled nodes (random, and around the solute's excluded core), to 1e-12 of each row
    block's scale."""
    spec, dp = config5
    o = Oracle.from_spec(spec)
    x = perturbed(spec, host(dp.init_fields(dev(spec.initial_x()))), 109)
    nodes = solute_sample(spec, 5)
    Fo = o.residual_nodes(x, nodes)
    Fg, _ = dp.fill_residual(dev(x))
    Fg = host(Fg)[:, nodes]
    C = spec.ncomp
    xs = x[:, nodes]
    for a, b in ((0, C), (C, C + spec.nrb), (C + spec.nrb, spec.nunk)):
        # a rho-bar or mean-field row is x minus a stencil sum of the same size (|u * rho| reaches 14
        # here): the tolerance is relative to that size
        scale = max(np.abs(Fo[a:b]).max(), np.abs(xs[a:b]).max(), 1.0)
        assert np.abs(Fg[a:b] - Fo[a:b]).max() < RTOL * scale, (a, b, np.abs(Fg[a:b] - Fo[a:b]).max(), scale)


@pytest.mark.gpu
def test_config5_full_size_jvp_and_adjoint_identity(config5):
    spec, dp = config5
    print(f"\n128^3 solute: fmt_fft {dp.info.fmt_fft} mf_fft {dp.info.mf_fft} xreg {dp.info.fft_xreg} "
          f"regpass {dp.info.fft_regpass} precond {dp.info.precond}", flush=True)
    x = dp.init_fields(dev(spec.initial_x()))
    xh = host(x)
    # init_fields and the fill agree on the rho-bar and mean-field rows of the initial state
    F0, _ = dp.fill_residual(x)
    F0 = host(F0)
    C = spec.ncomp
    assert np.abs(F0[C:]).max() < 1e-10 * np.abs(xh[C:]).max()
    x = perturbed(spec, xh, 107)
    rng = np.random.default_rng(13)
    v = rng.uniform(-1.0, 1.0, x.shape) * np.abs(x) * 0.1
    h = 1e-6
    Fp, _ = dp.fill_residual(dev(x + h * v))
    Fm, _ = dp.fill_residual(dev(x - h * v))
    fd = (host(Fp) - host(Fm)) / (2 * h)
    Jv = host(dp.apply_jacobian(dev(x), dev(v)))
    assert rel(Jv, fd) < 1e-6
    u = rng.uniform(-1.0, 1.0, x.shape)
    w = rng.uniform(-1.0, 1.0, x.shape)
    Ju = host(dp.apply_jacobian(dev(x), dev(u)))
    JTw = host(dp.apply_jacobian_transpose(dev(x), dev(w)))
    a, b = float((w * Ju).sum()), float((JTw * u).sum())
    assert abs(a - b) < 1e-12 * np.abs(w * Ju).sum(), (a, b)
    Ju32 = host(dp.apply_jacobian(dev(x), dev(u), fp32=True))
    assert rel(Ju32, Ju) < 1e-5


@pytest.mark.gpu
def test_config5_uniform_bulk_is_a_zero_of_the_residual():
    """Without the solute, the uniform bulk state (rho_b, the bulk rho-bars, u * rho_b) is a zero
    of F on the 128^3 three-component grid: every stencil sum, FFT product and local term at once."""
    from tramonto_amd.device import DeviceProblem
    spec = dataclasses.replace(solute_3d(), nwall=0, vext=None, zero_tf=None)
    dp = DeviceProblem(spec)
    C, N = spec.ncomp, spec.nnodes
    x0 = np.zeros((spec.nunk, N))
    for c in range(C):
        x0[c] = spec.rho_b[c]
    x = host(dp.init_fields(dev(x0)))
    bulk = dp.bulk()
    assert np.abs(x[C:C + 4] - bulk[:4, None]).max() < 1e-12 * np.abs(bulk[:4]).max()
    F, sumsq = dp.fill_residual(dev(x))
    F = host(F)
    scale = max(np.abs(x).max(), 1.0)
    assert np.abs(F).max() < 1e-11 * scale, np.abs(F).max()
