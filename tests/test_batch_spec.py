"""Host-side checks of the batch layout (BatchSpec) for 1D pores and 2D pieces: the validation the
device batch relies on, the packing order, and the state-point choice of the 2D builder."""
import dataclasses

import numpy as np
import pytest
from scipy.integrate import quad

from tramonto_amd.batch import BatchSpec, rough_slit_pore_batch_2d, slit_pore_batch
from tramonto_amd.problem import IN_WALL, PERIODIC, REFLECT, rough_slit_pore_2d


def _piece(ny=31, **kw):
    return rough_slit_pore_2d(seed=3, nx=32, ny=ny, dx=0.1, eps_w=1.0, rho_b=0.4, eps_ff=0.5, **kw)


def test_2d_pieces_pack_along_y_in_row_order():
    b = BatchSpec([_piece(31), _piece(41)])
    assert list(b.nodes) == [32 * 31, 32 * 41] and b.nnodes == 32 * 72
    v = b.packed("vext")
    assert v.shape == (b.nnodes, 1)
    # piece 1's first row follows piece 0's last row: node i + j nx of piece 1 sits at 31 nx + i + j nx
    assert np.array_equal(v[32 * 31:, 0], b.pores[1].vext[:, 0])
    x = b.initial_x()
    parts = b.split(x)
    assert [q.shape[1] for q in parts] == [32 * 31, 32 * 41]
    assert np.array_equal(np.concatenate(parts, axis=1), x)


@pytest.mark.parametrize("change", ["ndim", "nx", "ybc", "xbc", "xbc_mixed", "dx"])
def test_2d_batch_rejects_pieces_that_do_not_share_the_layout(change):
    a = _piece(31)
    if change == "ndim":
        other = slit_pore_batch(npores=1, seed=1).pores[0]
    elif change == "nx":
        other = rough_slit_pore_2d(seed=3, nx=48, ny=31, dx=0.1, eps_w=1.0, rho_b=0.4, eps_ff=0.5)
    elif change == "ybc":
        other = dataclasses.replace(_piece(31), bc=[[PERIODIC, PERIODIC], [IN_WALL, REFLECT]])
    elif change == "xbc":
        other = dataclasses.replace(_piece(31), bc=[[IN_WALL, IN_WALL], [IN_WALL, IN_WALL]])
    elif change == "xbc_mixed":
        a = dataclasses.replace(a, bc=[[PERIODIC, IN_WALL], [IN_WALL, IN_WALL]])
        other = dataclasses.replace(_piece(31), bc=[[PERIODIC, IN_WALL], [IN_WALL, IN_WALL]])
    else:
        other = rough_slit_pore_2d(seed=3, nx=32, ny=31, dx=0.05, eps_w=1.0, rho_b=0.4, eps_ff=0.5)
    with pytest.raises(ValueError):
        BatchSpec([a, other])


def test_walled_x_pieces_are_accepted():
    a = dataclasses.replace(_piece(31), bc=[[IN_WALL, IN_WALL], [IN_WALL, IN_WALL]])
    b = dataclasses.replace(_piece(41), bc=[[IN_WALL, IN_WALL], [IN_WALL, IN_WALL]])
    assert BatchSpec([a, b]).nnodes == 32 * 72


def _mean_field_kTc(cut):
    """kT_c / eps of the bulk fluid: ideal + Percus-Yevick (compressibility) hard spheres of diameter 1,
    the bulk limit of the Rosenfeld functional, plus the mean field -a rho^2 / 2 of the WCA-split
    Lennard-Jones attraction cut at `cut` (u_att = -eps inside 2^(1/6), the LJ tail outside), with
    and without the shift by u(cut).  The spinodal kT(rho) = a rho (1 - eta)^4 / (1 + 2 eta)^2
    peaks at the critical point."""
    rm = 2.0 ** (1.0 / 6.0)
    lj = lambda r: 4.0 * (r ** -12 - r ** -6)  # noqa: E731
    out = []
    for s in (0.0, lj(cut)):
        a = -4.0 * np.pi * (quad(lambda r: r * r * (-1.0 - s), 0.0, rm)[0]
                            + quad(lambda r: r * r * (lj(r) - s), rm, cut)[0])
        rho = np.linspace(1e-3, 1.0, 20001)
        eta = np.pi / 6.0 * rho
        out.append(float(np.max(a * rho * (1.0 - eta) ** 4 / (1.0 + 2.0 * eta) ** 2)))
    return out


def test_2d_builder_state_points_are_supercritical():
    """The 2D batch fluid (eps_ff / kT = 0.5) lies above the mean-field critical temperature, while
    eps_ff / kT = 1 puts its bulk densities 0.3-0.45 inside the spinodal (a bulk state no Newton start
    near rho_b settles into).  Check of the critical-point search: the maximum of
    eta (1 - eta)^4 / (1 + 2 eta)^2 solves 1 - 7 eta - 6 eta^2 = 0, eta_c = (sqrt(73) - 7) / 12."""
    b = rough_slit_pore_batch_2d(npores=8, seed=11)
    eps = float(b.pores[0].eps_ff[0, 0])
    assert eps == 0.5
    tcs = _mean_field_kTc(2.5)
    assert 1.1 < min(tcs) and max(tcs) < 1.4
    assert 1.0 / eps > max(tcs)
    rho = np.array([p.rho_b[0] for p in b.pores])
    assert rho.min() >= 0.3 and rho.max() <= 0.5
    # eps_ff / kT = 1 (kT = 1 < kT_c): dP/drho < 0 at rho = 0.4
    rm = 2.0 ** (1.0 / 6.0)
    a = -4.0 * np.pi * (quad(lambda r: -r * r, 0.0, rm)[0] + quad(lambda r: r * r * 4.0 * (r ** -12 - r ** -6), rm, 2.5)[0])
    eta = np.pi / 6.0 * 0.4
    assert (1.0 + 2.0 * eta) ** 2 / (1.0 - eta) ** 4 - a * 0.4 < 0.0
    rho = np.linspace(1e-3, 1.0, 200001)
    eta = np.pi / 6.0 * rho
    g = rho * (1.0 - eta) ** 4 / (1.0 + 2.0 * eta) ** 2
    assert abs(np.pi / 6.0 * rho[np.argmax(g)] - (np.sqrt(73.0) - 7.0) / 12.0) < 1e-5
