"""Host-side pieces of the adjoint harness (no GPU needed)."""
import numpy as np

from oracle.oracle import Oracle
from tramonto_amd.adjoint import bulk_betamu, dbetamu_drho_b, node_volumes
from tramonto_amd.problem import rough_wall_2d, small_case, solute_3d


def test_bulk_betamu_matches_oracle_constants():
    for case in ("1d_pore_mix", "2d_rough_wall", "3d_mix_bulk"):
        spec = small_case(case)
        b = Oracle.from_spec(spec).bulk()
        C = spec.ncomp
        ref = np.log(spec.rho_b[:C]) + b[20:20 + C] + (b[24:24 + C] if spec.type_attr >= 0 else 0.0)
        assert np.allclose(bulk_betamu(spec), ref, rtol=0, atol=1e-13)


def test_dbetamu_is_ideal_gas_at_low_density():
    spec = small_case("1d_hw")
    spec.rho_b = [1e-6]
    d = dbetamu_drho_b(spec)[0, 0]
    assert abs(d * 1e-6 - 1.0) < 1e-4    # d ln(rho)/d rho = 1/rho dominates


def test_node_volumes_integrate_the_box():
    sp = small_case("2d_rough_wall")
    w = node_volumes(sp)
    assert np.isclose(w.sum(), (sp.nx[0] - 1) * sp.dx[0] * sp.nx[1] * sp.dx[1])
    assert np.isclose(node_volumes(small_case("3d_periodic_small")).sum(), (12 * 0.25) ** 3)


def test_rough_wall_excluded_set_independent_of_eps_w():
    a = rough_wall_2d(nx=40, ny=32, dx=0.1, cut=1.5, wall_cut=2.0, eps_w=1.2)
    b = rough_wall_2d(nx=40, ny=32, dx=0.1, cut=1.5, wall_cut=2.0, eps_w=1.8)
    assert np.array_equal(a.zero_tf, b.zero_tf)
    free = a.zero_tf[:, 0] == 0
    assert np.allclose(a.vext[free, 0] / 1.2, b.vext[free, 0] / 1.8)


def test_solute_bulk_packing_fraction():
    sp = solute_3d(n=16, dx=0.25, radius=1.0, cut_fac=1.2, wall_cut=2.0)
    eta = sum(np.pi / 6 * s ** 3 * r for s, r in zip(sp.hs_diam, sp.rho_b))
    assert abs(eta - 0.35) < 1e-12
