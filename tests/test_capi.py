"""The C ABI library (CPU-side checks, no GPU needed).

* libtramonto_mi355x.so loads and exports every function include/*.h declares.
* The library's host-side stencil builder and bulk constants agree with the
  oracle bit for bit (two independent implementations of calc_stencils and
  the bulk thermodynamics).
"""
import ctypes as ct
import os
import re
import subprocess

import numpy as np
import pytest

from oracle.oracle import Oracle
from tramonto_amd import _lib
from tramonto_amd.problem import small_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(inc, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "for", "return", "sizeof"):
                names.append(name)
    return names


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) > 40
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert hasattr(L, n)
    assert set(_lib.DECLARED_SYMBOLS) <= set(names)


def host_stencil(spec, t, i=0, j=0):
    L = _lib.lib()
    p = _lib.fill_params(_lib.TmiParams(), spec.as_params())
    n = ct.c_int()
    _lib.check(L.tmi_host_stencil(ct.byref(p), t, i, j, ct.byref(n), None, None), "tmi_host_stencil")
    off = np.zeros((n.value, 3), dtype=np.int32)
    w = np.zeros(n.value)
    _lib.check(L.tmi_host_stencil(ct.byref(p), t, i, j, ct.byref(n), off.ctypes.data_as(ct.c_void_p),
                                  w.ctypes.data_as(ct.c_void_p)), "tmi_host_stencil")
    return off, w


@pytest.mark.parametrize("case", ["1d_hw", "1d_reflect_wca", "1d_pore_mix", "2d_periodic", "2d_wall_reflect",
                                  "3d_periodic_small", "3d_mix_bulk"])
def test_host_stencils_match_oracle_bitwise(case):
    sp = small_case(case)
    o = Oracle.from_spec(sp)
    for t in range(3 if sp.type_attr >= 0 else 2):
        for i in range(sp.ncomp):
            for j in range(sp.ncomp if t == 2 else 1):
                og, wg = host_stencil(sp, t, i, j)
                oo, wo = o.stencil(t, i, j)
                assert np.array_equal(og, oo)
                assert np.array_equal(wg, wo)


@pytest.mark.parametrize("case", ["1d_hw", "1d_pore_mix", "3d_periodic_wb", "3d_mix_bulk"])
def test_host_bulk_constants_match_oracle(case):
    sp = small_case(case)
    o = Oracle.from_spec(sp)
    out = np.zeros(60)
    p = _lib.fill_params(_lib.TmiParams(), sp.as_params())
    _lib.check(_lib.lib().tmi_host_bulk(ct.byref(p), out.ctypes.data_as(ct.c_void_p)), "tmi_host_bulk")
    ref = o.bulk()
    assert np.allclose(out, ref, rtol=1e-14, atol=1e-14)


def test_rejects_unsupported_problems():
    sp = small_case("1d_hw")
    d = sp.as_params()
    d["lhard_surf"] = 1
    n = ct.c_int()
    for attr in (-1, 0):   # hard surfaces are served with and without attraction
        d["type_attr"] = attr
        p = _lib.fill_params(_lib.TmiParams(), d)
        assert _lib.lib().tmi_host_stencil(ct.byref(p), 0, 0, 0, ct.byref(n), None, None) == 0
    d["lhard_surf"] = 0
    for func in (1, 3):   # FMT2, FMT4
        d["type_func"] = func
        p = _lib.fill_params(_lib.TmiParams(), d)
        assert _lib.lib().tmi_host_stencil(ct.byref(p), 0, 0, 0, ct.byref(n), None, None) == 0
    d["type_func"] = 4   # no such functional
    p = _lib.fill_params(_lib.TmiParams(), d)
    rc = _lib.lib().tmi_host_stencil(ct.byref(p), 0, 0, 0, ct.byref(n), None, None)
    assert rc == -2
    assert b"type_func" in _lib.lib().tmi_last_error()


@pytest.mark.parametrize("case", ["2d_periodic", "2d_wall_reflect", "2d_rough_wall", "3d_periodic_small",
                                  "3d_periodic_wb", "3d_mix_bulk", "3d_solute_small"])
def test_host_stencils_reflection_symmetric(case):
    """The FFT sums store each stencil spectrum for k_y, k_z >= 0 only: that needs the
    weights at o and at o with one component negated to agree (to 1e-13 of the largest weight)
    for every axis, which the sphere and shell quadratures give."""
    sp = small_case(case)
    for t in range(3 if sp.type_attr >= 0 else 2):
        for i in range(sp.ncomp):
            for j in range(sp.ncomp if t == 2 else 1):
                off, w = host_stencil(sp, t, i, j)
                at = {}
                for o, v in zip(map(tuple, off), w):
                    at[o] = at.get(o, 0.0) + v
                wmax = np.abs(w).max()
                for o, v in at.items():
                    for a in range(sp.ndim):
                        r = list(o)
                        r[a] = -r[a]
                        assert abs(at.get(tuple(r), np.inf) - v) <= 1e-13 * wmax, (t, i, j, o, a)


def test_library_exports_every_reference_abi_name():
    """Every LinProbMgr / ParameterList name the reference's C driver and NOX group call
    (fixture made by tests/golden/make_abi_names.py from the reference sources) is exported,
    so the reference links against this library unchanged."""
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_abi_names.json")) as f:
        names = json.load(f)["hard_sphere_path"]
    assert "dft_basic_lin_prob_mgr_create" in names and "dft_parameterlist_set_all_aztec" in names
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names + ["fill_resid_and_matrix_control"] if n not in exported]
    assert not missing, missing
    declared = set(declared_functions())
    assert set(names) <= declared


def test_reference_abi_names_fixture_is_current():
    """The fixture lists what the reference sources call today (skipped where the reference is absent)."""
    import json
    import sys
    src = "/root/reference/src"
    if not os.path.isdir(src):
        pytest.skip("reference sources not present")
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_abi_names
    path = os.path.join(ROOT, "tests", "golden", "reference_abi_names.json")
    before = open(path).read()
    make_abi_names.main()
    assert open(path).read() == before


def test_linprobmgr_block_structure_without_gpu():
    """finalizeblockstructure: -1 while a set call is missing, 0 once, 1 after; the basic
    manager needs maps only; destructors return void (dft_basic_lin_prob_mgr_wrapper.h:46)."""
    L = _lib.lib()
    vp = lambda a: a.ctypes.data_as(ct.c_void_p)
    gids = np.arange(10, dtype=np.int32)
    plist = L.dft_parameterlist_create()
    assert plist
    L.dft_parameterlist_set_int(plist, b"Max_iter", 7)
    L.dft_parameterlist_set_double(plist, b"Tol", 1e-6)
    opts = np.zeros(47, dtype=np.int32)
    pars = np.zeros(30)
    opts[6], opts[10], pars[0] = 123, 45, 1e-7
    L.dft_parameterlist_set_all_aztec(plist, vp(opts), vp(pars))
    hs = L.dft_hardsphere_lin_prob_mgr_create(4, plist, 0)
    assert L.dft_linprobmgr_finalizeblockstructure(hs) == -1
    assert L.dft_linprobmgr_setnodalrowmap(hs, 10, vp(gids)) == 0
    assert L.dft_linprobmgr_setnodalcolmap(hs, 10, vp(gids)) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(hs) == -1          # no equation IDs yet
    ind, dens = np.array([1, 2, 3], dtype=np.int32), np.array([0], dtype=np.int32)
    assert L.dft_hardsphere_lin_prob_mgr_setindnonlocalequationids(hs, 3, vp(ind)) == 0
    assert L.dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids(hs, 0, None) == 0
    assert L.dft_hardsphere_lin_prob_mgr_setdensityequationids(hs, 1, vp(dens)) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(hs) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(hs) == 1
    assert L.dft_linprobmgr_setupsolver(hs) == -4                      # no problem registered: state error
    assert L.dft_linprobmgr_insertonematrixvalue(hs, 0, 10, 0, 0, 1.0) == -1   # owned node out of range
    assert L.dft_linprobmgr_getmatrixvalue(hs, 9, 0, 0, 0) == 0.0
    dbg = L.dft_hardsphere_lin_prob_mgr_create_debug(4, plist, 0)
    basic = L.dft_basic_lin_prob_mgr_create(4, plist, 0)
    assert L.dft_linprobmgr_setnodalrowmap(basic, 10, vp(gids)) == 0
    assert L.dft_linprobmgr_setnodalcolmap(basic, 10, vp(gids)) == 0
    assert L.dft_linprobmgr_finalizeblockstructure(basic) == 0
    assert L.dft_hardsphere_lin_prob_mgr_destruct(hs) is None
    assert L.dft_linprobmgr_destruct(dbg) is None
    assert L.dft_linprobmgr_destruct(basic) is None
    L.dft_parameterlist_destruct(plist)
