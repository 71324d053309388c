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
    d["type_func"] = 1   # FMT2
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
