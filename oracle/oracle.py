"""ctypes wrapper of the CPU oracle (oracle/_build/libtmo.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only as the checker.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libtmo.so")
NC = 4


class TmoParams(ct.Structure):
    _fields_ = [
        ("ndim", ct.c_int), ("ncomp", ct.c_int), ("nx", ct.c_int * 3), ("dx", ct.c_double * 3),
        ("bc", (ct.c_int * 2) * 3), ("type_func", ct.c_int), ("type_attr", ct.c_int), ("nwall", ct.c_int),
        ("lhard_surf", ct.c_int), ("vext_max", ct.c_double), ("hs_diam", ct.c_double * NC),
        ("sigma_ff", (ct.c_double * NC) * NC), ("eps_ff", (ct.c_double * NC) * NC),
        ("cut_ff", (ct.c_double * NC) * NC), ("rho_b", ct.c_double * NC),
    ]


class TmoNewtonOpts(ct.Structure):
    _fields_ = [("max_iter", ct.c_int), ("nl_rel_tol", ct.c_double), ("nl_abs_tol", ct.c_double),
                ("update_scaling", ct.c_double), ("resid_tol", ct.c_double), ("gmres_restart", ct.c_int),
                ("gmres_max_iter", ct.c_int), ("gmres_tol", ct.c_double)]


class TmoNewtonStats(ct.Structure):
    _fields_ = [("iters", ct.c_int), ("converged", ct.c_int), ("resid_norm", ct.c_double),
                ("gmres_iters_total", ct.c_int)]


_L: Optional[ct.CDLL] = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ct.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            build()
        L = ct.CDLL(LIB)
        vp, i, d = ct.c_void_p, ct.c_int, ct.c_double
        L.tmo_create.restype = vp
        L.tmo_create.argtypes = [ct.POINTER(TmoParams), vp, vp]
        L.tmo_destroy.argtypes = [vp]
        L.tmo_set_hw_masks.argtypes = [vp, vp]
        L.tmo_last_error.restype = ct.c_char_p
        L.tmo_nunk.argtypes = [vp]
        L.tmo_nnodes.argtypes = [vp]
        L.tmo_unk_layout.argtypes = [vp] + [ct.POINTER(i)] * 5
        L.tmo_stencil_length.argtypes = [vp, i, i, i]
        L.tmo_stencil_get.argtypes = [vp, i, i, i, vp, vp]
        L.tmo_bulk.argtypes = [vp, vp]
        L.tmo_fmt_dphi.argtypes = [vp, vp, vp]
        L.tmo_fmt_phi.argtypes = [vp, vp]
        L.tmo_fmt_phi.restype = d
        L.tmo_fmt_d2_delta.argtypes = [vp, vp, vp, vp, i, vp]
        L.tmo_fmt_d2_theta.argtypes = [vp, vp, vp]
        L.tmo_residual.argtypes = [vp, vp, vp]
        L.tmo_residual.restype = d
        L.tmo_gauss_rule.argtypes = [i, vp, vp]
        L.tmo_gauss_rule.restype = None
        L.tmo_set_stencil_rule.argtypes = [i, i, d]
        L.tmo_set_stencil_rule.restype = None
        L.tmo_residual_nodes.argtypes = [vp, vp, vp, i, vp]
        L.tmo_residual_nodes.restype = d
        L.tmo_jacobian_coo.argtypes = [vp, vp, vp, vp, vp, ct.c_long]
        L.tmo_jacobian_coo.restype = ct.c_long
        L.tmo_init_fields.argtypes = [vp, vp]
        L.tmo_newton.argtypes = [vp, vp, ct.POINTER(TmoNewtonOpts), ct.POINTER(TmoNewtonStats)]
        L.tmo_observables.argtypes = [vp, vp, vp, vp, d, vp]
        _L = L
    return _L


def _fill(struct, d: Dict):
    names = {f[0] for f in struct._fields_}
    for k, v in d.items():
        if k not in names:   # layout options of the device library (att_in_a22): the oracle keeps MF rows
            continue
        f = getattr(struct, k)
        if isinstance(v, (list, tuple)):
            if isinstance(v[0], (list, tuple)):
                for a, row in enumerate(v):
                    for b, e in enumerate(row):
                        f[a][b] = e
            else:
                for a, e in enumerate(v):
                    f[a] = e
        else:
            setattr(struct, k, v)
    return struct


def _p(a: np.ndarray) -> ct.c_void_p:
    return a.ctypes.data_as(ct.c_void_p)


class Oracle:
    """CPU restatement of the reference fill for one problem."""

    def __init__(self, params: Dict, vext: Optional[np.ndarray] = None, zero_tf: Optional[np.ndarray] = None):
        L = lib()
        self.params = _fill(TmoParams(), params)
        self.vext = None if vext is None else np.ascontiguousarray(vext, dtype=np.float64)
        self.zero = None if zero_tf is None else np.ascontiguousarray(zero_tf, dtype=np.uint8)
        self.h = L.tmo_create(ct.byref(self.params), None if self.vext is None else _p(self.vext),
                              None if self.zero is None else _p(self.zero))
        if not self.h:
            raise ValueError("oracle: " + L.tmo_last_error().decode())
        self.nunk = L.tmo_nunk(self.h)
        self.N = L.tmo_nnodes(self.h)
        a = [ct.c_int() for _ in range(5)]
        L.tmo_unk_layout(self.h, *[ct.byref(v) for v in a])
        self.iu_rho, self.iu_rb, self.n_rb, self.iu_mf, self.n_mf = [v.value for v in a]

    @classmethod
    def from_spec(cls, spec):
        o = cls(spec.as_params(), spec.vext, spec.zero_tf)
        if getattr(spec, "hw_mask", None) is not None:
            o.hw_mask = np.ascontiguousarray(spec.hw_mask, dtype=np.uint8)
            lib().tmo_set_hw_masks(o.h, _p(o.hw_mask))
        return o

    def __del__(self):
        try:
            if self.h:
                lib().tmo_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def stencil(self, isten, icomp=0, jcomp=0):
        L = lib()
        n = L.tmo_stencil_length(self.h, isten, icomp, jcomp)
        off = np.zeros((n, 3), dtype=np.int32)
        w = np.zeros(n)
        L.tmo_stencil_get(self.h, isten, icomp, jcomp, _p(off), _p(w))
        return off, w

    def bulk(self) -> np.ndarray:
        out = np.zeros(60)
        lib().tmo_bulk(self.h, _p(out))
        return out

    def residual(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        F = np.zeros_like(x)
        lib().tmo_residual(self.h, _p(x), _p(F))
        return F

    def residual_nodes(self, x: np.ndarray, nodes: np.ndarray) -> np.ndarray:
        """F[:, nodes] ([nunk][len(nodes)]) without filling the other nodes' rows."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        nodes = np.ascontiguousarray(nodes, dtype=np.int32)
        F = np.zeros((self.nunk, len(nodes)))
        lib().tmo_residual_nodes(self.h, _p(x), _p(nodes), len(nodes), _p(F))
        return F

    def jacobian_coo(self, x: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float64)
        L = lib()
        n = L.tmo_jacobian_coo(self.h, _p(x), None, None, None, 0)
        r = np.zeros(n, dtype=np.int32)
        c = np.zeros(n, dtype=np.int32)
        v = np.zeros(n)
        L.tmo_jacobian_coo(self.h, _p(x), _p(r), _p(c), _p(v), n)
        return r, c, v

    def jacobian_dense(self, x: np.ndarray) -> np.ndarray:
        """Dense J with += semantics (small problems only)."""
        r, c, v = self.jacobian_coo(x)
        n = self.nunk * self.N
        J = np.zeros((n, n))
        np.add.at(J, (r, c), v)
        return J

    def init_fields(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        lib().tmo_init_fields(self.h, _p(x))
        return x

    def newton(self, x: np.ndarray, max_iter=50, nl_rel_tol=1e-3, nl_abs_tol=1e-8, update_scaling=0.1,
               resid_tol=1e-10, gmres_restart=100, gmres_max_iter=2000, gmres_tol=1e-10):
        x = np.ascontiguousarray(x, dtype=np.float64).copy()
        o = TmoNewtonOpts(max_iter, nl_rel_tol, nl_abs_tol, update_scaling, resid_tol, gmres_restart, gmres_max_iter,
                          gmres_tol)
        st = TmoNewtonStats()
        lib().tmo_newton(self.h, _p(x), ct.byref(o), ct.byref(st))
        return x, dict(iters=st.iters, converged=bool(st.converged), resid_norm=st.resid_norm,
                       gmres_iters=st.gmres_iters_total)

    def observables(self, x: np.ndarray, nel_hit: np.ndarray, nel_hit2: np.ndarray, scale: float = 1.0) -> Dict:
        """Free-energy terms and adsorption of the profile x (calc_free_energy, calc_adsorption):
        nel_hit, nel_hit2 are the [ncomp][N] element counts of the reference's setup_integrals."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        h1 = np.ascontiguousarray(nel_hit, dtype=np.int32)
        h2 = np.ascontiguousarray(nel_hit2, dtype=np.int32)
        out = np.zeros(19)
        lib().tmo_observables(self.h, _p(x), _p(h1), _p(h2), float(scale), _p(out))
        C = self.params.ncomp
        keys = ("omega", "omega_s", "ideal", "ideal_b", "mu", "mu_b", "vext", "hs", "hs_b", "att", "att_b")
        r = {k: float(v) for k, v in zip(keys, out[:11])}
        r["ads"] = out[11:11 + C].copy()
        r["ads_b"] = out[15:15 + C].copy()
        return r

    # pointwise FMT pieces (n order)
    def dphi(self, n: np.ndarray) -> np.ndarray:
        n = np.ascontiguousarray(n, dtype=np.float64)
        out = np.zeros(10)
        lib().tmo_fmt_dphi(self.h, _p(n), _p(out))
        return out[: 4 + 2 * self.params.ndim]

    def phi(self, n: np.ndarray) -> float:
        n = np.ascontiguousarray(n, dtype=np.float64)
        return lib().tmo_fmt_phi(self.h, _p(n))

    def d2_delta(self, n, offset, sign, icomp=0):
        n = np.ascontiguousarray(n, dtype=np.float64)
        off = np.ascontiguousarray(offset, dtype=np.int32)
        sg = np.ascontiguousarray(sign, dtype=np.float64)
        out = np.zeros(10)
        lib().tmo_fmt_d2_delta(self.h, _p(n), _p(off), _p(sg), icomp, _p(out))
        return out[: 4 + 2 * self.params.ndim]

    def d2_theta(self, n):
        n = np.ascontiguousarray(n, dtype=np.float64)
        out = np.zeros(10)
        lib().tmo_fmt_d2_theta(self.h, _p(n), _p(out))
        return out[: 4 + 2 * self.params.ndim]
