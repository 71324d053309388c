"""Host-side handle of a device problem: the Python mirror of the C ABI.

Method names follow the reference interface they stand in for:

* ``fill_residual``   -- fill_resid_and_matrix_control(x, iter, TRUE)   (src/dft_fill_control.c:39)
* ``apply_jacobian``  -- dft_linprobmgr_applymatrix / Schur ApplyGlobal    (src/dft_Schur_Epetra_Operator.cpp:149)
* ``init_fields``     -- calc_init_rho_bar + calc_init_mf_attract          (src/dft_guess_HSFMT.c:70, src/dft_guess_MFATT.c:66)
* ``newton``          -- newton_solver                                     (src/dft_newton.c:131)

Device memory and streams come from PyTorch; the arithmetic is the HIP
library's.  Tensors are float64, shape [nunk, nnodes], on cuda:0.
"""
from __future__ import annotations

import copy
import ctypes as ct
from typing import Dict, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .problem import ProblemSpec

PREC_FP64, PREC_MIXED = 0, 1
CONV_RESID, CONV_UPDATE = 0, 1
PRECOND_DIAG, PRECOND_BULK = 0, 1
JAC_EXACT, JAC_REFERENCE = 0, 1


def _stream() -> ct.c_void_p:
    return ct.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: torch.Tensor) -> ct.c_void_p:
    if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous():
        raise ValueError("expected a contiguous float64 CUDA tensor")
    return ct.c_void_p(t.data_ptr())


class DeviceProblem:
    def __init__(self, spec: ProblemSpec):
        L = _lib.lib()
        self.spec = copy.copy(spec)   # parameter updates (set_rho_b, set_vext) change this copy only
        self.params = _lib.fill_params(_lib.TmiParams(), spec.as_params())
        C, N = spec.ncomp, spec.nnodes
        vext = None if spec.vext is None else np.ascontiguousarray(spec.vext, dtype=np.float64).reshape(N, C)
        zero = None if spec.zero_tf is None else np.ascontiguousarray(spec.zero_tf, dtype=np.uint8).reshape(N, C)
        self._keep = (vext, zero)
        h = ct.c_void_p()
        _lib.check(L.tmi_problem_create(ct.byref(self.params),
                                        None if vext is None else vext.ctypes.data_as(ct.c_void_p),
                                        None if zero is None else zero.ctypes.data_as(ct.c_void_p),
                                        ct.byref(h)), "tmi_problem_create")
        self.handle = h
        if getattr(spec, "hw_mask", None) is not None:
            self._hw_mask = np.ascontiguousarray(spec.hw_mask, dtype=np.uint8).reshape(C, N)
            _lib.check(L.tmi_set_hard_surface_masks(self.handle, self._hw_mask.ctypes.data_as(ct.c_void_p)),
                       "tmi_set_hard_surface_masks")
        info = _lib.TmiInfo()
        _lib.check(L.tmi_problem_info(h, ct.byref(info)), "tmi_problem_info")
        self.info = info
        self.nunk, self.nnodes = info.nunk, info.nnodes

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().tmi_problem_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ tables
    def stencil(self, isten: int, icomp: int = 0, jcomp: int = 0) -> Tuple[np.ndarray, np.ndarray]:
        L = _lib.lib()
        n = ct.c_int()
        _lib.check(L.tmi_stencil_get(self.handle, isten, icomp, jcomp, ct.byref(n), None, None), "tmi_stencil_get")
        off = np.zeros((n.value, 3), dtype=np.int32)
        w = np.zeros(n.value)
        _lib.check(L.tmi_stencil_get(self.handle, isten, icomp, jcomp, ct.byref(n), off.ctypes.data_as(ct.c_void_p),
                                     w.ctypes.data_as(ct.c_void_p)), "tmi_stencil_get")
        return off, w

    def bulk(self) -> np.ndarray:
        out = np.zeros(60)
        _lib.check(_lib.lib().tmi_bulk_get(self.handle, out.ctypes.data_as(ct.c_void_p)), "tmi_bulk_get")
        return out

    def row_tuning(self) -> Dict:
        """Autotuned launch shapes of the 1D row kernels: {precision: {kind: (per_lane, lanes, ms)}}."""
        shape = np.zeros(12, dtype=np.int32)
        ms = np.zeros(6)
        _lib.check(_lib.lib().tmi_row_tuning(self.handle, shape.ctypes.data_as(ct.c_void_p),
                                             ms.ctypes.data_as(ct.c_void_p)), "tmi_row_tuning")
        out = {}
        for p, pname in enumerate(("fp64", "fp32")):
            out[pname] = {k: (int(shape[(p * 3 + i) * 2]), int(shape[(p * 3 + i) * 2 + 1]), float(ms[p * 3 + i]))
                          for i, k in enumerate(("rb_rows", "mf_rows", "el_rows"))}
        return out

    def set_row_tuning(self, per_lane: int, lanes: int, prec: Optional[int] = None, kind: Optional[int] = None):
        """Forces a launch shape of the 1D row kernels (all precisions / kinds when None)."""
        for p in ((0, 1) if prec is None else (prec,)):
            for k in ((0, 1, 2) if kind is None else (kind,)):
                _lib.check(_lib.lib().tmi_row_tuning_set(self.handle, p, k, per_lane, lanes), "tmi_row_tuning_set")

    def set_rho_b(self, rho_b) -> None:
        """New bulk densities, in place (chemical potentials and bulk boundary values follow)."""
        a = np.ascontiguousarray(np.asarray(rho_b, dtype=np.float64)[: self.spec.ncomp])
        _lib.check(_lib.lib().tmi_problem_set_rho_b(self.handle, a.ctypes.data_as(ct.c_void_p)), "tmi_problem_set_rho_b")
        self.spec.rho_b = [float(v) for v in a]

    def set_vext(self, vext: np.ndarray) -> None:
        """New external field [nnodes, ncomp], in place (zero-density flags unchanged)."""
        a = np.ascontiguousarray(np.asarray(vext, dtype=np.float64).reshape(self.nnodes, self.spec.ncomp))
        _lib.check(_lib.lib().tmi_problem_set_vext(self.handle, a.ctypes.data_as(ct.c_void_p)), "tmi_problem_set_vext")
        self.spec.vext = a

    def set_preconditioner(self, kind: int) -> None:
        """Schur GMRES preconditioner: PRECOND_DIAG (the A22 diagonal) or PRECOND_BULK
        (diagonal times the inverse bulk Schur symbol, periodic FFT grids)."""
        _lib.check(_lib.lib().tmi_problem_set_preconditioner(self.handle, int(kind)), "tmi_problem_set_preconditioner")
        self.info.precond = int(kind)

    def set_jacobian(self, mode: int) -> None:
        """Jacobian of the products: JAC_EXACT (the exact derivative) or JAC_REFERENCE (the
        reference's assembled entries, inexact for FMT2/FMT4 and White Bear across REFLECT)."""
        _lib.check(_lib.lib().tmi_problem_set_jacobian(self.handle, int(mode)), "tmi_problem_set_jacobian")
        self.info.jacobian = int(mode)

    def observables(self, x: torch.Tensor, nel_hit: np.ndarray, nel_hit2: Optional[np.ndarray] = None,
                    scale: float = 1.0) -> Dict:
        """Free-energy terms and adsorption of the profile x, summed on the device
        (tmi_compute_observables): nel_hit / nel_hit2 are the [ncomp][N] element counts of the
        reference's setup_integrals (nel_hit2 defaults to nel_hit, as without hard surfaces)."""
        h1 = np.ascontiguousarray(nel_hit, dtype=np.int32)
        h2 = h1 if nel_hit2 is None else np.ascontiguousarray(nel_hit2, dtype=np.int32)
        if h1.shape != (self.spec.ncomp, self.nnodes) or h2.shape != h1.shape:
            raise ValueError("element counts must be [ncomp][nnodes]")
        o = _lib.TmiObservables()
        _lib.check(_lib.lib().tmi_compute_observables(self.handle, _ptr(x), h1.ctypes.data_as(ct.c_void_p),
                                                      h2.ctypes.data_as(ct.c_void_p), float(scale), ct.byref(o),
                                                      _stream()), "tmi_compute_observables")
        C = self.spec.ncomp
        r = {k: float(getattr(o, k)) for k in ("omega", "omega_s", "ideal", "ideal_b", "mu", "mu_b", "vext", "hs",
                                                "hs_b", "att", "att_b")}
        r["ads"] = np.array(o.ads[:C])
        r["ads_b"] = np.array(o.ads_b[:C])
        return r

    # ------------------------------------------------------------ device ops
    def force(self, x: torch.Tensor, weights: np.ndarray) -> np.ndarray:
        """Wall forces of the profile x, summed on the device (tmi_compute_force): weights
        [nsum][ncomp][nnodes], e.g. examples.force_weights."""
        W = np.ascontiguousarray(weights, dtype=np.float64)
        if W.ndim != 3 or W.shape[1:] != (self.spec.ncomp, self.nnodes):
            raise ValueError("weights must be [nsum][ncomp][nnodes]")
        Wd = torch.from_numpy(W).to(x.device)
        out = np.zeros(W.shape[0])
        _lib.check(_lib.lib().tmi_compute_force(self.handle, _ptr(x), _ptr(Wd), W.shape[0],
                                                out.ctypes.data_as(ct.c_void_p), _stream()), "tmi_compute_force")
        return out

    def zeros(self) -> torch.Tensor:
        return torch.zeros((self.nunk, self.nnodes), dtype=torch.float64, device="cuda")

    def fill_residual(self, x: torch.Tensor, F: Optional[torch.Tensor] = None) -> Tuple[torch.Tensor, float]:
        F = self.zeros() if F is None else F
        s = ct.c_double()
        _lib.check(_lib.lib().tmi_fill_residual(self.handle, _ptr(x), _ptr(F), ct.byref(s), _stream()),
                   "tmi_fill_residual")
        return F, s.value

    def fill_residual_device_norm(self, x: torch.Tensor, F: torch.Tensor, sumsq: torch.Tensor) -> None:
        """The residual fill with sum F^2 written to the device scalar `sumsq` (float64, one element)
        instead of returned to the host: no host synchronisation (tmi_fill_residual_noalloc)."""
        if sumsq.dtype != torch.float64 or sumsq.numel() < 1 or not sumsq.is_cuda:
            raise ValueError("sumsq must be a float64 device tensor with at least one element")
        _lib.check(_lib.lib().tmi_fill_residual_noalloc(self.handle, _ptr(x), _ptr(F), _ptr(sumsq), _stream()),
                   "tmi_fill_residual_noalloc")

    def fill_residual_timed(self, x: torch.Tensor, reps: int = 10) -> np.ndarray:
        """Mean per-stage milliseconds of `reps` fills (HIP events on the launch stream)."""
        F = self.zeros()
        ms = np.zeros(5)
        _lib.check(_lib.lib().tmi_fill_residual_timed(self.handle, _ptr(x), _ptr(F), reps,
                                                      ms.ctypes.data_as(ct.c_void_p), _stream()),
                   "tmi_fill_residual_timed")
        return ms

    def fft_timed(self, which: int = 0, reps: int = 10) -> Dict:
        """One kernel family of a residual fill in FFT form (HIP events on the launch stream):
        which = 0 the plain axis passes, 1 the fused x passes, 2 the density product's rows pass
        (wrapping grids); mean ms per fill, algorithmic bytes per fill, launches per fill."""
        ms, nbytes, launches = ct.c_double(), ct.c_double(), ct.c_int()
        _lib.check(_lib.lib().tmi_fft_timed(self.handle, which, reps, ct.byref(ms), ct.byref(nbytes),
                                            ct.byref(launches), _stream()), "tmi_fft_timed")
        return dict(ms=ms.value, bytes=nbytes.value, launches=launches.value)

    def apply_jacobian(self, x: torch.Tensor, v: torch.Tensor, out: Optional[torch.Tensor] = None,
                       fp32: bool = False) -> torch.Tensor:
        out = self.zeros() if out is None else out
        fn = _lib.lib().tmi_apply_jacobian_fp32 if fp32 else _lib.lib().tmi_apply_jacobian
        _lib.check(fn(self.handle, _ptr(x), _ptr(v), _ptr(out), _stream()), "tmi_apply_jacobian")
        return out

    def apply_jacobian_transpose(self, x: torch.Tensor, w: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        out = self.zeros() if out is None else out
        _lib.check(_lib.lib().tmi_apply_jacobian_transpose(self.handle, _ptr(x), _ptr(w), _ptr(out), _stream()),
                   "tmi_apply_jacobian_transpose")
        return out

    def init_fields(self, x: torch.Tensor) -> torch.Tensor:
        _lib.check(_lib.lib().tmi_init_fields(self.handle, _ptr(x), _stream()), "tmi_init_fields")
        return x

    def schur_solve(self, x: torch.Tensor, F: torch.Tensor, rtol=1e-10, restart=100, maxit=1000,
                    precision=PREC_FP64) -> Tuple[torch.Tensor, int]:
        dx = self.zeros()
        its = ct.c_int()
        _lib.check(_lib.lib().tmi_schur_solve(self.handle, _ptr(x), _ptr(F), _ptr(dx), rtol, restart, maxit, precision,
                                              ct.byref(its), _stream()), "tmi_schur_solve")
        return dx, its.value

    def newton(self, x: torch.Tensor, max_iter=50, resid_tol=1e-8, nl_rel_tol=1e-3, nl_abs_tol=1e-8,
               update_scaling=0.1, gmres_restart=60, gmres_max_iter=600, gmres_tol=1e-6,
               precision=PREC_FP64, verbose=False, convergence=CONV_RESID) -> Dict:
        """newton_solver (src/dft_newton.c:131-196) on the device; convergence CONV_RESID stops when
        ||F||_2 < resid_tol, CONV_UPDATE on the reference's weighted update norm <= 1."""
        o = _lib.TmiNewtonOpts(max_iter, nl_rel_tol, nl_abs_tol, update_scaling, resid_tol, gmres_restart,
                               gmres_max_iter, gmres_tol, precision, 1 if verbose else 0, convergence)
        st = _lib.TmiNewtonStats()
        _lib.check(_lib.lib().tmi_newton_solve(self.handle, _ptr(x), ct.byref(o), ct.byref(st), _stream()),
                   "tmi_newton_solve")
        return dict(iters=st.iters, converged=bool(st.converged), gmres_iters=st.gmres_iters_total,
                    resid_norm=st.resid_norm, seconds=st.seconds)

    def initial_state(self) -> torch.Tensor:
        x = torch.from_numpy(self.spec.initial_x()).to("cuda")
        return self.init_fields(x.contiguous())
