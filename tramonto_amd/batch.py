"""Batches of independent 1D or 2D problems on one device ("pores").

A batch packs P problems that share the fluid (components, functional,
potentials, grid spacing) into one device problem: unknown rows hold pore
0's nodes, then pore 1's, and so on.  Each pore keeps its own node count,
bulk densities, V_ext and zero-density flags, and the stencil kernels see
every pore between its own walls (1D: walls at both ends; 2D: pieces of a
common x width, walled at both y ends, stacked along y).
``DeviceBatch.newton`` solves all pores at once with a per-pore
Newton-Krylov (tmi_batch_newton_solve).

``slit_pore_batch`` builds the uneven slit-pore workload: binary Rosenfeld +
WCA fluid in P pores of width H ~ U(4, 30) sigma between LJ 9-3 walls with
eps_w ~ U(0.5, 2) and bulk densities rho_b,i ~ U(0.05, 0.35), all seeded.
"""
from __future__ import annotations

import ctypes as ct
import dataclasses
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .device import PREC_FP64, DeviceProblem, _ptr, _stream
from .problem import IN_WALL, PERIODIC, ProblemSpec, SplitMix64, rough_slit_pore_2d, slit_pore_1d


@dataclasses.dataclass
class BatchSpec:
    pores: List[ProblemSpec]

    def __post_init__(self):
        if not self.pores:
            raise ValueError("empty batch")
        f = self.pores[0]
        if f.ndim not in (1, 2):
            raise ValueError("batches hold 1D or 2D problems")
        for p in self.pores:
            if p.ndim != f.ndim or p.ncomp != f.ncomp or p.dx != f.dx or p.type_func != f.type_func \
                    or p.type_attr != f.type_attr or list(p.hs_diam) != list(f.hs_diam):
                raise ValueError("pores of a batch must share the fluid and the grid spacing")
            for a, b in ((p.sigma_ff, f.sigma_ff), (p.eps_ff, f.eps_ff), (p.cut_ff, f.cut_ff)):
                if (a is None) != (b is None) or (a is not None and not np.array_equal(a, b)):
                    raise ValueError("pores of a batch must share the pair potentials")
            if p.bc[p.ndim - 1] != [IN_WALL, IN_WALL]:
                raise ValueError("batched pores have walls (IN_WALL) at both ends of their last axis")
            if p.ndim == 2 and (p.nx[0] != f.nx[0] or p.bc[0] != f.bc[0] or p.bc[0][0] != p.bc[0][1]
                                or p.bc[0][0] not in (PERIODIC, IN_WALL)):
                raise ValueError("2D pieces share nx and a periodic or walled x boundary")

    @property
    def npores(self) -> int:
        return len(self.pores)

    @property
    def ncomp(self) -> int:
        return self.pores[0].ncomp

    @property
    def nodes(self) -> np.ndarray:
        return np.array([p.nnodes for p in self.pores], dtype=np.int32)

    @property
    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.nodes)]).astype(np.int64)

    @property
    def nnodes(self) -> int:
        return int(self.nodes.sum())

    @property
    def nunk(self) -> int:
        return self.pores[0].nunk

    def packed(self, name: str) -> Optional[np.ndarray]:
        """[N, C] node-major array of the pores' vext or zero_tf."""
        C = self.ncomp
        parts = []
        for p in self.pores:
            a = getattr(p, name)
            if a is None:
                a = np.zeros((p.nnodes, C), dtype=np.uint8 if name == "zero_tf" else np.float64)
            parts.append(np.asarray(a).reshape(p.nnodes, C))
        return np.ascontiguousarray(np.concatenate(parts, axis=0))

    def initial_x(self) -> np.ndarray:
        return np.concatenate([p.initial_x() for p in self.pores], axis=1)

    def split(self, x: np.ndarray) -> List[np.ndarray]:
        o = self.offsets
        return [x[:, o[i]:o[i + 1]] for i in range(self.npores)]


def slit_pore_batch(npores: int = 256, seed: int = 9, h_range=(4.0, 30.0), dx: float = 0.02, sig=(1.0, 0.8),
                    eps_range=(0.5, 2.0), rho_range=(0.05, 0.35), cut_fac: float = 2.5,
                    eps_ff: float = 0.7) -> BatchSpec:
    rng = SplitMix64(seed)
    H = rng.uniform(npores, *h_range)
    eps = rng.uniform(npores, *eps_range)
    rho = rng.uniform(npores * len(sig), *rho_range).reshape(npores, len(sig))
    pores = [slit_pore_1d(H=float(np.round(H[i] / dx) * dx), dx=dx, sig=sig, rho_b=tuple(rho[i]), eps_w=float(eps[i]),
                          cut_fac=cut_fac, seed=seed * 1000 + i, eps_ff=eps_ff) for i in range(npores)]
    return BatchSpec(pores)


def rough_slit_pore_batch_2d(npores: int = 16, seed: int = 11, nx: int = 64, h_range=(4.0, 10.0), dx: float = 0.1,
                             eps_range=(0.5, 1.5), rho_range=(0.3, 0.5), cut: float = 2.5,
                             eps_ff: float = 0.5) -> BatchSpec:
    """Uneven 2D state points: P rough slit pores (rough_slit_pore_2d) of a common width nx dx along
    x (periodic), heights H ~ U(h_range) sigma, wall strengths eps_w ~ U(eps_range) and bulk
    densities rho_b ~ U(rho_range), all seeded.  The fluid-fluid eps_ff/kT = 0.5 keeps the fluid
    above its mean-field critical temperature: at eps_ff/kT = 1 these narrow pores condense and a
    Newton start from the gas-like Boltzmann profile diverges (the oracle's as well)."""
    rng = SplitMix64(seed)
    H = rng.uniform(npores, *h_range)
    eps = rng.uniform(npores, *eps_range)
    rho = rng.uniform(npores, *rho_range)
    pores = [rough_slit_pore_2d(seed=seed * 1000 + i, nx=nx, ny=int(np.round(H[i] / dx)) + 1, dx=dx,
                                eps_w=float(eps[i]), rho_b=float(rho[i]), cut=cut, eps_ff=eps_ff)
             for i in range(npores)]
    return BatchSpec(pores)


class DeviceBatch(DeviceProblem):
    """A batch on the device; the single-problem methods apply to the packed batch."""

    def __init__(self, batch: BatchSpec):
        L = _lib.lib()
        self.batch = batch
        self.spec = batch.pores[0]
        self.params = _lib.fill_params(_lib.TmiParams(), batch.pores[0].as_params())
        C = batch.ncomp
        nodes = np.ascontiguousarray(batch.nodes, dtype=np.int32)
        rho_b = np.ascontiguousarray(np.array([p.rho_b[:C] for p in batch.pores], dtype=np.float64))
        vext = batch.packed("vext")
        zero = batch.packed("zero_tf").astype(np.uint8)
        self._keep = (nodes, rho_b, vext, zero)
        h = ct.c_void_p()
        _lib.check(L.tmi_batch_create(ct.byref(self.params), batch.npores, nodes.ctypes.data_as(ct.c_void_p),
                                      rho_b.ctypes.data_as(ct.c_void_p), vext.ctypes.data_as(ct.c_void_p),
                                      zero.ctypes.data_as(ct.c_void_p), ct.byref(h)), "tmi_batch_create")
        self.handle = h
        info = _lib.TmiInfo()
        _lib.check(L.tmi_problem_info(h, ct.byref(info)), "tmi_problem_info")
        self.info = info
        self.nunk, self.nnodes = info.nunk, info.nnodes

    def initial_state(self) -> torch.Tensor:
        x = torch.from_numpy(self.batch.initial_x()).to("cuda").contiguous()
        return self.init_fields(x)

    def newton(self, x: torch.Tensor, max_iter=50, resid_tol=1e-8, nl_rel_tol=1e-3, nl_abs_tol=1e-8,
               update_scaling=0.1, gmres_restart=60, gmres_max_iter=600, gmres_tol=1e-6,
               precision=PREC_FP64, verbose=False) -> Dict:
        o = _lib.TmiNewtonOpts(max_iter, nl_rel_tol, nl_abs_tol, update_scaling, resid_tol, gmres_restart,
                               gmres_max_iter, gmres_tol, precision, 1 if verbose else 0)
        st = _lib.TmiNewtonStats()
        P = self.batch.npores
        its = np.zeros(P, dtype=np.int32)
        conv = np.zeros(P, dtype=np.int32)
        gm = np.zeros(P, dtype=np.int32)
        rn = np.zeros(P)
        _lib.check(_lib.lib().tmi_batch_newton_solve(self.handle, _ptr(x), ct.byref(o), ct.byref(st),
                                                     its.ctypes.data_as(ct.c_void_p), conv.ctypes.data_as(ct.c_void_p),
                                                     gm.ctypes.data_as(ct.c_void_p), rn.ctypes.data_as(ct.c_void_p),
                                                     _stream()), "tmi_batch_newton_solve")
        return dict(iters=st.iters, converged=bool(st.converged), gmres_iters=st.gmres_iters_total,
                    resid_norm=st.resid_norm, seconds=st.seconds, pore_iters=its, pore_converged=conv.astype(bool),
                    pore_gmres=gm, pore_resid=rn)
