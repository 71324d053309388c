"""Differentiable equilibrium profiles: torch.autograd.Function over the C ABI.

Two functions make the device fill differentiable:

* :class:`Residual` -- F(x) with the backward pass J(x)^T g, computed by the
  hand-written transpose-Jacobian kernels (``tmi_apply_jacobian_transpose``).
* :class:`EquilibriumProfile` -- x*(theta), the converged solution of
  F(x; theta) = 0 (device Newton-Krylov forward).  Its backward pass is the
  implicit-function adjoint: solve J(x*)^T lam = dL/dx*, then
  dL/dtheta_k = -lam . dF/dtheta_k.

With a scalar observable computed in torch from x* (the adsorption, say),
``Gamma.backward()`` yields dGamma/dtheta: the sensitivities LOCA-style
continuation in a parameter needs (the reference reaches them only through
repeated solves, src/dft_continuation.c, src/NOXLOCA_Tramonto_Group.cpp).

The transpose solve runs restarted GMRES on the transpose Schur complement.
With the density unknowns d and the nonlocal unknowns n (rho-bar and
mean-field rows), J = [[J_dd, J_dn], [J_nd, -I]] (the reference's A22, A21,
A12, A11 = -I; src/dft_Schur_Epetra_Operator.cpp:86-175), so J^T lam = g reduces to

    (J_dd + J_nd^T J_dn^T) lam_d = g_d + J_nd^T g_n,     lam_n = J_dn^T lam_d - g_n,

each application costing two transpose products; J_dd is diagonal (1/rho, or
1 on zero-density rows) and preconditions from the right.
"""
from __future__ import annotations

from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .device import PREC_FP64, DeviceProblem
from .problem import PERIODIC, ProblemSpec


# --------------------------------------------------------------------------- residual
class Residual(torch.autograd.Function):
    """F = residual(x) on the device; backward: J(x)^T grad_F."""

    @staticmethod
    def forward(ctx, x: torch.Tensor, dp: DeviceProblem) -> torch.Tensor:
        xc = x.detach().contiguous()
        F, _ = dp.fill_residual(xc)
        ctx.dp = dp
        ctx.save_for_backward(xc)
        return F

    @staticmethod
    def backward(ctx, gF: torch.Tensor):
        (x,) = ctx.saved_tensors
        return ctx.dp.apply_jacobian_transpose(x, gF.contiguous()), None


def residual(x: torch.Tensor, dp: DeviceProblem) -> torch.Tensor:
    return Residual.apply(x, dp)


# --------------------------------------------------------------------------- transpose solve
def _gmres(apply, b: torch.Tensor, precond, tol: float, restart: int, maxit: int) -> Tuple[torch.Tensor, int, float]:
    """Right-preconditioned restarted GMRES (classical Gram-Schmidt, applied twice)."""
    x = torch.zeros_like(b)
    bnorm = float(torch.linalg.vector_norm(b))
    if bnorm == 0.0:
        return x, 0, 0.0
    total = 0
    r = b.clone()
    rel = 1.0
    while total < maxit:
        beta = float(torch.linalg.vector_norm(r))
        rel = beta / bnorm
        if rel <= tol:
            break
        m = min(restart, maxit - total)
        V = torch.zeros((m + 1, b.numel()), dtype=b.dtype, device=b.device)
        Z = torch.zeros((m, b.numel()), dtype=b.dtype, device=b.device)
        H = torch.zeros((m + 1, m), dtype=torch.float64)
        V[0] = r.reshape(-1) / beta
        k_done = 0
        for k in range(m):
            Z[k] = precond(V[k].reshape(b.shape)).reshape(-1)
            w = apply(Z[k].reshape(b.shape)).reshape(-1)
            for _ in range(2):
                h = V[: k + 1] @ w
                w = w - V[: k + 1].T @ h
                H[: k + 1, k] += h.cpu()
            H[k + 1, k] = float(torch.linalg.vector_norm(w))
            k_done = k + 1
            total += 1
            if H[k + 1, k] > 0:
                V[k + 1] = w / H[k + 1, k]
            e1 = torch.zeros(k + 2, dtype=torch.float64)
            e1[0] = beta
            y = torch.linalg.lstsq(H[: k + 2, : k + 1], e1).solution
            res = float(torch.linalg.vector_norm(H[: k + 2, : k + 1] @ y - e1))
            if res / bnorm <= tol or H[k + 1, k] == 0:
                break
        y = torch.linalg.lstsq(H[: k_done + 1, :k_done], torch.cat([torch.tensor([beta], dtype=torch.float64),
                                                                    torch.zeros(k_done, dtype=torch.float64)])).solution
        x = x + (Z[:k_done].T @ y.to(b.device)).reshape(b.shape)
        r = b - apply(x)
    return x, total, float(torch.linalg.vector_norm(r)) / bnorm


def transpose_solve(dp: DeviceProblem, x: torch.Tensor, g: torch.Tensor, tol: float = 1e-11, restart: int = 80,
                    maxit: int = 2000) -> Tuple[torch.Tensor, dict]:
    """lam with J(x)^T lam = g (g, lam: [nunk, N] float64 on the device)."""
    C = dp.spec.ncomp
    zero_rows = torch.from_numpy(np.zeros((C, dp.nnodes), dtype=bool) if dp.spec.zero_tf is None else
                                 (np.asarray(dp.spec.zero_tf).reshape(dp.nnodes, C).T != 0)).to(x.device)
    if dp.spec.nwall > 0 and dp.spec.vext is not None:
        vx = torch.from_numpy(np.asarray(dp.spec.vext, dtype=np.float64).reshape(dp.nnodes, C).T.copy()).to(x.device)
        zero_rows = zero_rows | (vx == dp.spec.vext_max)
    dinv = torch.where(zero_rows, torch.ones_like(x[:C]), x[:C])    # (J_dd)^-1

    def vjp_d(ld):          # J^T [ld; 0]
        w = torch.zeros_like(x)
        w[:C] = ld
        return dp.apply_jacobian_transpose(x, w)

    def vjp_n(un):          # J^T [0; un]
        w = torch.zeros_like(x)
        w[C:] = un
        return dp.apply_jacobian_transpose(x, w)

    def schur_t(ld):
        r = vjp_d(ld)
        return r[:C] + vjp_n(r[C:])[:C]

    rhs = g[:C] + vjp_n(g[C:])[:C]
    lam_d, its, rel = _gmres(schur_t, rhs.contiguous(), lambda v: dinv * v, tol, restart, maxit)
    lam = torch.zeros_like(x)
    lam[:C] = lam_d
    lam[C:] = vjp_d(lam_d)[C:] - g[C:]
    check = dp.apply_jacobian_transpose(x, lam) - g
    info = dict(gmres_iters=its, schur_rel_resid=rel,
                rel_resid=float(torch.linalg.vector_norm(check) / torch.linalg.vector_norm(g)))
    return lam, info


# --------------------------------------------------------------------------- parameters
class ParametricModel:
    """A family of problems F(x; theta).

    ``make_spec(theta)`` returns the ProblemSpec at parameter vector theta
    (numpy float64).  dF/dtheta at fixed x is, by default, the central
    difference of the device residual between the problems at theta +- h_k
    (relative step ``rel_step``); a model may override ``dF_dtheta`` with a
    closed form.
    """

    def __init__(self, make_spec: Callable[[np.ndarray], ProblemSpec], rel_step: float = 1e-6,
                 newton_kwargs: Optional[dict] = None, in_place: bool = False):
        """in_place: theta changes only the bulk densities and V_ext (not the zero-density flags),
        so dF/dtheta can update the solved problem in place (tmi_problem_set_rho_b / _set_vext)
        instead of building the problems at theta +- h."""
        self.make_spec = make_spec
        self.rel_step = rel_step
        self.in_place = in_place
        self.newton_kwargs = dict(max_iter=60, resid_tol=1e-11, gmres_tol=1e-9, precision=PREC_FP64)
        if newton_kwargs:
            self.newton_kwargs.update(newton_kwargs)

    def problem(self, theta: np.ndarray) -> DeviceProblem:
        return DeviceProblem(self.make_spec(np.asarray(theta, dtype=np.float64)))

    def _apply(self, dp: DeviceProblem, theta: np.ndarray) -> None:
        spec = self.make_spec(np.asarray(theta, dtype=np.float64))
        if (spec.zero_tf is None) != (dp.spec.zero_tf is None) or (
                spec.zero_tf is not None and not np.array_equal(spec.zero_tf, dp.spec.zero_tf)):
            raise ValueError("theta changes the zero-density flags: not an in-place parameter")
        dp.set_rho_b(spec.rho_b)
        if spec.nwall > 0:
            dp.set_vext(spec.vext)

    def dF_dtheta(self, theta: np.ndarray, x: torch.Tensor, dp: Optional[DeviceProblem] = None) -> List[torch.Tensor]:
        out = []
        inplace = self.in_place and dp is not None
        for k in range(theta.size):
            h = self.rel_step * max(abs(theta[k]), 1.0)
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            if inplace:
                self._apply(dp, tp)
                Fp, _ = dp.fill_residual(x)
                self._apply(dp, tm)
                Fm, _ = dp.fill_residual(x)
            else:
                Fp, _ = self.problem(tp).fill_residual(x)
                Fm, _ = self.problem(tm).fill_residual(x)
            out.append((Fp - Fm) / (2 * h))
        if inplace:
            self._apply(dp, theta)
        return out


class EquilibriumProfile(torch.autograd.Function):
    """x*(theta) with F(x*; theta) = 0; backward by the implicit-function adjoint."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, model: ParametricModel, x_init: Optional[torch.Tensor] = None):
        th = theta.detach().cpu().numpy().astype(np.float64)
        dp = model.problem(th)
        x = dp.initial_state() if x_init is None else x_init.detach().clone().contiguous()
        st = dp.newton(x, **model.newton_kwargs)
        if not st["converged"]:
            raise RuntimeError(f"Newton did not converge: {st}")
        ctx.model, ctx.dp, ctx.theta = model, dp, th
        model.last_newton = st
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, gx: torch.Tensor):
        (x,) = ctx.saved_tensors
        lam, info = transpose_solve(ctx.dp, x, gx.contiguous())
        ctx.model.last_adjoint = info
        dF = ctx.model.dF_dtheta(ctx.theta, x, ctx.dp)
        grad = torch.tensor([-float((lam * d).sum()) for d in dF], dtype=torch.float64)
        return grad, None, None


def equilibrium_profile(theta: torch.Tensor, model: ParametricModel,
                        x_init: Optional[torch.Tensor] = None) -> torch.Tensor:
    return EquilibriumProfile.apply(theta, model, x_init)


# --------------------------------------------------------------------------- observables
def node_volumes(spec: ProblemSpec) -> np.ndarray:
    """Integration weight of every node: the product over dimensions of dx
    (halved at a non-periodic domain edge), i.e. the trapezoid rule."""
    w = np.ones(1)
    for d in range(spec.ndim):
        n = spec.nx[d]
        wd = np.full(n, spec.dx[d])
        if spec.bc[d][0] != PERIODIC:
            wd[0] *= 0.5
            wd[-1] *= 0.5
        w = np.multiply.outer(wd, w).reshape(-1) if d > 0 else wd
    # node index i + j*nx + k*nx*ny: x fastest
    return w.reshape(-1)


def excess_adsorption(x: torch.Tensor, spec: ProblemSpec, rho_b: torch.Tensor, area: float) -> torch.Tensor:
    """Gamma_c = sum_i w_i (rho_c(i) - rho_b,c [i not excluded]) / area, summed over c."""
    C = spec.ncomp
    w = torch.from_numpy(node_volumes(spec)).to(x.device)
    acc = torch.zeros((), dtype=torch.float64, device=x.device)
    for c in range(C):
        free = torch.ones(spec.nnodes, dtype=torch.float64, device=x.device)
        if spec.zero_tf is not None:
            free = free - torch.from_numpy(np.asarray(spec.zero_tf)[:, c].astype(np.float64)).to(x.device)
        acc = acc + (w * (x[c] - rho_b[c].to(x.device) * free)).sum()
    return acc / area


def bulk_betamu(spec: ProblemSpec) -> np.ndarray:
    """beta mu_c of the bulk state: ln rho_b,c + beta mu_hs,ex + beta mu_att
    (chempot_FMT_hs, src/dft_thermo_hs.c:136-154; chempot_att,
    src/dft_thermo_att.c:86-97), from the library's host-side bulk constants."""
    import ctypes as ct

    from . import _lib
    params = _lib.fill_params(_lib.TmiParams(), spec.as_params())
    out = np.zeros(60)
    _lib.check(_lib.lib().tmi_host_bulk(ct.byref(params), out.ctypes.data_as(ct.c_void_p)), "tmi_host_bulk")
    C = spec.ncomp
    att = out[24:24 + C] if spec.type_attr >= 0 else 0.0
    return np.log(np.asarray(spec.rho_b[:C])) + out[20:20 + C] + att


def dbetamu_drho_b(spec: ProblemSpec, rel_step: float = 1e-6) -> np.ndarray:
    """d(beta mu_c)/d rho_b,c' (C x C) by central differences of bulk_betamu."""
    import copy
    C = spec.ncomp
    D = np.zeros((C, C))
    for j in range(C):
        h = rel_step * spec.rho_b[j]
        sp, sm = copy.copy(spec), copy.copy(spec)
        sp.rho_b = list(spec.rho_b)
        sm.rho_b = list(spec.rho_b)
        sp.rho_b[j] += h
        sm.rho_b[j] -= h
        D[:, j] = (bulk_betamu(sp) - bulk_betamu(sm)) / (2 * h)
    return D
