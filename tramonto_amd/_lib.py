"""ctypes binding of libtramonto_mi355x.so (the C ABI in include/tramonto_mi355x.h).

The shared library is built in-tree (tramonto_amd/lib/).  There is no
fallback: if the library is missing or fails to load, every entry point
raises.
"""
from __future__ import annotations

import ctypes as ct
import os
from typing import Optional

NC = 4
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libtramonto_mi355x.so")


class TmiParams(ct.Structure):
    _fields_ = [
        ("ndim", ct.c_int), ("ncomp", ct.c_int),
        ("nx", ct.c_int * 3), ("dx", ct.c_double * 3),
        ("bc", (ct.c_int * 2) * 3),
        ("type_func", ct.c_int), ("type_attr", ct.c_int), ("nwall", ct.c_int), ("lhard_surf", ct.c_int),
        ("vext_max", ct.c_double),
        ("hs_diam", ct.c_double * NC),
        ("sigma_ff", (ct.c_double * NC) * NC), ("eps_ff", (ct.c_double * NC) * NC), ("cut_ff", (ct.c_double * NC) * NC),
        ("rho_b", ct.c_double * NC),
        ("att_in_a22", ct.c_int),
    ]


class TmiInfo(ct.Structure):
    _fields_ = [
        ("nunk", ct.c_int), ("nnodes", ct.c_int), ("iu_rho", ct.c_int), ("iu_rb", ct.c_int), ("n_rb", ct.c_int),
        ("iu_mf", ct.c_int), ("n_mf", ct.c_int), ("sten_len", (ct.c_int * (NC * NC)) * 3), ("mf_fft", ct.c_int),
        ("fmt_fft", ct.c_int), ("precond", ct.c_int), ("fft_fold", ct.c_int), ("fft_xreg", ct.c_int),
        ("fft_regpass", ct.c_int), ("jacobian", ct.c_int),
    ]


class TmiObservables(ct.Structure):
    _fields_ = [("omega", ct.c_double), ("omega_s", ct.c_double), ("ideal", ct.c_double), ("ideal_b", ct.c_double),
                ("mu", ct.c_double), ("mu_b", ct.c_double), ("vext", ct.c_double), ("hs", ct.c_double),
                ("hs_b", ct.c_double), ("att", ct.c_double), ("att_b", ct.c_double),
                ("ads", ct.c_double * NC), ("ads_b", ct.c_double * NC)]


class TmiNewtonOpts(ct.Structure):
    _fields_ = [
        ("max_iter", ct.c_int), ("nl_rel_tol", ct.c_double), ("nl_abs_tol", ct.c_double),
        ("update_scaling", ct.c_double), ("resid_tol", ct.c_double), ("gmres_restart", ct.c_int),
        ("gmres_max_iter", ct.c_int), ("gmres_tol", ct.c_double), ("precision", ct.c_int), ("verbose", ct.c_int),
        ("convergence", ct.c_int),
    ]


class TmiNewtonStats(ct.Structure):
    _fields_ = [("iters", ct.c_int), ("converged", ct.c_int), ("gmres_iters_total", ct.c_int),
                ("resid_norm", ct.c_double), ("seconds", ct.c_double)]


def fill_params(struct, d: dict):
    for k, v in d.items():
        f = getattr(struct, k)
        if isinstance(v, (list, tuple)):
            if isinstance(v[0], (list, tuple)):
                for i, row in enumerate(v):
                    for j, e in enumerate(row):
                        f[i][j] = e
            else:
                for i, e in enumerate(v):
                    f[i] = e
        else:
            setattr(struct, k, v)
    return struct


_lib: Optional[ct.CDLL] = None

# exported symbols declared in include/tramonto_mi355x.h
DECLARED_SYMBOLS = [
    "tmi_last_error", "tmi_version", "tmi_problem_create", "tmi_problem_destroy", "tmi_problem_info",
    "tmi_stencil_get", "tmi_bulk_get", "tmi_host_stencil", "tmi_host_bulk", "tmi_apply_jacobian_fp32", "tmi_fill_residual_timed", "tmi_fill_residual_noalloc", "tmi_fft_timed", "tmi_fill_residual", "tmi_apply_jacobian", "tmi_apply_jacobian_transpose",
    "tmi_init_fields", "tmi_set_hard_surface_masks", "tmi_compute_force", "tmi_newton_solve", "tmi_schur_solve", "tmi_set_active_problem",
    "tmi_batch_create", "tmi_batch_info", "tmi_batch_newton_solve", "tmi_row_tuning", "tmi_row_tuning_set",
    "tmi_problem_set_rho_b", "tmi_problem_set_vext", "tmi_problem_set_preconditioner", "tmi_problem_set_jacobian", "tmi_compute_observables",
    "fill_resid_and_matrix_control", "tmi_linprobmgr_last_solve",
    "dft_parameterlist_create", "dft_parameterlist_destruct", "dft_parameterlist_set_int", "dft_parameterlist_set_double",
    "dft_parameterlist_set_int_array", "dft_parameterlist_set_double_array", "dft_parameterlist_set_all_aztec",
    "dft_hardsphere_lin_prob_mgr_create", "dft_hardsphere_lin_prob_mgr_create_debug",
    "dft_hardsphere_lin_prob_mgr_destruct", "dft_basic_lin_prob_mgr_create",
    "dft_hardsphere_lin_prob_mgr_setindnonlocalequationids",
    "dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids", "dft_hardsphere_lin_prob_mgr_setdensityequationids",
    "dft_hardsphere_lin_prob_mgr_seta22blockisdiagonal",
    "dft_linprobmgr_destruct", "dft_linprobmgr_setnodalrowmap", "dft_linprobmgr_setnodalcolmap",
    "dft_linprobmgr_setcoarsenednodeslist", "dft_linprobmgr_finalizeblockstructure",
    "dft_linprobmgr_initializeproblemvalues", "dft_linprobmgr_insertrhsvalue", "dft_linprobmgr_insertonematrixvalue",
    "dft_linprobmgr_insertmultinodematrixvalues", "dft_linprobmgr_insertmultiphysicsmatrixvalues",
    "dft_linprobmgr_finalizeproblemvalues", "dft_linprobmgr_setrhs", "dft_linprobmgr_getrhs", "dft_linprobmgr_getlhs",
    "dft_linprobmgr_setupsolver", "dft_linprobmgr_solve", "dft_linprobmgr_applymatrix", "dft_linprobmgr_importr2c",
    "dft_linprobmgr_importnodalr2c", "dft_linprobmgr_importsingleunknownr2c", "dft_linprobmgr_getmatrixvalue",
    "dft_linprobmgr_writeMatrix",
]


def lib() -> ct.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libtramonto_mi355x.so not built (expected {LIB_PATH}); run __graft_entry__.build()")
    L = ct.CDLL(LIB_PATH)
    vp, i, d = ct.c_void_p, ct.c_int, ct.c_double
    L.tmi_last_error.restype = ct.c_char_p
    L.tmi_problem_create.argtypes = [ct.POINTER(TmiParams), vp, vp, ct.POINTER(vp)]
    L.tmi_problem_destroy.argtypes = [vp]
    L.tmi_problem_info.argtypes = [vp, ct.POINTER(TmiInfo)]
    L.tmi_stencil_get.argtypes = [vp, i, i, i, ct.POINTER(i), vp, vp]
    L.tmi_bulk_get.argtypes = [vp, vp]
    L.tmi_host_stencil.argtypes = [ct.POINTER(TmiParams), i, i, i, ct.POINTER(i), vp, vp]
    L.tmi_host_bulk.argtypes = [ct.POINTER(TmiParams), vp]
    L.tmi_fill_residual.argtypes = [vp, vp, vp, ct.POINTER(d), vp]
    L.tmi_fill_residual_noalloc.argtypes = [vp, vp, vp, vp, vp]
    L.tmi_apply_jacobian.argtypes = [vp, vp, vp, vp, vp]
    L.tmi_fill_residual_timed.argtypes = [vp, vp, vp, i, vp, vp]
    L.tmi_fft_timed.argtypes = [vp, i, i, ct.POINTER(d), ct.POINTER(d), ct.POINTER(i), vp]
    L.tmi_apply_jacobian_fp32.argtypes = [vp, vp, vp, vp, vp]
    L.tmi_apply_jacobian_transpose.argtypes = [vp, vp, vp, vp, vp]
    L.tmi_init_fields.argtypes = [vp, vp, vp]
    L.tmi_set_hard_surface_masks.argtypes = [vp, vp]
    L.tmi_compute_force.argtypes = [vp, vp, vp, ct.c_int, vp, vp]
    L.tmi_newton_solve.argtypes = [vp, vp, ct.POINTER(TmiNewtonOpts), ct.POINTER(TmiNewtonStats), vp]
    L.tmi_schur_solve.argtypes = [vp, vp, vp, vp, d, i, i, i, ct.POINTER(i), vp]
    L.tmi_set_active_problem.argtypes = [vp, vp]
    L.tmi_row_tuning.argtypes = [vp, vp, vp]
    L.tmi_row_tuning_set.argtypes = [vp, i, i, i, i]
    L.tmi_problem_set_rho_b.argtypes = [vp, vp]
    L.tmi_problem_set_vext.argtypes = [vp, vp]
    L.tmi_problem_set_preconditioner.argtypes = [vp, i]
    L.tmi_problem_set_jacobian.argtypes = [vp, i]
    L.tmi_compute_observables.argtypes = [vp, vp, vp, vp, d, ct.POINTER(TmiObservables), vp]
    L.tmi_batch_create.argtypes = [ct.POINTER(TmiParams), i, vp, vp, vp, vp, ct.POINTER(vp)]
    L.tmi_batch_info.argtypes = [vp, ct.POINTER(i), vp]
    L.tmi_batch_newton_solve.argtypes = [vp, vp, ct.POINTER(TmiNewtonOpts), ct.POINTER(TmiNewtonStats), vp, vp, vp,
                                         vp, vp]
    L.fill_resid_and_matrix_control.argtypes = [vp, i, i]
    L.fill_resid_and_matrix_control.restype = d
    for name in ("dft_hardsphere_lin_prob_mgr_create", "dft_hardsphere_lin_prob_mgr_create_debug",
                 "dft_basic_lin_prob_mgr_create"):
        getattr(L, name).argtypes = [i, vp, i]   # the communicator is MPICH's int MPI_Comm, never read
        getattr(L, name).restype = vp
    for name in ("dft_linprobmgr_destruct", "dft_hardsphere_lin_prob_mgr_destruct", "dft_parameterlist_destruct"):
        getattr(L, name).argtypes = [vp]
        getattr(L, name).restype = None
    L.dft_parameterlist_create.argtypes = []
    L.dft_parameterlist_create.restype = vp
    L.dft_parameterlist_set_int.argtypes = [vp, ct.c_char_p, i]
    L.dft_parameterlist_set_int.restype = None
    L.dft_parameterlist_set_double.argtypes = [vp, ct.c_char_p, d]
    L.dft_parameterlist_set_double.restype = None
    L.dft_parameterlist_set_all_aztec.argtypes = [vp, vp, vp]
    L.dft_parameterlist_set_all_aztec.restype = None
    L.tmi_linprobmgr_last_solve.argtypes = [vp, ct.POINTER(i), ct.POINTER(d)]
    for name in ("dft_hardsphere_lin_prob_mgr_setindnonlocalequationids",
                 "dft_hardsphere_lin_prob_mgr_setdepnonlocalequationids",
                 "dft_hardsphere_lin_prob_mgr_setdensityequationids", "dft_linprobmgr_setnodalrowmap",
                 "dft_linprobmgr_setnodalcolmap", "dft_linprobmgr_setcoarsenednodeslist"):
        getattr(L, name).argtypes = [vp, i, vp]
    L.dft_hardsphere_lin_prob_mgr_seta22blockisdiagonal.argtypes = [vp, i]
    for name in ("dft_linprobmgr_finalizeblockstructure",
                 "dft_linprobmgr_initializeproblemvalues", "dft_linprobmgr_finalizeproblemvalues",
                 "dft_linprobmgr_setupsolver", "dft_linprobmgr_solve"):
        getattr(L, name).argtypes = [vp]
    L.dft_linprobmgr_insertrhsvalue.argtypes = [vp, i, i, d]
    L.dft_linprobmgr_insertonematrixvalue.argtypes = [vp, i, i, i, i, d]
    L.dft_linprobmgr_setrhs.argtypes = [vp, vp]
    L.dft_linprobmgr_getrhs.argtypes = [vp, vp]
    L.dft_linprobmgr_getlhs.argtypes = [vp, vp]
    L.dft_linprobmgr_applymatrix.argtypes = [vp, vp, vp]
    L.dft_linprobmgr_importr2c.argtypes = [vp, vp, vp]
    L.dft_linprobmgr_insertmultinodematrixvalues.argtypes = [vp, i, i, i, vp, vp, i]
    L.dft_linprobmgr_insertmultiphysicsmatrixvalues.argtypes = [vp, i, i, vp, i, vp, i]
    L.dft_linprobmgr_importnodalr2c.argtypes = [vp, vp, vp]
    L.dft_linprobmgr_importsingleunknownr2c.argtypes = [vp, vp, vp]
    L.dft_linprobmgr_writeMatrix.argtypes = [vp, ct.c_char_p, ct.c_char_p, ct.c_char_p]
    L.dft_linprobmgr_getmatrixvalue.argtypes = [vp, i, i, i, i]
    L.dft_linprobmgr_getmatrixvalue.restype = d
    _lib = L
    return L


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().tmi_last_error()
        raise RuntimeError(f"{what} failed (code {rc}): {msg.decode() if msg else ''}")
