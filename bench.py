"""Benchmark of the nonlocal-density fill and Newton-Krylov solve on one MI355X.

Headline workload: the 3D 128^3 grid at dx = 0.1 sigma, three-component White Bear FMT + WCA
mean-field mixture (sigma 1.0/0.9/0.7, cut-off 3 sigma_i, kT = 1) around a spherical solute of
radius 4 sigma with an LJ 9-3 surface potential, bulk packing fraction 0.35 split by seeded
Dirichlet weights (tramonto_amd.problem.solute_3d): 2,097,152 nodes, 16 unknowns per node, the
largest single-GPU configuration of the benchmark set.  Inputs are resident in HBM before the
timed region starts.

A step is one FP64 residual fill over the whole grid (rows and sum F^2 computed on the device);
`value` is the fill throughput in Gnode/s over K timed steps.  The same JSON line carries the fill
time when each fill also returns sum F^2 to the host, the Jacobian-vector product throughput (FP64
and FP32), the Newton time-to-solve to ||F||_2 < 1e-8 for the plain FP64 path and the mixed path
(FP32 inner Krylov products, FP64 residual/update), the per-stage kernel times of one fill (HIP
events), the roofline of the dominant kernel (the longest launch of the fill: the density rows pass
at 128^3), and the CPU baseline (the oracle's OpenMP fill on the host cores, on a bounded sample of
nodes).  Extra keys: the 64^3 periodic Rosenfeld + WCA workload, the batched 1D slit pores, the
batched 2D rough slit pores and the 2D adjoint workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-newton] [--no-cpu]
                       [--no-batch] [--no-adjoint] [--no-periodic] [--fill-only] [--copy-probe]
                       [--workload solute128|periodic64]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
COPY_PROBE_BYTES = 512 << 20
# HBM bytes per launch, fetch + write in MiB, of the density product's rows pass in a fill (the dominant
# kernel on wrapping grids): rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of this script with
# --fill-only --copy-probe, that launch's counters scaled by their readings on the 512 MiB copy (x2.0 and
# x1.0).  Keys: workload.
ROWS_TRAFFIC_MIB = {"periodic64": 35.7 + 24.6, "solute128": 534.6 + 234.3}
# The same for the fused x pass (mean per launch) and the plain FFT axis passes, where they are the
# dominant kernel
XPROD_TRAFFIC_MIB = {"periodic64": (7.5 + 12.0 + 14.9 + 4.0) / 2, "solute128": None}
AXIS_TRAFFIC_MIB = {"periodic64": None, "solute128": None}
# HBM bytes of one whole fill in FFT form (every launch of the fill, FETCH + WRITE in MiB, from the
# same counter runs): the fill kernels' measured traffic, against the fill's live time
FILL_TRAFFIC_MIB = {"periodic64": 245.6, "solute128": 4595.7}


def stencil_flops_per_node(dp, spec):
    """Direct-form flops per node of one residual fill (2 per stencil point and product)."""
    nd, C = spec.ndim, spec.ncomp
    L = dp.info.sten_len
    fwd = sum(L[1][c] + (1 + nd) * L[0][c] for c in range(C))          # theta + delta (scalar + vector)
    adj = sum(L[1][c] + (1 + nd) * L[0][c] for c in range(C))
    mf = sum(L[2][q] for q in range(C * C)) if spec.type_attr >= 0 else 0
    return 2.0 * (fwd + adj + mf), 2.0 * fwd, 2.0 * adj, 2.0 * mf


def cpu_baseline(spec, x_np, label, target_s=10.0):
    """The oracle's OpenMP FP64 fill on the host cores, on a bounded sample: the pointwise FMT
    derivatives of every node (as the reference's FMT1stDeriv_switch over all box nodes) plus the
    residual rows of a random node sample sized to about target_s seconds, scaled to the grid."""
    import numpy as np

    from oracle.oracle import Oracle
    o = Oracle.from_spec(spec)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    rng = np.random.default_rng(17)
    N = spec.nnodes
    t0 = time.perf_counter()
    o.residual_nodes(x_np, np.array([0], dtype=np.int32))          # derivatives of every node + one row
    t_deriv = time.perf_counter() - t0
    n = min(N, 256)
    while True:
        nodes = rng.choice(N, n, replace=False).astype(np.int32)
        t0 = time.perf_counter()
        o.residual_nodes(x_np, nodes)
        t = time.perf_counter() - t0
        rows = max(t - t_deriv, 1e-9)
        if n == N or t >= 0.5 * target_s:
            break
        n = int(min(N, max(2 * n, n * 0.9 * target_s / max(t, 1e-3))))
    t_full = t_deriv + rows * N / n
    return dict(value=N / t_full / 1e9, unit="Gnode/s", cores=threads, kind="port",
                sample=f"{label}: pointwise derivatives of all {N} nodes ({t_deriv:.2f} s) plus the residual rows of "
                       f"{n} random nodes ({rows:.2f} s) by the OpenMP C oracle, scaled to the grid: "
                       f"{t_full:.1f} s per FP64 fill")


def bench_batched_pores(npores=256):
    """The uneven slit-pore workload: every pore solved to ||F_p|| < 1e-8, (a) all
    pores at once by the batched per-pore Newton-Krylov (graph replay and direct
    launches, FP64 and mixed), (b) pore by pore with the plain single-problem
    FP64 path.  Problem set-up is outside the timed regions."""
    from tramonto_amd.batch import slit_pore_batch

    return bench_batch(slit_pore_batch(npores=npores),
                       f"{npores} uneven 1D slit pores, H ~ U(4,30) sigma at dx = 0.02 sigma, binary "
                       "(sigma 1.0/0.8) Rosenfeld FMT + WCA (cut-off 2.5 sigma_i, eps_ff/kT = 0.7), LJ 9-3 walls "
                       "eps_w ~ U(0.5,2), rho_b,i ~ U(0.05,0.35)")


def bench_batched_2d(npores=16):
    """Uneven 2D state points: 16 rough slit pores of 64 columns (periodic x) stacked along y, solved
    as bench_batched_pores solves the 1D pores."""
    from tramonto_amd.batch import rough_slit_pore_batch_2d

    return bench_batch(rough_slit_pore_batch_2d(npores=npores, seed=11),
                       f"{npores} uneven 2D rough slit pores, 64 columns (periodic x) at dx = 0.1 sigma, heights "
                       "H ~ U(4,10) sigma between rough LJ 9-3 walls (eps_w ~ U(0.5,1.5)), Rosenfeld FMT + WCA "
                       "(cut-off 2.5 sigma, eps_ff/kT = 0.5), rho_b ~ U(0.3,0.5)")


def bench_batch(b, workload):
    """A batch (BatchSpec) solved all at once and pore by pore; see bench_batched_pores."""
    import os as _os

    import numpy as np
    import torch

    from tramonto_amd.batch import DeviceBatch
    from tramonto_amd.device import PREC_FP64, PREC_MIXED, DeviceProblem

    npores = b.npores
    db = DeviceBatch(b)
    x0 = db.initial_state()
    F = db.zeros()
    ssd = torch.zeros(1, dtype=torch.float64, device="cuda")
    for _ in range(3):
        db.fill_residual_device_norm(x0, F, ssd)
        db.fill_residual(x0, F)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):   # sum F^2 on the device, as the headline fill
        db.fill_residual_device_norm(x0, F, ssd)
    e1.record()
    torch.cuda.synchronize()
    fill_ms = e0.elapsed_time(e1) / 20
    e0.record()
    for _ in range(20):
        db.fill_residual(x0, F)
    e1.record()
    torch.cuda.synchronize()
    fill_ms_host = e0.elapsed_time(e1) / 20
    out = {"workload": workload,
           "pores": npores, "nodes": int(b.nnodes), "pore_nodes_min_max": [int(b.nodes.min()), int(b.nodes.max())],
           "fill_ms": fill_ms, "fill_gnode_s": b.nnodes / (fill_ms * 1e-3) / 1e9, "fill_ms_host_norm": fill_ms_host,
           "fill_stage_ms": dict(zip(["pad_rho", "rb_rows", "mf_rows", "fmt_point+pad_adj", "el_rows"],
                                     [float(v) for v in db.fill_residual_timed(x0, 10)]))}
    sols = {}
    for name, prec, graph in (("batched_fp64_graph", PREC_FP64, "1"), ("batched_fp64_direct", PREC_FP64, "0"),
                              ("batched_mixed_graph", PREC_MIXED, "1"), ("batched_mixed_direct", PREC_MIXED, "0")):
        _os.environ["TMI_BATCH_GRAPH"] = graph
        xw = x0.clone()
        db.newton(xw, max_iter=2, resid_tol=1e-8, gmres_tol=1e-6, precision=prec)   # buffers and graphs
        x = x0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = db.newton(x, max_iter=60, resid_tol=1e-8, gmres_tol=1e-6, precision=prec)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out[name] = {"seconds": wall, "newton_iters": st["iters"], "pores_converged": int(st["pore_converged"].sum()),
                     "gmres_iters_total": st["gmres_iters"], "max_pore_resid": float(st["pore_resid"].max())}
        sols[name] = x
    _os.environ.pop("TMI_BATCH_GRAPH", None)
    # plain path: one single-problem device Newton per pore
    dps = [DeviceProblem(p) for p in b.pores]
    xs = [dp.initial_state() for dp in dps]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nconv, iters = 0, 0
    for dp, x in zip(dps, xs):
        st = dp.newton(x, max_iter=60, resid_tol=1e-8, gmres_tol=1e-6)
        nconv += int(st["converged"])
        iters += st["iters"]
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    out["sequential_fp64"] = {"seconds": seq, "pores_converged": nconv, "newton_iters_total": iters}
    xb = sols["batched_fp64_graph"].cpu().numpy()
    xm = sols["batched_mixed_graph"].cpu().numpy()
    C = b.ncomp
    out["max_profile_diff_batched_vs_sequential"] = max(
        float(np.abs(xp[:C] - x.cpu().numpy()[:C]).max()) for xp, x in zip(b.split(xb), xs))
    out["max_profile_diff_mixed_vs_fp64"] = float(np.abs(xb[:C] - xm[:C]).max())
    out["speedup_batched_fp64_vs_sequential"] = seq / out["batched_fp64_graph"]["seconds"]
    out["speedup_batched_mixed_vs_sequential"] = seq / out["batched_mixed_graph"]["seconds"]
    return out


def bench_adjoint_2d():
    """The rough-wall workload: 256 x 256 at dx = 0.05 sigma; Newton to
    ||F|| < 1e-8 (timed with the problem set-up), then dGamma/d(eps_w) and
    dGamma/d(rho_b) from one backward pass of the implicit-function adjoint
    (transpose Schur GMRES with the device transpose products);
    dGamma/d(beta mu) by the bulk equation of state.  Timed after one untimed
    forward + backward pass."""
    import numpy as np
    import torch

    from tramonto_amd.adjoint import (ParametricModel, dbetamu_drho_b, equilibrium_profile,
                                      excess_adsorption)
    from tramonto_amd.problem import rough_wall_2d

    model = ParametricModel(lambda th: rough_wall_2d(eps_w=float(th[0]), rho_b=float(th[1])),
                            newton_kwargs=dict(resid_tol=1e-8, gmres_tol=1e-7), in_place=True)
    theta0 = np.array([1.5, 0.5])
    spec = model.make_spec(theta0)
    area = spec.nx[1] * spec.dx[1]
    # one untimed forward + backward first: kernel code loading and the transpose buffers
    th_w = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    excess_adsorption(equilibrium_profile(th_w, model), spec, th_w[1:2], area).backward()
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = equilibrium_profile(theta, model)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gam = excess_adsorption(x, spec, theta[1:2], area)
    gam.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    dmu = dbetamu_drho_b(spec)[0, 0]
    return {"workload": "2D 256x256 at dx = 0.05 sigma, Rosenfeld FMT + WCA (eps = 1, cut-off 2.5 sigma, kT = 1), "
                        "rough LJ 9-3 wall eps_w = 1.5 (amplitude 0.5 sigma, correlation length 2 sigma), "
                        "rho_b = 0.5",
            "nodes": spec.nnodes, "newton_seconds": t1 - t0, "newton": model.last_newton,
            "adjoint_seconds": t2 - t1, "adjoint": model.last_adjoint,
            "gamma": float(gam.detach()), "dgamma_deps_w": float(theta.grad[0]),
            "dgamma_drho_b": float(theta.grad[1]), "dgamma_dbetamu": float(theta.grad[1]) / dmu}


def bench_workload(spec, label, args, newton_tol=1e-8, newton=True):
    """Fill throughput, per-stage times, JVP throughput, the dominant kernel's roofline and the FP64 and
    mixed Newton time-to-solve of one workload.  Set-up is outside every timed region."""
    import torch

    from tramonto_amd.device import PREC_FP64, PREC_MIXED, DeviceProblem

    dp = DeviceProblem(spec)
    x0 = dp.initial_state()
    torch.cuda.synchronize()
    N = spec.nnodes
    F = dp.zeros()
    ssd = torch.zeros(1, dtype=torch.float64, device="cuda")
    # the timed fill: residual rows and sum F^2 on the device (tmi_fill_residual_noalloc, no host read
    # per fill); the same fill returning sum F^2 to the host each time is timed beside it
    for _ in range(args.warmup):
        dp.fill_residual_device_norm(x0, F, ssd)
        dp.fill_residual(x0, F)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        dp.fill_residual_device_norm(x0, F, ssd)
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_fill = e0.elapsed_time(e1) / args.steps
    e0.record()
    for _ in range(args.steps):
        _, ss_host = dp.fill_residual(x0, F)
    e1.record()
    torch.cuda.synchronize()
    ms_fill_host = e0.elapsed_time(e1) / args.steps
    if abs(float(ssd.item()) - ss_host) > 1e-12 * max(ss_host, 1e-300):
        raise RuntimeError(f"device and host sums of squares differ: {float(ssd.item())} vs {ss_host}")
    out = {"nodes": N, "unknowns_per_node": spec.nunk, "fill_ms": ms_fill, "fill_gnode_s": N / (ms_fill * 1e-3) / 1e9,
           "fill_ms_host_norm": ms_fill_host, "wall_s_timed": wall}
    stages = dp.fill_residual_timed(x0, reps=max(3, args.steps // 5))
    jvp = {}
    v = torch.rand_like(x0)
    Jv = dp.zeros()
    for fp32 in (() if args.fill_only else (False, True)):
        for _ in range(args.warmup):
            dp.apply_jacobian(x0, v, Jv, fp32=fp32)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(args.steps):
            dp.apply_jacobian(x0, v, Jv, fp32=fp32)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / args.steps
        jvp["fp32" if fp32 else "fp64"] = dict(ms=ms, gnode_s=N / (ms * 1e-3) / 1e9)
    out["jvp"] = jvp
    flops_node, f_fwd, f_adj, f_mf = stencil_flops_per_node(dp, spec)
    C = spec.ncomp
    bytes_node = 8.0 * (2 * spec.nunk + C) + C
    out["flops_per_node_direct"] = flops_node
    out["bytes_per_node_residual"] = bytes_node
    out["fill_hbm_frac"] = (N * bytes_node / (ms_fill * 1e-3)) / (PEAK_HBM_GBS * 1e9)
    ft = FILL_TRAFFIC_MIB.get(label) if dp.info.fmt_fft and dp.info.mf_fft else None
    if ft:
        gbs = ft * (1 << 20) / (ms_fill * 1e-3) / 1e9
        out["fill_traffic"] = {"mib_per_fill": ft, "achieved_gb_s": gbs, "frac": gbs / PEAK_HBM_GBS,
                               "note": "HBM bytes of all the fill's launches (rocprofv3 FETCH_SIZE + WRITE_SIZE per "
                                       "launch, scaled by a 512 MiB copy) over the live fill time; fill_hbm_frac "
                                       "counts only the compulsory bytes (x, F, V_ext, zero flags)"}
    out["mean_field_form"] = "fft" if dp.info.mf_fft else "direct"
    out["fmt_form"] = ["direct", "fft", "fft, x pass fused with the products"][dp.info.fmt_fft]
    out["fft_xreg"] = dp.info.fft_xreg
    out["precond"] = ["A22 diagonal", "bulk symbol"][dp.info.precond]
    out["stencil_lengths"] = {"delta": [dp.info.sten_len[0][c] for c in range(C)],
                              "theta": [dp.info.sten_len[1][c] for c in range(C)],
                              "uatt": [dp.info.sten_len[2][q] for q in range(C * C)]}
    if dp.info.fmt_fft:
        # every stencil sum is an FFT convolution.  Dominant kernel: the longer launch of the density
        # product's last pass on a wrapping grid (inverse z transform of the output pairs fused with the
        # rho-bar and mean-field rows) and the fused x pass when it is in use, else the axis passes; each
        # timed with HIP events around its launches in a fill's order (tmi_fft_timed)
        out["fill_stage_ms"] = dict(zip(["pad_rho", "rb_mf_rows_fft", "mf_rows_direct", "fmt_point", "el_rows_fft"],
                                        [float(v) for v in stages]))
        fused = dp.info.fmt_fft == 2
        reps = max(5, args.steps // 5)
        fp = dp.fft_timed(0, reps)
        try:
            fr = dp.fft_timed(2, reps)
        except RuntimeError:
            fr = None
        axis = {"ms_per_fill": fp["ms"], "launches_per_fill": fp["launches"],
                "achieved_gb_s": fp["bytes"] / (fp["ms"] * 1e-3) / 1e9,
                "frac": fp["bytes"] / (fp["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS,
                "kernels": "k_fft_strided<double, +-1, 0, R1, R2> (y, z) and k_fft_xline (x), plain mode"}
        fx = dp.fft_timed(1, reps) if fused else None
        if fr is not None and fx is not None and fx["ms"] / fx["launches"] > fr["ms"] / fr["launches"]:
            fr = None   # the fused x pass is the longer launch
        if fr is not None:
            ft, tr = fr, ROWS_TRAFFIC_MIB.get(label)
            kernel = "k_fft_strided<double, 1, 2, 16, 8> (density rows pass)" if label == "solute128" else \
                "k_fft_strided<double, 1, 2, R1, R2> (density rows pass)"
            note = ("last inverse z pass of the density product fused with the rho-bar and mean-field rows: output "
                    "pairs read once, each row's unknowns read and residual written once (algorithmic), HIP events "
                    "on the launch stream")
        elif fused:
            ft = fx
            tr = XPROD_TRAFFIC_MIB.get(label)
            kernel = "k_fft_xprod_reg" if dp.info.fft_xreg else "k_fft_xprod"
            note = ("fused x passes of one fill: input pairs and distinct spectrum tables read, output pairs "
                    "written (algorithmic), HIP events on the launch stream")
        else:
            ft, tr, kernel = fp, AXIS_TRAFFIC_MIB.get(label), "k_fft_strided / k_fft_xline"
            note = "FFT axis passes of one fill (HIP events on the launch stream), HBM bound"
        achieved = ft["bytes"] / (ft["ms"] * 1e-3) / 1e9
        out["roofline"] = {
            "bound": "hbm", "kernel": kernel,
            "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
            "traffic": tr * (1 << 20) if tr else None,
            "launch_ms": ft["ms"] / ft["launches"], "launches_per_fill": ft["launches"],
            "algorithmic_bytes_per_launch": ft["bytes"] / ft["launches"], "share_of_fill": ft["ms"] / ms_fill,
            "traffic_note": ("HBM bytes per launch of that kernel in a fill, from rocprofv3 --pmc "
                             "FETCH_SIZE/WRITE_SIZE scaled by a 512 MiB copy"),
            "note": note, "axis_passes": axis}
    else:
        names = ["pad_rho", "rb_rows", "mf_rows", "fmt_point+pad_adj", "el_rows"]
        out["fill_stage_ms"] = dict(zip(names, [float(v) for v in stages]))
        per_stage_flops = [0.0, f_fwd, 0.0 if dp.info.mf_fft else f_mf, 0.0, f_adj]
        priced = [i for i in range(5) if per_stage_flops[i] > 0]
        dom = max(priced, key=lambda i: stages[i])
        achieved = per_stage_flops[dom] * N / (stages[dom] * 1e-3) / 1e12
        out["roofline"] = {"bound": "mfma", "kernel": names[dom], "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                           "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                           "note": "direct-form stencil sum, FP64 FMA bound (MI355X FP64 vector peak = FP64 matrix "
                                   "peak = 78.6 TFLOP/s)"}
    if newton and not (args.no_newton or args.fill_only):
        res = {}
        sols = {}
        for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
            xw = x0.clone()
            dp.newton(xw, max_iter=1, resid_tol=newton_tol, gmres_tol=1e-6, precision=prec)   # scratch, code load
            xs = x0.clone()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = dp.newton(xs, max_iter=60, resid_tol=newton_tol, gmres_tol=1e-6, precision=prec)
            torch.cuda.synchronize()
            st["wall_s"] = time.perf_counter() - t0
            res[name] = st
            sols[name] = xs
        res["mixed_vs_fp64_time_ratio"] = res["fp64"]["wall_s"] / res["mixed"]["wall_s"]
        res["max_profile_diff_mixed_vs_fp64"] = float((sols["fp64"] - sols["mixed"])[: spec.ncomp].abs().max())
        out["newton"] = res
    out["_x0"] = x0
    return out


WORKLOADS = {
    "solute128": ("3D 128^3 grid at dx = 0.1 sigma, 3-component White Bear FMT + WCA mean field (sigma 1.0/0.9/0.7, "
                  "eps = 1, cut-off 3 sigma_i, kT = 1), spherical solute of radius 4 sigma with LJ 9-3 surface, "
                  "eta_b = 0.35 split by seeded Dirichlet weights, periodic box"),
    "periodic64": ("3D periodic Rosenfeld FMT + WCA mean field, 64^3 grid, dx = 0.1 sigma, cut-off 2.5 sigma, "
                   "kT = 1, seeded field clip(0.6 + 0.2 sum of 8 random-phase modes of wavelength 2-6 sigma, "
                   "0.05, 0.9)"),
}


def make_spec(name):
    from tramonto_amd.problem import periodic_3d, solute_3d
    return solute_3d() if name == "solute128" else periodic_3d()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="solute128")
    ap.add_argument("--no-newton", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched workloads (1D slit pores, 2D pieces)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 2D rough-wall adjoint workload")
    ap.add_argument("--no-periodic", action="store_true", help="skip the 64^3 periodic workload (extra key)")
    ap.add_argument("--fill-only", action="store_true",
                    help="only the fill throughput, its stages and the roofline (profiling runs)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first run a plain 512 MiB device copy three times (calibrates FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.copy_probe:
        a = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    spec = make_spec(args.workload)
    head = bench_workload(spec, args.workload, args)
    x0 = head.pop("_x0")
    cpu = None
    if not (args.no_cpu or args.fill_only):
        cpu = cpu_baseline(spec, x0.cpu().numpy(), args.workload)
    del x0
    torch.cuda.empty_cache()

    extra = {}
    if args.workload == "solute128" and not (args.no_periodic or args.fill_only):
        e = bench_workload(make_spec("periodic64"), "periodic64", args)
        e.pop("_x0")
        e["workload"] = WORKLOADS["periodic64"]
        extra["periodic_64"] = e
    if not (args.no_batch or args.fill_only):
        extra["batched_pores"] = bench_batched_pores()
        extra["batched_2d"] = bench_batched_2d()
    if not (args.no_adjoint or args.fill_only):
        extra["adjoint_2d"] = bench_adjoint_2d()

    out = {
        "metric": "Newton time-to-solve & fill Gnodes/s, mixed/batched vs plain FP64, res<1e-8",
        "value": head["fill_gnode_s"],
        "unit": "Gnode/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["fill_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "nodes": spec.nnodes, "unknowns_per_node": spec.nunk,
                   "components": spec.ncomp, "grid": list(spec.nx)},
        "roofline": head.pop("roofline"),
        "cpu_baseline": cpu,
        "newton": head.pop("newton", {}),
    }
    out.update(head)
    out.update(extra)
    print(json.dumps(out, default=float))


if __name__ == "__main__":
    main()
