#!/usr/bin/env python
"""Throughput of the PISO pressure-velocity step on one MI355X.

Workload: 3-D lid-driven cavity on an octree (unit cube, 32^3 root blocks,
two refinement levels in cells within 0.1 of any wall, 2:1 balanced,
1,125,384 cells), Re = 1000, PISO with 2 correctors at CFL 0.5.  One "step"
is one PISO time step.  The run is made twice from the same initial state:
plain FP64 and with the FP32-SpMV PCG inside the FP64 refinement loop; the
headline value is the mixed-precision rate when its final U and p are
within 1e-6 relative L2 of the plain FP64 run.

Also measured: the dominant kernel (the PCG SpMV) timed with HIP events on
the solver's stream, against the HBM roofline, and the FP64 CPU oracle on
the host cores as a bounded sample of the same step; the 2-D SIMPLE cavity
(64x64, 500 iterations) on the GPU beside the oracle; the serving stream;
plain FP64 against mixed on the 21.0M-cell bluff body in a tunnel (a few
steps each, ``--bluff-steps``).
"""

from __future__ import annotations

import argparse
import copy
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
CASE = "cavity3d_octree_piso.yaml"


def _log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


INNER_RTOL = None  # --inner-rtol: refinement trigger of the mixed solve (None: the config default)


def _cfg(mixed: bool) -> dict:
    import yaml

    from gridfoam_amd.config import CASES

    cfg = yaml.safe_load((CASES / CASE).read_text())
    for k in ("p", "pFinal"):
        cfg["simulator"]["fvSolution"]["solvers"][k]["mixed_precision"] = bool(mixed)
        if mixed and INNER_RTOL is not None:
            cfg["simulator"]["fvSolution"]["solvers"][k]["inner_rel_tolerance"] = INNER_RTOL
    return cfg


def run_gpu(mixed: bool, steps: int, warmup: int):
    import torch

    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config

    grid, alg = create_case(load_config(_cfg(mixed)), device="cuda:0")
    for _ in range(warmup):
        alg.step()
    torch.cuda.synchronize()
    p_iters = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        alg.step()
        p_iters += sum(s.iterations for s in alg.p_solve_history[-alg.n_correctors:])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    return grid, alg, dt, p_iters


def bluff_measure(steps: int, warmup: int = 1, check_every: int = 10, mixed: bool = True,
                  keep: bool = False) -> dict:
    """The 3-D bluff body in a tunnel (21.0M cells, gridfoam_amd/cases/bluff3d_piso.yaml):
    ``steps`` timed PISO steps with the FP32-in-FP64 pressure solve (plain
    FP64 with ``mixed`` False) and the FP64 continuity check of the
    corrected flux every ``check_every`` steps, inside the timed region; the
    grid build and ``warmup`` steps before it.  ``keep``: the final U and p
    (host arrays) under "_U" / "_p"."""
    import torch
    import yaml

    from gridfoam_amd.algorithms import compute_continuity_error, create_case
    from gridfoam_amd.config import CASES, load_config
    from gridfoam_amd.forces import ForceEvaluator

    cfg = yaml.safe_load((CASES / "bluff3d_piso.yaml").read_text())
    cfg["simulator"]["fvSolution"]["solvers"]["p"]["mixed_precision"] = bool(mixed)
    if mixed and INNER_RTOL is not None:
        cfg["simulator"]["fvSolution"]["solvers"]["p"]["inner_rel_tolerance"] = INNER_RTOL
    torch.cuda.reset_peak_memory_stats()
    t0 = time.perf_counter()
    grid, alg = create_case(load_config(cfg), device="cuda:0")
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    _log(f"bluff body: {grid.num_cells} cells built in {t_build:.1f} s; {warmup} warmup + {steps} timed steps")
    for _ in range(warmup):
        alg.step()
    torch.cuda.synchronize()
    p_iters, checks = 0, []
    t0 = time.perf_counter()
    for k in range(steps):
        alg.step()
        it = sum(s.iterations for s in alg.p_solve_history[-alg.n_correctors:])
        p_iters += it
        msg = f"bluff body: step {k + 1}, {it} p-CG iterations"
        if (k + 1) % check_every == 0:
            checks.append(compute_continuity_error(alg.phi, grid.dt)[0])
            msg += f", continuity error {checks[-1]:.3e}"
        _log(msg)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cd = float(ForceEvaluator(grid).evaluate(turbulence=float(alg.nu)).Cd)
    C = grid.num_cells
    _log("bluff body: SpMV roofline on its mesh")
    pm = laplacian_operator(grid)
    roof = {"fp32": spmv_roofline(grid, pm, fp32=True, launches=50),
            "fp64": spmv_roofline(grid, pm, fp32=False, launches=50)}
    kept = {"_U": alg.U.data.detach().cpu().numpy(), "_p": alg.p.data.detach().cpu().numpy(),
            "_fluid": fluid_cells(grid)} if keep else {}
    return {**kept, "workload": "3-D bluff body in a tunnel: 20x4x3 root, surface level 4, wake box level 3, "
                                "3 leaf refinements (7 levels), PISO 2 correctors, Re=1e4",
            "pressure_solve": "f32-in-f64" if mixed else "f64",
            "cells": C, "immersed_faces": grid.num_immersed_faces, "steps": steps, "warmup": warmup,
            "value": round(C * steps / dt / 1e6, 3), "unit": "Mcell-steps/s",
            "ms_per_step": round(dt / steps * 1e3, 1), "p_cg_iters_per_step": round(p_iters / steps, 1),
            "continuity_error_local_fp64_every_10": checks, "Cd_after": round(cd, 6),
            "grid_build_s": round(t_build, 2),
            "peak_allocated_GiB": round(torch.cuda.max_memory_allocated() / 2 ** 30, 2), "roofline": roof}


def adjoint_measure(steps: int, every: int = 10, fd_coefs=(1,), h: float = 1e-4, tol=None) -> dict:
    """The Fourier-body channel (205,600 cells, PIMPLE 2x2, FP64,
    gridfoam_amd/cases/channel2d_fourier_pimple.yaml, its own solver
    tolerances) from its initial state: the drag coefficient after ``steps``
    steps and its gradient with respect to the body's 16 outline
    coefficients, through checkpointed steps (a checkpoint every ``every``
    steps); the forward without a graph for comparison; central differences
    (step ``h``) of the coefficients in ``fd_coefs`` on grids cut with
    perturbed outlines, a check at the resolution the solves allow.  ``tol``:
    relative tolerance of both solvers instead of the case's."""
    import copy

    import torch
    import yaml

    from gridfoam_amd.adjoint import TransientStep, checkpointed_rollout, transient_state
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import CASES, ForceCoeffConfig, load_config
    from gridfoam_amd.forces import ForceEvaluator
    from gridfoam_amd.shape import FourierShape

    base = yaml.safe_load((CASES / "channel2d_fourier_pimple.yaml").read_text())
    if tol is not None:
        for k, it in (("U", 2000), ("p", 8000)):
            base["simulator"]["fvSolution"]["solvers"][k].update({"tolerance": 0.0, "rel_tolerance": tol,
                                                                   "max_iter": it})
    force = ForceCoeffConfig.model_validate({
        "patches": ["body"], "rho": 1.0, "magU_ref": 1.0, "A_ref": 1.0, "L_ref": 1.0,
        "local_coord": {"drag_dir": [1.0, 0.0, 0.0], "lift_dir": [0.0, 1.0, 0.0],
                        "center_of_rotation": [5.0, 4.0, 0.25]}})

    def case(coef=None):
        cfg = copy.deepcopy(base)
        if coef is not None:
            body = cfg["fluxel"]["immersed_bodies"][0]
            body["a"], body["b"] = [float(v) for v in coef[:8]], [float(v) for v in coef[8:]]
        return create_case(load_config(cfg), device="cuda:0")

    def drag(alg, grid, state0, grad):
        s = tuple(t.clone() for t in state0)
        step = TransientStep(alg)
        calls = [0]

        def logged(st):
            calls[0] += 1
            if calls[0] % 50 == 0:
                _log(f"adjoint: step call {calls[0]}")
            return step(st)

        with torch.set_grad_enabled(grad):
            params = grid.shape.tensors() if grad and grid.shape is not None else None
            out = checkpointed_rollout(logged, s, steps, every=every if grad else None, params=params)
            alg.U.data, alg.p.data = out[0], out[3]
            alg.U._state = alg.p._state = None
            return ForceEvaluator(grid, config=force).evaluate(turbulence=float(alg.nu)).Cd

    grid, alg = case()
    state0 = transient_state(alg)
    shape = FourierShape(grid, "body")
    c0 = shape.coefficients.detach().cpu().tolist()
    _log(f"adjoint: forward without a graph, {steps} PIMPLE steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cd_plain = float(drag(alg, grid, state0, False))
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    _log(f"adjoint: {t_plain:.1f} s; checkpointed forward and backward")
    torch.cuda.reset_peak_memory_stats()
    mem0 = torch.cuda.memory_allocated()
    with shape:
        t0 = time.perf_counter()
        cd = drag(alg, grid, state0, True)
        torch.cuda.synchronize()
        t_fwd = time.perf_counter() - t0
        (g,) = torch.autograd.grad(cd, shape.coefficients)
        torch.cuda.synchronize()
        t_bwd = time.perf_counter() - t0 - t_fwd
    peak = (torch.cuda.max_memory_allocated() - mem0) / 2 ** 30
    _log(f"adjoint: forward {t_fwd:.1f} s, backward {t_bwd:.1f} s, Cd {float(cd):.10f}")
    fd = {}
    for j in fd_coefs:
        vals = []
        for sgn in (1.0, -1.0):
            c = list(c0)
            c[j] += sgn * h
            gj, aj = case(c)
            same = all(torch.equal(getattr(gj, k), getattr(grid, k)) for k in (
                "ap_is_immersed_faces", "ap_owner_near_boundary", "ap_neighbour_near_boundary"))
            if not same:
                raise RuntimeError("perturbed outline changed the cut topology")
            vals.append(float(drag(aj, gj, state0, False)))
            del gj, aj
        fd[j] = (vals[0] - vals[1]) / (2 * h)
        _log(f"adjoint: coefficient {j}: adjoint {float(g[j]):.8e}, central difference {fd[j]:.8e}")
    return {"workload": "2-D channel around the Fourier body (205,600 cells), PIMPLE 2x2 from the initial state, "
                        "drag coefficient gradient w.r.t. the 16 outline coefficients",
            "solver_tolerances": ("the case's: U 1e-8 absolute, p 1e-8 relative" if tol is None else
                                  f"U and p {tol:g} relative"), "fd_step": h,
            "steps": steps, "checkpoint_every": every, "Cd": round(float(cd), 10), "Cd_no_graph": round(cd_plain, 10),
            "forward_no_graph_s": round(t_plain, 2), "forward_checkpointed_s": round(t_fwd, 2),
            "backward_s": round(t_bwd, 2), "gradient_vs_forward": round((t_fwd + t_bwd) / t_plain, 2),
            "peak_extra_GiB": round(peak, 3), "dCd_dcoef": [float(v) for v in g.cpu()],
            "central_differences": {str(j): v for j, v in fd.items()},
            "rel_err_vs_central_differences": {str(j): abs(float(g[j]) - v) / abs(v) for j, v in fd.items()}}


def bluff_compare(steps: int, mode: str = "both") -> dict:
    """bluff_measure in plain FP64 and mixed precision (``mode`` "both") from
    the same initial state: rates, their ratio and the relative L2 of the
    mixed U and p against the plain ones over the cells connected to the
    fluid (and over all cells)."""
    import numpy as np

    res = {}
    for m in (("plain", "mixed") if mode == "both" else (mode,)):
        res[m] = bluff_measure(steps, mixed=m == "mixed", keep=mode == "both")
        _log(f"bluff body {m}: {res[m]['ms_per_step']} ms/step")
    if mode == "both":
        P, M = res["plain"], res["mixed"]
        res["speedup_mixed_vs_fp64"] = round(P["ms_per_step"] / M["ms_per_step"], 3)
        fl = P["_fluid"]

        def rel(k, mask):
            x, y = M["_" + k][mask], P["_" + k][mask]
            return float(np.linalg.norm(x - y) / np.linalg.norm(y))

        res["fluid_cells"] = int(fl.sum())
        res["mixed_vs_fp64_rel_l2"] = {k: rel(k, fl) for k in ("U", "p")}
        # all cells: the pressure inside the body is fixed by each solver's iteration alone
        res["mixed_vs_fp64_rel_l2_all_cells"] = {k: rel(k, slice(None)) for k in ("U", "p")}
        for r in (P, M):
            del r["_U"], r["_p"], r["_fluid"]
    return res


def fluid_cells(grid):
    """Mask of the cells connected to cell 0 through the internal faces (host).
    Immersed faces couple nothing, so the cells inside a body form regions
    whose pressure has only zero-gradient conditions: it is fixed up to a
    constant by each solver's iteration alone and is left out of the
    comparisons."""
    import numpy as np
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components

    C, F = grid.num_cells, grid.num_internal_faces
    o = grid.owner_i32[:F].cpu().numpy()
    n = grid.neighbour_i32[:F].cpu().numpy()
    adj = sp.coo_matrix((np.ones(F, np.int8), (o, n)), shape=(C, C)).tocsr()
    _, lab = connected_components(adj, directed=False)
    return lab == lab[0]


def laplacian_operator(grid, seed: int = 0):
    """Pressure-type operator for the SpMV measurement: -laplacian(gamma, p)
    with all-Neumann walls, gamma ~ U(0.5, 1.5) (torch generator seed 0),
    the reference cell's diagonal doubled as set_reference_value does."""
    import torch

    from gridfoam_amd.fields import CellField, NeumannBC
    from gridfoam_amd.fv import fvm

    gen = torch.Generator(device="cpu").manual_seed(seed)
    gamma = (0.5 + torch.rand(grid.num_cells, generator=gen, dtype=torch.float64)).to(grid.device)
    P = CellField(grid, "p_spmv", ())
    P.add_boundary_conditions({k: NeumannBC(0.0) for k in
                               ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")})
    pm = fvm.assemble(P, lap=-1.0, gamma=gamma, corrected=False)
    pm.diag[0] = 2 * pm.diag[0]
    return pm


def uniform_grid(n: int):
    """Uniform n^3 octree over the unit cube (no refinement)."""
    from gridfoam_amd.config import load_config
    from gridfoam_amd.grid import create_grid

    cfg = _cfg(False)
    cfg["fluxel"]["root_resolution"] = [n, n, n]
    cfg["fluxel"]["refinement_regions"] = []
    return create_grid(load_config(cfg), device="cuda:0")


def spmv_roofline(grid, pmat, fp32: bool, launches: int = 200) -> dict:
    """Average duration of the PCG direction/SpMV kernel k_cg_dir (the
    solver's own kernel on the solver's layout, launched on its own by
    gf_pcg_dir_bench), timed with HIP events on the stream it is launched on."""
    import ctypes

    import torch

    from gridfoam_amd import _native as nat

    gen = torch.Generator(device="cpu").manual_seed(1)
    x = torch.randn(grid.num_cells, generator=gen, dtype=torch.float64).to(device=grid.device)
    y = torch.empty_like(x)
    scratch = torch.empty(int(nat.lib().gf_pcg_dir_bench_bytes(grid.mesh_ref, int(fp32))), dtype=torch.uint8,
                          device=grid.device)

    def bench(phase: int, reps: int = 1) -> None:
        nat.call("gf_pcg_dir_bench", grid.mesh_ref, ctypes.c_int32(phase), ctypes.c_int32(int(fp32)),
                 ctypes.c_int32(reps), nat.ptr(pmat._diag), nat.ptr(pmat._upper), nat.ptr(pmat._lower), nat.ptr(x),
                 nat.ptr(y), nat.ptr(scratch), nat.stream())

    bench(0)
    bench(1, 10)
    s = torch.cuda.current_stream()  # the stream gf_pcg_dir_bench launches on
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    bench(1, launches)
    e1.record(s)
    torch.cuda.synchronize()
    avg_s = e0.elapsed_time(e1) / 1e3 / launches
    bench(2)
    # check: y = A x against the matrix's own FP64 multiply
    ref = pmat.multiply(x)
    err = float(torch.linalg.norm(y - ref) / torch.linalg.norm(ref))
    C, F = grid.num_cells, grid.num_internal_faces
    if fp32:
        # Algorithmic bytes of one fused mixed-precision launch k_cg_fused: the
        # symmetric LDU SpMV values (one FP32 value per face, 4F, two int32
        # indices per face, 8F) plus ten FP32 vectors of C: row sum and M^-1,
        # r, Ap, x and p_old read; x, r, p and Ap written.  40C + 12F.
        B = 40 * C + 12 * F
    else:
        # Algorithmic bytes of one k_cg_dir launch: the symmetric LDU SpMV
        # (diagonal 8C, one off-diagonal value per face 8F, two int32 indices
        # per face 8F, x read and Ap written 16C) plus the direction update it
        # carries (p_old read, p written: 16C).  40C + 16F.
        B = 40 * C + 16 * F
    achieved = B / avg_s / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": _traffic(fp32, C),
            "kernel": f"k_cg_fused<{os.environ.get('GRIDFOAM_FUSED_TB', '1024')}, 1>" if fp32 else "k_cg_dir<double, 1, false>",
            "cells": C, "bricks": int(grid.mesh.n_bricks), "brick_cells_frac": round(64 * grid.mesh.n_bricks / C, 4),
            "bytes_per_launch": int(B), "avg_us": round(avg_s * 1e6, 2), "rel_err_vs_fp64": err}


def stream_copy(a, b) -> None:
    import ctypes

    from gridfoam_amd import _native as nat

    nat.call("gf_stream_copy", nat.ptr(a), nat.ptr(b), ctypes.c_int64(a.numel() * a.element_size()), nat.stream())


def copy_peak(nbytes: int = 1 << 30, launches: int = 50) -> float:
    """Rate of the library's plain copy kernel (read + write bytes / s) on
    1 GiB, a measured HBM reference beside the vendor peak."""
    import torch

    a = torch.ones(nbytes // 8, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    stream_copy(a, b)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(launches):
        stream_copy(a, b)
    e1.record(s)
    torch.cuda.synchronize()
    return round(nbytes * launches / (e0.elapsed_time(e1) / 1e3) / 1e9, 1)


def _traffic(fp32: bool, cells: int):
    """HBM bytes per SpMV launch from the committed rocprofv3 counter runs,
    for the mesh those runs measured (None for any other mesh)."""
    p = ROOT / "profiles" / "spmv_traffic.json"
    if not p.exists():
        return None
    try:
        d = json.loads(p.read_text())
    except (ValueError, OSError):
        return None
    prec = "fp32" if fp32 else "fp64"
    for key, rec in d.items():
        if key.split("_")[0] == prec and rec.get("cells") == cells:
            return rec.get("bytes_per_launch")
    return None


def _cpu_model() -> str:
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(steps: int) -> dict:
    """The FP64 oracle (the CPU restatement of the reference's path) on the
    host: `steps` complete PISO steps of the same case from the same initial
    state, no iteration cap and no scaling; setup (mesh + algorithm) timed
    apart from the step loop, as solve_s of tests/profile/case.py:149-205.
    The Krylov multiplies run through a CSR copy on N = len(sched_getaffinity)
    torch threads; numpy's elementwise work is single-threaded (BLAS limited
    to one thread so it does not compete with torch's pool)."""
    import threadpoolctl
    import torch

    from oracle import fv_oracle as fo

    n = len(os.sched_getaffinity(0))
    torch.set_num_threads(n)
    fo.use_threaded_spmv(n)
    try:
        with threadpoolctl.threadpool_limits(1, user_api="blas"):
            t0 = time.perf_counter()
            case = fo.Case(_cfg(False))
            setup_s = time.perf_counter() - t0
            its = []
            t0 = time.perf_counter()
            for k in range(steps):
                case.step()
                its.append(int(case.stats["p"][0].iterations))
                _log(f"cpu baseline: oracle step {k + 1}/{steps} done")
            solve_s = time.perf_counter() - t0
    finally:
        fo.use_threaded_spmv(0)
    C = case.g.num_cells
    return {"value": round(C * steps / solve_s / 1e6, 5), "unit": "Mcell-steps/s", "cores": n, "kind": "port",
            "cpu_count": os.cpu_count(), "cpu_model": _cpu_model(), "steps": steps,
            "setup_s": round(setup_s, 2), "solve_s": round(solve_s, 2),
            "sample": (f"{steps} complete PISO steps of the same {C}-cell case from the same initial state "
                       f"(no iteration cap, no scaling; last pressure solve per step {its} CG iterations); "
                       f"FP64 numpy oracle, Krylov SpMV on {n} torch threads (CSR), numpy elementwise "
                       f"single-threaded")}


def config1_measure(cpu: bool, cpu_iters: int = 100) -> dict:
    """The reference's CPU-runnable case: the 2-D lid-driven cavity, 64x64x1,
    Re 100, SIMPLE, 500 iterations, FP64.  GPU: the 500 iterations timed
    (synchronised at both ends), the final U and p against the oracle's
    500-iteration state (the committed golden fixture).  CPU: the FP64 oracle
    timed over the first `cpu_iters` iterations on the host."""
    import numpy as np
    import torch
    import yaml

    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import CASES, load_config

    cfg = yaml.safe_load((CASES / "cavity2d_simple.yaml").read_text())
    n = int(round(cfg["simulator"]["control"]["endTime"] / cfg["simulator"]["control"]["deltaT"]))
    grid, alg = create_case(load_config(cfg), device="cuda:0")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        alg.step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    C = grid.num_cells
    out = {"workload": "2-D lid-driven cavity 64x64x1, Re 100, SIMPLE, 500 iterations, FP64", "cells": C,
           "iterations": n, "solve_s": round(dt, 3), "value": round(C * n / dt / 1e6, 4), "unit": "Mcell-steps/s"}
    gold = ROOT / "tests" / "golden" / "cavity2d_simple_500.npz"
    if gold.exists():
        d = np.load(gold)
        U, p = alg.U.data.cpu().numpy(), alg.p.data.cpu().numpy()
        out["rel_l2_vs_oracle"] = {"U": float(np.linalg.norm(U - d["U"]) / np.linalg.norm(d["U"])),
                                   "p": float(np.linalg.norm(p - d["p"]) / np.linalg.norm(d["p"]))}
    if cpu:
        from oracle import fv_oracle as fo

        case = fo.Case(cfg)
        t0 = time.perf_counter()
        for _ in range(cpu_iters):
            case.step()
        tc = (time.perf_counter() - t0) / cpu_iters
        out["cpu_baseline"] = {"value": round(C / tc / 1e6, 5), "unit": "Mcell-steps/s", "cores": 1, "kind": "port",
                               "sample": f"the first {cpu_iters} SIMPLE iterations of the same case "
                                         f"({tc * 1e3:.0f} ms each); numpy FP64, 1 thread"}
    return out


def serve_measure(n_requests: int, slots: int, budget: int = 30, seed: int = 0, lanes: int = 2) -> dict:
    """Serving stream of random steady requests (lid-driven cube on a base
    grid n^3, n log-uniform in [16, 96], 1-4 box obstacles with 0-2
    refinement levels around them, Re U(100, 1000); SIMPLE until U and p
    residuals < 1e-6 or `budget` iterations, FP32-in-FP64 pressure solves).
    Measured: one-at-a-time plain FP64 solves/s through the SIMPLE driver;
    the batched server's capacity with every request queued at t = 0
    (requests below 48^3 root cells batched `slots` at a time on one
    stream, larger ones alone in two lanes on their own streams, all
    concurrently); p50/p99 latency with Poisson arrivals at 80% of that
    capacity; the measured device footprint per batched cell."""
    import numpy as np
    import torch

    from gridfoam_amd.batching import BatchServer, base_cells
    from gridfoam_amd.serving import latency_stats, random_case, solve_case

    large = 48 ** 3
    rng = np.random.default_rng(seed)
    cases = [random_case(rng, max_iterations=budget) for _ in range(n_requests)]
    solve_case(random_case(np.random.default_rng(1), n_range=(16, 16), max_iterations=2))  # library, kernels
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, c in enumerate(cases):
        pl = copy.deepcopy(c)
        pl["simulator"]["fvSolution"]["solvers"]["p"]["mixed_precision"] = False
        solve_case(pl)
        _log(f"serving: one at a time {k + 1}/{n_requests} ({base_cells(c)} root cells), "
             f"{time.perf_counter() - t0:.1f} s")
    t_seq = time.perf_counter() - t0
    server = BatchServer(max_cases=slots, large_cells=large, large_lanes=lanes, max_bytes=64 << 30)
    server.on_result = lambda r: _log(f"serving: request {r.rid} done at {r.end:.1f} s "
                                      f"({r.iterations} SIMPLE iterations, converged {r.converged})")
    t0 = time.perf_counter()
    sat = server.run([(k, c, 0.0) for k, c in enumerate(cases)])
    t_sat = time.perf_counter() - t0
    _log(f"serving: all {n_requests} queued at once served in {t_sat:.1f} s; Poisson arrivals next")
    cap = n_requests / t_sat
    arr = np.cumsum(rng.exponential(1.0 / (0.8 * cap), size=n_requests))
    poi = server.run([(k, c, float(a)) for k, (c, a) in enumerate(zip(cases, arr))])
    lat = latency_stats(poi)
    return {"requests": n_requests, "batch_slots": slots, "alone_lanes": lanes, "large_root_cells": large,
            "requests_alone_lane": int(sum(base_cells(c) >= large for c in cases)),
            "simple_budget": budget, "converged": int(sum(r.converged for r in sat)),
            "one_at_a_time_plain_solves_per_s": round(n_requests / t_seq, 3),
            "served_capacity_solves_per_s": round(cap, 3),
            "served_vs_one_at_a_time_plain": round(t_seq / t_sat, 3),
            "poisson_load": 0.8, "poisson_solves_per_s": round(lat["solves_per_s"], 3),
            "p50_ms": round(lat["p50_ms"], 1), "p99_ms": round(lat["p99_ms"], 1),
            "batch_bytes_per_cell": round(server.bytes_per_cell, 1) if server.bytes_per_cell else None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=("both", "plain", "mixed"), default="both")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--cpu-steps", type=int, default=2, help="complete oracle PISO steps timed for the CPU baseline")
    ap.add_argument("--no-large", action="store_true", help="skip the 256^3 SpMV measurement")
    ap.add_argument("--no-config1", action="store_true", help="skip the 64x64 SIMPLE cavity measurement")
    ap.add_argument("--spmv-only", choices=("fp32", "fp64"), default=None,
                    help="only launch the SpMV kernel (for counter runs) and exit")
    ap.add_argument("--spmv-mesh", default="cavity",
                    help="mesh of --spmv-only: cavity (the octree cavity), uniformN (N^3) or bluff (21.0M cells)")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests in the serving measurement (0: skip)")
    ap.add_argument("--serve-slots", type=int, default=16, help="cases per batch in the serving measurement")
    ap.add_argument("--serve-only", action="store_true", help="only run the serving measurement")
    ap.add_argument("--serve-lanes", type=int, default=2, help="alone lanes for requests of >= 48^3 root cells")
    ap.add_argument("--serve-budget", type=int, default=30, help="SIMPLE iteration budget per served request")
    ap.add_argument("--bluff-steps", type=int, default=5,
                    help="timed PISO steps per mode of the 21.0M-cell bluff-body comparison (0: skip)")
    ap.add_argument("--bluff-only", action="store_true", help="only run the bluff-body measurement")
    ap.add_argument("--adjoint-steps", type=int, default=0,
                    help="only run the channel shape-gradient measurement over this many PIMPLE steps")
    ap.add_argument("--adjoint-tol", type=float, default=None,
                    help="relative solver tolerance of the shape-gradient measurement (default: the case's)")
    ap.add_argument("--adjoint-h", type=float, default=1e-4, help="central-difference step of that measurement")
    ap.add_argument("--inner-rtol", type=float, default=None,
                    help="refinement trigger of the mixed solve (default: the solver config's inner_rel_tolerance)")
    a = ap.parse_args()
    global INNER_RTOL
    INNER_RTOL = a.inner_rtol
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not part of this build", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    torch.cuda.set_device(0)
    if a.adjoint_steps > 0:
        print(json.dumps({"shape_adjoint": adjoint_measure(a.adjoint_steps, h=a.adjoint_h, tol=a.adjoint_tol)}),
              flush=True)
        return 0
    if a.bluff_only:
        print(json.dumps({"bluff_body": bluff_compare(a.bluff_steps or 10, a.mode)}), flush=True)
        return 0
    if a.serve_only:
        print(json.dumps({"serving": serve_measure(a.serve_requests, a.serve_slots, budget=a.serve_budget,
                                                    lanes=a.serve_lanes)}), flush=True)
        return 0
    if a.spmv_only:
        from gridfoam_amd.algorithms import create_case
        from gridfoam_amd.config import load_config

        if a.spmv_mesh.startswith("uniform"):
            grid = uniform_grid(int(a.spmv_mesh[len("uniform"):]))
        elif a.spmv_mesh == "bluff":
            import yaml

            from gridfoam_amd.config import CASES
            from gridfoam_amd.grid import create_grid

            grid = create_grid(load_config(yaml.safe_load((CASES / "bluff3d_piso.yaml").read_text())),
                               device="cuda:0")
        else:
            grid, _ = create_case(load_config(_cfg(False)), device="cuda:0")
        r = spmv_roofline(grid, laplacian_operator(grid), fp32=a.spmv_only == "fp32", launches=20)
        # calibration for the HBM counters: k_stream_copy reads 256 MiB and writes 256 MiB
        x = torch.ones(64 << 20, dtype=torch.float32, device="cuda:0")
        y = torch.empty_like(x)
        for _ in range(3):
            stream_copy(x, y)
        torch.cuda.synchronize()
        print(json.dumps({"spmv_only": r}), flush=True)
        return 0
    out = {}
    for mode in (("plain", "mixed") if a.mode == "both" else (a.mode,)):
        _log(f"{mode}: {a.warmup} warmup + {a.steps} timed steps")
        grid, alg, dt, iters = run_gpu(mode == "mixed", a.steps, a.warmup)
        from gridfoam_amd.algorithms import compute_continuity_error
        from gridfoam_amd.fv import fvc

        out[mode] = dict(grid=grid, alg=alg, dt=dt, iters=iters,
                         U=alg.U.data.detach().cpu().numpy(), p=alg.p.data.detach().cpu().numpy(),
                         gradp=fvc.grad(alg.p).cpu().numpy(),
                         continuity=compute_continuity_error(alg.phi, grid.dt)[0])
        _log(f"{mode}: {dt:.2f} s, {dt / a.steps * 1e3:.1f} ms/step, {iters / a.steps:.0f} p-CG iterations/step")
    head = "mixed" if "mixed" in out else "plain"
    rel = None
    if "mixed" in out and "plain" in out:
        P, M = out["plain"], out["mixed"]
        rel = {"U": float(np.linalg.norm(M["U"] - P["U"]) / np.linalg.norm(P["U"])),
               "p": float(np.linalg.norm(M["p"] - P["p"]) / np.linalg.norm(P["p"])),
               "grad_p": float(np.linalg.norm(M["gradp"] - P["gradp"]) / np.linalg.norm(P["gradp"]))}
        if max(rel.values()) > 1e-6:
            _log(f"mixed precision error {rel} exceeds 1e-6: headline falls back to plain FP64")
            head = "plain"
    H = out[head]
    grid = H["grid"]
    C = grid.num_cells
    value = C * a.steps / H["dt"] / 1e6
    _log("SpMV roofline on the cavity mesh")
    roof = spmv_roofline(grid, laplacian_operator(grid), fp32=(head == "mixed"))
    line = {
        "metric": "Mcell-steps/s at <=1e-6 rel err vs plain FP64",
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(H["dt"] / a.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-in-f64" if head == "mixed" else "f64",
        "data": "synthetic (lid-driven cavity initial state: U=0, p=0)",
        "config": {"workload": "3-D lid-driven cavity, octree 32^3 root + 2 wall levels, PISO 2 correctors, Re=1000",
                   "cells": C, "internal_faces": grid.num_internal_faces, "parallelism": "single GPU"},
        "roofline": roof,
    }
    for mode, key in (("plain", "plain_fp64"), ("mixed", "mixed")):
        if mode in out:
            line[key] = {"value": round(C * a.steps / out[mode]["dt"] / 1e6, 3),
                         "ms_per_step": round(out[mode]["dt"] / a.steps * 1e3, 3),
                         "p_cg_iters_per_step": round(out[mode]["iters"] / a.steps, 1),
                         "continuity_error_local_fp64": out[mode]["continuity"]}
    roof["copy_peak_GBs"] = copy_peak()
    if not a.no_large:
        _log("SpMV roofline on a uniform 256^3 mesh (working set above the 256 MB MALL)")
        big = uniform_grid(256)
        pm = laplacian_operator(big)
        line["roofline_256cubed"] = {"fp32": spmv_roofline(big, pm, fp32=True, launches=50),
                                     "fp64": spmv_roofline(big, pm, fp32=False, launches=50)}
        del big, pm
    if not a.no_config1:
        _log("2-D SIMPLE cavity 64x64, 500 iterations")
        line["cavity2d_simple_64"] = config1_measure(cpu=not a.no_cpu)
    if a.serve_requests > 0:
        _log(f"serving: {a.serve_requests} requests, batches of {a.serve_slots}")
        line["serving"] = serve_measure(a.serve_requests, a.serve_slots, budget=a.serve_budget, lanes=a.serve_lanes)
    if a.bluff_steps > 0:
        _log(f"bluff body: plain FP64 and mixed, {a.bluff_steps} steps each")
        line["bluff_body"] = bluff_compare(a.bluff_steps)
    if rel is not None:
        line["mixed_vs_fp64_rel_l2"] = rel
        line["speedup_mixed_vs_fp64"] = round(out["plain"]["dt"] / out["mixed"]["dt"], 3)
    if not a.no_cpu and a.cpu_steps > 0:
        _log(f"cpu baseline: FP64 oracle, {a.cpu_steps} complete PISO steps")
        line["cpu_baseline"] = cpu_baseline(a.cpu_steps)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
