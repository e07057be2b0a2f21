"""Continuous-batching solve server: many cases in flight on one GPU.

Requests (a case configuration each, e.g. a design-optimisation evaluation)
arrive over time.  A request is admitted as soon as a slot is free and its
estimated device footprint fits the HBM budget; while admitted it keeps its
grid, fields and operators resident on the device and runs on a stream of
its own, so the small, latency-bound kernels of different cases overlap on
the GPU.  Requests that arrive while others run are admitted as slots and
memory free up: there is no batch boundary.  The kernels are deterministic
and never read another case's memory, so a result is bit-identical whether
its case ran alone or in the mix.

Slots are threads with a HIP stream each by default (ctypes releases the
interpreter lock inside every native call, including the Krylov solves'
device-state polls); worker processes (one interpreter and HIP context per
slot) are the alternative.  Streams keep the default priority.
"""

from __future__ import annotations

import copy
import math
import multiprocessing as mp
import queue as queue_mod
import threading
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from .algorithms import create_case
from .config import CASES, load_config

BYTES_PER_CELL = 4096  # device footprint of one cell of a SIMPLE case (grid, fields, operators, workspace)


@dataclass
class Request:
    rid: int
    config: dict
    arrival: float = 0.0  # seconds after the start of the run
    iterations: int = 5   # SIMPLE outer iterations


@dataclass
class Result:
    rid: int
    cells: int
    U: torch.Tensor
    p: torch.Tensor
    p_iterations: int
    arrival: float
    start: float
    end: float

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def estimate_cells(cfg: dict) -> int:
    """Upper estimate of the leaf count: base cells plus 8^level - 1 extra cells
    for every base cell that a refinement box touches."""
    f = cfg["fluxel"]
    nx, ny, nz = f["root_resolution"]
    lo, hi = np.asarray(f["domain"]["lower"], float), np.asarray(f["domain"]["upper"], float)
    h = (hi - lo) / np.asarray([nx, ny, nz], float)
    n = nx * ny * nz
    extra = 0
    for r in f.get("refinement_regions", []) or []:
        span = np.clip(np.asarray(r["max"]) - np.asarray(r["min"]), 0, None)
        touched = np.prod(np.minimum(np.ceil(span / h) + 1, [nx, ny, nz]))
        extra += int(touched) * (8 ** int(r["level"]) - 1)
    return int(n + extra)


def footprint_bytes(cfg: dict) -> int:
    return estimate_cells(cfg) * BYTES_PER_CELL


def random_case(rng: np.random.Generator, *, n_range=(16, 96), mixed: bool = True, residual: float = 1e-6,
                max_iterations: int = 2000, obstacles: bool = True) -> dict:
    """A serving request: a lid-driven cube on a base grid n^3 with n
    log-uniform in n_range, 1-4 axis-aligned box obstacles (side U(0.05,
    0.2), centre uniform in the interior) with no-slip walls, 0-2 refinement
    levels around them, Re U(100, 1000), steady SIMPLE until every
    residualControl entry (U, p) is below ``residual`` or ``max_iterations``
    outer iterations ran; FP32-in-FP64 pressure solves when ``mixed``."""
    cfg = copy.deepcopy(load_config_dict("cavity2d_simple.yaml"))
    n = int(math.floor(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
    n = max(n_range[0], min(n_range[1], n))
    f = cfg["fluxel"]
    f["root_resolution"] = [n, n, n]
    f["domain"] = {"lower": [0.0, 0.0, 0.0], "upper": [1.0, 1.0, 1.0]}
    level = int(rng.integers(0, 3))
    regions, bodies = [], []
    for k in range(int(rng.integers(1, 5))):
        side = rng.uniform(0.05, 0.2)
        c = rng.uniform(side / 2, 1 - side / 2, size=3)
        lo, hi = [float(v) for v in c - side / 2], [float(v) for v in c + side / 2]
        if level > 0:
            regions.append({"name": f"box{k}", "min": lo, "max": hi, "level": level})
        if obstacles:
            bodies.append({"name": f"box{k}", "kind": "box", "min": lo, "max": hi})
    f["refinement_regions"] = regions
    f["immersed_bodies"] = bodies
    sim = cfg["simulator"]
    sim["properties"]["transport"]["nu"] = 1.0 / rng.uniform(100.0, 1000.0)
    sim["control"]["endTime"] = float(max_iterations)
    cond = sim["conditions"]
    cond["U"]["boundary"]["wall"]["patches"] = ["x_minus", "x_plus", "y_minus", "z_minus", "z_plus"]
    cond["U"]["boundary"].pop("empty", None)
    cond["p"]["boundary"]["wall"]["patches"] = ["x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus"]
    cond["p"]["boundary"].pop("empty", None)
    names = [b["name"] for b in bodies]
    if names:
        cond["U"]["boundary"]["obstacles"] = {"type": "dirichlet", "patches": names, "value": [0.0, 0.0, 0.0]}
        cond["p"]["boundary"]["obstacles"] = {"type": "neumann", "patches": names, "value": 0.0}
    alg = sim["fvSolution"]["algorithm"]
    alg["residualControl"] = {"U": residual, "p": residual}
    sol = sim["fvSolution"]["solvers"]
    sol["p"]["mixed_precision"] = bool(mixed)
    sol["p"]["rel_tolerance"] = 1e-6
    return cfg


def load_config_dict(name: str) -> dict:
    import yaml

    return yaml.safe_load((CASES / name).read_text())


def run_case(cfg: dict, iterations: int) -> tuple:
    """Build the case on the current stream, run its SIMPLE iterations and
    return (cells, U, p, pressure iterations) with U, p on the host."""
    grid, alg = create_case(load_config(cfg), device="cuda:0")
    p_its = 0
    for _ in range(iterations):
        alg.step()
        p_its += sum(s.iterations for s in alg.solve_stats["p"])
    U, p = alg.U.data.to("cpu"), alg.p.data.to("cpu")  # synchronises this stream only
    return grid.num_cells, U, p, p_its


def solve_case(cfg: dict) -> tuple:
    """One request alone, as the SIMPLE driver runs it (simple.py:241-275):
    iterations until every residualControl entry is met or the budget
    endTime / deltaT is spent.  Returns (cells, U, p, iterations, converged,
    pressure iterations) with U, p on the host."""
    config = load_config(cfg)
    grid, alg = create_case(config, device="cuda:0")
    ctl = config.simulator.control
    budget = max(1, int(round(ctl.endTime / ctl.deltaT)))
    its, converged, p_its = 0, False, 0
    while its < budget:
        alg.step()
        its += 1
        p_its += sum(s.iterations for s in alg.solve_stats["p"])
        if alg.has_simulation_converged():
            converged = True
            break
    return grid.num_cells, alg.U.data.to("cpu"), alg.p.data.to("cpu"), its, converged, p_its


def _worker_main(wid: int, req_q, res_q) -> None:
    """Slot process: owns its HIP context; runs one request at a time."""
    torch.cuda.set_device(0)
    warm = random_case(np.random.default_rng(0), n_range=(8, 8))
    run_case(warm, 1)
    res_q.put(("ready", wid))
    while True:
        item = req_q.get()
        if item is None:
            break
        rid, cfg, iterations = item
        start = time.time()
        try:
            cells, U, p, its = run_case(cfg, iterations)
        except BaseException as exc:  # report and keep serving
            res_q.put(("error", wid, rid, repr(exc)))
            continue
        res_q.put(("done", wid, rid, cells, U.numpy(), p.numpy(), its, start, time.time()))


class SolveServer:
    """Admits requests by free slot and HBM footprint; each admitted case runs
    in a slot of its own (a thread with its own stream by default, or a
    worker process)."""

    def __init__(self, slots: int = 4, hbm_budget: Optional[int] = None, mode: str = "thread"):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        if mode not in ("process", "thread"):
            raise ValueError("mode must be 'process' or 'thread'")
        self.slots = slots
        self.mode = mode
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(0)
            hbm_budget = int(0.6 * free)
        self.hbm_budget = int(hbm_budget)
        self._streams = [torch.cuda.Stream(device=0) for _ in range(slots)] if mode == "thread" else []
        self._procs: List = []
        if mode == "process":
            ctx = mp.get_context("spawn")
            self._res_q = ctx.Queue()
            self._req_qs = [ctx.Queue() for _ in range(slots)]
            self._procs = [ctx.Process(target=_worker_main, args=(w, self._req_qs[w], self._res_q), daemon=True)
                           for w in range(slots)]
            for pr in self._procs:
                pr.start()
            ready = 0
            while ready < slots:
                msg = self._res_q.get(timeout=600)
                if msg[0] == "ready":
                    ready += 1

    def close(self) -> None:
        for q in getattr(self, "_req_qs", []):
            q.put(None)
        for pr in self._procs:
            pr.join(timeout=60)
        self._procs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def run(self, requests: Sequence[Request]) -> List[Result]:
        """Serve requests that arrive at their ``arrival`` offsets (seconds
        after this call); returns the results in completion order."""
        if self.mode == "process":
            return self._run_processes(requests)
        return self._run_threads(requests)

    def _run_processes(self, requests: Sequence[Request]) -> List[Result]:
        order = sorted(requests, key=lambda r: (r.arrival, r.rid))
        t0 = time.time()
        pending: List[Request] = []
        free = list(range(self.slots))
        in_use, running, nxt = 0, {}, 0
        results: List[Result] = []
        by_id = {r.rid: r for r in order}
        while len(results) < len(order):
            now = time.time() - t0
            while nxt < len(order) and order[nxt].arrival <= now:
                pending.append(order[nxt])
                nxt += 1
            while free and pending:
                pick = next((r for r in pending
                             if not running or in_use + footprint_bytes(r.config) <= self.hbm_budget), None)
                if pick is None:
                    break
                pending.remove(pick)
                w = free.pop()
                need = footprint_bytes(pick.config)
                in_use += need
                running[pick.rid] = (w, need)
                self._req_qs[w].put((pick.rid, pick.config, pick.iterations))
            wait = 0.05 if nxt >= len(order) else max(0.0, min(0.05, order[nxt].arrival - (time.time() - t0)))
            try:
                msg = self._res_q.get(timeout=wait) if wait > 0 else self._res_q.get_nowait()
            except queue_mod.Empty:
                continue
            if msg[0] == "error":
                raise RuntimeError(f"request {msg[2]} failed in slot {msg[1]}: {msg[3]}")
            _, w, rid, cells, U, p, its, start, end = msg
            _, need = running.pop(rid)
            in_use -= need
            free.append(w)
            req = by_id[rid]
            results.append(Result(rid, cells, torch.from_numpy(U), torch.from_numpy(p), its, req.arrival,
                                  start - t0, end - t0))
        return results

    def _run_threads(self, requests: Sequence[Request]) -> List[Result]:
        """Serve requests that arrive at their ``arrival`` offsets (seconds
        after this call); returns the results in completion order."""
        order = sorted(requests, key=lambda r: (r.arrival, r.rid))
        lock = threading.Condition()
        pending: List[Request] = []
        in_use = [0]
        running = [0]
        arrived = [0]
        results: List[Result] = []
        errors: List[BaseException] = []
        t0 = time.perf_counter()

        def dispatcher():
            for r in order:
                delay = r.arrival - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
                with lock:
                    pending.append(r)
                    arrived[0] += 1
                    lock.notify_all()

        def admissible(r: Request) -> bool:
            need = footprint_bytes(r.config)
            return running[0] == 0 or in_use[0] + need <= self.hbm_budget

        def worker(k: int):
            torch.cuda.set_device(0)
            stream = self._streams[k]
            while True:
                with lock:
                    while True:
                        if errors:
                            return
                        pick = next((r for r in pending if admissible(r)), None)
                        if pick is not None:
                            pending.remove(pick)
                            need = footprint_bytes(pick.config)
                            in_use[0] += need
                            running[0] += 1
                            break
                        if arrived[0] == len(order) and not pending:
                            return
                        lock.wait(0.05)
                start = time.perf_counter() - t0
                try:
                    with torch.cuda.stream(stream):
                        cells, U, p, p_its = run_case(pick.config, pick.iterations)
                except BaseException as exc:  # surface the first failure to the caller
                    with lock:
                        errors.append(exc)
                        lock.notify_all()
                    return
                end = time.perf_counter() - t0
                res = Result(pick.rid, cells, U, p, p_its, pick.arrival, start, end)
                with lock:
                    in_use[0] -= need
                    running[0] -= 1
                    results.append(res)
                    lock.notify_all()

        threads = [threading.Thread(target=dispatcher, daemon=True)]
        threads += [threading.Thread(target=worker, args=(k,), daemon=True) for k in range(self.slots)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        torch.cuda.synchronize()
        return results


def latency_stats(results: Sequence[Result]) -> Dict[str, float]:
    lat = np.sort(np.asarray([r.latency for r in results]))
    span = max(r.end for r in results) - min(r.arrival for r in results)
    return {"requests": len(results), "solves_per_s": len(results) / span,
            "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99))}

