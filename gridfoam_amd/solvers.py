"""Linear solvers behind the reference's plugin API (src/gridfoam/solvers/).

``LinearSolver.solve(eq)`` keeps the reference contract (base.py:70-137):
primal solve without autograd, then an implicit adjoint attached with an
``autograd.Function`` whose backward solves ``A^T lambda = x_bar`` with the
same solver and assembles LDU/source gradients (adjoint/attach.py:21-115,
adjoint/ldu_grads.py:11-53).  Both the primal and the transpose solve, and
the gradient assembly, run on the HIP library; the components of a vector
field are solved in one batched launch sequence.
"""

from __future__ import annotations

import ctypes
import logging
import os
from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import torch

from . import _native as nat
from .config import NormType, PreconditionerType, SolverConfig, SolverType
from .fvmatrix import Equation, FvMatrix

logger = logging.getLogger(__name__)

# Iterations queued between two polls of the device state (one launch graph
# per chunk); GRIDFOAM_CHECK_EVERY_CG / _CG_MIXED / _BICGSTAB override them
# for tuning.  The mixed-precision chunk also carries the FP64 refinement
# launches, so a longer one queues fewer of them (on the 1.1M-cell cavity:
# 59.8 ms per PISO step at 16, 57.9 at 24 and 32; the plain step is 80.7 ms
# at 16, 24 and 32).
_CHECK_EVERY_CG = int(os.environ.get("GRIDFOAM_CHECK_EVERY_CG", "16"))
_CHECK_EVERY_CG_MIXED = int(os.environ.get("GRIDFOAM_CHECK_EVERY_CG_MIXED", "32"))
_CHECK_EVERY_BICGSTAB = int(os.environ.get("GRIDFOAM_CHECK_EVERY_BICGSTAB", "4"))


@dataclass(frozen=True)
class SolveStats:  # solvers/base.py:15-47
    solver: str
    initial_residual: float
    final_residual: float
    iterations: int
    converged: bool


@dataclass(frozen=True)
class SolveResult:  # solvers/base.py:50-67
    solution: torch.Tensor
    stats: Tuple[SolveStats, ...]


def residual_threshold(atol: float, rtol: float, ref_norm: torch.Tensor) -> torch.Tensor:
    """solvers/base.py:140-162 -- max(atol, rtol * ref_norm)."""
    return torch.maximum(torch.full_like(ref_norm, atol), rtol * ref_norm)


class _Workspace:
    """Krylov workspace per (grid, nrhs), allocated once and reused.  The
    library caches launch graphs per workspace; they are released with it."""

    def __init__(self):
        self._bufs: Dict[tuple, torch.Tensor] = {}

    def get(self, grid, nrhs: int, cfg: nat.GfKrylovCfg) -> torch.Tensor:
        nbytes = int(nat.lib().gf_krylov_workspace_bytes(grid.mesh_ref, nrhs, ctypes.byref(cfg)))
        if nbytes < 0:
            raise RuntimeError("gf_krylov_workspace_bytes failed")
        key = (id(grid), nrhs)
        buf = self._bufs.get(key)
        if buf is None or buf.numel() < nbytes:
            if buf is not None:
                _release(buf)
            buf = torch.empty(nbytes, dtype=torch.uint8, device=grid.device)
            self._bufs[key] = buf
        return buf

    def __del__(self):
        for buf in self._bufs.values():
            _release(buf)


def _release(buf: torch.Tensor) -> None:
    try:
        nat.lib().gf_krylov_release(ctypes.c_void_p(buf.data_ptr()))
    except Exception:  # interpreter shutdown: the library may already be gone
        pass


class LinearSolver:
    """solvers/base.py:70-137."""

    atol: float
    rtol: float
    grad_mode: str = "adjoint"

    def __init__(self, config: SolverConfig):
        self.config = config
        self.solver_name = config.method.value
        self.precon_type = config.preconditioner
        self.atol = config.tolerance
        self.rtol = config.rel_tolerance
        self.max_iter = config.max_iter
        self.norm_order = config.norm_type.to_norm_order()
        self.max_restart = config.max_restart
        self.log_interval = config.log_interval
        self._ws = _Workspace()

    # -- native configuration --------------------------------------------------
    def _cfg(self, mixed: Optional[bool] = None) -> nat.GfKrylovCfg:
        c = self.config
        k = nat.GfKrylovCfg()
        k.method = 0 if c.method == SolverType.CG else 1
        k.precond = 1 if c.preconditioner == PreconditionerType.JACOBI else 0
        k.norm = 1 if c.norm_type == NormType.L_inf else 0
        k.max_iter = c.max_iter
        k.max_restart = c.max_restart
        k.atol = c.tolerance
        k.rtol = c.rel_tolerance
        k.mixed = int(c.mixed_precision if mixed is None else mixed)
        k.inner_rtol = c.inner_rel_tolerance
        if k.method == 1:
            k.check_every = _CHECK_EVERY_BICGSTAB
        else:
            k.check_every = _CHECK_EVERY_CG_MIXED if k.mixed else _CHECK_EVERY_CG
        return k

    def _native_solve(self, A: FvMatrix, b: torch.Tensor, x0: torch.Tensor, transpose: bool,
                      mixed: Optional[bool] = None) -> Tuple[torch.Tensor, Tuple[SolveStats, ...]]:
        g = A.grid
        nrhs = 1 if b.dim() == 1 else b.shape[1]
        if nrhs > 1:
            mixed = False  # the FP32 inner solve takes one right-hand side; components solve in FP64
        cfg = self._cfg(mixed)
        # a symmetric operator (upper == lower, e.g. the pressure Laplacian)
        # lets a one-rhs PCG run on the brick layout of the solver row order
        cfg.symmetric = int(cfg.method == 0 and nrhs == 1 and torch.equal(A._upper, A._lower))
        ws = self._ws.get(g, nrhs, cfg)
        x = x0.detach().clone().contiguous()
        bc = b.detach().contiguous()
        stats = (ctypes.c_double * (4 * nrhs))()
        nat.call("gf_krylov_solve", g.mesh_ref, ctypes.byref(cfg), ctypes.c_int32(nrhs), ctypes.c_int32(int(transpose)),
                 nat.ptr(A._diag), nat.ptr(A._upper), nat.ptr(A._lower), nat.ptr(bc),
                 nat.ptr(x), nat.ptr(ws), stats, nat.stream())
        out = []
        for r in range(nrhs):
            s = SolveStats(self.solver_name, float(stats[4 * r]), float(stats[4 * r + 1]),
                           int(stats[4 * r + 2]), bool(stats[4 * r + 3]))
            if not s.converged:
                logger.warning("%s reached max_iter=%d without convergence", self.solver_name, self.max_iter)
            out.append(s)
        return x, tuple(out)

    # -- reference API -----------------------------------------------------------
    def solve(self, eq: Equation) -> SolveResult:
        if self.grad_mode == "unrolled":
            raise NotImplementedError("unrolled differentiation is not supported by the native solver")
        A = eq.fv_matrix
        with torch.no_grad():
            result = self._solve_primal(eq)
        solution = attach_implicit_adjoint(A, result.solution, solve_transpose=self.solve_transpose)
        return SolveResult(solution=solution, stats=result.stats)

    def _solve_primal(self, eq: Equation) -> SolveResult:
        A = eq.fv_matrix
        x, stats = self._native_solve(A, A.source, eq.target.data, transpose=False)
        return SolveResult(solution=x, stats=stats)

    def solve_transpose(self, A: FvMatrix, rhs: torch.Tensor) -> torch.Tensor:
        """Zero initial guess, same tolerances (cg.py:61-73).  ``A`` is the
        primal matrix; the transpose is applied by swapping upper/lower
        while packing, so no transposed copy is made."""
        x, _ = self._native_solve(A, rhs, torch.zeros_like(rhs), transpose=True)
        return x


class CGSolver(LinearSolver):
    """solvers/cg.py:17-174 (Jacobi-preconditioned conjugate gradient)."""


class BiCGSTABSolver(LinearSolver):
    """solvers/bicgstab.py:23-272 (right-preconditioned, with restarts)."""


def create_solver(config: SolverConfig) -> LinearSolver:
    """solvers/factory.py:9-36."""
    if config.method == SolverType.CG:
        return CGSolver(config)
    if config.method == SolverType.BiCGSTAB:
        if config.mixed_precision:
            raise ValueError("mixed_precision is implemented for the cg pressure solve only")
        return BiCGSTABSolver(config)
    raise ValueError(f"Unknown solver method: {config.method}")


def resolve_solver(solvers: Dict[str, LinearSolver], field_name: str, *, is_final: bool = False) -> LinearSolver:
    """solvers/resolver.py:46-80 -- ``<name>Final`` fallback."""
    if is_final:
        for key in (f"{field_name}Final", f"{field_name}_final"):
            if key in solvers:
                return solvers[key]
    if field_name not in solvers:
        raise KeyError(f"Solver for {field_name!r} not found")
    return solvers[field_name]


# ---------------------------------------------------------------------------
# Implicit adjoint (solvers/adjoint/attach.py:21-115, ldu_grads.py:11-53)
# ---------------------------------------------------------------------------
class _ImplicitAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, diag, upper, lower, source, sol, A, solve_transpose):  # noqa: D401
        ctx.A = A
        ctx.solve_transpose = solve_transpose
        ctx.save_for_backward(sol)
        return sol.clone()

    @staticmethod
    def backward(ctx, grad_out):
        (x,) = ctx.saved_tensors
        A: FvMatrix = ctx.A
        lam = ctx.solve_transpose(A, grad_out.detach().contiguous())
        g = A.grid
        nc = 1 if x.dim() == 1 else x.shape[1]
        gd = torch.empty_like(A._diag)
        gu = torch.empty_like(A._upper)
        gl = torch.empty_like(A._lower)
        lam_c, x_c = lam.contiguous(), x.contiguous()
        nat.call("gf_ldu_grads", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(lam_c), nat.ptr(x_c),
                 nat.ptr(gd), nat.ptr(gu), nat.ptr(gl), nat.stream())
        F = g.num_internal_faces
        return gd, gu[:F], gl[:F], lam, None, None, None


def attach_implicit_adjoint(A: FvMatrix, solution: torch.Tensor, *, solve_transpose) -> torch.Tensor:
    needs = any(t.requires_grad for t in (A._diag, A._upper, A._lower, A._source))
    if not needs:
        return solution
    return _ImplicitAdjoint.apply(A._diag, A.upper, A.lower, A._source, solution, A, solve_transpose)
