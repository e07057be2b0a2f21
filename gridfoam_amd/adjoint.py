"""Checkpointed transient rollouts for gradients through many time steps.

A transient run whose step is differentiable -- native assembly
(``gf_assemble`` / ``gf_assemble_vjp``) plus a Krylov solve (transpose solve
as its backward) -- keeps the autograd graph of every step by default, which
grows with the step count.  ``checkpointed_rollout`` keeps only the state at
segment boundaries and recomputes each segment's steps in the backward pass
(the steps are deterministic: fixed reduction orders in every kernel), so the
memory held is one segment's graph plus the boundary states.

``ScalarTransport`` is such a step: implicit (Euler or BDF2) transport of a
scalar by a given face flux with a diffusivity, ``ddt(q) + div(phi, q) -
laplacian(gamma, q) = 0`` (fv/fvm/ddt.py, fv/fvm/div.py, fv/fvm/laplacian.py),
differentiable with respect to the initial field, the diffusivity, the flux
and the boundary values.
"""

from __future__ import annotations

from typing import Callable, Optional, Sequence, Tuple, Union

import torch
from torch.utils.checkpoint import checkpoint

from .config import DivScheme, SolverConfig
from .fields import CellField, FaceField
from .fv import fvm
from .fvmatrix import Equation
from .solvers import LinearSolver, create_solver

State = Tuple[torch.Tensor, ...]


def checkpointed_rollout(step: Callable[[State], State], state: State, n_steps: int,
                         every: Optional[int] = None) -> State:
    """Advance ``state`` by ``n_steps`` calls of ``step``.  With ``every``,
    the steps run in segments of that many under activation checkpointing:
    the backward pass recomputes a segment from its saved input state."""
    if n_steps < 0:
        raise ValueError("n_steps must be >= 0")
    if every is not None and every < 1:
        raise ValueError("every must be >= 1")

    def segment(k: int):
        def run(*s):
            out = tuple(s)
            for _ in range(k):
                out = step(out)
            return out
        return run

    done = 0
    while done < n_steps:
        k = n_steps - done if every is None else min(every, n_steps - done)
        if every is None or not torch.is_grad_enabled():
            state = segment(k)(*state)
        else:
            state = checkpoint(segment(k), *state, use_reentrant=False)
        done += k
    return state


class ScalarTransport:
    """One implicit transport step of a scalar field ``q`` (with its boundary
    conditions) by the face flux ``phi``; the state is ``(q^n,)`` for Euler
    and ``(q^n, q^{n-1})`` for BDF2 (constant time step).  The explicit parts
    (hanging-face corrections) use the previous level, as the reference's
    deferred corrections do.  The step sets ``q``'s data and history, so
    after a backward pass that recomputed segments ``q`` holds the last
    recomputed step's values."""

    def __init__(self, q: CellField, phi: FaceField, gamma: Union[float, torch.Tensor], *,
                 scheme: DivScheme = DivScheme.UPWIND, bdf2: bool = False,
                 solver: Optional[LinearSolver] = None):
        if q.component_shape != ():
            raise ValueError("ScalarTransport advances a scalar field")
        self.q, self.phi, self.gamma = q, phi, gamma
        self.scheme, self.bdf2 = scheme, bdf2
        self.solver = solver if solver is not None else create_solver(SolverConfig(
            method="bicgstab", preconditioner="jacobi", tolerance=0.0, rel_tolerance=1e-13, max_iter=2000))

    def __call__(self, state: State) -> State:
        q = self.q
        old = state[0]
        older = state[1] if self.bdf2 else None
        abc = (1.5, 2.0, 0.5) if self.bdf2 else (1.0, 1.0, 0.0)
        q.old_data, q.older_data = old, older
        q.data = old
        m = fvm.assemble(q, ddt=1.0, div=1.0, lap=-1.0, phi=self.phi, gamma=self.gamma, abc=abc,
                         scheme=self.scheme)
        new = self.solver.solve(Equation(q, m)).solution
        return (new, old) if self.bdf2 else (new,)


def rollout_scalar_transport(q: CellField, phi: FaceField, gamma, q0: torch.Tensor, n_steps: int, *,
                             every: Optional[int] = None, bdf2: bool = False,
                             scheme: DivScheme = DivScheme.UPWIND,
                             solver: Optional[LinearSolver] = None) -> torch.Tensor:
    """q after ``n_steps`` implicit steps from ``q0`` (BDF2 starts from
    ``q^{-1} = q0``), checkpointed in segments of ``every`` steps."""
    step = ScalarTransport(q, phi, gamma, scheme=scheme, bdf2=bdf2, solver=solver)
    state: Sequence[torch.Tensor] = (q0, q0) if bdf2 else (q0,)
    return checkpointed_rollout(step, tuple(state), n_steps, every)[0]
