"""Per-step CSV diagnostics of a run: ``continuity_error.csv`` and
``solver_info.csv``.

The file names, headers, column order and cadence are the reference's output
format (post/diagnostics/collector.py:74-93 for the headers, :190-264 for
which steps produce a row): a file gets a row at every step that is a
multiple of its ``writeInterval`` and at the run's last step; each row is
flushed as it is written, so a file can be read while the run goes on.

Design: each enabled file is an independent ``_Stream`` -- an open CSV file,
its interval and a function that turns one step's data into one row.  The
collector owns a list of streams and fans every recorded step out to them;
adding another diagnostic is one more stream.  The continuity errors
(post/diagnostics/continuity.py:12-49) come from a callable that evaluates
them in FP64 on the GPU (``algorithms.compute_continuity_error``).
"""

from __future__ import annotations

import csv
from pathlib import Path
from typing import Callable, Dict, List, Optional, Sequence, Tuple

from .config import PostProcessingConfig
from .solvers import SolveStats

_STAT_SUFFIXES = ("solver", "initial", "final", "iters", "converged")
_VECTOR_FIELDS = {"U": 3}  # component counts fixed by the algorithms


def solver_info_layout(fields: Sequence[str], components: Callable[[str], int]) -> List[Tuple[str, str, int]]:
    """(column prefix, field name, component) for every statistics group:
    ``p`` for a scalar field, ``U_0``..``U_2`` for a vector field."""
    out: List[Tuple[str, str, int]] = []
    for name in fields:
        n = components(name)
        out.extend([(name, name, 0)] if n == 1 else [(f"{name}_{k}", name, k) for k in range(n)])
    return out


class _Stream:
    """One CSV file: header written and flushed on creation, then one flushed
    row per due step."""

    def __init__(self, path: Path, header: Sequence[str], interval: int, last_step: int,
                 make_row: Callable[[int, object, Dict[str, Tuple[SolveStats, ...]]], list]):
        self.interval = max(1, int(interval))
        self.last_step = last_step
        self.make_row = make_row
        self.fh = path.open("w", newline="")
        self.out = csv.writer(self.fh)
        self.emit(list(header))

    def due(self, step: int) -> bool:
        return step == self.last_step or step % self.interval == 0

    def emit(self, row: list) -> None:
        self.out.writerow(row)
        self.fh.flush()

    def close(self) -> None:
        self.fh.flush()
        self.fh.close()


class DiagnosticsCollector:
    """Writes the enabled diagnostics of a run as CSV, one row per due step."""

    def __init__(self, *, grid, config: PostProcessingConfig, output_dir: Path, n_steps: int, delta_t: float,
                 phase: Optional[str] = None,
                 continuity: Optional[Callable[[object, float], Tuple[float, float]]] = None) -> None:
        if continuity is None:
            from .algorithms import compute_continuity_error as continuity
        self.grid = grid
        self.phase = phase
        self.dt = float(delta_t)
        self.n_steps = int(n_steps)
        self.continuity = continuity
        self.mass_drift = 0.0  # running sum of the global continuity error
        self.streams: List[_Stream] = []
        out = Path(output_dir)
        out.mkdir(parents=True, exist_ok=True)
        if config.continuityError is not None:
            self.streams.append(_Stream(out / "continuity_error.csv", ("time", "local", "global", "cumulative"),
                                        config.continuityError.writeInterval, self.n_steps, self._continuity_row))
        if config.solverInfo is not None:
            fields = config.solverInfo.fields if config.solverInfo.fields is not None else ["U", "p"]
            self.groups = solver_info_layout(fields, self._components)
            header = ["time"] + [f"{pre}_{suf}" for pre, _, _ in self.groups for suf in _STAT_SUFFIXES]
            self.streams.append(_Stream(out / "solver_info.csv", header, config.solverInfo.writeInterval,
                                        self.n_steps, self._solver_row))

    @classmethod
    def from_config(cls, grid, post_processing: Optional[PostProcessingConfig], output_dir: Path, *, n_steps: int,
                    delta_t: float, phase: Optional[str] = None, **kw) -> Optional["DiagnosticsCollector"]:
        """A collector when at least one diagnostic is configured, else None."""
        pp = post_processing
        if pp is None or (pp.continuityError is None and pp.solverInfo is None):
            return None
        return cls(grid=grid, config=pp, output_dir=output_dir, n_steps=n_steps, delta_t=delta_t, phase=phase, **kw)

    def _components(self, name: str) -> int:
        if name in _VECTOR_FIELDS:
            return _VECTOR_FIELDS[name]
        sim = getattr(self.grid, "sim_config", None)
        cond = sim.conditions.get(name) if sim is not None else None
        if cond is not None and isinstance(cond.internal, (list, tuple)):
            return len(cond.internal)
        return 1

    def _continuity_row(self, step: int, phi, stats) -> list:
        local, glob = self.continuity(phi, self.dt)
        self.mass_drift += glob
        return [step * self.dt, local, glob, self.mass_drift]

    def _solver_row(self, step: int, phi, stats) -> list:
        row: list = [step * self.dt]
        for _, name, comp in self.groups:
            per_comp = stats.get(name)
            if per_comp is None or comp >= len(per_comp):
                row.extend([""] * len(_STAT_SUFFIXES))
            else:
                s = per_comp[comp]
                row.extend([s.solver, s.initial_residual, s.final_residual, s.iterations, int(s.converged)])
        return row

    def record_step(self, step: int, phi, solve_stats: Dict[str, Tuple[SolveStats, ...]]) -> None:
        """Write the rows of 1-based step ``step`` to every file it is due in."""
        for s in self.streams:
            if s.due(step):
                s.emit(s.make_row(step, phi, solve_stats))

    def close(self) -> None:
        """Flush and close every file; calling it again does nothing."""
        streams, self.streams = self.streams, []
        for s in streams:
            s.close()

    def __enter__(self) -> "DiagnosticsCollector":
        return self

    def __exit__(self, *exc) -> None:
        self.close()
