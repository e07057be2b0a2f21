"""Streaming CSV diagnostics: continuity_error.csv and solver_info.csv.

Restates src/gridfoam/post/diagnostics/collector.py:27-312 (same files,
headers, row layout, write intervals and flush-per-row behaviour).  The
continuity errors (post/diagnostics/continuity.py:12-49) are computed in FP64
on the GPU by ``algorithms.compute_continuity_error``.
"""

from __future__ import annotations

import csv
from dataclasses import dataclass
from pathlib import Path
from types import TracebackType
from typing import Callable, Dict, List, Optional, TextIO, Tuple

from .config import PostProcessingConfig
from .solvers import SolveStats

# Fields whose component count is fixed by the algorithms (velocity is a 3-vector).
_KNOWN_COMPONENTS = {"U": 3}


@dataclass(frozen=True)
class SolverInfoColumn:
    """One solver_info column group (collector.py:21-27)."""

    csv_key: str
    field_name: str
    component_index: int


class DiagnosticsCollector:
    """Append one row per recorded step to each enabled CSV file and flush it
    at once, so memory use stays constant (collector.py:30-120)."""

    def __init__(self, *, grid, config: PostProcessingConfig, output_dir: Path, n_steps: int, delta_t: float,
                 phase: Optional[str] = None,
                 continuity: Optional[Callable[[object, float], Tuple[float, float]]] = None) -> None:
        self._grid = grid
        self._config = config
        self._output_dir = Path(output_dir)
        self._n_steps = int(n_steps)
        self._delta_t = float(delta_t)
        self._phase = phase
        self._cumulative_continuity = 0.0
        self._closed = False
        if continuity is None:
            from .algorithms import compute_continuity_error as continuity
        self._continuity = continuity
        self._continuity_file: Optional[TextIO] = None
        self._continuity_writer = None
        self._solver_info_file: Optional[TextIO] = None
        self._solver_info_writer = None
        self._solver_info_columns: List[SolverInfoColumn] = []

        self._output_dir.mkdir(parents=True, exist_ok=True)
        if config.continuityError is not None:
            self._continuity_file = (self._output_dir / "continuity_error.csv").open("w", newline="")
            self._continuity_writer = csv.writer(self._continuity_file)
            self._continuity_writer.writerow(["time", "local", "global", "cumulative"])
            self._continuity_file.flush()
        if config.solverInfo is not None:
            self._solver_info_columns = self._build_solver_info_columns(config)
            self._solver_info_file = (self._output_dir / "solver_info.csv").open("w", newline="")
            self._solver_info_writer = csv.writer(self._solver_info_file)
            header = ["time"]
            for c in self._solver_info_columns:
                header += [f"{c.csv_key}_solver", f"{c.csv_key}_initial", f"{c.csv_key}_final",
                           f"{c.csv_key}_iters", f"{c.csv_key}_converged"]
            self._solver_info_writer.writerow(header)
            self._solver_info_file.flush()

    @classmethod
    def from_config(cls, grid, post_processing: Optional[PostProcessingConfig], output_dir: Path, *, n_steps: int,
                    delta_t: float, phase: Optional[str] = None, **kw) -> Optional["DiagnosticsCollector"]:
        """collector.py:101-142: None unless a diagnostic is enabled."""
        if post_processing is None:
            return None
        if post_processing.continuityError is None and post_processing.solverInfo is None:
            return None
        return cls(grid=grid, config=post_processing, output_dir=output_dir, n_steps=n_steps, delta_t=delta_t,
                   phase=phase, **kw)

    # ---- columns (collector.py:144-186)
    def _resolve_solver_info_fields(self, config: PostProcessingConfig) -> List[str]:
        if config.solverInfo.fields is not None:
            return list(config.solverInfo.fields)
        return ["U", "p"]

    def _resolve_field_num_components(self, field_name: str) -> int:
        if field_name in _KNOWN_COMPONENTS:
            return _KNOWN_COMPONENTS[field_name]
        sim = getattr(self._grid, "sim_config", None)
        cond = sim.conditions.get(field_name) if sim is not None else None
        if cond is not None:
            return len(cond.internal) if isinstance(cond.internal, (list, tuple)) else 1
        return 1

    def _build_solver_info_columns(self, config: PostProcessingConfig) -> List[SolverInfoColumn]:
        cols: List[SolverInfoColumn] = []
        for name in self._resolve_solver_info_fields(config):
            n = self._resolve_field_num_components(name)
            if n == 1:
                cols.append(SolverInfoColumn(name, name, 0))
            else:
                cols += [SolverInfoColumn(f"{name}_{k}", name, k) for k in range(n)]
        return cols

    def _should_record(self, step: int, write_interval: int) -> bool:
        return step % write_interval == 0 or step == self._n_steps

    # ---- rows (collector.py:188-262)
    def record_step(self, step: int, phi, solve_stats: Dict[str, Tuple[SolveStats, ...]]) -> None:
        """Append the rows of step `step` (1-based) when its interval matches."""
        if self._closed:
            return
        t = step * self._delta_t
        cc = self._config.continuityError
        if cc is not None and self._continuity_writer is not None and self._should_record(step, cc.writeInterval):
            local, glob = self._continuity(phi, self._delta_t)
            self._cumulative_continuity += glob
            self._continuity_writer.writerow([t, local, glob, self._cumulative_continuity])
            self._continuity_file.flush()
        si = self._config.solverInfo
        if si is not None and self._solver_info_writer is not None and self._should_record(step, si.writeInterval):
            row: list = [t]
            for c in self._solver_info_columns:
                stats = solve_stats.get(c.field_name)
                if stats is None or c.component_index >= len(stats):
                    row += ["", "", "", "", ""]
                    continue
                s = stats[c.component_index]
                row += [s.solver, s.initial_residual, s.final_residual, s.iterations, int(s.converged)]
            self._solver_info_writer.writerow(row)
            self._solver_info_file.flush()

    def close(self) -> None:
        """Flush and close both files (collector.py:285-299); idempotent."""
        if self._closed:
            return
        for f in (self._continuity_file, self._solver_info_file):
            if f is not None:
                f.flush()
                f.close()
        self._continuity_file = self._solver_info_file = None
        self._continuity_writer = self._solver_info_writer = None
        self._closed = True

    def __enter__(self) -> "DiagnosticsCollector":
        return self

    def __exit__(self, exc_type: Optional[type], exc: Optional[BaseException], tb: Optional[TracebackType]) -> None:
        self.close()
