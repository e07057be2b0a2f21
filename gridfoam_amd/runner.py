"""Run a case from its YAML file: steps, diagnostics CSVs and VTU output.

Restates src/gridfoam/runner.py:46-194 (``manual_run``, ``manual_step``,
``all_run``) over this package's grid and algorithms.  Potential-flow
initialisation and force coefficients (runner.py:96-99, 125-144) need
immersed bodies and are not part of this build: configurations asking for
them are rejected.
"""

from __future__ import annotations

import logging
from dataclasses import dataclass
from pathlib import Path
from typing import Iterator, Union

from .algorithms import AlgorithmBase, create_algorithm
from .config import ControlConfig, load_config
from .diagnostics import DiagnosticsCollector
from .grid import OctreeGrid, create_grid
from .vtu import HexMesh, save_export_fields_as_vtu, to_unstructured_grid

logger = logging.getLogger(__name__)


@dataclass
class StepData:
    """One step of a run (runner.py:22-43)."""

    step: int
    control: ControlConfig
    algorithm: AlgorithmBase
    ugrid: HexMesh


def _output(control: ControlConfig) -> tuple:
    out = control.output or {}
    if "output_dir" not in out or "base_name" not in out:
        raise ValueError("control.output needs output_dir and base_name")
    return Path(out["output_dir"]), str(out["base_name"])


def manual_run(config_path: Union[str, Path, dict], device: str = "cuda") -> OctreeGrid:
    """runner.py:46-64: configuration -> grid (the simulator settings attached)."""
    config = load_config(config_path)
    grid = create_grid(config, device=device)
    grid.sim_config = config.simulator
    return grid


def manual_step(config_path: Union[str, Path, dict], device: str = "cuda") -> Iterator[StepData]:
    """runner.py:67-147: yields every step with the algorithm not yet advanced;
    the caller calls ``algorithm.step()``.  Diagnostics record each step."""
    grid = manual_run(config_path, device)
    sim = grid.sim_config
    algorithm = create_algorithm(grid)
    output_dir, _ = _output(sim.control)
    output_dir.mkdir(parents=True, exist_ok=True)
    ugrid = to_unstructured_grid(grid)
    n_steps = int(sim.control.endTime / sim.control.deltaT)
    diagnostics = DiagnosticsCollector.from_config(grid, sim.post_processing, output_dir, n_steps=n_steps,
                                                   delta_t=sim.control.deltaT)
    if diagnostics is not None:
        algorithm.attach_diagnostics(diagnostics)
    try:
        for step in range(1, n_steps + 1):
            yield StepData(step=step, control=sim.control, algorithm=algorithm, ugrid=ugrid)
        # runner.py:132-143: force coefficients of the body patches at the end
        # of a run that went through every step (coefficients.csv; the
        # surface-mesh VTU of the sampled forces is not written)
        if sim.post_processing is not None and sim.post_processing.forceCoeff is not None:
            from .forces import ForceEvaluator  # noqa: PLC0415

            ev = ForceEvaluator(grid)
            ev.evaluate(grid, time=float(sim.control.endTime))
            ev.write_csv(output_dir / "coefficients.csv")
    finally:
        if diagnostics is not None:
            diagnostics.close()


def all_run(config_path: Union[str, Path, dict], device: str = "cuda") -> int:
    """runner.py:159-194: advance every step, stop on a converged steady
    SIMPLE run, write U and p as VTU every writeInterval steps, at the last
    step and on a converged exit.  Returns the number of steps taken."""
    taken = 0
    for sd in manual_step(config_path, device):
        alg = sd.algorithm
        alg.step()
        taken = sd.step
        converged = alg.has_simulation_converged()
        n_steps = int(sd.control.endTime / sd.control.deltaT)
        if converged or sd.step % sd.control.writeInterval == 0 or sd.step == n_steps:
            output_dir, base_name = _output(sd.control)
            save_export_fields_as_vtu(alg.grid, str(output_dir / f"{base_name}_{sd.step:04d}.vtu"), ugrid=sd.ugrid,
                                      fields=(alg.U, alg.p))
            logger.info("step=%4d", sd.step)
        if converged:
            logger.info("Simulation converged at step=%d", sd.step)
            break
    return taken
