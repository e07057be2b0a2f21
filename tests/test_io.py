"""Diagnostics CSV and VTU output on the host (no GPU).

The collector cases follow tests/unit/test_diagnostics_collector.py of the
reference (intervals, final rows, per-component statistics, flush before
close); the continuity errors are injected here and computed on the GPU in
the GPU tests.  The VTU cases check the hexahedra against the geometry they
must have: corner coordinates from centre +- half size in VTK node order and
a positive signed volume equal to the cell volume.
"""

from __future__ import annotations

import csv
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from gridfoam_amd.config import ContinuityErrorConfig, PostProcessingConfig, SolverInfoConfig
from gridfoam_amd.diagnostics import DiagnosticsCollector
from gridfoam_amd.solvers import SolveStats
from gridfoam_amd.vtu import VTK_HEXAHEDRON, read_vtu, save_export_fields_as_vtu, to_unstructured_grid, write_vtu

GRID = SimpleNamespace(sim_config=None)


def _rows(path):
    with path.open() as f:
        return list(csv.reader(f))


def test_collector_writes_continuity_rows_on_interval(tmp_path):
    """Header, rows at steps 2 and 4 (interval 2) and the off-interval last step."""
    calls = []

    def cont(phi, dt):
        calls.append(dt)
        return 0.25 * dt, -0.5 * dt

    cfg = PostProcessingConfig(continuityError=ContinuityErrorConfig(writeInterval=2))
    col = DiagnosticsCollector.from_config(GRID, cfg, tmp_path / "out", n_steps=5, delta_t=1.0, continuity=cont)
    assert col is not None
    with col:
        for step in range(1, 6):
            col.record_step(step, None, {})
    rows = _rows(tmp_path / "out" / "continuity_error.csv")
    assert rows[0] == ["time", "local", "global", "cumulative"]
    assert len(rows) == 4
    assert [r[0] for r in rows[1:]] == ["2.0", "4.0", "5.0"]
    assert [float(r[3]) for r in rows[1:]] == [-0.5, -1.0, -1.5]  # running sum of the global error
    assert len(calls) == 3


def test_collector_writes_solver_info_with_flush(tmp_path):
    """Readable before close; close adds no row."""
    cfg = PostProcessingConfig(solverInfo=SolverInfoConfig(fields=["p"], writeInterval=1))
    stats = {"p": (SolveStats("cg", 1.0, 1e-4, 3, True),)}
    col = DiagnosticsCollector.from_config(GRID, cfg, tmp_path, n_steps=1, delta_t=0.5)
    col.record_step(1, None, stats)
    rows = _rows(tmp_path / "solver_info.csv")
    assert rows[0] == ["time", "p_solver", "p_initial", "p_final", "p_iters", "p_converged"]
    assert rows[1][0] == "0.5" and rows[1][1] == "cg" and rows[1][4] == "3" and rows[1][5] == "1"
    col.close()
    col.close()
    assert len(_rows(tmp_path / "solver_info.csv")) == 2


def test_collector_writes_vector_solver_info_per_component(tmp_path):
    cfg = PostProcessingConfig(solverInfo=SolverInfoConfig(fields=["U"], writeInterval=1))
    stats = {"U": (SolveStats("bicgstab", 1.0, 1e-3, 2, True), SolveStats("bicgstab", 0.5, 2e-3, 3, True),
                   SolveStats("bicgstab", 0.2, 3e-3, 4, False))}
    with DiagnosticsCollector.from_config(GRID, cfg, tmp_path, n_steps=1, delta_t=1.0) as col:
        col.record_step(1, None, stats)
    rows = _rows(tmp_path / "solver_info.csv")
    assert rows[0][1:6] == ["U_0_solver", "U_0_initial", "U_0_final", "U_0_iters", "U_0_converged"]
    assert len(rows[0]) == 1 + 3 * 5
    assert [rows[1][1 + 5 * k + 3] for k in range(3)] == ["2", "3", "4"]
    assert [rows[1][1 + 5 * k + 4] for k in range(3)] == ["1", "1", "0"]


def test_collector_defaults_and_missing_stats(tmp_path):
    """Default fields U and p; a field without statistics leaves empty cells;
    nothing enabled -> no collector."""
    assert DiagnosticsCollector.from_config(GRID, None, tmp_path, n_steps=1, delta_t=1.0) is None
    assert DiagnosticsCollector.from_config(GRID, PostProcessingConfig(), tmp_path, n_steps=1, delta_t=1.0) is None
    cfg = PostProcessingConfig(solverInfo=SolverInfoConfig())
    with DiagnosticsCollector.from_config(GRID, cfg, tmp_path, n_steps=3, delta_t=1.0) as col:
        col.record_step(1, None, {"p": (SolveStats("cg", 1.0, 0.1, 7, False),)})
    rows = _rows(tmp_path / "solver_info.csv")
    assert rows[0][1] == "U_0_solver" and rows[0][16] == "p_solver"
    assert rows[1][1:16] == [""] * 15
    assert rows[1][16:] == ["cg", "1.0", "0.1", "7", "0"]


def _hex_volume(p: np.ndarray) -> float:
    """Signed volume of a VTK hexahedron (nodes 0-3 bottom, 4-7 top) from six
    tetrahedra of the cube split along the 0-6 diagonal."""
    tets = [(0, 1, 2, 6), (0, 2, 3, 6), (0, 3, 7, 6), (0, 7, 4, 6), (0, 4, 5, 6), (0, 5, 1, 6)]
    v = 0.0
    for a, b, c, d in tets:
        v += np.dot(np.cross(p[b] - p[a], p[c] - p[a]), p[d] - p[a]) / 6.0
    return v


def _fake_grid():
    centers = torch.tensor([[0.25, 0.25, 0.25], [0.75, 0.125, 0.125], [0.625, 0.625, 0.875]], dtype=torch.float64)
    sizes = torch.tensor([[0.5, 0.5, 0.5], [0.5, 0.25, 0.25], [0.25, 0.25, 0.25]], dtype=torch.float64)
    return SimpleNamespace(cell_centers=centers, cell_sizes=sizes, num_cells=3)


def test_vtu_hexahedra_match_cell_geometry(tmp_path):
    g = _fake_grid()
    U = SimpleNamespace(name="U", export=True, data=torch.arange(9, dtype=torch.float64).reshape(3, 3))
    p = SimpleNamespace(name="p", export=True, data=torch.tensor([1.0, -2.0, 3.5], dtype=torch.float64))
    hidden = SimpleNamespace(name="hidden", export=False, data=torch.zeros(3, dtype=torch.float64))
    path = tmp_path / "case_0001.vtu"
    save_export_fields_as_vtu(g, str(path), fields=(U, p, hidden))
    m = read_vtu(str(path))
    assert m.points.shape == (24, 3)
    np.testing.assert_array_equal(m.connectivity, np.arange(24))
    np.testing.assert_array_equal(m.offsets, [8, 16, 24])
    np.testing.assert_array_equal(m.types, [VTK_HEXAHEDRON] * 3)
    order = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)]
    for c in range(3):
        ctr, half = g.cell_centers[c].numpy(), g.cell_sizes[c].numpy() / 2
        for k, s in enumerate(order):
            np.testing.assert_array_equal(m.points[8 * c + k], ctr + np.array(s) * half)
        vol = float(np.prod(g.cell_sizes[c].numpy()))
        assert _hex_volume(m.points[8 * c:8 * c + 8]) == pytest.approx(vol, rel=1e-14)
    assert set(m.cell_data) == {"U", "p"}
    np.testing.assert_array_equal(m.cell_data["U"], U.data.numpy())
    np.testing.assert_array_equal(m.cell_data["p"], p.data.numpy())


def test_vtu_reuses_grid_and_round_trips_dtypes(tmp_path):
    g = _fake_grid()
    ug = to_unstructured_grid(g)
    ug.cell_data["level"] = np.array([0, 1, 2], dtype=np.int32)
    write_vtu(ug, str(tmp_path / "a.vtu"))
    m = read_vtu(str(tmp_path / "a.vtu"))
    assert m.cell_data["level"].dtype == np.int32
    np.testing.assert_array_equal(m.cell_data["level"], [0, 1, 2])
    np.testing.assert_array_equal(m.points, ug.points)


def test_manual_algorithm_config_loads():
    """meta/config.py:515-521: the manual algorithm type validates."""
    from gridfoam_amd.config import ManualAlgorithm, load_config
    from tests.cases import small_cavity2d

    cfg = small_cavity2d(8)
    cfg["simulator"]["fvSolution"]["algorithm"] = {"type": "manual"}
    assert isinstance(load_config(cfg).simulator.fvSolution.algorithm, ManualAlgorithm)
