"""GPU: SIMPLE residualControl, the runner, diagnostics CSVs and VTU output
against the FP64 oracle on the same case (runner.py:46-194,
post/diagnostics/collector.py, io/vtu.py, simple.py:241-285)."""

from __future__ import annotations

import csv
import sys

import numpy as np
import pytest

from oracle import fv_oracle as fo
from tests.cases import small_cavity2d

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _oracle_history(cfg, steps):
    case = fo.Case(cfg)
    out = []
    for _ in range(steps):
        case.step()
        out.append(dict(case.simple_cur))
    return out


# Residual parity: the U and p entries are O(1) normalised residuals (1e-8);
# the phi entry is |div phi| after the correction, i.e. what is left of the
# pressure solve's own residual (1e-8 of the right-hand side), so it carries
# the rounding differences of two converged solves: 1e-3.
_RES_TOL = {"p": 1e-8, "U": 1e-8, "phi": 1e-3}


def test_simple_residual_control_matches_oracle():
    """Recorded U, p and phi residuals within _RES_TOL of the oracle's every
    step, and the same convergence flag every step for tolerances (the
    medians of the oracle's values) that no residual comes within 10x
    _RES_TOL of."""
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config

    steps = 12
    cfg = small_cavity2d(16)
    cfg["simulator"]["fvSolution"]["algorithm"]["residualControl"] = {"p": 0.0, "U": 0.0, "phi": 0.0}
    hist = _oracle_history(cfg, steps)
    tol = {k: float(np.median([h[k] for h in hist])) for k in hist[0]}
    for k, t in tol.items():
        assert min(abs(h[k] - t) / t for h in hist) > 10 * _RES_TOL[k], k
    cfg["simulator"]["fvSolution"]["algorithm"]["residualControl"] = {
        "p": tol["p"], "U": {"tolerance": tol["U"], "rel_tolerance": 0.5}, "phi": tol["phi"]}
    _say("simple residualControl cavity2d 16x16")
    grid, alg = create_case(load_config(cfg))
    case = fo.Case(cfg)
    flags = []
    for _ in range(steps):
        alg.step()
        case.step()
        for k in ("p", "U", "phi"):
            cur, ini = alg._current_residuals[k], alg._initial_residuals[k]
            assert abs(cur - case.simple_cur[k]) <= _RES_TOL[k] * abs(case.simple_cur[k]), (k, cur, case.simple_cur[k])
            assert abs(ini - case.simple_init[k]) <= _RES_TOL[k] * abs(case.simple_init[k]), (k, ini)
        assert alg.has_converged() == case.simple_converged()
        assert alg.has_simulation_converged() == alg.has_converged()
        flags.append(case.simple_converged())
    assert any(flags) and not all(flags)  # the run crosses the tolerances


def _post_cfg(tmp_path, end, write_interval, residual_control=None):
    cfg = small_cavity2d(16)
    sim = cfg["simulator"]
    sim["control"]["endTime"] = float(end)
    sim["control"]["writeInterval"] = write_interval
    sim["control"]["output"] = {"output_dir": str(tmp_path), "base_name": "cavity"}
    sim["post_processing"] = {"continuityError": {"writeInterval": 1}, "solverInfo": {"writeInterval": 2}}
    if residual_control is not None:
        sim["fvSolution"]["algorithm"]["residualControl"] = residual_control
    return cfg


def test_runner_diagnostics_and_vtu_match_oracle(tmp_path):
    """manual_step over 5 SIMPLE iterations: continuity_error.csv rows equal
    the oracle's errors of each step (the local error is what is left of the
    pressure solve's residual, so to _RES_TOL["phi"]; the global one, a sum
    of boundary fluxes of a closed cavity, to 1e-12 of the flux scale),
    solver_info.csv rows carry each recorded step's statistics,
    and the VTU written at the last step holds the oracle-matching U and p on
    hexahedra with the oracle's cell geometry."""
    from gridfoam_amd.runner import manual_step
    from gridfoam_amd.vtu import read_vtu, save_export_fields_as_vtu

    cfg = _post_cfg(tmp_path, 5, 5)
    case = fo.Case(cfg)
    expected, stats_rows = [], {}
    _say("runner manual_step cavity2d 16x16, 5 steps")
    alg = None
    for sd in manual_step(cfg):
        alg = sd.algorithm
        alg.step()
        case.step()
        expected.append(fo.continuity_error(case.g, case.phi, 1.0))
        stats_rows[sd.step] = {k: tuple(v) for k, v in alg.solve_stats.items()}
        if sd.step == 5:
            save_export_fields_as_vtu(alg.grid, str(tmp_path / "cavity_0005.vtu"), ugrid=sd.ugrid,
                                      fields=(alg.U, alg.p))
    with (tmp_path / "continuity_error.csv").open() as f:
        rows = list(csv.reader(f))
    assert rows[0] == ["time", "local", "global", "cumulative"] and len(rows) == 6
    scale = float(np.abs(case.phi.single).max())
    cum = 0.0
    for r, (loc, glob) in zip(rows[1:], expected):
        assert abs(float(r[1]) - loc) <= _RES_TOL["phi"] * abs(loc) + 1e-14 * scale
        assert abs(float(r[2]) - glob) <= 1e-12 * scale
        cum += float(r[2])
        assert float(r[3]) == pytest.approx(cum, rel=1e-12, abs=1e-300)
    with (tmp_path / "solver_info.csv").open() as f:
        info = list(csv.reader(f))
    assert [r[0] for r in info[1:]] == ["2.0", "4.0", "5.0"]  # interval 2 plus the last step
    p_col = info[0].index("p_iters")
    for r in info[1:]:
        st = stats_rows[int(float(r[0]))]["p"][0]
        assert int(r[p_col]) == st.iterations
    m = read_vtu(str(tmp_path / "cavity_0005.vtu"))
    g = case.g
    signs = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                      [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=np.float64)
    corners = (g.cell_centers[:, None, :] + signs[None] * (g.cell_sizes[:, None, :] / 2)).reshape(-1, 3)
    np.testing.assert_allclose(m.points, corners, rtol=0, atol=1e-14)
    assert _rel(m.cell_data["U"], case.U.data) < 1e-10
    assert _rel(m.cell_data["p"], case.p.data) < 1e-10
    np.testing.assert_array_equal(m.cell_data["U"], _np(alg.U.data))


def test_all_run_stops_on_converged_simple(tmp_path):
    """all_run ends at the first step whose residuals meet residualControl
    (the oracle's step), writing the VTU of that step."""
    from gridfoam_amd.runner import all_run

    probe = small_cavity2d(16)
    probe["simulator"]["fvSolution"]["algorithm"]["residualControl"] = {"phi": 0.0}
    hist = [h["phi"] for h in _oracle_history(probe, 8)]
    stop = next(k for k in range(2, 8) if hist[k] < hist[0])  # phi drops below its first value
    tol = float(np.sqrt(hist[stop] * min(hist[:stop])))  # between step `stop` and every earlier step
    assert all(h > tol * (1 + 1e-2) for h in hist[:stop]) and hist[stop] < tol * (1 - 1e-2)
    cfg = _post_cfg(tmp_path, 50, 100, residual_control={"phi": tol})
    _say(f"all_run cavity2d 16x16 residualControl stop at {stop + 1}")
    taken = all_run(cfg)
    assert taken == stop + 1
    assert (tmp_path / f"cavity_{stop + 1:04d}.vtu").exists()
    assert not (tmp_path / f"cavity_{stop + 2:04d}.vtu").exists()


def test_manual_algorithm_gives_a_grid_and_no_step_sequence():
    """runner.py:46-96, algorithms/factory.py:37-38: a ``manual`` algorithm
    loads, manual_run builds its grid, and asking for a step sequence raises."""
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config
    from gridfoam_amd.runner import manual_run, manual_step

    cfg = small_cavity2d(8)
    cfg["simulator"]["fvSolution"]["algorithm"] = {"type": "manual"}
    _say("manual algorithm")
    grid = manual_run(cfg)
    assert grid.num_cells == 64 and grid.sim_config.fvSolution.algorithm.type == "manual"
    with pytest.raises(ValueError, match="no step sequence"):
        create_case(load_config(cfg))
    with pytest.raises(ValueError, match="no step sequence"):
        next(iter(manual_step(cfg)))
