"""GPU: ForceEvaluator (post/forces/evaluator.py:22-180) on the HIP path
against the oracle's force integration, and the reference's controlled
pressure and flow scenarios (tests/e2e/test_force_coeffs.py) on analytic
bodies."""

from __future__ import annotations

import copy
import math
import sys

import numpy as np
import pytest
import torch

from oracle import fv_oracle as fo
from tests.cases import case_dict, fourier_channel

pytestmark = pytest.mark.gpu

FORCE = {"patches": ["body"], "rho": 1.0, "magU_ref": 1.0, "A_ref": 1.0, "L_ref": 1.0,
         "local_coord": {"drag_dir": [1.0, 0.0, 0.0], "lift_dir": [0.0, 0.0, 1.0],
                         "center_of_rotation": [0.5, 0.5, 0.5]}}
CUBE = dict(name="body", kind="box", min=[0.4, 0.4, 0.4], max=[0.6, 0.6, 0.6])
PLATE = dict(name="body", kind="box", min=[0.3, 0.3, 0.49], max=[0.7, 0.7, 0.51])


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _body_case(body, n, p_data):
    from gridfoam_amd.config import FluxelConfig, ForceCoeffConfig
    from gridfoam_amd.fields import CellField, DirichletBC, NeumannBC
    from gridfoam_amd.forces import ForceEvaluator
    from gridfoam_amd.grid import OctreeGrid

    fx = {"domain": {"lower": [0, 0, 0], "upper": [1, 1, 1]}, "root_resolution": [n, n, n],
          "immersed_bodies": [body]}
    g = OctreeGrid(FluxelConfig.model_validate(fx), dt=1.0)
    U = CellField(g, "U", (3,))
    p = CellField(g, "p", ())
    U.add_boundary_conditions({"body": DirichletBC([0.0, 0.0, 0.0])})
    p.add_boundary_conditions({"body": NeumannBC(0.0)})
    p.data = torch.as_tensor(p_data(g.cell_centers[: g.num_cells]), dtype=torch.float64, device=g.device)
    ev = ForceEvaluator(g, config=ForceCoeffConfig.model_validate(copy.deepcopy(FORCE)))
    return g, ev


@pytest.mark.parametrize("body", ["cube", "flat_plate"])
def test_uniform_pressure_gives_zero_force_on_closed_body(body):
    """test_force_coeffs.py:105-152."""
    _say(f"uniform pressure {body}")
    g, ev = _body_case(CUBE if body == "cube" else PLATE, 16, lambda c: torch.full_like(c[:, 0], 2.5))
    co = ev.evaluate(g, time=0.0, turbulence=0.1)
    for k in ("Cd", "Cs", "Cl", "CmRoll", "CmPitch", "CmYaw"):
        assert abs(float(getattr(co, k))) < 1e-10, k


def test_thin_sheet_pressure_jump_converges_and_matches_oracle():
    """test_force_coeffs.py:155-240 through the HIP path; each grid's lift
    equals the oracle's."""
    errors = []
    for n in (16, 32, 64):
        _say(f"thin sheet n={n}")
        g, ev = _body_case(PLATE, n, lambda c: torch.where(c[:, 2] > 0.5, 1.0, 2.0))
        cl = float(ev.evaluate(g, time=0.0, turbulence=0.0).Cl)
        expected = 1.0 * 0.16 / 0.5
        errors.append(abs(cl - expected) / expected)
        o = fo.Grid([0, 0, 0], [1, 1, 1], [n, n, n], bodies=[PLATE])
        U = fo.Field("U", np.zeros((o.num_cells, 3)), {"body": fo.BC("dirichlet", np.zeros(3))})
        p = fo.Field("p", np.where(o.cell_centers[:, 2] > 0.5, 1.0, 2.0), {"body": fo.BC("neumann", 0.0)})
        f, m = fo.patch_force(o, U, p, 0.0, "body", 1.0, [0.5, 0.5, 0.5])
        assert abs(cl - f[2] / 0.5) <= 1e-12 * abs(cl)
    assert errors[-1] < 0.10 and errors[-1] <= errors[0] + 1e-12, errors


def test_fourier_channel_coefficients_match_oracle():
    """Drag, lift and moments on the Fourier body after 3 PIMPLE steps: host
    ForceEvaluator on the GPU fields against the oracle's integration of the
    oracle's fields (relative 1e-9)."""
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config
    from gridfoam_amd.forces import ForceEvaluator

    cfg = fourier_channel(root=(40, 16, 1), level=1)
    force = copy.deepcopy(FORCE)
    force["local_coord"] = {"drag_dir": [1.0, 0.0, 0.0], "lift_dir": [0.0, 1.0, 0.0],
                            "center_of_rotation": [5.0, 4.0, 0.0]}
    force["A_ref"] = 1.0 * cfg["fluxel"]["domain"]["upper"][2]
    cfg["simulator"]["post_processing"] = {"forceCoeff": force}
    _say("fourier channel forces")
    grid, alg = create_case(load_config(cfg))
    case = fo.Case(cfg)
    for _ in range(3):
        alg.step()
        case.step()
    ev = ForceEvaluator(grid)
    co = ev.evaluate(grid, time=0.15)
    nu = cfg["simulator"]["properties"]["transport"]["nu"]
    f, m = fo.patch_force(case.g, case.U, case.p, nu, "body", 1.0, [5.0, 4.0, 0.0])
    oc = fo.force_coeffs(f, m, fo.orthonormal_coord([1, 0, 0], lift_dir=[0, 1, 0]), 0.5 * force["A_ref"],
                         0.5 * force["A_ref"])
    _say(f"Cd gpu {float(co.Cd):.10f} oracle {oc['Cd']:.10f}; Cl {float(co.Cl):.3e} / {oc['Cl']:.3e}")
    assert float(co.Cd) > 0.0
    scale = max(abs(oc["Cd"]), 1e-30)
    for k in ("Cd", "Cs", "Cl", "CmRoll", "CmPitch", "CmYaw"):
        assert abs(float(getattr(co, k)) - oc[k]) <= 1e-9 * scale, (k, float(getattr(co, k)), oc[k])


def test_cube_low_re_drag_with_simple(tmp_path):
    """test_force_coeffs.py:243-303: a Re ~ 12 cube in a slip channel after
    20 SIMPLE iterations has a finite drag coefficient in (0.1, 20) and
    |Cl| < 1; the coefficients go to a CSV with the reference's columns."""
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config
    from gridfoam_amd.forces import CSV_HEADER, ForceEvaluator

    cfg = copy.deepcopy(case_dict("cavity2d_simple.yaml"))
    cfg["fluxel"].update({"domain": {"lower": [-2.0, -1.0, 0.0], "upper": [3.0, 1.0, 2.0]},
                          "root_resolution": [40, 16, 16],
                          "immersed_bodies": [dict(name="body", kind="box", min=[0.2, -0.3, 0.7],
                                                   max=[0.8, 0.3, 1.3])]})
    sim = cfg["simulator"]
    sim["control"].update({"deltaT": 1.0, "endTime": 20.0, "writeInterval": 20})
    sim["properties"]["transport"]["nu"] = 0.05
    sim["fvSolution"]["solvers"]["U"].update({"max_iter": 500, "rel_tolerance": 0.1, "tolerance": 1e-8})
    sim["fvSolution"]["solvers"]["p"].update({"max_iter": 500, "rel_tolerance": 0.01, "tolerance": 1e-7})
    slip = ["y_minus", "y_plus", "z_minus", "z_plus"]
    sim["conditions"] = {
        "U": {"internal": [0.0, 0.0, 0.0], "boundary": {
            "inlet": {"type": "dirichlet", "patches": ["x_minus"], "value": [1.0, 0.0, 0.0]},
            "outlet": {"type": "inlet_outlet", "patches": ["x_plus"], "value": [0.0, 0.0, 0.0]},
            "slip": {"type": "slip", "patches": slip},
            "body": {"type": "dirichlet", "patches": ["body"], "value": [0.0, 0.0, 0.0]}}},
        "p": {"internal": 0.0, "boundary": {
            "inlet": {"type": "neumann", "patches": ["x_minus"], "value": 0.0},
            "outlet": {"type": "dirichlet", "patches": ["x_plus"], "value": 0.0},
            "slip": {"type": "neumann", "patches": slip, "value": 0.0},
            "body": {"type": "neumann", "patches": ["body"], "value": 0.0}}}}
    force = copy.deepcopy(FORCE)
    force.update({"A_ref": 0.36, "L_ref": 0.6})
    force["local_coord"]["center_of_rotation"] = [0.5, 0.0, 1.0]
    sim["post_processing"] = {"forceCoeff": force}
    _say("cube SIMPLE 20 iterations")
    grid, alg = create_case(load_config(cfg))
    ev = ForceEvaluator(grid)
    for _ in range(20):
        alg.step()
    co = ev.evaluate(grid, time=20.0)
    cd, cl = float(co.Cd), float(co.Cl)
    _say(f"cube Cd {cd:.4f} Cl {cl:.3e}")
    assert math.isfinite(cd) and math.isfinite(cl)
    assert 0.1 < cd < 20.0
    assert abs(cl) < 1.0
    out = tmp_path / "forceCoeffs.csv"
    ev.write_csv(out)
    rows = out.read_text().splitlines()
    assert rows[0].split(",") == CSV_HEADER and len(rows) == 2


def test_runner_writes_force_coefficients_at_the_end_of_the_run(tmp_path):
    """runner.py:132-143: with post_processing.forceCoeff the runner evaluates
    the body's coefficients after the last step into coefficients.csv; they
    equal the evaluator's on the same run driven step by step."""
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config
    from gridfoam_amd.forces import CSV_HEADER, ForceEvaluator
    from gridfoam_amd.runner import all_run

    cfg = fourier_channel(root=(40, 16, 1), level=1)
    sim = cfg["simulator"]
    sim["control"]["endTime"] = 3 * sim["control"]["deltaT"]
    sim["control"]["writeInterval"] = 1000
    sim["control"]["output"] = {"output_dir": str(tmp_path), "base_name": "channel"}
    force = copy.deepcopy(FORCE)
    force["local_coord"]["center_of_rotation"] = [5.0, 4.0, 0.25]
    sim["post_processing"] = {"forceCoeff": force}
    _say("runner: Fourier channel 3 PIMPLE steps with forceCoeff")
    assert all_run(copy.deepcopy(cfg)) == 3
    rows = (tmp_path / "coefficients.csv").read_text().splitlines()
    assert rows[0].split(",") == CSV_HEADER and len(rows) == 2
    grid, alg = create_case(load_config(cfg))
    for _ in range(3):
        alg.step()
    co = ForceEvaluator(grid).evaluate(grid, time=float(sim["control"]["endTime"]))
    vals = [float(v) for v in rows[1].split(",")]
    assert vals[0] == pytest.approx(float(sim["control"]["endTime"]))
    assert vals[1] == pytest.approx(float(co.Cd), rel=1e-12, abs=1e-14)
    assert vals[4] == pytest.approx(float(co.Cl), rel=1e-12, abs=1e-14)
