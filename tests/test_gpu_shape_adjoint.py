"""GPU: the drag gradient with respect to the 16 outline coefficients of the
Fourier body in the 2-D channel (205,600 cells, 2,144 immersed faces, PIMPLE
2x2, FP64), through ``FourierShape`` and checkpointed PIMPLE steps, against
central differences of runs on grids cut with perturbed coefficients
(post/forces/evaluator.py:22-180 for the drag coefficient)."""

from __future__ import annotations

import copy
import sys

import pytest
import torch

from tests.cases import case_dict

pytestmark = pytest.mark.gpu

FORCE = {"patches": ["body"], "rho": 1.0, "magU_ref": 1.0, "A_ref": 1.0, "L_ref": 1.0,
         "local_coord": {"drag_dir": [1.0, 0.0, 0.0], "lift_dir": [0.0, 1.0, 0.0],
                         "center_of_rotation": [5.0, 4.0, 0.25]}}
STEPS, EVERY, WARM = 20, 10, 5
H = 1e-6  # moves the outline by at most R h = 5e-7; the cut topology's margin is ~1.2e-4


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _config(coef=None) -> dict:
    cfg = copy.deepcopy(case_dict("channel2d_fourier_pimple.yaml"))
    sol = cfg["simulator"]["fvSolution"]["solvers"]
    sol["U"].update({"tolerance": 0.0, "rel_tolerance": 1e-13, "max_iter": 2000})
    sol["p"].update({"tolerance": 0.0, "rel_tolerance": 1e-13, "max_iter": 8000})
    if coef is not None:
        body = cfg["fluxel"]["immersed_bodies"][0]
        body["a"], body["b"] = [float(v) for v in coef[:8]], [float(v) for v in coef[8:]]
    return cfg


def _case(coef=None, warm: int = 0):
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config

    grid, alg = create_case(load_config(_config(coef)))
    for _ in range(warm):
        alg.step()
    return grid, alg


def _evaluator(grid):
    from gridfoam_amd.config import ForceCoeffConfig
    from gridfoam_amd.forces import ForceEvaluator

    return ForceEvaluator(grid, config=ForceCoeffConfig.model_validate(copy.deepcopy(FORCE)))


def _drag(alg, ev, state0, grad: bool):
    from gridfoam_amd.adjoint import TransientStep, checkpointed_rollout

    step = TransientStep(alg)
    s = tuple(t.clone() for t in state0)
    with torch.set_grad_enabled(grad):
        params = alg.grid.shape.tensors() if grad and alg.grid.shape is not None else None
        out = checkpointed_rollout(step, s, STEPS, every=EVERY, params=params)
        alg.U.data, alg.p.data = out[0], out[3]
        alg.U._state = None
        alg.p._state = None
        return ev.evaluate(turbulence=float(alg.nu)).Cd


def _same_topology(g, g0) -> bool:
    return all(torch.equal(getattr(g, k), getattr(g0, k)) for k in (
        "ap_is_immersed_faces", "ap_owner_near_boundary", "ap_neighbour_near_boundary",
        "ap_owner_bnd_patch_id", "ap_neighbour_bnd_patch_id", "bnd_owner_i32"))


def test_shape_mode_leaves_the_forward_unchanged():
    from gridfoam_amd.adjoint import TransientStep, transient_state
    from gridfoam_amd.shape import FourierShape

    _say("Fourier channel: 2 PIMPLE steps with and without FourierShape")
    grid, alg = _case(warm=1)
    state0 = transient_state(alg)
    step = TransientStep(alg)
    with torch.no_grad():
        ref = step(tuple(t.clone() for t in state0))
    shape = FourierShape(grid, "body")
    with shape:
        out = step(tuple(t.clone() for t in state0))
        assert out[0].requires_grad and out[3].requires_grad
    for a, b in zip(out, ref):
        assert torch.equal(a.detach(), b)
    assert grid.shape is None and not grid.ap_dist_owner_to_bnd.requires_grad


def test_drag_gradient_wrt_the_16_outline_coefficients_matches_central_differences():
    from gridfoam_amd.adjoint import transient_state
    from gridfoam_amd.shape import FourierShape

    _say(f"Fourier channel shape adjoint: drag after {STEPS} PIMPLE steps, checkpoints every {EVERY}")
    grid, alg = _case(warm=WARM)
    state0 = transient_state(alg)
    ev = _evaluator(grid)
    shape = FourierShape(grid, "body")
    c0 = shape.coefficients.detach().cpu().tolist()
    with shape:
        cd = _drag(alg, ev, state0, True)
        (g,) = torch.autograd.grad(cd, shape.coefficients)
    g = g.cpu()
    _say(f"Cd = {float(cd.detach()):.10f}; adjoint dCd/dc = {[f'{v:.6e}' for v in g.tolist()]}")
    fd = torch.zeros(16, dtype=torch.float64)
    for j in range(16):
        vals = []
        for sgn in (1.0, -1.0):
            c = list(c0)
            c[j] += sgn * H
            gj, aj = _case(c)
            assert gj.num_cells == grid.num_cells and _same_topology(gj, grid), j
            vals.append(float(_drag(aj, _evaluator(gj), state0, False)))
        fd[j] = (vals[0] - vals[1]) / (2 * H)
        _say(f"coefficient {j}: adjoint {float(g[j]):.8e}, central difference {float(fd[j]):.8e}")
    big = float(fd.abs().max())
    assert big > 1e-2
    assert float(torch.linalg.vector_norm(g - fd)) <= 1e-5 * float(torch.linalg.vector_norm(fd))
    for j in range(16):
        assert abs(float(g[j] - fd[j])) <= 1e-5 * max(abs(float(fd[j])), 1e-2 * big), j
