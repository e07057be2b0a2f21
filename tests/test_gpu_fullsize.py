"""GPU: the 3-D octree lid-driven cavity at its full size (32^3 root blocks,
two refinement levels within 0.1 of every wall, 1,125,384 cells, Re 1000,
PISO with 2 correctors, linear convection, CFL 0.5).

Plain FP64 after 2 steps is compared with the oracle's state frozen in
tests/golden/cavity3d_octree_piso_2.npz (the oracle needs about 3 minutes
per step on the host, so it runs once, in tests/golden/make_golden.py): U and
p at 20,000 sampled cells, the L2 norms of the whole fields and two
whole-field functionals sum(p r_p), sum(U r_U) with fixed random weights,
all to 1e-10.  The mixed-precision run (FP32 PCG inside FP64 refinement) is
compared with the plain FP64 run to 1e-6 relative L2 on U and p.

The 2-D channel around the Fourier body (205,600 cells, 2,144 immersed
faces, PIMPLE) is checked the same way against
tests/golden/channel2d_fourier_pimple_2.npz, with its AP payload and the
corrected flux on both sides of every immersed face.
"""

from __future__ import annotations

import copy
import pathlib
import sys

import numpy as np
import pytest
import torch

from tests.cases import case_dict

pytestmark = pytest.mark.gpu

GOLDEN = pathlib.Path(__file__).parent / "golden" / "cavity3d_octree_piso_2.npz"
GOLDEN_CHANNEL = pathlib.Path(__file__).parent / "golden" / "channel2d_fourier_pimple_2.npz"


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _rel(a, b) -> float:
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _run(mixed: bool, steps: int = 2):
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config

    cfg = copy.deepcopy(case_dict("cavity3d_octree_piso.yaml"))
    for k in ("p", "pFinal"):
        cfg["simulator"]["fvSolution"]["solvers"][k]["mixed_precision"] = mixed
    grid, alg = create_case(load_config(cfg))
    its = []
    for _ in range(steps):
        alg.step()
        its.append([s.iterations for s in alg.solve_stats["p"]])
    torch.cuda.synchronize()
    out = (grid.num_cells, alg.U.data.cpu().numpy(), alg.p.data.cpu().numpy(), its)
    del alg, grid
    return out


@pytest.fixture(scope="module")
def plain():
    _say("full cavity 1,125,384 cells: plain FP64, 2 PISO steps")
    return _run(False)


def test_full_size_plain_fp64_matches_the_oracle_after_two_steps(plain):
    C, U, p, its = plain
    g = np.load(GOLDEN)
    assert C == int(g["cells"]) == 1125384 and int(g["steps"]) == 2
    idx = g["idx"]
    assert _rel(U[idx], g["U"]) < 1e-10
    assert _rel(p[idx], g["p"]) < 1e-10
    assert abs(np.linalg.norm(U) - float(g["U_norm"])) <= 1e-10 * float(g["U_norm"])
    assert abs(np.linalg.norm(p) - float(g["p_norm"])) <= 1e-10 * float(g["p_norm"])
    rng = np.random.default_rng(1)
    r_p = rng.standard_normal(C)
    r_U = rng.standard_normal((C, 3))
    # |f_gpu - f_oracle| <= |field_gpu - field_oracle| |r| (Cauchy-Schwarz): 1e-10 relative
    assert abs(float(np.dot(p, r_p)) - float(g["f_p"])) <= 1e-10 * np.linalg.norm(p) * np.linalg.norm(r_p)
    assert abs(float(np.sum(U * r_U)) - float(g["f_U"])) <= 1e-10 * np.linalg.norm(U) * np.linalg.norm(r_U)
    assert abs(its[-1][-1] - int(g["p_iterations"])) <= 1


def test_full_size_mixed_precision_matches_plain_fp64(plain):
    _, U64, p64, its64 = plain
    _say("full cavity 1,125,384 cells: mixed precision, 2 PISO steps")
    C, U, p, its = _run(True)
    eu, ep = _rel(U, U64), _rel(p, p64)
    _say(f"mixed vs plain: U {eu:.3e} p {ep:.3e}; CG iterations {its} vs {its64}")
    assert eu < 1e-6 and ep < 1e-6


def test_full_size_fourier_channel_pimple_matches_the_oracle_after_two_steps():
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config
    from gridfoam_amd.fv import fvc

    _say("Fourier channel 205,600 cells: plain FP64, 2 PIMPLE steps")
    grid, alg = create_case(load_config(case_dict("channel2d_fourier_pimple.yaml")))
    g = np.load(GOLDEN_CHANNEL)
    C = grid.num_cells
    assert C == int(g["cells"]) == 205600 and grid.num_immersed_faces == int(g["immersed"]) == 2144
    np.testing.assert_array_equal(grid.ap_owner_near_boundary.cpu().numpy(), g["owner_near"])
    np.testing.assert_array_equal(grid.ap_neighbour_near_boundary.cpu().numpy(), g["neighbour_near"])
    np.testing.assert_allclose(grid.ap_dist_owner_to_bnd.cpu().numpy(), g["dist_owner"], rtol=0, atol=1e-13)
    np.testing.assert_allclose(grid.ap_dist_neighbour_to_bnd.cpu().numpy(), g["dist_neighbour"], rtol=0, atol=1e-13)
    for _ in range(int(g["steps"])):
        alg.step()
    torch.cuda.synchronize()
    U, p = alg.U.data.cpu().numpy(), alg.p.data.cpu().numpy()
    idx = g["idx"]
    assert _rel(U[idx], g["U"]) < 1e-10
    assert _rel(p[idx], g["p"]) < 1e-10
    assert abs(np.linalg.norm(U) - float(g["U_norm"])) <= 1e-10 * float(g["U_norm"])
    assert abs(np.linalg.norm(p) - float(g["p_norm"])) <= 1e-10 * float(g["p_norm"])
    rng = np.random.default_rng(1)
    r_p = rng.standard_normal(C)
    r_U = rng.standard_normal((C, 3))
    assert abs(float(np.dot(p, r_p)) - float(g["f_p"])) <= 1e-10 * np.linalg.norm(p) * np.linalg.norm(r_p)
    assert abs(float(np.sum(U * r_U)) - float(g["f_U"])) <= 1e-10 * np.linalg.norm(U) * np.linalg.norm(r_U)
    B = grid.num_domain_bnd_faces
    assert _rel(alg.phi.bnd_data[B:].cpu().numpy(), g["phi_immersed"]) < 1e-10
    div_max = float(fvc.div(alg.phi).abs().max())
    _say(f"max |div phi| gpu {div_max:.3e} oracle {float(g['div_max']):.3e}")
    assert div_max < 10.0 * float(g["div_max"]) + 1e-14
    its = sum(s.iterations for s in alg.solve_stats["p"])
    assert abs(its - int(g["p_iterations"])) <= 1
