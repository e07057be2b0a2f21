"""GPU: the 3-D bluff body in a tunnel.

- The octree with surface refinement to target_level (meta/config.py:100-106)
  is bit-exact against the oracle's build.
- The axis-projected cut of the bluff body (rounded nose edges, rear slant,
  struts) matches the oracle.
- Two PISO steps on a coarse version of the case match the oracle's driver
  (relative L2 of U and p <= 1e-10).
- The full case's grid (21.0M cells) has the leaf count the oracle's
  pre-leaf build predicts."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from oracle import fv_oracle as fo
from tests.cases import bluff_body, case_dict

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _np(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


def _rel(a, b) -> float:
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _grids(cfg: dict):
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.grid import OctreeGrid

    fx = cfg["fluxel"]
    g = OctreeGrid(FluxelConfig.model_validate(fx), dt=0.01)
    o = fo.grid_from_cfg({"fluxel": fx, "simulator": {"control": {"deltaT": 0.01}}})
    return g, o


GRIDS = {
    "surface4": dict(target_level=4),
    "surface3_wake2_leaf1": dict(target_level=3, wake_level=2, n_leaf=1),
    "surface5_wake3": dict(root=(20, 4, 3), target_level=5, wake_level=3),
}


@pytest.mark.parametrize("name", list(GRIDS))
def test_surface_refined_octree_and_bluff_cut_match_oracle(name):
    cfg = bluff_body(**GRIDS[name])
    _say(f"bluff grid {name}")
    g, o = _grids(cfg)
    ht = g.topology_host()
    t = o.topo
    assert g.num_cells == len(t.level) and g.lmax == t.lmax
    np.testing.assert_array_equal(ht["level"], t.level)
    np.testing.assert_array_equal(ht["ijk"], t.ijk)
    F_all = g.num_all_internal_faces
    assert F_all == len(t.owner)
    np.testing.assert_array_equal(ht["bnd_owner"], t.bnd_owner)
    np.testing.assert_array_equal(ht["bnd_dir"], t.bnd_dir)
    # the cut
    assert g.num_immersed_faces == o.num_immersed > 0
    np.testing.assert_array_equal(_np(g.ap_is_immersed_faces), o.is_immersed)
    np.testing.assert_array_equal(_np(g.ap_owner_near_boundary), o.ap.owner_near)
    np.testing.assert_array_equal(_np(g.ap_neighbour_near_boundary), o.ap.neighbour_near)
    np.testing.assert_array_equal(_np(g.ap_owner_bnd_patch_id), o.ap.owner_patch)
    h = float(o.cell_sizes.min())
    np.testing.assert_allclose(_np(g.ap_dist_owner_to_bnd), o.ap.dist_owner, rtol=0, atol=1e-12 * h)
    np.testing.assert_allclose(_np(g.ap_dist_neighbour_to_bnd), o.ap.dist_neighbour, rtol=0, atol=1e-12 * h)
    np.testing.assert_allclose(_np(g.ap_owner_normal), o.ap.owner_normal, rtol=0, atol=1e-8)
    np.testing.assert_allclose(_np(g.ap_neighbour_normal), o.ap.neighbour_normal, rtol=0, atol=1e-8)
    F = g.num_internal_faces
    np.testing.assert_array_equal(_np(g.owner_i32[:F]), o.owner)
    np.testing.assert_array_equal(_np(g.neighbour_i32[:F]), o.neighbour)
    # the surface is resolved at the target level: the hit lies in one of a
    # cut face's two cells, whose centre is then within half its diagonal
    # of the surface
    lev = _np(g.level)[: g.num_cells]
    Bd, Fi = g.num_domain_bnd_faces, g.num_immersed_faces
    own, nbr = _np(g.bnd_owner[Bd:Bd + Fi]), _np(g.bnd_owner[Bd + Fi:Bd + 2 * Fi])
    finest = GRIDS[name]["target_level"] + GRIDS[name].get("n_leaf", 0)
    assert np.all(np.maximum(lev[own], lev[nbr]) == finest)


def test_bluff_piso_steps_match_oracle():
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config

    cfg = bluff_body(target_level=4, mixed=False)
    sol = cfg["simulator"]["fvSolution"]["solvers"]  # solves far below the compared differences
    sol["U"].update({"tolerance": 1e-13, "rel_tolerance": 0.0, "max_iter": 2000})
    sol["p"].update({"tolerance": 0.0, "rel_tolerance": 1e-13, "max_iter": 20000})
    _say("bluff body PISO, 2 steps, surface level 4")
    grid, alg = create_case(load_config(cfg))
    case = fo.Case(cfg)
    assert grid.num_cells == case.g.num_cells and grid.num_immersed_faces == case.g.num_immersed > 0
    for _ in range(2):
        alg.step()
        case.step()
    assert _rel(_np(alg.U.data), case.U.data) < 1e-10
    assert _rel(_np(alg.p.data), case.p.data) < 1e-10
    assert _rel(_np(alg.phi.bnd_data), case.phi.bnd) < 1e-10


def test_full_case_grid_has_the_predicted_size():
    """21.0M cells: the oracle's build before the 3 leaf refinements times 512."""
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.grid import OctreeGrid

    cfg = case_dict("bluff3d_piso.yaml")
    fx = dict(cfg["fluxel"])
    pre = fo.build_octree(fx["domain"]["lower"], fx["domain"]["upper"], fx["root_resolution"],
                          [(r["min"], r["max"], r["level"]) for r in fx["refinement_regions"]], 0,
                          fx["immersed_bodies"], fx["target_level"])
    _say(f"bluff full grid: {len(pre.level)} leaves before leaf refinement")
    g = OctreeGrid(FluxelConfig.model_validate(fx), dt=cfg["simulator"]["control"]["deltaT"])
    assert g.num_cells == 512 * len(pre.level) == 21010432
    assert g.lmax == 7 and g.num_immersed_faces > 100000
    lev = _np(g.level)[: g.num_cells]
    np.testing.assert_array_equal(np.bincount(lev)[3:], 512 * np.bincount(pre.level))
