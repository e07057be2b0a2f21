"""GPU: the gf_mesh face-order contract at the boundary that replaces
GridBase's face lists (core/grid/base.py:19-275).

A grid handed over by another mesher (fluxel's cell and face order is not
specified) may list its cells and faces in any order.  ``OctreeGrid(...,
topology=...)`` brings the faces into the contract (grouped by owner,
ascending; a stable sort), and ``gf_mesh_validate`` rejects a gf_mesh that
breaks it with GF_EINVAL instead of assembling silently wrong operators.
"""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch

from oracle import fv_oracle as fo
from tests.cases import FLUXEL, fluxel_block

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b) -> float:
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _shuffled_topology(t: dict, seed: int):
    """Cells and faces in random order: cell i of the result is cell cp[i]
    of ``t``, face j is face fp[j], domain face k is domain face bp[k]."""
    rng = np.random.default_rng(seed)
    C, F, B = len(t["level"]), len(t["owner"]), len(t["bnd_owner"])
    cp, fp, bp = rng.permutation(C), rng.permutation(F), rng.permutation(B)
    inv = np.empty(C, np.int64)
    inv[cp] = np.arange(C)
    out = {"level": t["level"][cp], "ijk": t["ijk"][cp],
           "owner": inv[t["owner"][fp]], "neighbour": inv[t["neighbour"][fp]], "axis": t["axis"][fp],
           "bnd_owner": inv[t["bnd_owner"][bp]], "bnd_dir": t["bnd_dir"][bp]}
    return out, cp, fp, bp


@pytest.mark.parametrize("name", ["refined3d", "walls8", "odd"])
def test_shuffled_cells_and_faces_give_the_oracle_assembly_and_solve(name):
    from gridfoam_amd.config import FluxelConfig, SolverConfig
    from gridfoam_amd.fields import CellField, NeumannBC
    from gridfoam_amd.fv import fvm
    from gridfoam_amd.fvmatrix import Equation
    from gridfoam_amd.grid import OctreeGrid
    from gridfoam_amd.solvers import create_solver

    _say(f"face order {name}")
    f = FLUXEL[name]
    o = fo.Grid(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"], dt=0.01)
    t = {k: np.asarray(getattr(o.topo, k)).astype(np.int64) for k in
         ("level", "ijk", "owner", "neighbour", "axis", "bnd_owner", "bnd_dir")}
    ts, cp, fp, bp = _shuffled_topology(t, seed=7)
    g = OctreeGrid(FluxelConfig.model_validate(fluxel_block(name)), dt=0.01, topology=ts)
    # new face k is given face face_order[k], i.e. oracle face fp[face_order[k]]
    fmap = fp[_np(g.face_order)]
    bmap = bp[_np(g.bnd_face_order)]
    np.testing.assert_array_equal(np.diff(_np(g.owner)) >= 0, True)
    inv = np.empty(len(cp), np.int64)
    inv[cp] = np.arange(len(cp))
    np.testing.assert_array_equal(_np(g.owner), inv[o.owner[fmap]])
    np.testing.assert_array_equal(_np(g.domain_bnd_owner), inv[o.bnd_owner[bmap]])

    rng = np.random.default_rng(3)
    gamma = 0.5 + rng.random(g.num_cells)
    b = rng.standard_normal(g.num_cells)
    b -= b.mean()
    P = CellField(g, "p", ())
    P.add_boundary_conditions({k: NeumannBC(0.0) for k in fo.PATCHES})
    oP = fo.Field("p", np.zeros(g.num_cells), {k: fo.BC("neumann", 0.0) for k in fo.PATCHES})
    pm = fvm.assemble(P, lap=-1.0, gamma=torch.tensor(gamma[cp], device=g.device), corrected=True)
    opm = -fo.fvm_laplacian(o, gamma, oP, True)
    assert _rel(_np(pm.upper), opm.upper[fmap]) < 1e-14
    assert _rel(_np(pm.lower), opm.lower[fmap]) < 1e-14
    assert _rel(_np(pm.diag), opm.diag[cp]) < 1e-13
    assert _rel(_np(pm.source), opm.source[cp]) < 1e-12
    if g.num_hanging_faces:
        assert _rel(_np(pm.face_flux_correction), opm.ffc[fmap]) < 1e-12
    # reference cell 0 of the oracle is cell inv[0] here
    pm.diag[int(inv[0])] = 2 * pm.diag[int(inv[0])]
    pm.source = torch.tensor(b[cp], device=g.device)
    opm.diag[0] *= 2
    opm.source = b.copy()
    cfg = SolverConfig(method="cg", preconditioner="jacobi", tolerance=0.0, rel_tolerance=1e-10, max_iter=2000)
    res = create_solver(cfg).solve(Equation(P, pm))
    ox, ost = fo.solve(opm, np.zeros(g.num_cells), fo.SolverCfg("cg", "jacobi", 0.0, 1e-10, 2000))
    assert res.stats[0].converged
    assert _rel(_np(res.solution), ox[cp]) < 1e-8
    assert abs(res.stats[0].iterations - ost[0].iterations) <= 1


def test_validation_rejects_a_mesh_whose_faces_are_not_grouped_by_owner():
    from gridfoam_amd import _native as nat
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.grid import OctreeGrid

    _say("face order validation")
    g = OctreeGrid(FluxelConfig.model_validate(fluxel_block("refined3d")), dt=0.01)
    flags = ctypes.c_int32(-1)
    nat.call("gf_mesh_validate", g.mesh_ref, ctypes.byref(flags), nat.stream())
    assert flags.value == 0
    F, B = g.num_internal_faces, g.num_domain_bnd_faces
    rng = np.random.default_rng(11)
    fp = torch.tensor(rng.permutation(F), device=g.device)
    bp = torch.tensor(rng.permutation(B), device=g.device)
    own = g.owner_i32[:F][fp].contiguous()
    nbr = g.neighbour_i32[:F][fp].contiguous()
    ax = g.axis_i32[:F][fp].contiguous()
    bown = g.bnd_owner_i32[:B][bp].contiguous()
    bdir = g.bnd_dir_i32[:B][bp].contiguous()
    m = nat.GfMesh()
    ctypes.pointer(m)[0] = g.mesh  # copy of the valid mesh, then shuffled face lists
    m.owner, m.neighbour, m.axis = nat.ptr(own), nat.ptr(nbr), nat.ptr(ax)
    m.bnd_owner, m.bnd_dir = nat.ptr(bown), nat.ptr(bdir)
    with pytest.raises(RuntimeError, match="grouped by owner"):
        nat.call("gf_mesh_validate", ctypes.byref(m), ctypes.byref(flags), nat.stream())
    assert flags.value & 2 and flags.value & 8
    # swapping owner and neighbour breaks the +axis side rule
    m2 = nat.GfMesh()
    ctypes.pointer(m2)[0] = g.mesh
    m2.owner, m2.neighbour = g.mesh.neighbour, g.mesh.owner
    m2.own_row = None
    with pytest.raises(RuntimeError, match="neighbour not on the"):
        nat.call("gf_mesh_validate", ctypes.byref(m2), ctypes.byref(flags), nat.stream())
    assert flags.value & 16
