"""GPU: the Krylov solves at the edges the reference's solver tests exercise
(solvers/base.py:140-162, cg.py:75-174, bicgstab.py:76-272): two cells, a
27-cell grid (one partial slice of 64 rows), odd ragged octrees, the
zero-iteration exit and the strict threshold, an exhausted iteration budget,
the L-infinity norm and three components sharing one matrix, each against
the FP64 oracle on the same inputs."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from oracle import fv_oracle as fo
from tests.cases import FLUXEL, fluxel_block

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _grid(name: str):
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.grid import OctreeGrid

    _say(f"grid {name}")
    g = OctreeGrid(FluxelConfig.model_validate(fluxel_block(name)), dt=0.01)
    f = FLUXEL[name]
    o = fo.Grid(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"], dt=0.01)
    return g, o


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _systems(g, o, shift: float, seed: int, comps=()):
    """-laplacian(gamma) + shift * V on both sides (SPD for shift > 0; the
    reference cell's diagonal doubled otherwise), random right-hand side."""
    from gridfoam_amd.fields import CellField, NeumannBC
    from gridfoam_amd.fv import fvm

    rng = np.random.default_rng(seed)
    P = CellField(g, "q", comps)
    P.add_boundary_conditions({k: NeumannBC(0.0 if comps == () else [0.0, 0.0, 0.0]) for k in
                               ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")})
    zero = 0.0 if comps == () else np.zeros(3)
    oP = fo.Field("q", np.zeros((g.num_cells, *comps)), {k: fo.BC("neumann", zero) for k in fo.PATCHES})
    gamma = 0.5 + rng.random(g.num_cells)
    b = rng.standard_normal((g.num_cells, *comps))
    m = fvm.assemble(P, lap=-1.0, gamma=torch.tensor(gamma, device=g.device), corrected=False)
    om = -fo.fvm_laplacian(o, gamma, oP, False)
    V = _np(g.cell_volumes)
    if shift > 0:
        m.diag = m.diag + shift * g.cell_volumes
        om.diag = om.diag + shift * V
    else:
        m.diag[0] = 2 * m.diag[0]
        om.diag[0] *= 2
    m.source = torch.tensor(b, device=g.device)
    om.source = b.copy()
    return P, m, om


def _solve_both(P, m, om, method, mixed=False, **kw):
    from gridfoam_amd.config import SolverConfig
    from gridfoam_amd.fvmatrix import Equation
    from gridfoam_amd.solvers import create_solver

    cfg = dict(method=method, preconditioner="jacobi", tolerance=0.0, rel_tolerance=1e-10, max_iter=500)
    cfg.update(kw)
    res = create_solver(SolverConfig(mixed_precision=mixed, **cfg)).solve(Equation(P, m))
    ocfg = fo.SolverCfg(method, cfg["preconditioner"], cfg["tolerance"], cfg["rel_tolerance"], cfg["max_iter"],
                        cfg.get("norm_type", "l_2"))
    ox, ost = fo.solve(om, np.zeros_like(om.source), ocfg)
    return res, ox, ost


@pytest.mark.parametrize("name", ["pair", "cube3", "odd"])
@pytest.mark.parametrize("method,mixed", [("cg", False), ("cg", True), ("bicgstab", False)])
def test_small_and_ragged_grids_match_oracle(name, method, mixed):
    g, o = _grid(name)
    _say(f"edge solve {name} {method} mixed={mixed} cells={g.num_cells}")
    # BiCGSTAB gets the shifted operator: on the pinned Neumann operator of
    # "odd" it exhausts its restarts, in the oracle too
    P, m, om = _systems(g, o, shift=0.0 if method == "cg" else 1.0, seed=7)
    res, ox, ost = _solve_both(P, m, om, method, mixed)
    st = res.stats[0]
    assert st.converged and ost[0].converged
    assert _rel(_np(res.solution), ox) < (1e-6 if mixed else 1e-8)
    assert abs(st.initial_residual - ost[0].initial_residual) <= 1e-12 * ost[0].initial_residual
    if method == "cg" and not mixed:
        assert abs(st.iterations - ost[0].iterations) <= 1
    elif method == "bicgstab":
        # BiCGSTAB's residual history is not monotone: on the ill-conditioned
        # 1752-cell operator the rounding of the two sides' reductions moves
        # the exit by a few iterations (the solution still agrees to 1e-8)
        assert abs(st.iterations - ost[0].iterations) <= max(1, ost[0].iterations // 10)


@pytest.mark.parametrize("mixed", [False, True])
def test_threshold_is_strict_and_zero_iteration_exit(mixed):
    """Identity matrix, source 1 on 16 cells: |r0| = 4.  atol 4 + 1e-12
    exits before the first iteration with x0 untouched; atol 4 is not met
    by |r0| = 4 (strict <), one iteration solves it."""
    from gridfoam_amd.config import SolverConfig
    from gridfoam_amd.fields import CellField
    from gridfoam_amd.fvmatrix import Equation, FvMatrix
    from gridfoam_amd.solvers import create_solver

    g, _ = _grid("uniform2d")
    P = CellField(g, "q_id", ())
    m = FvMatrix(P)
    m.diag = torch.ones(g.num_cells, dtype=torch.float64, device=g.device)
    m.source = torch.ones(g.num_cells, dtype=torch.float64, device=g.device)
    _say(f"threshold semantics mixed={mixed}")
    for atol, iters in ((4.0 + 1e-12, 0), (4.0, 1)):
        cfg = SolverConfig(method="cg", preconditioner="none", tolerance=atol, rel_tolerance=0.0, max_iter=10,
                           mixed_precision=mixed)
        res = create_solver(cfg).solve(Equation(P, m))
        st = res.stats[0]
        assert st.iterations == iters and st.converged
        assert st.initial_residual == pytest.approx(4.0, rel=1e-15)
        expect = 0.0 if iters == 0 else 1.0
        torch.testing.assert_close(res.solution, torch.full_like(m.source, expect), atol=1e-12, rtol=0)


def test_exhausted_budget_reports_unconverged_and_matches_oracle_iterate():
    """max_iter = 5 on the walls8 pressure operator: five iterations, not
    converged, and the fifth iterate and its residual equal the oracle's."""
    g, o = _grid("walls8")
    P, m, om = _systems(g, o, shift=0.0, seed=11)
    _say("cg budget exhaustion")
    res, ox, ost = _solve_both(P, m, om, "cg", max_iter=5, rel_tolerance=1e-12)
    st = res.stats[0]
    assert st.iterations == 5 and not st.converged
    assert ost[0].iterations == 5 and not ost[0].converged
    assert _rel(_np(res.solution), ox) < 1e-10
    assert st.final_residual == pytest.approx(ost[0].final_residual, rel=1e-8)


@pytest.mark.parametrize("method", ["cg", "bicgstab"])
def test_linf_norm_matches_oracle(method):
    g, o = _grid("refined3d")
    P, m, om = _systems(g, o, shift=0.0, seed=13)
    _say(f"l_inf norm {method}")
    res, ox, ost = _solve_both(P, m, om, method, norm_type="l_inf", rel_tolerance=1e-9)
    st = res.stats[0]
    assert st.converged and ost[0].converged
    assert abs(st.iterations - ost[0].iterations) <= 1
    assert st.initial_residual == pytest.approx(ost[0].initial_residual, rel=1e-12)
    assert _rel(_np(res.solution), ox) < 1e-8


def test_three_component_cg_matches_per_component_oracle():
    """One launch sequence for three components sharing the matrix
    (solvers/krylov.py:34-82) equals three separate oracle solves."""
    g, o = _grid("refined3d")
    P, m, om = _systems(g, o, shift=1.0, seed=17, comps=(3,))
    _say("cg three components")
    res, ox, ost = _solve_both(P, m, om, "cg")
    assert len(res.stats) == 3
    x = _np(res.solution)
    for r in range(3):
        assert res.stats[r].converged
        assert abs(res.stats[r].iterations - ost[r].iterations) <= 1
        assert _rel(x[:, r], ox[:, r]) < 1e-8


@pytest.mark.parametrize("name", ["odd", "walls8"])
def test_sliced_ell_layout(name):
    """The operator's sliced-ELL arrays, bit for bit: each slice of 64
    Morton-ordered rows is as wide as its longest row; row i holds the cells
    of its face neighbours in ascending order with the face codes (f: upper,
    -f-1: lower), then padding (its own row, code INT32_MIN)."""
    g, o = _grid(name)
    _say(f"operator layout {name}")
    C = g.num_cells
    own, nbr = o.owner, o.neighbour
    length = np.bincount(own, minlength=C) + np.bincount(nbr, minlength=C)
    ns = (C + 63) // 64
    lens = np.zeros(ns * 64, np.int64)
    lens[:C] = length
    widths = lens.reshape(ns, 64).max(axis=1)
    ptr = _np(g.sell_ptr).astype(np.int64)
    np.testing.assert_array_equal(np.diff(ptr), 64 * widths)
    col = _np(g.sell_col).astype(np.int64)
    code = _np(g.sell_code).astype(np.int64)
    assert g.mesh.nnz_pad == ptr[-1] == col.size
    nb_of = [[] for _ in range(C)]
    for f, (a, b) in enumerate(zip(own, nbr)):
        nb_of[a].append((b, f))
        nb_of[b].append((a, -f - 1))
    for s in range(ns):
        W = widths[s]
        for lane in range(64):
            i = s * 64 + lane
            if i >= C:
                continue
            ent = sorted(nb_of[i])
            e = ptr[s] + np.arange(W) * 64 + lane
            got = list(zip(col[e], code[e]))
            assert got[:len(ent)] == ent, (i, got, ent)
            assert all(c == i and k == np.iinfo(np.int32).min for c, k in got[len(ent):]), (i, got)


def test_assembly_with_a_grad_requiring_diffusivity_equals_the_plain_assembly():
    """A diffusivity that requires gradients goes through the differentiable
    assembly node; its arrays equal the plain assembly's bit for bit."""
    from gridfoam_amd.fields import CellField
    from gridfoam_amd.fv import fvm

    g, _ = _grid("refined3d")
    P = CellField(g, "q_grad", ())
    gamma = torch.linspace(0.5, 1.5, g.num_cells, dtype=torch.float64, device=g.device).requires_grad_()
    _say("assembly with grad-requiring diffusivity")
    m = fvm.assemble(P, lap=-1.0, gamma=gamma)
    assert m.diag.requires_grad and m.upper.requires_grad
    m2 = fvm.assemble(P, lap=-1.0, gamma=gamma.detach())
    assert torch.equal(m.diag, m2.diag) and torch.equal(m.upper, m2.upper) and torch.equal(m.source, m2.source)
    assert torch.equal(m.face_flux_correction, m2.face_flux_correction)
