"""GPU: batched steady SIMPLE (gridfoam_amd.batching).

Several serving cases run as one block-diagonal mesh.  The segmented PCG
keeps each case's CG scalars, threshold and stop, so a case must take the
same iteration count as its own CG and reach the same solution; the batched
SIMPLE iteration must give a case the same fields whatever other cases share
its batch (bit for bit), and the fields of the case's own SIMPLE driver to
rounding; the server must retire each case with the state and iteration
count at which it stops running alone."""

from __future__ import annotations

import copy
import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _laplacians(grids, seed, shifts):
    """A diagonally shifted variable-coefficient Laplacian per grid and a
    random right-hand side."""
    from gridfoam_amd.config import DomainBoundaryPatch
    from gridfoam_amd.fields import CellField, NeumannBC
    from gridfoam_amd.fv import fvm

    gen = torch.Generator().manual_seed(seed)
    out = []
    for k, (g, shift) in enumerate(zip(grids, shifts)):
        q = CellField(g, f"q{k}", ())
        q.add_boundary_conditions({p: NeumannBC(0.0) for p in DomainBoundaryPatch})
        gam = (torch.rand(g.num_cells, dtype=torch.float64, generator=gen) + 0.5).to(g.device)
        m = fvm.assemble(q, lap=-1.0, gamma=gam, corrected=False)
        m.diag = m.diag.clone() * (1.0 + shift) + 1e-3
        m.source = torch.randn(g.num_cells, dtype=torch.float64, generator=gen).to(g.device)
        out.append((q, m))
    return out


def _combined(grids, mats):
    from gridfoam_amd.batching import concat_grids
    from gridfoam_amd.fields import CellField
    from gridfoam_amd.fvmatrix import FvMatrix

    g, _ = concat_grids(grids, grids[0].sim_config)
    M = FvMatrix(CellField(g, "qq", ()))
    M.diag = torch.cat([m.diag for _, m in mats])
    F = g.num_internal_faces
    M._upper[:F] = torch.cat([m.upper for _, m in mats])
    M._lower[:F] = torch.cat([m.lower for _, m in mats])
    M.source = torch.cat([m.source for _, m in mats])
    rows = [0]
    for k in grids:
        rows.append(rows[-1] + k.num_cells)
    seg_row = torch.tensor(rows, dtype=torch.int32, device=g.device)
    row_seg = torch.repeat_interleave(torch.arange(len(grids), dtype=torch.int32, device=g.device),
                                      torch.tensor([k.num_cells for k in grids], device=g.device))
    return M, rows, seg_row, row_seg


def _box(n, lo=(0.0, 0.0, 0.0)):
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.grid import OctreeGrid

    return OctreeGrid(FluxelConfig.model_validate({"domain": {"lower": list(lo), "upper": [1, 1, 1]},
                                                   "root_resolution": [n, n, n]}), dt=1.0)


@pytest.mark.parametrize("rtol,max_iter", [(1e-10, 5000), (1e-10, 25)])
def test_segmented_cg_follows_each_case_cg(rtol, max_iter):
    """Four segments of different sizes and conditioning: every segment's
    iteration count, convergence flag and solution equal those of the
    single-case Jacobi PCG on its own matrix (1e-12 relative), also when
    the iteration budget stops them."""
    from gridfoam_amd.batching import SegmentedCG
    from gridfoam_amd.config import SolverConfig
    from gridfoam_amd.fvmatrix import Equation
    from gridfoam_amd.solvers import create_solver

    _say(f"segmented cg rtol {rtol} max_iter {max_iter}")
    grids = [_box(n) for n in (9, 17, 12, 24)]
    mats = _laplacians(grids, 3, [0.05, 0.002, 0.3, 0.01])
    M, rows, seg_row, row_seg = _combined(grids, mats)
    x, st = SegmentedCG(0.0, rtol, max_iter).solve(M, M.source, torch.zeros_like(M.source), seg_row, rows, row_seg)
    cg = create_solver(SolverConfig(method="cg", preconditioner="jacobi", tolerance=0.0, rel_tolerance=rtol,
                                    max_iter=max_iter))
    counts = []
    for k, (q, m) in enumerate(mats):
        r = cg.solve(Equation(q, m))
        s = r.stats[0]
        xs = x[rows[k]:rows[k + 1]]
        rel = float(torch.linalg.vector_norm(xs - r.solution) / torch.linalg.vector_norm(r.solution))
        _say(f"segment {k}: {st.iterations[k]} / {s.iterations} iterations, rel {rel:.2e}")
        assert st.iterations[k] == s.iterations
        assert st.converged[k] == s.converged
        assert rel < 1e-12
        assert abs(st.initial_residual[k] - s.initial_residual) <= 1e-12 * s.initial_residual
        counts.append(st.iterations[k])
    if max_iter > 100:
        assert len(set(counts)) > 1, "segments should stop at different iterations"
    else:
        assert all(c == max_iter for c in counts) and not any(st.converged)


def _convection(grids, seed, shifts):
    """An upwind convection-diffusion operator per grid (random face flux and
    diffusivity, diagonal shifted): the shape of a momentum matrix."""
    from gridfoam_amd.config import DivScheme, DomainBoundaryPatch
    from gridfoam_amd.fields import CellField, FaceField, NeumannBC
    from gridfoam_amd.fv import fvm

    gen = torch.Generator().manual_seed(seed)
    out = []
    for k, (g, shift) in enumerate(zip(grids, shifts)):
        q = CellField(g, f"c{k}", ())
        q.add_boundary_conditions({p: NeumannBC(0.0) for p in DomainBoundaryPatch})
        phi = FaceField(g, f"phi_c{k}")
        phi.single_data = (0.3 * torch.randn(g.num_internal_faces, dtype=torch.float64, generator=gen)).to(g.device)
        gam = (torch.rand(g.num_cells, dtype=torch.float64, generator=gen) * 0.05 + 0.01).to(g.device)
        m = fvm.assemble(q, div=1.0, lap=-1.0, phi=phi, gamma=gam, scheme=DivScheme.UPWIND, corrected=False)
        m.diag = m.diag.clone() * (1.0 + shift) + 1e-3
        out.append((q, m))
    return out


@pytest.mark.parametrize("rtol,max_iter", [(1e-10, 2000), (1e-10, 4)])
def test_segmented_bicgstab_follows_each_case_bicgstab(rtol, max_iter):
    """Three segments of an asymmetric operator with three right-hand sides
    each: every (segment, component) takes the iterations of the single-case
    BiCGSTAB on that component (to two iterations) and reaches its solution
    (1e-8 relative), also when the iteration budget stops it."""
    from gridfoam_amd.batching import SegmentedBiCGSTAB
    from gridfoam_amd.config import SolverConfig
    from gridfoam_amd.fvmatrix import Equation
    from gridfoam_amd.solvers import create_solver

    _say(f"segmented bicgstab rtol {rtol} max_iter {max_iter}")
    grids = [_box(n) for n in (10, 15, 12)]
    mats = _convection(grids, 7, [0.05, 0.01, 0.2])
    M, rows, _, _ = _combined(grids, mats)
    gen = torch.Generator().manual_seed(8)
    B = torch.randn((rows[-1], 3), dtype=torch.float64, generator=gen).to(M.grid.device)
    X0 = 0.1 * torch.randn((rows[-1], 3), dtype=torch.float64, generator=gen).to(M.grid.device)
    x, st = SegmentedBiCGSTAB(0.0, rtol, max_iter, 5).solve(M, B, X0, rows)
    solver = create_solver(SolverConfig(method="bicgstab", preconditioner="jacobi", tolerance=0.0,
                                        rel_tolerance=rtol, max_iter=max_iter))
    for k, (q, m) in enumerate(mats):
        a, b = rows[k], rows[k + 1]
        for c in range(3):
            m.source = B[a:b, c].contiguous()
            q.data = X0[a:b, c].contiguous()
            r = solver.solve(Equation(q, m))
            s = r.stats[0]
            j = 3 * k + c
            ref = r.solution
            rel = float(torch.linalg.vector_norm(x[a:b, c] - ref) / torch.linalg.vector_norm(ref))
            _say(f"segment {k} component {c}: {st.iterations[j]} / {s.iterations} iterations, rel {rel:.2e}")
            # BiCGSTAB's residual is not monotone: over ~150 iterations the
            # order of the dot products' sums can move the crossing of the
            # threshold by an iteration or two
            assert abs(st.iterations[j] - s.iterations) <= 2
            assert st.converged[j] == s.converged
            assert abs(st.initial_residual[j] - s.initial_residual) <= 1e-12 * s.initial_residual
            assert rel < 1e-8


def _requests(n, seed, n_range=(16, 22), residual=1e-6, max_iterations=2000):
    from gridfoam_amd.serving import random_case

    rng = np.random.default_rng(seed)
    return [random_case(rng, n_range=n_range, mixed=False, residual=residual, max_iterations=max_iterations)
            for _ in range(n)]


def test_batched_simple_is_independent_of_the_batch_and_matches_the_driver():
    """15 batched SIMPLE iterations of three cases with obstacles: each case's
    U and p equal those of the same case batched alone bit for bit, and
    those of its own SIMPLE driver to 1e-10 (relative)."""
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.batching import Batch, new_case
    from gridfoam_amd.config import load_config

    raws = _requests(3, 3)
    for r in raws:
        r["simulator"]["fvSolution"]["algorithm"]["residualControl"] = {}
        # momentum solved far below rounding of the fields: the driver's
        # BiCGSTAB and the segmented one then agree whatever their sums' order
        r["simulator"]["fvSolution"]["solvers"]["U"].update(tolerance=1e-14, rel_tolerance=1e-13)
    _say("batched simple: mix of 3")
    mix = Batch([new_case(k, load_config(r), r) for k, r in enumerate(raws)])
    assert sum(c.grid.num_immersed_faces for c in mix.cases) > 0
    for _ in range(15):
        mix.step()
    for k, r in enumerate(raws):
        _say(f"batched simple: case {k} alone and through its driver")
        one = Batch([new_case(k, load_config(r), r)])
        for _ in range(15):
            one.step()
        one.extract(0)
        mix.extract(k)
        assert torch.equal(mix.cases[k].U, one.cases[0].U)
        assert torch.equal(mix.cases[k].p, one.cases[0].p)
        assert mix.cases[k].p_iterations == one.cases[0].p_iterations
        _, alg = create_case(load_config(copy.deepcopy(r)))
        for _ in range(15):
            alg.step()
        for a, b in ((one.cases[0].U, alg.U.data), (one.cases[0].p, alg.p.data)):
            rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
            assert rel < 1e-10, rel


def test_server_retires_each_case_with_its_alone_result():
    """Five requests arriving over time through a server of at most three
    cases: every result is converged, took the iteration count of the case
    batched alone and holds its fields bit for bit."""
    from gridfoam_amd.batching import Batch, BatchServer, new_case
    from gridfoam_amd.config import load_config

    raws = _requests(5, 11, n_range=(12, 18), residual=1e-4, max_iterations=600)
    reqs = [(k, r, 0.02 * k) for k, r in enumerate(raws)]
    _say("batch server: 5 requests, 3 slots")
    res = {r.rid: r for r in BatchServer(max_cases=3, chunk=5).run(reqs)}
    assert sorted(res) == list(range(5))
    for k, r in enumerate(raws):
        _say(f"batch server: request {k} alone")
        one = Batch([new_case(k, load_config(r), r)])
        its = 0
        while True:
            conv = one.step()
            its += 1
            if conv or its >= 600:
                break
        one.extract(0)
        got = res[k]
        assert got.converged and conv
        assert got.iterations == its
        assert torch.equal(got.U, one.cases[0].U.cpu())
        assert torch.equal(got.p, one.cases[0].p.cpu())
        assert got.end >= got.start >= got.arrival - 1e-9


def test_two_lane_server_matches_alone_runs():
    """Requests at or above the size threshold run alone in two lanes on
    their own streams while the rest are batched on a third: a lane's result
    equals the SIMPLE driver run alone (FP64 requests: bit for bit), a
    batched one the case batched alone."""
    from gridfoam_amd.batching import Batch, BatchServer, base_cells, new_case
    from gridfoam_amd.config import load_config
    from gridfoam_amd.serving import random_case, solve_case

    rng = np.random.default_rng(21)
    raws = [random_case(rng, n_range=(12, 20), mixed=False, residual=1e-4, max_iterations=40) for _ in range(6)]
    large = 16 ** 3
    kinds = [base_cells(r) >= large for r in raws]
    assert any(kinds) and not all(kinds)
    _say(f"two-lane server: {sum(kinds)} alone, {len(kinds) - sum(kinds)} batched")
    res = {r.rid: r for r in BatchServer(max_cases=4, chunk=5, large_cells=large, large_lanes=2,
                                         max_bytes=8 << 30).run([(k, r, 0.0) for k, r in enumerate(raws)])}
    for k, r in enumerate(raws):
        got = res[k]
        if kinds[k]:
            cells, U, p, its, conv, _ = solve_case(r)
        else:
            one = Batch([new_case(k, load_config(r), r)])
            its = 0
            while True:
                conv = bool(one.step())
                its += 1
                if conv or its >= 40:
                    break
            one.extract(0)
            U, p = one.cases[0].U.cpu(), one.cases[0].p.cpu()
        assert got.iterations == its and got.converged == conv
        assert torch.equal(got.U, U) and torch.equal(got.p, p), k
