"""Pins of the oracle's axis-projected immersed boundary: the AP payload for
analytic bodies, the Gibou cell constraints, the symmetric fixed-value
elimination and the conservative closure of the pressure projection.

The plane cases restate the reference's own immersed tests
(tests/helpers/grids.py:40-70 builds the plane grid: a uniform [n, 1, 1]
grid of [0, L] x [0, L/10]^2 cut at x = (0.5 - 0.5/n + theta/n) L;
tests/integration/fv/test_immersed_dirichlet.py holds the assertions).  The
reference builds the plane from an STL through fluxel; here the same plane
is an analytic body.  The Fourier and box cuts are checked against what the
geometry fixes: the sign change of the outline along each cut ray, the box
faces' coordinates, and the near-boundary rule."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fv_oracle as fo
from tests.cases import box_cavity, fourier_channel

CG = fo.SolverCfg(method="cg", preconditioner="none", tolerance=1e-13, rel_tolerance=0.0, max_iter=1000)


def plane_grid(n: int, theta: float, scale: float = 1.0) -> fo.Grid:
    """helpers/grids.py:40-70."""
    x = (0.5 - 0.5 / n + theta / n) * scale
    g = fo.Grid([0, 0, 0], [scale, 0.1 * scale, 0.1 * scale], [n, 1, 1],
                bodies=[dict(name="plane", kind="plane", axis=0, offset=x)])
    assert g.num_immersed == (2 if theta == 0 else 1)
    return g


def scalar(g: fo.Grid, name: str, value) -> fo.Field:
    """test_immersed_dirichlet.py:28-45: zero-gradient sides, 0 and 1 at the
    x ends, ``value`` on the body."""
    bcs = {p: fo.BC("neumann", 0.0) for p in fo.PATCHES}
    bcs["x_minus"] = fo.BC("dirichlet", 0.0)
    bcs["x_plus"] = fo.BC("dirichlet", 1.0)
    bcs["plane"] = fo.BC("dirichlet", value)
    return fo.Field(name, np.zeros(g.num_cells), bcs)


@pytest.mark.parametrize("theta", [0.0, 0.03, 0.3, 0.97])
def test_plane_cuts_one_face_or_two_at_a_centre(theta):
    g = plane_grid(8, theta)
    C = g.num_cells
    assert g.num_faces == C - 1 - g.num_immersed
    assert g.num_bnd == g.num_domain_bnd + 2 * g.num_immersed
    up = slice(g.num_domain_bnd, g.num_domain_bnd + g.num_immersed)
    lo = slice(g.num_domain_bnd + g.num_immersed, g.num_bnd)
    # upper sides belong to the owner and face +x, lower sides -x
    assert np.all(g.bnd_Sf[up, 0] > 0) and np.all(g.bnd_Sf[lo, 0] < 0)
    assert np.all(g.bnd_owner[lo] == g.bnd_owner[up] + 1)
    # distances add up to the centre spacing
    np.testing.assert_allclose(g.ap.dist_owner + g.ap.dist_neighbour, 0.125, rtol=0, atol=1e-15)


def test_snapping_masks_and_distances_are_unit_invariant():
    """test_immersed_dirichlet.py:186-210."""
    a, b = plane_grid(8, 0.03), plane_grid(8, 0.03, 1000.0)
    np.testing.assert_array_equal(a.ap.owner_near, b.ap.owner_near)
    np.testing.assert_array_equal(a.ap.neighbour_near, b.ap.neighbour_near)
    np.testing.assert_allclose(a.ap.dist_owner, b.ap.dist_owner / 1000.0, atol=1e-7, rtol=0)
    assert float(a.ap.dist_owner[0]) < 0.01


@pytest.mark.parametrize("theta,count", [(0.3, 0), (0.03, 1), (0.0, 1)])
def test_gibou_rule_picks_the_near_cell(theta, count):
    """distance / dx <= dx / L (data_contracts.rst:190-194): at theta = 0.03
    the owner lies 0.03 h from the plane (<= h / L), at 0.3 it does not; at
    theta = 0 the plane runs through a centre (distance 0 on two sides)."""
    g = plane_grid(8, theta)
    q = scalar(g, "q", 0.7)
    cells, values = fo.immersed_constraints(g, q)
    assert len(cells) == count
    if count:
        assert cells.tolist() == [3]
        np.testing.assert_array_equal(values, [0.7])


def test_neumann_at_coincident_boundary_has_no_cell_constraint():
    """test_immersed_dirichlet.py:213-231."""
    g = plane_grid(8, 0.0)
    q = scalar(g, "q", 0.0)
    q.bcs["plane"] = fo.BC("neumann", 2.0)
    assert len(fo.immersed_constraints(g, q)[0]) == 0
    m = fo.fvm_laplacian(g, 1.0, q)
    assert np.isfinite(m.source).all()
    sg = fo.sn_grad_boundary(g, q)
    up = slice(g.num_domain_bnd, g.num_domain_bnd + g.num_immersed)
    np.testing.assert_array_equal(sg[up], np.full(g.num_immersed, 2.0))


def test_mixed_constraint_selection_tracks_flow_reversal():
    """test_immersed_dirichlet.py:86-110: inletOutlet constrains the near
    cell only while the immersed flux points into the fluid."""
    g = plane_grid(8, 0.03)
    q = fo.Field("q", np.zeros(g.num_cells), {"plane": fo.BC("inlet_outlet", np.asarray(0.7))})
    for flux, expected in ((1.0, 0), (-1.0, 1), (1.0, 0)):
        bnd = np.zeros(g.num_bnd)
        bnd[g.num_domain_bnd:] = flux
        phi = fo.FaceValues(np.zeros(g.num_faces), bnd)
        g.registry["phi"] = lambda phi=phi: phi
        cells, values = fo.immersed_constraints(g, q)
        assert len(cells) == expected
        np.testing.assert_array_equal(values, np.full(len(cells), 0.7))


def _gibou_error(n: int) -> float:
    h = 1.0 / n
    g = plane_grid(n, 0.4 * h)
    x_i = 0.5 - h / 2
    x_b = x_i + 0.4 * h * h
    q = scalar(g, "q", x_b ** 2)
    m = -fo.fvm_laplacian(g, 1.0, q)
    m.source = m.source - 2.0 * g.cell_volumes
    cells, values = fo.immersed_constraints(g, q)
    assert len(cells) == 1
    s = m.with_fixed_values(cells, values)
    np.testing.assert_array_equal(s.upper, s.lower)
    x, (st,) = fo.solve(s, np.zeros(g.num_cells), CG)
    assert st.converged
    np.testing.assert_allclose(x[cells], values, atol=1e-12, rtol=0)
    return float(np.abs(x - g.cell_centers[:, 0] ** 2).max())


def test_near_boundary_poisson_is_second_order():
    """test_immersed_dirichlet.py:113-144: a snapped Dirichlet solve stays
    symmetric and its error falls by more than 3x per halving of h."""
    e = [_gibou_error(n) for n in (8, 16, 32)]
    assert e[1] < 0.3 * e[0]
    assert e[2] < 0.3 * e[1]


@pytest.mark.parametrize("theta", [0.3, 0.03, 0.0])
def test_immersed_pressure_projection_conserves_every_cell(theta):
    """test_immersed_dirichlet.py:147-183: the corrected flux is divergence
    free in every cell (snapped cells through the area-weighted closure),
    and snapped pressure cells keep their prescribed value."""
    g = plane_grid(8, theta)
    p = scalar(g, "p", 0.7)
    r = np.linspace(0.5, 1.5, g.num_cells)
    rng = np.random.default_rng(7)
    pred = fo.FaceValues(rng.standard_normal(g.num_faces), np.zeros(g.num_bnd))
    pmat, (st,), _ = fo.solve_pressure(g, p, r, fo.fvc_div(g, pred), CG, None, 0, None)
    assert st.converged
    phi = fo.correct_phi(g, pred, pmat, p, r)
    np.testing.assert_allclose(fo.fvc_div(g, phi), 0.0, atol=1e-8, rtol=0)
    cells, values = fo.immersed_constraints(g, p)
    assert len(cells) == (0 if theta == 0.3 else 1)
    np.testing.assert_allclose(p.data[cells], values, atol=1e-12, rtol=0)
    assert np.isfinite(fo.lsq_grad(g, p)).all()
    assert np.isfinite(fo.sn_grad_boundary(g, p)).all()


def test_closure_is_what_makes_snapped_cells_conserve():
    """Without the closure the snapped cell's balance is left open: a
    planted omission of close_fixed_cell_flux is caught by the projection
    test above."""
    g = plane_grid(8, 0.03)
    p = scalar(g, "p", 0.7)
    r = np.linspace(0.5, 1.5, g.num_cells)
    pred = fo.FaceValues(np.random.default_rng(7).standard_normal(g.num_faces), np.zeros(g.num_bnd))
    pmat, _, _ = fo.solve_pressure(g, p, r, fo.fvc_div(g, pred), CG, None, 0, None)
    flux_p = (-pmat).flux(p.data)
    open_phi = fo.FaceValues(pred.single - flux_p, pred.bnd.copy())
    for b in fo.batches(g, p):
        f, rv, rg, _ = fo.bc_state(g, p, b)
        sng = fo.value_grad_coeff(b, f) * (rv - p.data[b.cells]) + (1.0 - f) * rg
        open_phi.bnd[b.mask] -= r[b.cells] * b.mag_Sf * sng
    assert abs(fo.fvc_div(g, open_phi)[3]) > 1e-3


def test_velocity_correction_keeps_prescribed_cell_values():
    """test_immersed_dirichlet.py:258-281."""
    g = plane_grid(8, 0.03)
    p = scalar(g, "p", 0.7)
    p.data = g.cell_centers[:, 0].copy()
    U = fo.Field("U", np.zeros((g.num_cells, 3)), {"plane": fo.BC("dirichlet", np.array([1.0, 2.0, 3.0]))})
    U.data = -fo.lsq_grad(g, p)
    fo.apply_immersed_values(g, U)
    cells, values = fo.immersed_constraints(g, U)
    assert len(cells) == 1
    np.testing.assert_array_equal(U.data[cells], values)


def test_fixed_value_elimination_solves_the_reduced_system():
    """fvmatrix.py:325-394 against a dense solve: fixed rows return their
    values, the free rows satisfy A_ff x_f = b_f - A_fc v, and the
    eliminated matrix stays symmetric."""
    g = fo.Grid([0, 0, 0], [1, 1, 1], [4, 4, 2], [([0.0, 0.0, 0.0], [0.5, 0.5, 1.0], 1)])
    C = g.num_cells
    rng = np.random.default_rng(3)
    m = fo.Matrix(g)
    m.upper = -rng.uniform(0.5, 1.5, g.num_faces)
    m.lower = m.upper.copy()
    m.diag = -(fo._scatter(g.owner, m.upper, C) + fo._scatter(g.neighbour, m.lower, C)) + 0.1
    m.source = rng.standard_normal(C)
    cells = np.array([0, 5, 17, C - 1])
    vals = rng.standard_normal(4)
    s = m.with_fixed_values(cells, vals)
    np.testing.assert_array_equal(s.upper, s.lower)
    A = np.diag(m.diag)
    np.add.at(A, (g.owner, g.neighbour), m.upper)
    np.add.at(A, (g.neighbour, g.owner), m.lower)
    free = np.setdiff1d(np.arange(C), cells)
    x_f = np.linalg.solve(A[np.ix_(free, free)], m.source[free] - A[np.ix_(free, cells)] @ vals)
    x = np.linalg.solve(np.diag(s.diag) + _offdiag(g, s), s.source)
    np.testing.assert_allclose(x[cells], vals, rtol=0, atol=1e-12)
    np.testing.assert_allclose(x[free], x_f, rtol=1e-10, atol=1e-12)


def _offdiag(g, m):
    A = np.zeros((g.num_cells, g.num_cells))
    np.add.at(A, (g.owner, g.neighbour), m.upper)
    np.add.at(A, (g.neighbour, g.owner), m.lower)
    return A


def test_zero_diagonal_row_gets_the_sign_of_the_diagonal_sum():
    g = fo.Grid([0, 0, 0], [1, 1, 1], [4, 1, 1])
    m = fo.Matrix(g)
    m.diag = np.array([-2.0, 0.0, -2.0, -1.0])
    m.upper = np.array([1.0, 1.0, 1.0])
    m.lower = m.upper.copy()
    s = m.with_fixed_values(np.array([1]), np.array([3.0]))
    assert s.diag[1] == -1.0 and s.source[1] == -3.0


FOURIER = dict(name="body", kind="fourier", center=[5.0, 4.0], radius=0.5,
               a=[0.02, -0.01, 0.015, 0.0, 0.01, -0.02, 0.005, 0.01],
               b=[-0.015, 0.02, 0.0, 0.01, -0.005, 0.01, 0.02, -0.01])


def _fourier_grid():
    return fo.Grid([0, 0, 0], [20, 8, 0.25], [80, 32, 1], [([3.5, 2.5, 0.0], [6.5, 5.5, 0.25], 2)],
                   bodies=[FOURIER])


def test_fourier_cut_rays_change_sign_exactly_once_at_the_hit():
    g = _fourier_grid()
    o, n, ax = g.topo.owner.astype(np.int64), g.topo.neighbour.astype(np.int64), g.topo.axis.astype(np.int64)
    C, S = g.cell_centers, g.cell_sizes
    s = g.cell_sizes
    cO, cN = C[o], C[n]
    P = (np.maximum(cO - s[o] / 2, cN - s[n] / 2) + np.minimum(cO + s[o] / 2, cN + s[n] / 2)) / 2

    def outline(k, t):
        x = np.where(ax[k] == 0, t, P[k, 0])
        y = np.where(ax[k] == 1, t, P[k, 1])
        return fo._fourier_outline(FOURIER, x, y)
    s0, s1 = C[o, ax], C[n, ax]
    allf = np.arange(len(o))
    g0, g1 = outline(allf, s0), outline(allf, s1)
    imm = g.is_immersed
    # cut exactly where the outline changes sign (or touches) along the ray
    np.testing.assert_array_equal(imm, (ax != 2) & (g0 * g1 <= 0))
    k = np.nonzero(imm)[0]
    hit = s0[k] + g.ap.dist_owner
    np.testing.assert_allclose(s1[k] - g.ap.dist_neighbour, hit, rtol=0, atol=1e-12)
    assert np.all(np.abs(outline(k, hit)) < 1e-12)
    L = 20.0
    dxo, dxn = S[o[k], ax[k]], S[n[k], ax[k]]
    np.testing.assert_array_equal(g.ap.owner_near, g.ap.dist_owner / dxo <= dxo / L)
    np.testing.assert_array_equal(g.ap.neighbour_near, g.ap.dist_neighbour / dxn <= dxn / L)
    # the outline is closed: every cut ray enters or leaves it once
    assert g.num_immersed > 100


def test_box_cut_hits_the_box_faces():
    box = dict(name="box", kind="box", min=[0.3, 0.35, 0.2], max=[0.55, 0.6, 0.7])
    g = fo.Grid([0, 0, 0], [1, 1, 1], [8, 8, 8], [([0.2, 0.2, 0.1], [0.7, 0.7, 0.8], 1)], bodies=[box])
    o, n, ax = g.topo.owner.astype(np.int64), g.topo.neighbour.astype(np.int64), g.topo.axis.astype(np.int64)
    C = g.cell_centers
    k = np.nonzero(g.is_immersed)[0]
    hit_o = C[o[k], ax[k]] + g.ap.dist_owner
    hit_n = C[n[k], ax[k]] - g.ap.dist_neighbour
    lo, hi = np.asarray(box["min"])[ax[k]], np.asarray(box["max"])[ax[k]]
    assert np.all(np.isclose(hit_o, lo, atol=1e-15) | np.isclose(hit_o, hi, atol=1e-15))
    assert np.all(np.isclose(hit_n, lo, atol=1e-15) | np.isclose(hit_n, hi, atol=1e-15))
    # a cell centre strictly inside and one strictly outside along an axis
    # ray through the box's cross-section is always a cut
    inside = np.all((C > np.asarray(box["min"])) & (C < np.asarray(box["max"])), axis=1)
    cross = inside[o] != inside[n]
    assert np.all(g.is_immersed[cross])
    assert g.num_immersed >= int(cross.sum()) > 0


def test_immersed_pimple_steps_conserve_mass_and_hold_the_wall():
    """PIMPLE steps on the Fourier-body channel: the corrected flux is
    divergence free to the solver tolerance and any snapped velocity cell
    carries the wall value."""
    case = fo.Case(fourier_channel(root=(40, 16, 1), level=1))
    case.step()
    case.step()
    g = case.g
    assert g.num_immersed > 20
    assert fo.continuity_max(g, case.phi) < 1e-6
    cells, values = fo.immersed_constraints(g, case.U)
    np.testing.assert_array_equal(case.U.data[cells], values)
    assert np.isfinite(case.U.data).all() and np.isfinite(case.p.data).all()


def test_snapped_box_simple_step_holds_the_wall():
    """SIMPLE around a box whose x faces run through cell centres (distance
    0, so those cells are snapped): the snapped velocity cells hold the wall
    value after the explicit correction and the pressure projection
    conserves mass in every cell to the solver tolerance."""
    case = fo.Case(box_cavity(16))
    g = case.g
    cells, _ = fo.immersed_constraints(g, case.U)
    assert len(cells) >= 8
    for _ in range(3):
        case.step()
    cells, values = fo.immersed_constraints(g, case.U)
    np.testing.assert_array_equal(case.U.data[cells], values)
    assert fo.continuity_max(g, case.phi) < 1e-7
    assert np.isfinite(case.U.data).all()
