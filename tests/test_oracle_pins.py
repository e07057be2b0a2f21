"""Pins for three oracle functions on the 3-D octree PISO cavity path:
adjust_phi (fv/adjust_phi.py:44-127), ddtCorr (fv/fvc/ddt_corr.py:13-78)
and linear convection (fv/fvm/div.py:15-73 with fv/schemes/div.py:36-142).

Each check is chosen so that a plausible slip in the restatement fails it:
- adjust_phi: the reference's two unit tests (tests/unit/test_adjust_phi.py:
  20-88) plus closed-form outlet values with a fixed-value outflow patch and
  the no-op when the pressure level is fixed;
- ddtCorr: the reference's test (tests/integration/algorithms/
  test_pressure_coupling.py:394-425: zero for a consistent old flux,
  bounded by |phiCorr|/dt otherwise), and closed forms: for phi0 =
  k U0_f.Sf the coupling is 1 - |k-1|/|k| and the correction
  (1 - |k-1|/|k|)(k-1) U0_f.Sf / dt (sign, coupling and 1/dt), BDF2's
  b, c weights on two levels, and, on a refined grid with a linear velocity,
  the hanging-face value G (C_f - C_w).S_f that only the two-point weight w
  (not 1 - w) gives;
- linear convection: exact for a linear field in a uniform flux on refined
  grids (A psi - b = V u.grad psi in every cell; a w <-> 1-w swap, a sign
  slip or a dropped hanging-face correction all fail), and the reference's
  negative control (tests/integration/fv/test_advection_bounds.py:154-163):
  the central scheme conserves mass to 1e-8 and overshoots a square pulse
  by more than 0.01 on uniform and refined grids.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fv_oracle as fo
from tests.test_oracle_tvd import pulse_metrics


# ---------------------------------------------------------------- adjust_phi
def _adj_channel():
    """tests/helpers/configs.py:466-520: 2 x 0.2 x 0.1 channel, 8 x 2 x 1 root cells."""
    return fo.Grid([0, 0, 0], [2.0, 0.2, 0.1], [8, 2, 1], [], 0, dt=0.01)


def _adj_fields(g, outlet_bc):
    U = fo.Field("U", np.zeros((g.num_cells, 3)),
                 {"x_minus": fo.BC("dirichlet", np.array([1.0, 0.0, 0.0])), "x_plus": outlet_bc})
    p = fo.Field("p", np.zeros(g.num_cells), {})
    return U, p


def test_adjust_phi_balances_domain_flux():
    """tests/unit/test_adjust_phi.py:20-53: an inletOutlet outlet is scaled
    until the domain-boundary flux sums to zero; the outlet faces then carry
    exactly the inflow spread over them (closed form n_in / n_out here)."""
    g = _adj_channel()
    U, p = _adj_fields(g, fo.BC("inlet_outlet", np.zeros(3)))
    phi = fo.FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
    inl, out = g.patch_mask("x_minus"), g.patch_mask("x_plus")
    phi.bnd[inl] = -1.0
    phi.bnd[out] = 0.8
    assert fo.adjust_phi(g, phi, U, p)
    assert abs(phi.bnd.sum()) < 1e-10
    np.testing.assert_allclose(phi.bnd[out], inl.sum() / out.sum(), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(phi.bnd[inl], -1.0)


def test_adjust_phi_skips_scaling_when_no_adjustable_outflow():
    """tests/unit/test_adjust_phi.py:56-88: with only fixed-value outlets phi is unchanged."""
    g = _adj_channel()
    U, p = _adj_fields(g, fo.BC("dirichlet", np.zeros(3)))
    phi = fo.FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
    phi.bnd[g.patch_mask("x_minus")] = -1.0
    phi.bnd[g.patch_mask("x_plus")] = 0.5
    before = phi.bnd.copy()
    assert not fo.adjust_phi(g, phi, U, p)
    np.testing.assert_array_equal(phi.bnd, before)


def test_adjust_phi_closed_form_with_fixed_outflow_and_inflow_on_the_outlet():
    """massCorr = (massIn - fixedOut) / adjustableOut: fixed-value outflow on
    y_plus is subtracted, inflow faces of the adjustable patch count as
    inflow and keep their value, only its outflow faces are scaled."""
    g = _adj_channel()
    U = fo.Field("U", np.zeros((g.num_cells, 3)),
                 {"x_minus": fo.BC("dirichlet", np.array([1.0, 0.0, 0.0])),
                  "y_plus": fo.BC("dirichlet", np.array([0.0, 0.1, 0.0])),
                  "x_plus": fo.BC("neumann", np.zeros(3))})
    p = fo.Field("p", np.zeros(g.num_cells), {})
    phi = fo.FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
    inl, top, out = g.patch_mask("x_minus"), g.patch_mask("y_plus"), g.patch_mask("x_plus")
    phi.bnd[inl] = -1.0                          # mass in 2
    phi.bnd[top] = 0.05                          # fixed out 8 * 0.05 = 0.4
    o = np.nonzero(out)[0]
    phi.bnd[o[0]] = -0.25                        # inflow on the adjustable patch: mass in 2.25
    phi.bnd[o[1]] = 0.5                          # adjustable out 0.5
    assert fo.adjust_phi(g, phi, U, p)
    corr = (2.0 + 0.25 - 8 * 0.05) / 0.5
    assert phi.bnd[o[1]] == pytest.approx(0.5 * corr, abs=1e-15)
    assert phi.bnd[o[0]] == -0.25
    np.testing.assert_array_equal(phi.bnd[top], 0.05)
    assert abs(phi.bnd.sum()) < 1e-12


def test_adjust_phi_is_a_no_op_when_pressure_is_fixed():
    """fv/adjust_phi.py:82-89: a Dirichlet pressure condition fixes the level."""
    g = _adj_channel()
    U, _ = _adj_fields(g, fo.BC("inlet_outlet", np.zeros(3)))
    p = fo.Field("p", np.zeros(g.num_cells), {"x_plus": fo.BC("dirichlet", 0.0)})
    phi = fo.FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
    phi.bnd[g.patch_mask("x_minus")] = -1.0
    phi.bnd[g.patch_mask("x_plus")] = 0.8
    before = phi.bnd.copy()
    assert not fo.adjust_phi(g, phi, U, p)
    np.testing.assert_array_equal(phi.bnd, before)


# ---------------------------------------------------------------- ddtCorr
def _ddt_grid(refined: bool):
    regions = [([0.75, 0.0, 0.0], [1.25, 0.5, 0.1], 1)] if refined else []
    return fo.Grid([0, 0, 0], [2, 0.5, 0.1], [8, 2, 1], regions, 0, dt=0.05)


def _two_point(g, U):
    w = g.geo.w[:, None]
    return np.sum((w * U[g.owner] + (1.0 - w) * U[g.neighbour]) * g.Sf, axis=1)


@pytest.mark.parametrize("refined", [False, True])
def test_ddt_corr_vanishes_for_consistent_old_flux_and_is_bounded(refined):
    """test_pressure_coupling.py:394-425 on the oracle."""
    g = _ddt_grid(refined)
    U = fo.Field("U", np.ones((g.num_cells, 3)), {})
    U.old = U.data.copy()
    phi0 = np.sum(fo.interpolate(g, U).single * g.Sf, axis=1)
    corr = fo.ddt_corr(g, U, phi0, (1.0, 1.0, 0.0))
    assert float(np.abs(corr).max()) < 1e-12
    phi0 = 1.5 * phi0
    corr = fo.ddt_corr(g, U, phi0, (1.0, 1.0, 0.0))
    phi_corr = phi0 - np.sum(fo.interpolate(g, U).single * g.Sf, axis=1)
    assert np.all(np.abs(corr) <= np.abs(phi_corr) / g.dt + 1e-12)
    assert float(np.abs(corr).max()) > 0.0


@pytest.mark.parametrize("k", [1.5, 0.6, -0.5, 3.0])
def test_ddt_corr_closed_form_for_a_scaled_old_flux(k):
    """phi0 = k U0_f.Sf: phiCorr = (k-1) U0_f.Sf, coupling 1 - min(|k-1|/|k|, 1);
    Euler: ddtCorr = coupling (k-1) U0_f.Sf / dt."""
    g = _ddt_grid(False)
    rng = np.random.default_rng(3)
    U = fo.Field("U", rng.standard_normal((g.num_cells, 3)), {})
    U.old = U.data.copy()
    base = _two_point(g, U.old)
    corr = fo.ddt_corr(g, U, k * base, (1.0, 1.0, 0.0))
    coupling = 1.0 - min(abs(k - 1.0) / abs(k), 1.0)
    np.testing.assert_allclose(corr, coupling * (k - 1.0) * base / g.dt, rtol=1e-12, atol=1e-12)
    if k == 1.5:  # the hand value: (1 - 1/3) * 0.5 = 1/3
        np.testing.assert_allclose(corr, base / (3.0 * g.dt), rtol=1e-12, atol=1e-12)


def test_ddt_corr_bdf2_weights_both_levels():
    """BDF2 (dt = dt0): a = 1.5, b = 2, c = 0.5.  phi0 = 2 U0_f.Sf, phi00 =
    U00_f.Sf + 0.3 |Sf|: history 2 (U0_f.Sf) - 0.5 (0.3 |Sf|), coupling from
    the previous level only: 1 - 1/2."""
    g = _ddt_grid(False)
    rng = np.random.default_rng(4)
    U = fo.Field("U", rng.standard_normal((g.num_cells, 3)), {})
    U.old = rng.standard_normal((g.num_cells, 3))
    U.older = rng.standard_normal((g.num_cells, 3))
    abc = fo.ddt_coefficients("backward", True, g.dt, g.dt)
    assert abc == pytest.approx((1.5, 2.0, 0.5))
    b0, b00 = _two_point(g, U.old), _two_point(g, U.older)
    magS = np.abs(g.Sf).sum(axis=1)
    corr = fo.ddt_corr(g, U, 2.0 * b0, abc, phi_older=b00 + 0.3 * magS)
    expect = 0.5 * (2.0 * b0 - 0.5 * 0.3 * magS) / g.dt
    np.testing.assert_allclose(corr, expect, rtol=1e-12, atol=1e-12)


def test_ddt_corr_uses_two_point_weights_on_hanging_faces():
    """Linear old velocity U = G x, old flux from the exact face-centre value
    U(C_f).S_f: the correction is G (C_f - C_w).S_f with C_w = w C_O +
    (1-w) C_N (zero where the centres line up, non-zero on the 2:1 faces,
    where w = 1/3 or 2/3)."""
    g = _ddt_grid(True)
    G = np.array([[0.3, -1.1, 0.2], [0.7, 0.4, -0.5], [-0.2, 0.9, 0.6]])
    U = fo.Field("U", g.cell_centers @ G.T, {})
    U.old = U.data.copy()
    phi0 = np.sum((g.face_centers @ G.T) * g.Sf, axis=1)
    w = g.geo.w[:, None]
    cw = w * g.cell_centers[g.owner] + (1.0 - w) * g.cell_centers[g.neighbour]
    pc = np.sum(((g.face_centers - cw) @ G.T) * g.Sf, axis=1)
    assert np.count_nonzero(np.abs(pc) > 1e-12) > 0
    coupling = 1.0 - np.minimum(np.abs(pc) / np.abs(phi0), 1.0)
    corr = fo.ddt_corr(g, U, phi0, (1.0, 1.0, 0.0))
    np.testing.assert_allclose(corr, coupling * pc / g.dt, rtol=1e-10, atol=1e-12)
    hang = np.zeros(g.num_faces, bool)
    hang[g.geo.hang_idx] = True
    assert np.all(np.abs(corr[~hang]) < 1e-12)
    assert set(np.round(g.geo.w[hang], 12)) <= {round(1 / 3, 12), round(2 / 3, 12)}


# ---------------------------------------------------------------- linear convection
_GRIDS = {
    "refined3d": lambda: fo.Grid([0, 0, 0], [1, 1, 1], [4, 4, 4], [([0.3, 0.3, 0.3], [0.7, 0.7, 0.7], 1)]),
    "walls3d": lambda: fo.Grid([0, 0, 0], [1, 1, 1], [4, 4, 4], [([0.0, 0.0, 0.0], [1.0, 0.2, 1.0], 2)]),
    "channel": lambda: _ddt_grid(True),
}


@pytest.mark.parametrize("name", sorted(_GRIDS))
@pytest.mark.parametrize("comps", [(), (3,)])
def test_linear_convection_is_exact_for_linear_fields(name, comps):
    """Uniform velocity u, flux u.S_f on every face, psi linear with the
    exact normal gradient on every patch: A psi - b = V (u . grad psi) in
    every cell, hanging faces included."""
    g = _GRIDS[name]()
    assert len(g.geo.hang_idx) > 0
    u = np.array([0.3, -0.7, 0.5])
    phi = fo.FaceValues(g.Sf @ u, g.bnd_Sf @ u)
    a = np.array([1.3, -0.4, 2.1]) if comps == () else np.array([[1.3, -0.4, 2.1], [0.2, 0.9, -1.5], [-0.6, 0.1, 0.8]])
    psi = g.cell_centers @ a.T + (0.7 if comps == () else np.array([0.7, -0.2, 0.4]))
    normals = {p: np.eye(3)[i // 2] * (2 * (i % 2) - 1) for i, p in enumerate(fo.PATCHES)}
    bcs = {p: fo.BC("neumann", a @ n) for p, n in normals.items()}
    fld = fo.Field("q", psi, bcs)
    m = fo.fvm_div(g, phi, fld, "linear")
    res = m.multiply(psi) - m.source
    expect = fo._bcast(g.cell_volumes, res) * (a @ u)
    np.testing.assert_allclose(res, expect, rtol=0, atol=1e-12 * float(np.abs(expect).max()))
    # the two-point face values alone are off on the hanging faces, so the
    # exactness above depends on the deferred hanging-face correction
    hang = g.geo.hang_idx
    base = fo.linear_face_values(g, psi)
    exact = g.face_centers @ a.T + (0.7 if comps == () else np.array([0.7, -0.2, 0.4]))
    assert np.abs(base[hang] - exact[hang]).max() > 1e-3


@pytest.mark.parametrize("refined", [False, True])
def test_linear_advection_exposes_unbounded_pulse(refined):
    """tests/integration/fv/test_advection_bounds.py:154-163 on the oracle."""
    m = pulse_metrics("linear", refined)
    assert m["relative_mass_error"] < 1e-8, m
    assert m["min"] < -0.01, m
    assert m["max"] > 1 + 0.01, m
