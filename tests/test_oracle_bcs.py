"""Oracle pins for the flux-dependent boundary conditions: inletOutlet
(boundaries/derived/inlet_outlet.py:61-123) and fixedFluxPressure
(boundaries/derived/fixed_flux_pressure.py:124-252).

The pins are the conditions' defining properties, chosen so that a sign or
scale mistake fails: inletOutlet reduces to the Dirichlet inlet value on an
all-inflow patch and to zero gradient on an all-outflow (or zero-flux)
patch, face by face on a mixed one; the fixedFluxPressure gradient makes the
corrected boundary flux of correct_phi equal the velocity boundary flux
exactly.  Whole channel runs are checked against the reference's own
pressure-coupling assertions (tests/integration/algorithms/
test_pressure_coupling.py:65-200): cell divergence below 1e-8, inlet flux
-|Sf|, impermeable walls, zero net flux with a referenced pressure, a
developing core velocity with no-slip walls."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fv_oracle as fo
from tests.cases import channel_flow


def _channel_grid(refined=True):
    regions = [([0.75, 0.0, 0.0], [1.25, 0.5, 0.1], 1)] if refined else []
    return fo.Grid([0, 0, 0], [2, 0.5, 0.1], [8, 2, 1], regions, 0, dt=0.05)


def _batch(g, fld, patch):
    return next(b for b in fo.batches(g, fld) if b.patch == patch)


def _phi(g, rng, outlet_signs):
    phi = fo.FaceValues(rng.standard_normal(g.num_faces), rng.standard_normal(g.num_bnd))
    m = g.patch_mask("x_plus")
    phi.bnd[m] = outlet_signs * (0.5 + rng.random(int(m.sum())))
    return phi


@pytest.mark.parametrize("signs", ["in", "out", "zero", "mixed"])
def test_inlet_outlet_face_values_follow_the_flux_sign(signs):
    g = _channel_grid()
    rng = np.random.default_rng(1)
    n = int(g.patch_mask("x_plus").sum())
    s = {"in": -np.ones(n), "out": np.ones(n), "zero": np.zeros(n),
         "mixed": np.where(np.arange(n) % 2 == 0, -1.0, 1.0)}[signs]
    phi = _phi(g, rng, s)
    g.registry["phi"] = lambda: phi
    v = np.array([0.3, -0.2, 0.1])
    U = fo.Field("U", rng.standard_normal((g.num_cells, 3)), {"x_plus": fo.BC("inlet_outlet", v)})
    b = _batch(g, U, "x_plus")
    f, rv, rg, psi_b = fo.bc_state(g, U, b)
    inflow = phi.bnd[b.mask] < 0
    np.testing.assert_array_equal(f, inflow.astype(float))
    np.testing.assert_array_equal(rg, 0.0)
    expect = np.where(inflow[:, None], v[None, :], U.data[b.cells])
    np.testing.assert_array_equal(psi_b, expect)
    assert fo.BC("inlet_outlet", v).assignable


@pytest.mark.parametrize("direction", ["in", "out"])
def test_inlet_outlet_convection_equals_dirichlet_or_zero_gradient(direction):
    """On an all-inflow patch the upwind convection operator equals the one
    with a Dirichlet inlet value; on an all-outflow patch the one with a
    zero-gradient condition."""
    g = _channel_grid()
    rng = np.random.default_rng(2)
    n = int(g.patch_mask("x_plus").sum())
    phi = _phi(g, rng, -np.ones(n) if direction == "in" else np.ones(n))
    g.registry["phi"] = lambda: phi
    v = np.array([0.7, 0.0, -0.4])
    data = rng.standard_normal((g.num_cells, 3))
    io = fo.Field("U", data, {"x_plus": fo.BC("inlet_outlet", v)})
    same = fo.Field("U", data, {"x_plus": fo.BC("dirichlet", v) if direction == "in"
                                else fo.BC("neumann", np.zeros(3))})
    for scheme in ("upwind", "linear"):
        a, b = fo.fvm_div(g, phi, io, scheme), fo.fvm_div(g, phi, same, scheme)
        np.testing.assert_array_equal(a.diag, b.diag)
        np.testing.assert_array_equal(a.source, b.source)
    a, b = fo.fvm_laplacian(g, 0.01, io), fo.fvm_laplacian(g, 0.01, same)
    np.testing.assert_array_equal(a.diag, b.diag)
    np.testing.assert_array_equal(a.source, b.source)


def test_inlet_outlet_without_registered_flux_raises():
    g = _channel_grid()
    U = fo.Field("U", np.zeros((g.num_cells, 3)), {"x_plus": fo.BC("inlet_outlet", np.zeros(3))})
    with pytest.raises(ValueError, match="phi"):
        fo.bc_state(g, U, _batch(g, U, "x_plus"))
    with pytest.raises(ValueError, match="shape"):
        g.registry["phi"] = lambda: fo.FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
        P = fo.Field("q", np.zeros(g.num_cells), {"x_plus": fo.BC("inlet_outlet", np.zeros(3))})
        fo.bc_state(g, P, _batch(g, P, "x_plus"))


def _ffp_setup(u_kind, seed=3):
    g = _channel_grid()
    rng = np.random.default_rng(seed)
    C = g.num_cells
    phi_h = fo.FaceValues(rng.standard_normal(g.num_faces), rng.standard_normal(g.num_bnd))
    phi = fo.FaceValues(rng.standard_normal(g.num_faces), rng.standard_normal(g.num_bnd))
    rAtU = 0.5 + rng.random(C)
    ubcs = {"x_minus": fo.BC("dirichlet", np.array([1.0, 0.2, 0.0])), "y_minus": fo.BC("neumann", np.zeros(3)),
            "y_plus": fo.BC("slip")}
    U = fo.Field("U", rng.standard_normal((C, 3)), {k: v for k, v in ubcs.items() if k in u_kind})
    p = fo.Field("p", rng.standard_normal(C), {k: fo.BC("fixed_flux_pressure") for k in
                                                ("x_minus", "y_minus", "y_plus")})
    p.bcs["x_plus"] = fo.BC("dirichlet", np.asarray(0.0))
    g.registry.update(U=lambda: U, phi=lambda: phi, phiHbyA=lambda: phi_h, rAtU=lambda: rAtU,
                      HbyA=lambda: rng.standard_normal((C, 3)))
    return g, U, p, phi, phi_h, rAtU


def test_fixed_flux_pressure_closed_form_and_corrected_flux():
    """ref_g = (phiHbyA_b - U_b . Sf) / (rAtU_P |Sf|) on every patch, and
    correct_phi with that gradient returns phi_b = U_b . Sf exactly (the
    flux the condition fixes); on the patch without a velocity condition
    the target is the registered boundary phi."""
    g, U, p, phi, phi_h, rAtU = _ffp_setup(("x_minus", "y_minus", "y_plus"))
    del U.bcs["y_plus"]
    pm = -fo.fvm_laplacian(g, rAtU, p, False)
    out = fo.correct_phi(g, phi_h, pm, p, rAtU)
    for patch in ("x_minus", "y_minus", "y_plus"):
        b = _batch(g, p, patch)
        f, rv, rg, _ = fo.bc_state(g, p, b)
        np.testing.assert_array_equal(f, 0.0)
        if patch in U.bcs:
            ub = fo.bc_state(g, U, _batch(g, U, patch))[3]
            target = np.sum(ub * b.Sf, axis=1)
        else:
            target = phi.bnd[b.mask]
        expect = (phi_h.bnd[b.mask] - target) / (rAtU[b.cells] * b.mag_Sf)
        np.testing.assert_allclose(rg, expect, rtol=1e-15, atol=0)
        np.testing.assert_allclose(out.bnd[b.mask], target, rtol=0, atol=1e-14 * np.abs(phi_h.bnd).max())
    # x_minus carries the Dirichlet velocity: the fixed flux is U_in . Sf
    b = _batch(g, p, "x_minus")
    np.testing.assert_allclose(out.bnd[b.mask], -b.mag_Sf, rtol=1e-14)


def test_fixed_flux_pressure_fallbacks():
    g, U, p, phi, phi_h, rAtU = _ffp_setup(("x_minus", "y_minus", "y_plus"))
    b = _batch(g, p, "y_minus")
    ub = fo.bc_state(g, U, _batch(g, U, "y_minus"))[3]
    target = np.sum(ub * b.Sf, axis=1)
    # rAU stands in for an unregistered rAtU
    del g.registry["rAtU"]
    g.registry["rAU"] = lambda: 2.0 * rAtU
    rg = fo.bc_state(g, p, b)[2]
    np.testing.assert_allclose(rg, (phi_h.bnd[b.mask] - target) / (2.0 * rAtU[b.cells] * b.mag_Sf), rtol=1e-15)
    # HbyA_P . Sf stands in for an unregistered phiHbyA
    H = np.random.default_rng(9).standard_normal((g.num_cells, 3))
    del g.registry["phiHbyA"]
    g.registry["HbyA"] = lambda: H
    rg = fo.bc_state(g, p, b)[2]
    pred = np.sum(H[b.cells] * b.Sf, axis=1)
    np.testing.assert_allclose(rg, (pred - target) / (2.0 * rAtU[b.cells] * b.mag_Sf), rtol=1e-14)
    # a zero coefficient (before the first corrector) gives zero gradient
    g.registry["rAU"] = lambda: np.zeros(g.num_cells)
    np.testing.assert_array_equal(fo.bc_state(g, p, b)[2], 0.0)
    # neither predicted flux registered, or no coefficient: zero gradient
    g.registry["rAU"] = lambda: rAtU
    del g.registry["HbyA"]
    np.testing.assert_array_equal(fo.bc_state(g, p, b)[2], 0.0)
    g.registry.clear()
    np.testing.assert_array_equal(fo.bc_state(g, p, b)[2], 0.0)
    with pytest.raises(ValueError, match="scalar"):
        V = fo.Field("V", np.zeros((g.num_cells, 3)), {"y_minus": fo.BC("fixed_flux_pressure")})
        fo.bc_state(g, V, _batch(g, V, "y_minus"))


def _div_max(case):
    return float(np.abs(fo.fvc_div(case.g, case.phi)).max())


def _assert_channel_fluxes(case):
    """test_pressure_coupling.py:70-97."""
    g = case.g
    assert _div_max(case) < 1e-8
    inlet = g.patch_mask("x_minus")
    np.testing.assert_allclose(case.phi.bnd[inlet], -np.abs(g.bnd_Sf[inlet][:, 0]), rtol=0, atol=1e-12)
    for pt in ("y_minus", "y_plus", "z_minus", "z_plus"):
        assert np.abs(case.phi.bnd[g.patch_mask(pt)]).max() < 1e-12


@pytest.mark.parametrize("walls", ["slip", "dirichlet"])
@pytest.mark.parametrize("p_outlet", ["dirichlet", "neumann"])
@pytest.mark.parametrize("algorithm", ["simple", "piso", "pimple"])
def test_channel_pressure_coupling_properties(algorithm, p_outlet, walls):
    case = fo.Case(channel_flow(algorithm, p_outlet=p_outlet, walls=walls))
    for _ in range(5 if algorithm == "simple" else 3):
        case.step()
    _assert_channel_fluxes(case)
    if p_outlet == "neumann":
        assert abs(float(case.phi.bnd.sum())) < 1e-10
    if algorithm == "simple" and walls == "dirichlet":
        assert float(case.U.data[:, 0].max()) > 1.05  # developing Poiseuille core


@pytest.mark.parametrize("algorithm", ["simple", "pimple"])
def test_channel_simplec_properties(algorithm):
    case = fo.Case(channel_flow(algorithm, walls="dirichlet", consistent=True))
    for _ in range(3):
        case.step()
    _assert_channel_fluxes(case)
    assert not np.allclose(case.rAtU_d, case.rAU_d)


def test_channel_outlet_switches_from_inflow_to_outflow():
    """A reversed initial velocity makes the outlet an inflow boundary for
    the first momentum predictor (the inlet value is taken there); after the
    first corrector the outlet flux is positive and stays so."""
    case = fo.Case(channel_flow("piso", u_internal=(-0.2, 0.0, 0.0)))
    g = case.g
    out = g.patch_mask("x_plus")
    assert (case.phi.bnd[out] < 0).all()
    b = _batch(g, case.U, "x_plus")
    np.testing.assert_array_equal(fo.bc_state(g, case.U, b)[3], 0.0)  # inlet value 0
    case.step()
    assert (case.phi.bnd[out] > 0).all()
    np.testing.assert_array_equal(fo.bc_state(g, case.U, b)[0], 0.0)
    _assert_channel_fluxes(case)


@pytest.mark.parametrize("algorithm", ["simple", "piso", "pimple"])
def test_fixed_flux_pressure_walls_keep_the_velocity_flux(algorithm):
    """fixedFluxPressure on walls with a zero-gradient velocity: the wall
    gradient is not zero, the flux stays divergence-free and the inlet flux
    stays -|Sf| while flow crosses the walls."""
    case = fo.Case(channel_flow(algorithm, ffp_walls=True, u_internal=(0.5, 0.3, 0.0)))
    g = case.g
    walls = g.patch_mask("y_minus") | g.patch_mask("y_plus")
    inlet = g.patch_mask("x_minus")
    for _ in range(3):
        case.step()
        rg = np.concatenate([fo.bc_state(g, case.p, _batch(g, case.p, pt))[2] for pt in ("y_minus", "y_plus")])
        assert np.abs(rg).max() > 1e-3
        assert _div_max(case) < 1e-8
        np.testing.assert_allclose(case.phi.bnd[inlet], -np.abs(g.bnd_Sf[inlet][:, 0]), rtol=0, atol=1e-12)
        assert np.abs(case.phi.bnd[walls]).max() > 1e-4
