"""GPU: the flux-dependent boundary conditions inletOutlet
(boundaries/derived/inlet_outlet.py:61-123) and fixedFluxPressure
(boundaries/derived/fixed_flux_pressure.py:124-252) against the FP64
oracle -- their face evaluation on registered fields, then whole SIMPLE,
SIMPLEC, PISO and PIMPLE runs of the reference's inlet/outlet channel
(tests/helpers/configs.py:335-461), with the reference's own flux
assertions (tests/integration/algorithms/test_pressure_coupling.py:65-200)
checked on the GPU result."""

from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from oracle import fv_oracle as fo
from tests.cases import channel_flow

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b) -> float:
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _grids():
    from gridfoam_amd.config import load_config
    from gridfoam_amd.grid import create_grid

    cfg = channel_flow("piso")
    g = create_grid(load_config(cfg))
    return g, fo.grid_from_cfg(cfg)


def _registered(g, o, seed, *, with_phi_hbya=True):
    """The same random coupling fields registered on both sides."""
    from gridfoam_amd.fields import CellField, FaceField

    rng = np.random.default_rng(seed)
    C, F, B = g.num_cells, g.num_internal_faces, g.num_domain_bnd_faces
    dev = g.device
    out = {}
    phi = FaceField(g, "phi")
    ph = fo.FaceValues(rng.standard_normal(F), rng.standard_normal(B))
    outlet = o.patch_mask("x_plus")
    ph.bnd[outlet] = np.where(np.arange(int(outlet.sum())) % 2 == 0, -1.0, 1.0) * (0.5 + rng.random(int(outlet.sum())))
    phi.single_data = torch.tensor(ph.single, device=dev)
    phi.domain_bnd_data = torch.tensor(ph.bnd, device=dev)
    out["phi"] = (phi, ph)
    if with_phi_hbya:
        phh = FaceField(g, "phiHbyA")
        oh = fo.FaceValues(rng.standard_normal(F), rng.standard_normal(B))
        phh.single_data = torch.tensor(oh.single, device=dev)
        phh.domain_bnd_data = torch.tensor(oh.bnd, device=dev)
        out["phiHbyA"] = (phh, oh)
    H = CellField(g, "HbyA", (3,))
    H.data = torch.tensor(rng.standard_normal((C, 3)), device=dev)
    out["HbyA"] = (H, _np(H.data))
    r = CellField(g, "rAtU")
    r.data = torch.tensor(0.5 + rng.random(C), device=dev)
    out["rAtU"] = (r, _np(r.data))
    reg = {"phi": lambda: out["phi"][1], "HbyA": lambda: out["HbyA"][1], "rAtU": lambda: out["rAtU"][1]}
    if with_phi_hbya:
        reg["phiHbyA"] = lambda: out["phiHbyA"][1]
    o.registry.clear()
    o.registry.update(reg)
    return out


def _velocity(g, o, seed):
    from gridfoam_amd.fields import CellField, DirichletBC, InletOutletBC, NeumannBC

    rng = np.random.default_rng(seed)
    U = CellField(g, "U", (3,))
    U.data = torch.tensor(rng.standard_normal((g.num_cells, 3)), device=g.device)
    v = [0.3, -0.2, 0.1]
    U.add_boundary_conditions({"x_minus": DirichletBC([1.0, 0.2, 0.0]), "y_minus": NeumannBC([0.0, 0.0, 0.0]),
                               "x_plus": InletOutletBC(v)})
    oU = fo.Field("U", _np(U.data), {"x_minus": fo.BC("dirichlet", np.array([1.0, 0.2, 0.0])),
                                     "y_minus": fo.BC("neumann", np.zeros(3)),
                                     "x_plus": fo.BC("inlet_outlet", np.array(v))})
    o.registry["U"] = lambda: oU
    return U, oU


def _pressure(g, o, seed):
    from gridfoam_amd.fields import CellField, DirichletBC, FixedFluxPressureBC

    rng = np.random.default_rng(seed)
    p = CellField(g, "p")
    p.data = torch.tensor(rng.standard_normal(g.num_cells), device=g.device)
    p.add_boundary_conditions({k: FixedFluxPressureBC() for k in ("x_minus", "y_minus", "y_plus")})
    p.add_boundary_conditions({"x_plus": DirichletBC(0.0)})
    op = fo.Field("p", _np(p.data), {k: fo.BC("fixed_flux_pressure") for k in ("x_minus", "y_minus", "y_plus")})
    op.bcs["x_plus"] = fo.BC("dirichlet", np.asarray(0.0))
    return p, op


def _oracle_state(o, fld):
    B = o.num_bnd
    cs = fld.cshape
    f, rv, rg, pb = np.zeros(B), np.zeros((B,) + cs), np.zeros((B,) + cs), np.zeros((B,) + cs)
    for b in fo.batches(o, fld):
        f[b.mask], rv[b.mask], rg[b.mask], pb[b.mask] = fo.bc_state(o, fld, b)
    return f, rv, rg, pb


def _compare_state(st, ost, B, nc, covered):
    f, rv, rg, pb = ost
    cov = covered
    np.testing.assert_array_equal(_np(st.frac[:B])[cov], f[cov])
    np.testing.assert_allclose(_np(st.ref_v[:B]).reshape(B, -1)[cov], rv.reshape(B, -1)[cov], rtol=0, atol=0)
    np.testing.assert_allclose(_np(st.ref_g[:B]).reshape(B, -1)[cov], rg.reshape(B, -1)[cov], rtol=1e-13,
                               atol=1e-13 * max(np.abs(rg).max(), 1e-300))
    np.testing.assert_allclose(_np(st.psi_b[:B]).reshape(B, -1)[cov], pb.reshape(B, -1)[cov], rtol=1e-14,
                               atol=1e-14 * np.abs(pb).max())


def test_inlet_outlet_and_fixed_flux_pressure_evaluation_match_oracle():
    """Mixed-sign outlet flux for inletOutlet; fixedFluxPressure on a patch
    with a Dirichlet velocity, one with a zero-gradient velocity and one
    without a velocity condition (target: the registered phi)."""
    g, o = _grids()
    _say(f"bc evaluation channel cells={g.num_cells}")
    B = g.num_domain_bnd_faces
    keep = _registered(g, o, 5)
    U, oU = _velocity(g, o, 6)
    p, op = _pressure(g, o, 7)
    st = U.boundary_state()
    cov = fo.covered_domain(o, oU)
    _compare_state(st, _oracle_state(o, oU), B, 3, cov)
    inflow = o.patch_mask("x_plus") & (keep["phi"][1].bnd < 0)
    assert inflow.any() and (o.patch_mask("x_plus") & ~inflow).any()
    np.testing.assert_array_equal(_np(st.frac[:B])[o.patch_mask("x_plus")],
                                  inflow[o.patch_mask("x_plus")].astype(float))
    stp = p.boundary_state()
    ost = _oracle_state(o, op)
    _compare_state(stp, ost, B, 1, fo.covered_domain(o, op))
    assert np.abs(ost[2]).max() > 0.1
    # flipping the flux sign on the device flips the faces on the next call
    keep["phi"][0].domain_bnd_data = -keep["phi"][0].domain_bnd_data
    keep["phi"][1].bnd[:] = -keep["phi"][1].bnd
    _compare_state(U.boundary_state(), _oracle_state(o, oU), B, 3, cov)
    _compare_state(p.boundary_state(), _oracle_state(o, op), B, 1, fo.covered_domain(o, op))


def test_fixed_flux_pressure_fallbacks_match_oracle():
    """HbyA_P . Sf for an unregistered phiHbyA; zero gradient for a zero
    coefficient and before any coupling field is registered."""
    g, o = _grids()
    _say("fixedFluxPressure fallbacks")
    B = g.num_domain_bnd_faces
    p, op = _pressure(g, o, 8)
    o.registry.clear()
    st = p.boundary_state()
    assert float(st.ref_g[:B].abs().max()) == 0.0
    keep = _registered(g, o, 9, with_phi_hbya=False)
    U, oU = _velocity(g, o, 10)
    ost = _oracle_state(o, op)
    assert np.abs(ost[2]).max() > 0.1
    _compare_state(p.boundary_state(), ost, B, 1, fo.covered_domain(o, op))
    keep["rAtU"][0].data.zero_()
    keep["rAtU"][1][:] = 0.0
    assert float(p.boundary_state().ref_g[:B].abs().max()) == 0.0
    del U


def _run_both(cfg_dict: dict, steps: int):
    from gridfoam_amd.algorithms import create_case
    from gridfoam_amd.config import load_config

    grid, alg = create_case(load_config(cfg_dict))
    case = fo.Case(cfg_dict)
    for _ in range(steps):
        alg.step()
        case.step()
    return grid, alg, case


def _assert_channel_fluxes(g, alg, o):
    """test_pressure_coupling.py:70-97 on the GPU result."""
    from gridfoam_amd.fv import fvc

    assert float(fvc.div(alg.phi).abs().max()) < 1e-8
    B = g.num_domain_bnd_faces
    phib = _np(alg.phi.domain_bnd_data)
    inlet = o.patch_mask("x_minus")
    np.testing.assert_allclose(phib[inlet], -np.abs(_np(g.domain_bnd_Sf[:B])[inlet, 0]), rtol=0, atol=1e-12)
    for pt in ("y_minus", "y_plus", "z_minus", "z_plus"):
        assert np.abs(phib[o.patch_mask(pt)]).max() < 1e-12


_RUNS = {
    "piso-dirichlet-slip": (dict(algorithm="piso"), 4),
    "piso-neumann-noslip": (dict(algorithm="piso", p_outlet="neumann", walls="dirichlet"), 4),
    "simple-dirichlet-noslip": (dict(algorithm="simple", walls="dirichlet"), 5),
    "simple-neumann-slip": (dict(algorithm="simple", p_outlet="neumann"), 5),
    "simplec-noslip": (dict(algorithm="simple", walls="dirichlet", consistent=True), 4),
    "pimple-dirichlet-noslip": (dict(algorithm="pimple", walls="dirichlet"), 3),
    "pimplec-neumann-noslip": (dict(algorithm="pimple", p_outlet="neumann", walls="dirichlet", consistent=True), 3),
    "piso-reversed-outlet": (dict(algorithm="piso", u_internal=(-0.2, 0.0, 0.0)), 3),
    "piso-ffp-walls": (dict(algorithm="piso", ffp_walls=True, u_internal=(0.5, 0.3, 0.0)), 4),
    "simplec-ffp-walls": (dict(algorithm="simple", ffp_walls=True, u_internal=(0.5, 0.3, 0.0), consistent=True), 4),
    "pimple-ffp-walls": (dict(algorithm="pimple", ffp_walls=True, u_internal=(0.5, 0.3, 0.0)), 3),
}


@pytest.mark.parametrize("name", list(_RUNS))
def test_channel_run_matches_oracle(name):
    kw, steps = _RUNS[name]
    cfg = channel_flow(**kw)
    _say(f"channel {name} {steps} steps")
    g, alg, case = _run_both(cfg, steps)
    assert _rel(_np(alg.U.data), case.U.data) < 1e-9
    assert _rel(_np(alg.p.data), case.p.data) < 1e-9
    assert _rel(_np(alg.phi.single_data), case.phi.single) < 1e-9
    assert _rel(_np(alg.phi.domain_bnd_data), case.phi.bnd) < 1e-9
    if kw.get("ffp_walls"):
        from gridfoam_amd.fv import fvc

        assert float(fvc.div(alg.phi).abs().max()) < 1e-8
        B = g.num_domain_bnd_faces
        walls = case.g.patch_mask("y_minus") | case.g.patch_mask("y_plus")
        rg = _np(alg.p.boundary_state().ref_g[:B]).reshape(-1)[walls]
        assert np.abs(rg).max() > 1e-3
        return
    _assert_channel_fluxes(g, alg, case.g)
    if kw.get("p_outlet") == "neumann":
        assert abs(float(alg.phi.domain_bnd_data.sum())) < 1e-10
    if kw["algorithm"] == "simple" and kw.get("walls") == "dirichlet" and not kw.get("consistent"):
        assert float(alg.U.data[:, 0].max()) > 1.05
    if kw.get("consistent"):
        assert not torch.allclose(alg.rAtU.data, alg.rAU.data)
