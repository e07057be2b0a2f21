"""Oracle pins for the TVD convection schemes (fv/schemes/div.py:145-461).

- Limiters: the textbook forms of van Leer, minmod, superbee and
  monotonized central (Sweby's diagram) at sample ratios; every limiter in
  Sweby's second-order TVD region (0 <= psi(r) <= min(2r, 2), psi = 0 for
  r <= 0, psi(1) = 1), and the symmetric ones satisfy psi(r)/r = psi(1/r).
- On a linear field the least-squares gradient is exact, so r = 1 and
  psi(r) = 1 (the face value is the linear one) on every face between cells
  with exact gradients (refined grids with hanging faces, scalar and vector
  fields, either flux sign).
- At a sharp step the ratio is <= 0 on the faces next to it, so the scheme
  falls back to upwind there.
- The reference's pulse test (tests/integration/fv/test_advection_bounds.py):
  a square pulse advected across a refinement interface stays within
  [0, 1] to 1e-10 and keeps its mass to 1e-8 for upwind and every TVD
  scheme."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import fv_oracle as fo
from tests.cases import FLUXEL

TVD = list(fo.TVD_LIMITERS)
R = np.array([-2.0, -0.5, 0.0, 0.1, 0.25, 0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 10.0])


def test_textbook_limiter_values():
    pos = np.maximum(R, 0.0)
    expect = {
        "vanleer": np.where(R > 0, 2 * R / (1 + R), 0.0),
        "minmod": np.minimum(pos, 1.0),
        "superbee": np.maximum(np.maximum(np.minimum(2 * pos, 1.0), np.minimum(pos, 2.0)), 0.0),
        "monotonized_central": np.minimum(np.minimum(2 * pos, (1 + pos) / 2), 2.0) * (R > 0),
        "limitedlinear": np.minimum(2 * pos, 1.0),
    }
    for name, e in expect.items():
        np.testing.assert_allclose(fo.TVD_LIMITERS[name](R), e, rtol=1e-15, atol=1e-300, err_msg=name)


@pytest.mark.parametrize("name", TVD)
def test_limiters_lie_in_the_second_order_tvd_region(name):
    lim = fo.TVD_LIMITERS[name]
    r = np.linspace(-3, 8, 2001)
    v = lim(r)
    assert np.all(v[r <= 0] == 0.0)
    assert np.all(v >= 0.0) and np.all(v <= np.minimum(2 * np.maximum(r, 0), 2.0) + 1e-15)
    assert lim(np.array([1.0]))[0] == pytest.approx(1.0, abs=1e-15)
    if name in ("vanleer", "minmod", "superbee", "monotonized_central"):
        rr = r[r > 0.05]
        np.testing.assert_allclose(lim(rr) / rr, lim(1.0 / rr), rtol=1e-12)


def _grid(name):
    f = FLUXEL[name]
    return fo.Grid(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"], dt=0.1)


def _linear_field(g, comps, kz):
    """A linear field without boundary conditions (uncovered faces act as
    zero gradient): the least-squares gradient is exact in every cell whose
    boundary faces are normal to a direction the field does not vary in."""
    c = g.cell_centers
    if comps == ():
        data = 0.3 + 1.2 * c[:, 0] - 0.7 * c[:, 1] + kz * 0.4 * c[:, 2]
    else:
        A = np.array([[1.0, -0.5, 0.2], [0.3, 0.8, -1.1], [-0.6, 0.1, 0.9]])
        A[:, 2] *= kz
        data = c @ A.T + np.array([0.1, -0.2, 0.3])
    return fo.Field("q", data, {})


@pytest.mark.parametrize("name", ["refined2d", "refined3d", "walls8"])
@pytest.mark.parametrize("comps", [(), (3,)])
def test_tvd_ratio_is_one_on_linear_fields(name, comps):
    """r = 1 (so psi(r) = 1 and the face value is the linear one) on every
    face between cells whose gradient is exact, for either flux sign."""
    g = _grid(name)
    twod = FLUXEL[name]["root"][2] == 1  # one root layer: the field does not vary in z
    fld = _linear_field(g, comps, 0.0 if twod else 1.0)
    has_bnd = np.zeros(g.num_cells, bool)
    varies = g.bnd_dir // 2 != 2 if twod else np.ones(g.num_bnd, bool)
    has_bnd[g.bnd_owner[varies]] = True
    inner = ~has_bnd[g.owner] & ~has_bnd[g.neighbour]
    if twod:  # z faces see no variation: r = -1 by the steep rule, psi = 0
        inner &= g.axis != 2
    assert inner.sum() > 10
    grad = fo.lsq_grad(g, fld)
    rng = np.random.default_rng(4)
    for up_own in (np.ones(g.num_faces, bool), rng.random(g.num_faces) < 0.5):
        r = fo.tvd_r(g, fld, grad, up_own)
        np.testing.assert_allclose(r[inner], 1.0, rtol=0, atol=1e-11)
        for s in TVD:
            np.testing.assert_allclose(fo.TVD_LIMITERS[s](r[inner]), 1.0, rtol=0, atol=3e-11, err_msg=s)


def test_tvd_is_upwind_next_to_a_step():
    """Flow across a sharp step: r <= 0 on the faces at the step (the upwind
    cell's centred gradient is half the jump), so every limiter is 0 there."""
    g = _grid("uniform2d")
    x = g.cell_centers[:, 0]
    fld = fo.Field("q", (x > 0.5).astype(float), {})
    phi = fo.FaceValues(np.ones(g.num_faces), np.zeros(g.num_bnd))
    r = fo.tvd_r(g, fld, fo.lsq_grad(g, fld), phi.single > 0)
    step = (x[g.owner] < 0.5) & (x[g.neighbour] > 0.5)
    assert step.sum() == 4 and np.all(r[step] <= 1e-12)
    for s in TVD:
        assert np.all(fo.TVD_LIMITERS[s](r[step]) <= 1e-12), s


def pulse_metrics(scheme: str, refined: bool) -> dict:
    """tests/integration/fv/test_advection_bounds.py:40-141 on the oracle."""
    n, slab = 16, 0.1
    dx = 1.0 / n
    dt = 0.1 * dx
    regions = [([0.5, 0.0, 0.0], [1.0, slab, slab], 1)] if refined else []
    g = fo.Grid([0, 0, 0], [1.0, slab, slab], [n, 1, 1], regions, 0, dt=dt)
    x = g.cell_centers[:, 0]
    q = fo.Field("q", ((x >= 4 * dx) & (x < 7 * dx)).astype(float),
                 {p: fo.BC("dirichlet", 0.0) if p == "x_minus" else fo.BC("neumann", 0.0) for p in fo.PATCHES})
    phi = fo.FaceValues(g.Sf[:, 0].copy(), g.bnd_Sf[:, 0].copy())
    cfg = fo.SolverCfg("bicgstab", "jacobi", 1e-13, 0.0, 1000)
    V = g.cell_volumes
    m0 = float((q.data * V).sum())
    lo, hi, err, out = 0.0, 1.0, 0.0, 0.0
    steps = math.ceil((8 + 1 - 4) / 0.1)
    for _ in range(steps):
        q.old = q.data.copy()
        mat = fo.fvm_ddt(g, q, (1.0, 1.0, 0.0)) + fo.fvm_div(g, phi, q, scheme)
        q.data, st = fo.solve(mat, q.data, cfg)
        assert st[0].converged
        lo, hi = min(lo, float(q.data.min())), max(hi, float(q.data.max()))
        out += dt * float((np.maximum(phi.bnd, 0.0) * q.data[g.bnd_owner]).sum())
        err = max(err, abs(float((q.data * V).sum()) + out - m0) / m0)
    assert float((q.data[x > 0.5] * V[x > 0.5]).sum()) > 0.5 * m0
    return {"min": lo, "max": hi, "relative_mass_error": err, "q": q.data}


@pytest.mark.parametrize("refined", [False, True])
@pytest.mark.parametrize("scheme", ["upwind"] + TVD)
def test_pulse_bounds_and_conservation(scheme, refined):
    m = pulse_metrics(scheme, refined)
    assert m["relative_mass_error"] < 1e-8, m
    assert m["min"] >= -1e-10 and m["max"] <= 1 + 1e-10, m
