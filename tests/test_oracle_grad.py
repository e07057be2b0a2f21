"""Oracle pins for the Green-Gauss ("linear") cell gradient
(fv/schemes/grad.py:21-44, fv/kernels/gauss_gradient.py:15-73), after the
reference's own tests (tests/integration/fv/test_grad_schemes.py): exact
for a linear field with the analytic normal derivative on every patch,
including hanging cells (and not exact there without the hanging-face
correction of the face values); zero for a constant field with
zero-gradient patches; the [C, component, derivative] layout of a vector
field's gradient with a nonsymmetric Jacobian."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fv_oracle as fo
from tests.cases import FLUXEL

G = np.array([2.0, -3.0, 5.0])
J = np.array([[1.0, 2.0, -3.0], [4.0, -2.0, 1.0], [0.5, 3.0, 2.0]])


def _grid(name):
    f = FLUXEL[name]
    return fo.Grid(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"])


def _normal(col):
    """Neumann values sign * gradient[axis] per patch (the analytic normal
    derivative of a linear field)."""
    out = {}
    for k, p in enumerate(fo.PATCHES):
        axis, sign = k // 2, 2.0 * (k % 2) - 1.0
        out[p] = fo.BC("neumann", sign * col(axis))
    return out


@pytest.mark.parametrize("name", ["refined2d", "refined3d", "walls8", "odd"])
@pytest.mark.parametrize("scheme", ["linear", "leastsquare"])
def test_gradients_exact_for_linear_fields(name, scheme):
    g = _grid(name)
    q = fo.Field("q", g.cell_centers @ G + 7.0, _normal(lambda a: G[a]))
    gr = fo.cell_grad(g, q, scheme)
    np.testing.assert_allclose(gr, np.broadcast_to(G, gr.shape), rtol=0, atol=1e-11)
    U = fo.Field("U", g.cell_centers @ J.T, _normal(lambda a: J[:, a]))
    gU = fo.cell_grad(g, U, scheme)
    assert gU.shape == (g.num_cells, 3, 3)
    np.testing.assert_allclose(gU, np.broadcast_to(J, gU.shape), rtol=0, atol=1e-11)


def test_gauss_needs_the_hanging_face_correction():
    g = _grid("refined3d")
    q = fo.Field("q", g.cell_centers @ G + 7.0, _normal(lambda a: G[a]))
    fv = fo.interpolate(g, q)
    fv.single = fo.linear_face_values(g, q.data)  # two-point face values only
    flux = fv.single[:, None] * g.Sf
    out = fo._scatter(g.owner, flux, g.num_cells) - fo._scatter(g.neighbour, flux, g.num_cells)
    out = (out + fo._scatter(g.bnd_owner, fv.bnd[:, None] * g.bnd_Sf, g.num_cells)) / g.cell_volumes[:, None]
    assert np.abs(out - G).max() > 1e-2


def test_constant_field_has_zero_gauss_gradient():
    g = _grid("refined2d")
    q = fo.Field("q", np.full(g.num_cells, 3.0), {p: fo.BC("neumann", 0.0) for p in fo.PATCHES})
    np.testing.assert_allclose(fo.gauss_grad(g, q), 0.0, atol=1e-12)


@pytest.mark.parametrize("name", ["refined2d", "refined3d", "odd"])
def test_reconstruct_recovers_uniform_and_linear_velocity(name):
    """tests/integration/fv/test_reconstruct.py:30-69: a uniform velocity and
    U = (x, 2y, -3z) (face fluxes from face-centre values) are recovered in
    every cell, hanging sub-faces included."""
    g = _grid(name)
    u = np.array([1.0, 0.5, -0.25])
    phi = fo.FaceValues(g.Sf @ u, g.bnd_Sf @ u)
    np.testing.assert_allclose(fo.reconstruct(g, phi), np.broadcast_to(u, (g.num_cells, 3)), rtol=1e-12, atol=1e-12)
    s = np.array([1.0, 2.0, -3.0])
    phi = fo.FaceValues(np.sum(g.face_centers * s * g.Sf, axis=1), np.sum(g.bnd_face_centers * s * g.bnd_Sf, axis=1))
    np.testing.assert_allclose(fo.reconstruct(g, phi), g.cell_centers * s, rtol=1e-12, atol=1e-12)
