"""Pins of the force-coefficient restatement (post/forces/*.py): the
coordinate bases and the controlled pressure scenarios of the reference's
tests/e2e/test_force_coeffs.py, on analytic bodies (the reference builds
the same cube and plate from STL files), and the body normals that stand in
for the surface-mesh faces."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import fv_oracle as fo


def test_drag_lift_coord_is_right_handed():
    """test_force_coeffs.py:45-74, for the oracle and the host basis."""
    from gridfoam_amd.config import DragLiftCoord
    from gridfoam_amd.forces import OrthonormalCoord

    for e1, e2, e3 in (fo.orthonormal_coord([2.0, 0.0, 0.0], lift_dir=[1.0, 0.0, 3.0]),
                       tuple(t.numpy() for t in (lambda c: (c.e1, c.e2, c.e3))(OrthonormalCoord.from_local_coord(
                           DragLiftCoord(drag_dir=[2.0, 0.0, 0.0], lift_dir=[1.0, 0.0, 3.0],
                                         center_of_rotation=[0.0, 0.0, 0.0]))))):
        for e in (e1, e2, e3):
            assert abs(np.linalg.norm(e) - 1.0) < 1e-12
        assert abs(e1 @ e2) < 1e-12 and abs(e2 @ e3) < 1e-12 and abs(e3 @ e1) < 1e-12
        np.testing.assert_allclose(np.cross(e1, e2), e3, atol=1e-12)
        np.testing.assert_allclose(e1, [1, 0, 0], atol=1e-12)
        np.testing.assert_allclose(e3, [0, 0, 1], atol=1e-12)


def test_drag_pitch_coord_is_right_handed():
    """test_force_coeffs.py:77-102."""
    from gridfoam_amd.config import DragPitchCoord
    from gridfoam_amd.forces import OrthonormalCoord

    c = OrthonormalCoord.from_local_coord(DragPitchCoord(drag_dir=[1.0, 0.0, 0.0], pitch_axis=[0.5, 1.0, 0.0],
                                                         center_of_rotation=[0.0, 0.0, 0.0]))
    for e1, e2, e3 in (fo.orthonormal_coord([1.0, 0.0, 0.0], pitch_axis=[0.5, 1.0, 0.0]),
                       (c.e1.numpy(), c.e2.numpy(), c.e3.numpy())):
        assert abs(e1 @ e2) < 1e-12 and abs(e2 @ e3) < 1e-12 and abs(e3 @ e1) < 1e-12
        np.testing.assert_allclose(np.cross(e1, e2), e3, atol=1e-12)
        np.testing.assert_allclose(e2, [0, 1, 0], atol=1e-12)


def test_coefficients_follow_the_axle_split():
    """coeffs.py:46-74 on a known force and moment."""
    coord = fo.orthonormal_coord([1, 0, 0], lift_dir=[0, 0, 1])
    c = fo.force_coeffs(np.array([2.0, 3.0, 4.0]), np.array([0.5, 0.25, 1.0]), coord, 2.0, 4.0)
    # e1 = x, e3 = z, e2 = e3 x e1 = y
    assert (c["Cd"], c["Cs"], c["Cl"]) == (1.0, 1.5, 2.0)
    assert (c["CmRoll"], c["CmPitch"], c["CmYaw"]) == (0.125, 0.0625, 0.25)
    assert c["Cd_f"] == 0.5 + 0.125 and c["Cl_r"] == 1.0 - 0.0625 and c["Cs_f"] == 0.75 + 0.25


CUBE = dict(name="body", kind="box", min=[0.4, 0.4, 0.4], max=[0.6, 0.6, 0.6])
PLATE = dict(name="body", kind="box", min=[0.3, 0.3, 0.49], max=[0.7, 0.7, 0.51])


def _body_grid(body, n):
    return fo.Grid([0, 0, 0], [1, 1, 1], [n, n, n], bodies=[body])


def _fields(g, p_data):
    U = fo.Field("U", np.zeros((g.num_cells, 3)), {"body": fo.BC("dirichlet", np.zeros(3))})
    p = fo.Field("p", np.asarray(p_data, np.float64), {"body": fo.BC("neumann", 0.0)})
    return U, p


@pytest.mark.parametrize("body", ["cube", "flat_plate"])
def test_uniform_pressure_gives_zero_force_on_closed_body(body):
    """test_force_coeffs.py:105-152: p = 2.5, U = 0 -> every coefficient 0."""
    g = _body_grid(CUBE if body == "cube" else PLATE, 16)
    U, p = _fields(g, np.full(g.num_cells, 2.5))
    f, m = fo.patch_force(g, U, p, 0.1, "body", 1.0, [0.5, 0.5, 0.5])
    c = fo.force_coeffs(f, m, fo.orthonormal_coord([1, 0, 0], lift_dir=[0, 0, 1]), 0.5, 0.5)
    assert all(abs(v) < 1e-10 for v in c.values())


def _jump(n):
    g = _body_grid(PLATE, n)
    U, p = _fields(g, np.where(g.cell_centers[:, 2] > 0.5, 1.0, 2.0))
    f, m = fo.patch_force(g, U, p, 0.0, "body", 1.0, [0.5, 0.5, 0.5])
    cl = fo.force_coeffs(f, m, fo.orthonormal_coord([1, 0, 0], lift_dir=[0, 0, 1]), 0.5, 0.5)["Cl"]
    expected = (2.0 - 1.0) * 0.4 * 0.4 / 0.5
    return abs(cl - expected) / expected


def test_thin_sheet_pressure_jump_converges_with_refinement():
    """test_force_coeffs.py:155-240: lift of a thin sheet with p 2 below and
    1 above is within 10% of (p_below - p_above) area / q on the finest grid
    and no worse than on the coarsest."""
    errors = [_jump(n) for n in (16, 32, 64)]
    assert errors[-1] < 0.10, errors
    assert errors[-1] <= errors[0] + 1e-12, errors


def test_fourier_normals_are_the_outline_gradient():
    """The body-outward normal at each Fourier hit is the normalised
    gradient of |P - c| - R r(theta), checked by central differences."""
    body = dict(name="body", kind="fourier", center=[5.0, 4.0], radius=0.5,
                a=[0.03, -0.02, 0.01, 0.0, 0.02, 0.0, -0.01, 0.01], b=[0.01, 0.02, -0.03, 0.0, 0.0, 0.01, 0.0, 0.02])
    g = fo.Grid([0, 0, 0], [10, 8, 0.25], [40, 32, 1], [([3.5, 2.5, 0.0], [6.5, 5.5, 0.25], 2)], bodies=[body])
    imm = g.is_immersed
    o = g.topo.owner.astype(np.int64)[imm]
    ax = g.topo.axis.astype(np.int64)[imm]
    hit = g.cell_centers[o, ax] + g.ap.dist_owner
    C = g.cell_centers
    P = C[o].copy()
    P[np.arange(len(o)), ax] = hit
    # the face centre's tangential coordinate is the ray's (finer side for 2:1 faces)
    nb = g.topo.neighbour.astype(np.int64)[imm]
    fine = np.where(g.cell_sizes[o, 0] <= g.cell_sizes[nb, 0], o, nb)
    t = 1 - ax  # the other in-plane axis
    P[np.arange(len(o)), t] = C[fine, t]
    h = 1e-7
    gx = (fo._fourier_outline(body, P[:, 0] + h, P[:, 1]) - fo._fourier_outline(body, P[:, 0] - h, P[:, 1])) / (2 * h)
    gy = (fo._fourier_outline(body, P[:, 0], P[:, 1] + h) - fo._fourier_outline(body, P[:, 0], P[:, 1] - h)) / (2 * h)
    gn = np.sqrt(gx * gx + gy * gy)
    np.testing.assert_allclose(g.ap.owner_normal[:, 0], gx / gn, atol=1e-6)
    np.testing.assert_allclose(g.ap.owner_normal[:, 1], gy / gn, atol=1e-6)
    np.testing.assert_allclose(g.ap.owner_normal, g.ap.neighbour_normal, atol=0)
    np.testing.assert_allclose(np.linalg.norm(g.ap.owner_normal, axis=1), 1.0, atol=1e-14)


def test_box_normals_point_out_of_the_box():
    g = _body_grid(CUBE, 16)
    imm = g.is_immersed
    o = g.topo.owner.astype(np.int64)[imm]
    ax = g.topo.axis.astype(np.int64)[imm]
    hit = g.cell_centers[o, ax] + g.ap.dist_owner
    centre = 0.5
    sign = np.sign(hit - centre)
    np.testing.assert_array_equal(g.ap.owner_normal[np.arange(len(o)), ax], sign)
    assert np.all(np.abs(g.ap.owner_normal).sum(axis=1) == 1.0)


def test_kept_sides_face_the_body():
    """Exactly one side of every cut face of a closed body is kept (the
    fluid side facing the body), so no sample is counted twice."""
    g = _body_grid(CUBE, 16)
    Bd, Fi = g.num_domain_bnd, g.num_immersed
    n_up = g.bnd_Sf[Bd:Bd + Fi] / np.linalg.norm(g.bnd_Sf[Bd:Bd + Fi], axis=1)[:, None]
    keep_up = np.sum(n_up * -g.ap.owner_normal, axis=1) > 0
    keep_lo = np.sum(-n_up * -g.ap.neighbour_normal, axis=1) > 0
    assert np.all(keep_up ^ keep_lo)
