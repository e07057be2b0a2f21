"""CPU: the oracle's bluff body (implicit function, axis-ray hits, normals)
and its surface refinement to target_level (meta/config.py:100-106), pinned
by closed forms on the body's faces, nose rounding, slant and struts, by the
1-Lipschitz bound the surface test relies on, and by the refinement's own
fixed-point property and surface coverage."""

from __future__ import annotations

import math

import numpy as np

from oracle import fv_oracle as fo

BODY = {"name": "body", "kind": "bluff", "min": [2.0, -0.195, 0.05], "max": [3.0, 0.195, 0.34],
        "radius": 0.1, "slant_angle": 25.0, "slant_length": 0.22,
        "struts": [[2.2, -0.12], [2.2, 0.12], [2.7, -0.12], [2.7, 0.12]],
        "strut_radius": 0.015, "strut_height": 0.05}


def G(x, y, z):
    return fo._bluff_outline(BODY, np.atleast_1d(np.float64(x)), np.atleast_1d(np.float64(y)),
                             np.atleast_1d(np.float64(z)))[0]


def test_closed_forms_on_the_surface_and_off_it():
    # side faces, bottom, top ahead of the slant: zero on the face, signed distance off it
    assert G(2.5, 0.195, 0.2) == 0.0 and abs(G(2.5, 0.295, 0.2) - 0.1) < 1e-15
    assert G(2.5, 0.0, 0.34) == 0.0 and abs(G(2.5, 0.0, 0.2) + 0.14) < 1e-15
    assert abs(G(2.5, 0.0, 0.05 - 0.02) - 0.02) < 1e-15
    # nose face (inside the rounding's inner lines): distance ahead of x0
    assert abs(G(1.9, 0.0, 0.2) - 0.1) < 1e-15
    # nose rounding: a quarter circle of radius 0.1 about (x0 + r, y1 - r)
    for phi in (0.3, 0.8, 1.2):
        x = 2.0 + 0.1 - 0.1 * math.cos(phi)
        y = 0.195 - 0.1 + 0.1 * math.sin(phi)
        assert abs(G(x, y, 0.2)) < 1e-15
        assert abs(G(x - 0.01 * math.cos(phi), y + 0.01 * math.sin(phi), 0.2) - 0.01) < 1e-12
    # the slant: a 25 degree plane through the top 0.22 ahead of the rear
    xs = 3.0 - 0.22
    for t in (0.05, 0.1, 0.2):
        assert abs(G(xs + t, 0.0, 0.34 - t * math.tan(math.radians(25.0)))) < 1e-15
    assert abs(G(3.0, 0.0, 0.34 - 0.22 * math.tan(math.radians(25.0)) - 0.05)) < 1e-15  # rear face below it
    # struts: a cylinder of radius 0.015 from the floor (z = 0) up
    assert abs(G(2.2 + 0.015, -0.12, 0.02)) < 1e-15
    assert abs(G(2.2, -0.12, 0.02) + 0.015) < 1e-15
    assert G(2.45, 0.0, 0.02) > 0.0  # between the struts under the body: fluid
    assert G(2.2, -0.12, -0.001) > 0.0  # below the floor level


def test_implicit_function_is_1_lipschitz():
    """|G(p) - G(q)| <= |p - q|: the surface refinement's test relies on it."""
    rng = np.random.default_rng(0)
    lo, hi = np.array([1.7, -0.4, -0.05]), np.array([3.3, 0.4, 0.5])
    p = lo + (hi - lo) * rng.random((200000, 3))
    q = p + 0.02 * rng.standard_normal(p.shape)
    gp = fo._bluff_outline(BODY, *p.T)
    gq = fo._bluff_outline(BODY, *q.T)
    assert np.all(np.abs(gp - gq) <= np.linalg.norm(p - q, axis=1) * (1 + 1e-12) + 1e-15)


def test_axis_ray_hits_lie_on_the_surface_with_outward_normals():
    rng = np.random.default_rng(1)
    n = 4000
    ax = rng.integers(0, 3, n)
    P = np.array([1.8, -0.3, 0.0]) + np.array([1.4, 0.6, 0.45]) * rng.random((n, 3))
    s0 = P[np.arange(n), ax] - 0.02
    s1 = s0 + 0.04
    h, h2 = fo._body_hits(BODY, ax, P, s0, s1)
    ok = ~np.isnan(h)
    assert ok.sum() > 100 and np.array_equal(h[ok], h2[ok])
    q = P[ok].copy()
    q[np.arange(ok.sum()), ax[ok]] = h[ok]
    assert np.abs(fo._bluff_outline(BODY, *q.T)).max() < 1e-12
    nrm = fo._body_normal(BODY, ax[ok], P[ok], h[ok])
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-12)
    # outward: a step along the normal leaves the body
    out = q + 1e-4 * nrm
    assert np.all(fo._bluff_outline(BODY, *out.T) > 0.0)


def test_surface_refinement_reaches_target_level_at_the_surface_and_nowhere_below_it():
    lower, upper, root, tl = [0.0, -1.0, 0.0], [10.0, 1.0, 1.5], [20, 4, 3], 3
    topo = fo.build_octree(lower, upper, root, (), 0, [BODY], tl)
    ext = np.asarray(upper) - np.asarray(lower)
    h = ext[None, :] / (np.asarray(root)[None, :] * (2.0 ** topo.level)[:, None])
    ctr = np.asarray(lower)[None, :] + (topo.ijk + 0.5) * h
    hd = 0.5 * np.linalg.norm(h, axis=1)
    g = fo._bluff_outline(BODY, *ctr.T)
    # fixed point: no leaf below the target level passes the surface test
    assert not np.any((topo.level < tl) & (np.abs(g) <= hd))
    # coverage: every sampled surface point lies in a leaf at the target level
    rng = np.random.default_rng(2)
    n = 3000
    ax = rng.integers(0, 3, n)
    P = np.array([1.8, -0.3, 0.0]) + np.array([1.4, 0.6, 0.45]) * rng.random((n, 3))
    s0 = P[np.arange(n), ax] - 0.05
    hit, _ = fo._body_hits(BODY, ax, P, s0, s0 + 0.1)
    ok = ~np.isnan(hit)
    q = P[ok].copy()
    q[np.arange(ok.sum()), ax[ok]] = hit[ok]
    lo_c, hi_c = ctr - h / 2, ctr + h / 2
    for pt in q[:400]:
        inside = np.all((lo_c <= pt) & (pt <= hi_c), axis=1)
        assert inside.any() and topo.level[inside].max() == tl
    assert topo.level.max() == tl and (topo.level == tl).sum() > 1000
    # the build without bodies is the unrefined root grid
    assert fo.build_octree(lower, upper, root, (), 0, [], tl).level.max() == 0
