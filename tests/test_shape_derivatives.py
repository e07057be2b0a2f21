"""CPU: the shape derivatives of a Fourier body's axis-ray hits and normals
(gridfoam_amd.shape) against central differences of the oracle's bisection
hits and analytic normals (oracle/fv_oracle.py _body_hits, _body_normal),
on rays through the outline of the 2-D channel's body."""

from __future__ import annotations

import numpy as np
import torch

from oracle import fv_oracle as fo

BODY = {"name": "body", "kind": "fourier", "center": [5.0, 4.0], "radius": 0.5,
        "a": [-0.046208236005, -0.007465017225, -0.021216333571, 0.019990187096,
              -0.017680500092, -0.025510938604, -0.012464491597, -0.017328832038],
        "b": [-0.025912542807, 0.030472632462, 0.00647321133, 0.04035452063,
              0.0227148468, -0.02453762678, 0.001427769424, 0.006760348061]}
H = 1e-6


def _rays(n: int, seed: int):
    """Rays along x or y through points near the outline: segments of length
    0.06 centred on the outline point at a random angle."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(-np.pi, np.pi, n)
    r = np.ones(n)
    for k in range(1, 9):
        r += BODY["a"][k - 1] * np.cos(k * th) + BODY["b"][k - 1] * np.sin(k * th)
    x = 5.0 + 0.5 * r * np.cos(th)
    y = 4.0 + 0.5 * r * np.sin(th)
    ax = rng.integers(0, 2, n)
    P = np.stack([x, y, np.full(n, 0.25)], axis=1)
    s = np.where(ax == 0, x, y)
    off = rng.uniform(-0.02, 0.02, n)
    return P, ax, s + off - 0.03, s + off + 0.03


def _with(coef: np.ndarray) -> dict:
    b = dict(BODY)
    b["a"], b["b"] = list(coef[:8]), list(coef[8:])
    return b


def _hits(coef: np.ndarray, P, ax, s0, s1) -> np.ndarray:
    h, _ = fo._body_hits(_with(coef), ax, P, s0, s1)
    return h


def _coef() -> np.ndarray:
    return np.array(BODY["a"] + BODY["b"], dtype=np.float64)


def _jacobian(fn, c: torch.Tensor, m: int) -> np.ndarray:
    rows = []
    out = fn(c)
    flat = out.reshape(-1)
    for i in range(flat.numel()):
        (g,) = torch.autograd.grad(flat[i], c, retain_graph=True)
        rows.append(g.numpy())
    return np.array(rows).reshape(*out.shape, m)


def test_hit_derivative_matches_central_differences_of_the_bisection():
    from gridfoam_amd.shape import hit_shift

    P, ax, s0, s1 = _rays(64, 3)
    c0 = _coef()
    hit = _hits(c0, P, ax, s0, s1)
    ok = ~np.isnan(hit)
    assert ok.sum() >= 48
    P, ax, s0, s1, hit = P[ok], ax[ok], s0[ok], s1[ok], hit[ok]
    c = torch.tensor(c0, requires_grad=True)
    Pt, axt, ht = torch.tensor(P), torch.tensor(ax), torch.tensor(hit)
    shift = hit_shift(c, 0.5, (5.0, 4.0), Pt, axt, ht)
    assert torch.equal(shift, torch.zeros_like(shift))  # zero in value
    J = _jacobian(lambda cc: hit_shift(cc, 0.5, (5.0, 4.0), Pt, axt, ht), c, 16)
    fd = np.zeros_like(J)
    for j in range(16):
        e = np.zeros(16)
        e[j] = H
        hp, hm = _hits(c0 + e, P, ax, s0, s1), _hits(c0 - e, P, ax, s0, s1)
        assert not (np.isnan(hp).any() or np.isnan(hm).any())
        fd[:, j] = (hp - hm) / (2 * H)
    scale = np.abs(fd).max()
    assert scale > 0.1
    np.testing.assert_allclose(J, fd, rtol=0, atol=1e-6 * scale)


def test_normal_derivative_matches_central_differences_of_the_analytic_normal():
    from gridfoam_amd.shape import _ray_points, hit_shift, outward_normal

    P, ax, s0, s1 = _rays(48, 5)
    c0 = _coef()
    hit = _hits(c0, P, ax, s0, s1)
    ok = ~np.isnan(hit)
    P, ax, s0, s1, hit = P[ok], ax[ok], s0[ok], s1[ok], hit[ok]
    Pt, axt, ht = torch.tensor(P), torch.tensor(ax), torch.tensor(hit)
    c = torch.tensor(c0, requires_grad=True)

    def normal(cc):
        s = ht + hit_shift(cc, 0.5, (5.0, 4.0), Pt, axt, ht)
        return outward_normal(cc, 0.5, (5.0, 4.0), *_ray_points(Pt, axt, s))[:, :2]

    n0 = normal(c).detach().numpy()
    np.testing.assert_allclose(n0, fo._body_normal(BODY, ax, P, hit)[:, :2], rtol=0, atol=1e-13)
    J = _jacobian(normal, c, 16)
    fd = np.zeros_like(J)
    for j in range(16):
        e = np.zeros(16)
        e[j] = H
        hp, hm = _hits(c0 + e, P, ax, s0, s1), _hits(c0 - e, P, ax, s0, s1)
        npl = fo._body_normal(_with(c0 + e), ax, P, hp)[:, :2]
        nmi = fo._body_normal(_with(c0 - e), ax, P, hm)[:, :2]
        fd[:, :, j] = (npl - nmi) / (2 * H)
    scale = np.abs(fd).max()
    assert scale > 0.1
    np.testing.assert_allclose(J, fd, rtol=0, atol=1e-6 * scale)


def test_sign_slip_in_the_hit_derivative_is_caught():
    """Negative control: the hit derivative with the sign of dG/ds flipped
    (a plausible slip) is far from the differences the tests above accept."""
    from gridfoam_amd.shape import _ray_points, outline_gradient, outline_value

    P, ax, s0, s1 = _rays(32, 7)
    c0 = _coef()
    hit = _hits(c0, P, ax, s0, s1)
    ok = ~np.isnan(hit)
    P, ax, s0, s1, hit = P[ok], ax[ok], s0[ok], s1[ok], hit[ok]
    Pt, axt, ht = torch.tensor(P), torch.tensor(ax), torch.tensor(hit)
    c = torch.tensor(c0, requires_grad=True)
    x, y = _ray_points(Pt, axt, ht)
    gx, gy = outline_gradient(c.detach(), 0.5, (5.0, 4.0), x, y)
    gs = torch.where(axt == 0, gx, gy)
    wrong = outline_value(c, 0.5, (5.0, 4.0), x, y) / gs  # sign slip
    (g,) = torch.autograd.grad(wrong.sum(), c)
    e = np.zeros(16)
    e[0] = H
    fd = ((_hits(c0 + e, P, ax, s0, s1) - _hits(c0 - e, P, ax, s0, s1)) / (2 * H)).sum()
    assert abs(float(g[0]) - fd) > 0.1 * abs(fd)
