"""Shape derivatives for an immersed Fourier body.

A ``fourier`` body's outline is r(theta) = R (1 + sum_k a_k cos k theta +
b_k sin k theta), k = 1..8. It reaches the discrete problem through two
continuous quantities of every immersed side that hits it (``gf_ap_cut``):

- the ray distance |d| from the cell centre to the hit along the face's axis;
- the body's outward normal at the hit.

Which faces are cut and which cells are snapped are discrete choices. The
derivative here is that of the discrete problem on the fixed cut topology.
Central differences reproduce it as long as a perturbation moves no cell
centre across the outline and no side across the snapping distance.

Hit derivative. Let G(x; c) = |x - x_c| - R r(theta(x); c) and let s* be
the hit on the axis ray (G(s*) = 0). By the implicit function theorem,
ds*/dc = -(dG/dc) / (dG/ds).

While a ``FourierShape`` is active (``with FourierShape(grid, "body") as
shape:``), the grid's immersed distances and normals are straight-through
values: v + (t - t.detach()). Their values are the kernel's, bit for bit;
their derivative with respect to ``shape.coefficients`` is that of t. The
operators that read |d| add a term that is zero in value and carries that
derivative:

- the Laplacian's boundary coefficient fraction / |d| on the immersed sides
  (``fvm.assemble``: diagonal and source, fv/boundary_ops.py:372-392);
- the least-squares gradient's boundary term (d / |d|^2)(psi_b - psi_P)
  (``fvc.grad``, fv/schemes/least_squares.py:29-101);
- the pressure-flux boundary coefficient (``grid.bnd_distances()``).

The force sampling reads the grid's distance and normal arrays directly
(post/forces/sampling.py:68-168). So forward values are unchanged, and one
backward pass gives the gradient of an objective with respect to all 16
coefficients (a_1..a_8, b_1..b_8) at once.

Not covered (they raise NotImplementedError):

- sides of the body on cells next to hanging faces, whose cell gradient
  enters the assembly's hanging corrections;
- conditions on the body whose boundary value psi_P + |d| g depends on |d|:
  a Neumann gradient that is not zero, or a mixed condition.
"""

from __future__ import annotations

from typing import Tuple

import torch

N_MODES = 8


def _modes(coef: torch.Tensor, th: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """(r(theta) - 1, r'(theta)) for coefficients [a_1..a_8, b_1..b_8]."""
    k = torch.arange(1, N_MODES + 1, dtype=torch.float64, device=th.device)
    kt = th[:, None] * k[None, :]
    c, s = torch.cos(kt), torch.sin(kt)
    a, b = coef[:N_MODES], coef[N_MODES:]
    return (c * a).sum(dim=1) + (s * b).sum(dim=1), (k * (b * c - a * s)).sum(dim=1)


def _ray_points(P: torch.Tensor, axis: torch.Tensor, s: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """(x, y) of the ray point with coordinate ``axis`` set to s."""
    return torch.where(axis == 0, s, P[:, 0]), torch.where(axis == 1, s, P[:, 1])


def outline_gradient(coef, radius: float, center, x, y) -> Tuple[torch.Tensor, torch.Tensor]:
    """dG/dx, dG/dy of G = |x - x_c| - R r(theta)."""
    dx, dy = x - center[0], y - center[1]
    rho2 = dx * dx + dy * dy
    rho = torch.sqrt(rho2)
    _, dr = _modes(coef, torch.atan2(dy, dx))
    Rdr = radius * dr
    return dx / rho + Rdr * dy / rho2, dy / rho - Rdr * dx / rho2


def outline_value(coef, radius: float, center, x, y) -> torch.Tensor:
    """G = |x - x_c| - R r(theta): negative inside the outline."""
    dx, dy = x - center[0], y - center[1]
    r1, _ = _modes(coef, torch.atan2(dy, dx))
    return torch.sqrt(dx * dx + dy * dy) - radius * (1.0 + r1)


def hit_shift(coef, radius: float, center, P, axis, s) -> torch.Tensor:
    """A zero-valued shift of the hits s (on the rays through P along
    ``axis``). Its derivative with respect to coef is
    ds*/dc = -(dG/dc) / (dG/ds)."""
    with torch.no_grad():
        x, y = _ray_points(P, axis, s)
        gx, gy = outline_gradient(coef.detach(), radius, center, x, y)
        gs = torch.where(axis == 0, gx, gy)
    x, y = _ray_points(P, axis, s)
    t = -outline_value(coef, radius, center, x, y) / gs
    return t - t.detach()


def outward_normal(coef, radius: float, center, x, y) -> torch.Tensor:
    """The unit body-outward normal grad G / |grad G| in the outline's plane."""
    gx, gy = outline_gradient(coef, radius, center, x, y)
    gn = torch.sqrt(gx * gx + gy * gy)
    return torch.stack([gx / gn, gy / gn, torch.zeros_like(gx)], dim=1)


class FourierShape:
    """The 16 outline coefficients of one immersed ``fourier`` body as a leaf
    tensor, ``coefficients`` = [a_1..a_8, b_1..b_8], zero-padded past the
    configured modes. Inside ``with shape:``, the grid's immersed geometry
    depends on it differentiably. Entering rebuilds that dependence, so each
    ``with`` block holds one forward pass and one backward pass."""

    _ARRAYS = ("ap_dist_owner_to_bnd", "ap_dist_neighbour_to_bnd", "ap_owner_normal", "ap_neighbour_normal")

    def __init__(self, grid, body: str):
        names = [b.name for b in getattr(grid, "immersed_bodies", [])]
        if body not in names:
            raise ValueError(f"no immersed body named {body!r}")
        idx = names.index(body)
        cfg = grid.immersed_bodies[idx]
        if cfg.kind != "fourier":
            raise ValueError(f"body {body!r} is a {cfg.kind}, not a fourier body")
        self.grid, self.body, self.index = grid, body, idx
        a = list(cfg.a) + [0.0] * (N_MODES - len(cfg.a))
        b = list(cfg.b) + [0.0] * (N_MODES - len(cfg.b))
        dev = grid.device
        self.coefficients = torch.tensor(a + b, dtype=torch.float64, device=dev, requires_grad=True)
        self.radius = float(cfg.radius)
        self.center = (float(cfg.center[0]), float(cfg.center[1]))
        Fi, Bd = grid.num_immersed_faces, grid.num_domain_bnd_faces
        self._sel_o = torch.nonzero(grid.ap_owner_bnd_patch_id == idx).flatten()
        self._sel_n = torch.nonzero(grid.ap_neighbour_bnd_patch_id == idx).flatten()
        ax = (grid.bnd_dir_i32[Bd:Bd + Fi] // 2).long()
        own, nbr = grid.bnd_owner[Bd:Bd + Fi], grid.bnd_owner[Bd + Fi:Bd + 2 * Fi]
        if bool(grid.hang_cell_mask[own[self._sel_o]].any()) or bool(grid.hang_cell_mask[nbr[self._sel_n]].any()):
            raise NotImplementedError("shape derivatives with body sides on cells next to hanging faces")
        cc = grid.cell_centers
        so, sn = self._sel_o, self._sel_n
        self._ax_o, self._ax_n = ax[so], ax[sn]
        self._P_o, self._P_n = grid.ap_face_centers[so], grid.ap_face_centers[sn]
        # the kernel's hits: owner centre + |d| and neighbour centre - |d| along the axis
        self._hit_o = cc[own[so], self._ax_o] + grid.ap_dist_owner_to_bnd[so]
        self._hit_n = cc[nbr[sn], self._ax_n] - grid.ap_dist_neighbour_to_bnd[sn]
        # this body's sides in the boundary list, their cells, axes and d/|d| signs
        self.sides = torch.cat([Bd + so, Bd + Fi + sn])
        self._cells = grid.bnd_owner[self.sides]
        self._axis = torch.cat([self._ax_o, self._ax_n])
        self._sign = torch.cat([torch.ones(so.numel(), dtype=torch.float64, device=dev),
                                -torch.ones(sn.numel(), dtype=torch.float64, device=dev)])
        self._base = None
        self.bnd_mag_d = None

    # ---- installation ---------------------------------------------------------
    def _build(self) -> None:
        g, c = self.grid, self.coefficients
        base = self._base
        R, ctr = self.radius, self.center
        ds_o = hit_shift(c, R, ctr, self._P_o, self._ax_o, self._hit_o)
        ds_n = hit_shift(c, R, ctr, self._P_n, self._ax_n, self._hit_n)
        d_o, d_n, n_o, n_n = (base[k] for k in self._ARRAYS)
        d_o_ad = d_o + torch.zeros_like(d_o).index_add(0, self._sel_o, ds_o)
        d_n_ad = d_n + torch.zeros_like(d_n).index_add(0, self._sel_n, -ds_n)
        nt_o = outward_normal(c, R, ctr, *_ray_points(self._P_o, self._ax_o, self._hit_o + ds_o))
        nt_n = outward_normal(c, R, ctr, *_ray_points(self._P_n, self._ax_n, self._hit_n + ds_n))
        n_o_ad = n_o + torch.zeros_like(n_o).index_add(0, self._sel_o, nt_o - nt_o.detach())
        n_n_ad = n_n + torch.zeros_like(n_n).index_add(0, self._sel_n, nt_n - nt_n.detach())
        Bd = g.num_domain_bnd_faces
        self.bnd_mag_d = torch.cat([g.bnd_mag_d[:Bd], d_o_ad, d_n_ad])
        for k, v in zip(self._ARRAYS, (d_o_ad, d_n_ad, n_o_ad, n_n_ad)):
            setattr(g, k, v)

    def __enter__(self) -> "FourierShape":
        g = self.grid
        if g.shape is not None:
            raise RuntimeError("another shape is active on this grid")
        self._base = {k: getattr(g, k) for k in self._ARRAYS}
        self._build()
        g.shape = self
        return self

    def tensors(self) -> Tuple[torch.Tensor, ...]:
        """The geometry tensor that time steps read (the boundary distances;
        the distance and normal arrays feed only the force sampling): the
        ``params`` of a graph-free checkpointed rollout."""
        return (self.bnd_mag_d,)

    def __exit__(self, *exc) -> None:
        g = self.grid
        for k, v in self._base.items():
            setattr(g, k, v)
        g.shape = None
        self.bnd_mag_d = None
        self._base = None

    # ---- zero-valued terms that carry the derivative --------------------------
    def _inv_d_delta(self, active: torch.Tensor) -> torch.Tensor:
        """1/|d| (shape-dependent) - 1/|d| (fixed) on this body's sides: zero,
        and zero where not ``active``."""
        g = self.grid
        one = torch.ones_like(g.bnd_mag_d[self.sides])
        d_ad = torch.where(active, self.bnd_mag_d[self.sides], one)
        d0 = torch.where(active, g.bnd_mag_d[self.sides], one)
        return torch.where(active, 1.0 / d_ad - 1.0 / d0, torch.zeros_like(one))

    def _check_conditions(self, st, rg: torch.Tensor, nc: int) -> None:
        k = self.sides
        if k.numel() == 0:
            return
        cov = st.covered[k].bool()
        gnz = rg.detach()[k].reshape(-1, nc).abs().amax(dim=1) != 0
        if bool((cov & (st.frac[k] != 1.0) & gnz).any()):
            raise NotImplementedError("shape derivatives with a nonzero Neumann gradient on the body")

    def add_laplacian_terms(self, mat, field, st, rv, rg, lap: float, gamma) -> None:
        """The Laplacian's boundary coefficient on the body's sides
        (-lap gamma |Sf| fraction / |d| on the diagonal, times the condition
        value on the source). The terms are zero in value."""
        g, k = self.grid, self.sides
        if k.numel() == 0:
            return
        nc = field.nc
        self._check_conditions(st, rg, nc)
        fr = st.frac[k]
        active = st.covered[k].bool() & (fr != 0) & ~(g.bnd_snap[k].bool() & (fr == 1.0))
        cells = self._cells
        if isinstance(gamma, torch.Tensor) and gamma.dim() > 0:
            gb = gamma[cells]
        else:
            gb = gamma
        cv = -lap * (gb * g.bnd_mag_sf[k] * fr * self._inv_d_delta(active))
        C = g.num_cells
        dcorr = torch.zeros(C, dtype=torch.float64, device=g.device).index_add(0, cells, cv)
        scorr = torch.zeros((C, nc), dtype=torch.float64, device=g.device).index_add(
            0, cells, cv[:, None] * rv[k].reshape(-1, nc))
        mat._diag = mat._diag + dcorr
        mat._source = mat._source + scorr.reshape(mat._source.shape)

    def add_lsq_terms(self, field, st, out: torch.Tensor) -> torch.Tensor:
        """The least-squares gradient's boundary term on the body's sides:
        ata_inv_P (d / |d|^2)(psi_b - psi_P), with d = +-|d| e_axis. The
        term is zero in value."""
        g, k = self.grid, self.sides
        if k.numel() == 0:
            return out
        nc = field.nc
        cells = self._cells
        cov = st.covered[k].bool()
        dpsi = (st.psi_b[k].reshape(-1, nc) - field.data.detach().reshape(g.num_cells, nc)[cells])
        w = self._sign * self._inv_d_delta(cov)                      # [n]
        e = torch.nn.functional.one_hot(self._axis, 3).to(torch.float64)  # [n, 3]
        atb = dpsi[:, :, None] * (w[:, None] * e)[:, None, :]        # [n, nc, 3]
        A = g.lsq_ata_inv[cells].reshape(-1, 3, 3)
        corr_k = (A[:, None, :, :] * atb[:, :, None, :]).sum(dim=-1)  # [n, nc, 3]
        corr = torch.zeros((g.num_cells, nc, 3), dtype=torch.float64, device=g.device).index_add(0, cells, corr_k)
        return out + corr.reshape(out.shape)
