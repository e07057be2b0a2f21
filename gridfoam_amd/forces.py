"""Force and moment coefficients on immersed bodies.

Restates src/gridfoam/post/forces/ (evaluator.py:22-180, integration.py,
sampling.py, coeffs.py, coord.py).  The force is integrated over the
axis-projected sides of the immersed faces, as in the reference: every cut
face contributes one sample per side, the side whose outward direction
(from its fluid cell into the body) runs against the body's outward surface
normal is kept, and each kept sample carries

    pressure force  rho p_b Sf
    viscous force   -rho nu (grad U + grad U^T) Sf,  grad U = dU/dn (x) n,
                    dU/dn = (U_b - U_P) / (|d| n_hat . n)

with n the surface normal pointing into the body, U_b, p_b the boundary
values of the fields' conditions (the cell value where a patch has none)
and the moment arm running from the centre of rotation to the sample's
boundary point (cell centre + |d| n_hat).  The surface normal at each hit
comes from the analytic body (``gf_ap_cut``); the reference reads it from
the surface-mesh face the ray hit.  All arithmetic runs on the device in
FP64 over the [F_i] immersed-side arrays.
"""

from __future__ import annotations

import csv
import pathlib
from dataclasses import dataclass
from typing import List, Optional, Tuple

import torch

from .config import DragLiftCoord, DragPitchCoord, ForceCoeffConfig


def _unit(v: torch.Tensor) -> torch.Tensor:
    n = torch.linalg.vector_norm(v)
    if float(n) == 0.0:
        raise ValueError("vector is zero vector")
    return v / n


@dataclass(frozen=True)
class OrthonormalCoord:
    """post/forces/coord.py:21-74: e1 drag / roll, e2 side / pitch, e3 lift / yaw."""

    e1: torch.Tensor
    e2: torch.Tensor
    e3: torch.Tensor

    @classmethod
    def from_local_coord(cls, lc) -> "OrthonormalCoord":
        f64 = dict(dtype=torch.float64)
        e1 = _unit(torch.tensor(lc.drag_dir, **f64))
        if isinstance(lc, DragLiftCoord):
            lift = torch.tensor(lc.lift_dir, **f64)
            e3 = _unit(lift - e1 * (lift @ e1))
            e2 = torch.linalg.cross(e3, e1)
        elif isinstance(lc, DragPitchCoord):
            pitch = torch.tensor(lc.pitch_axis, **f64)
            e2 = _unit(pitch - e1 * (pitch @ e1))
            e3 = torch.linalg.cross(e1, e2)
        else:
            raise ValueError(f"unknown force coordinate mode {lc!r}")
        return cls(e1, e2, e3)

    def to(self, device) -> "OrthonormalCoord":
        return OrthonormalCoord(self.e1.to(device), self.e2.to(device), self.e3.to(device))


@dataclass(frozen=True)
class ForceCoeffs:
    """post/forces/coeffs.py:9-74."""

    time: float
    Cd: torch.Tensor
    Cd_f: torch.Tensor
    Cd_r: torch.Tensor
    Cl: torch.Tensor
    Cl_f: torch.Tensor
    Cl_r: torch.Tensor
    CmPitch: torch.Tensor
    CmRoll: torch.Tensor
    CmYaw: torch.Tensor
    Cs: torch.Tensor
    Cs_f: torch.Tensor
    Cs_r: torch.Tensor

    @classmethod
    def from_force_moment(cls, *, time: float, force: torch.Tensor, moment: torch.Tensor, coord: OrthonormalCoord,
                          force_scale: float, moment_scale: float) -> "ForceCoeffs":
        Cd, Cs, Cl = (torch.sum(force * e) / force_scale for e in (coord.e1, coord.e2, coord.e3))
        CmRoll, CmPitch, CmYaw = (torch.sum(moment * e) / moment_scale for e in (coord.e1, coord.e2, coord.e3))
        return cls(time=time, Cd=Cd, Cd_f=0.5 * Cd + CmRoll, Cd_r=0.5 * Cd - CmRoll, Cl=Cl, Cl_f=0.5 * Cl + CmPitch,
                   Cl_r=0.5 * Cl - CmPitch, CmPitch=CmPitch, CmRoll=CmRoll, CmYaw=CmYaw, Cs=Cs, Cs_f=0.5 * Cs + CmYaw,
                   Cs_r=0.5 * Cs - CmYaw)


CSV_HEADER = ["time", "Cd", "Cd_f", "Cd_r", "Cl", "Cl_f", "Cl_r", "CmPitch", "CmRoll", "CmYaw", "Cs", "Cs_f", "Cs_r"]


@dataclass
class SideForces:
    """Per-sample forces of one side block of one patch (the data the
    reference accumulates onto surface-mesh faces)."""

    patch: str
    upper: bool
    face: torch.Tensor            # [K] immersed-face index of each kept sample
    pressure_force: torch.Tensor  # [K, 3]
    viscous_force: torch.Tensor   # [K, 3]
    points: torch.Tensor          # [K, 3] boundary points


def _boundary_values(field, B0: int, sel: torch.Tensor) -> torch.Tensor:
    """sampling.py:28-65: f ref_v + (1 - f)(psi_P + |d| ref_g) on the
    selected boundary faces, the cell value where no condition owns them
    (the boundary state holds exactly that); autograd-connected to the
    field and its condition values when they require grad."""
    from .fv import boundary_values

    return boundary_values(field)[B0 + sel].reshape((sel.numel(), *field.component_shape))


def integrate_patch(grid, *, p, U, nu_eff, patch: str, rho: float, center_of_rotation: torch.Tensor,
                    collect: Optional[List[SideForces]] = None) -> Tuple[torch.Tensor, torch.Tensor]:
    """integration.py:106-187 with sampling.py:68-168: (force, moment) of one
    immersed patch."""
    dev = grid.device
    f64 = dict(dtype=torch.float64, device=dev)
    force = torch.zeros(3, **f64)
    moment = torch.zeros(3, **f64)
    Fi, Bd = grid.num_immersed_faces, grid.num_domain_bnd_faces
    if Fi == 0:
        return force, moment
    up_mask, lo_mask = grid.ap_get_patch_mask(patch)
    imm = grid.ap_is_immersed_faces
    F_all = grid.num_all_internal_faces
    # owner / neighbour / area vector of every immersed face, in immersed-face order
    own_i = grid.bnd_owner[Bd:Bd + Fi]
    nbr_i = grid.bnd_owner[Bd + Fi:Bd + 2 * Fi]
    Sf_i = grid.bnd_Sf[Bd:Bd + Fi]
    assert int(imm.sum()) == Fi and F_all >= Fi
    if torch.is_tensor(nu_eff):
        nu = nu_eff.expand(grid.num_cells) if nu_eff.dim() == 0 else nu_eff
    else:
        nu = torch.full((grid.num_cells,), float(nu_eff), **f64)
    sides = ((True, up_mask, own_i, Sf_i, grid.ap_dist_owner_to_bnd, grid.ap_owner_normal, Bd),
             (False, lo_mask, nbr_i, -Sf_i, grid.ap_dist_neighbour_to_bnd, grid.ap_neighbour_normal, Bd + Fi))
    for upper, mask, cells_all, Sf_all, dist_all, nout_all, B0 in sides:
        sel = torch.nonzero(mask).flatten()
        if sel.numel() == 0:
            continue
        cells, Sf, mag_d = cells_all[sel], Sf_all[sel], dist_all[sel]
        n_hat = Sf / torch.linalg.vector_norm(Sf, dim=1)[:, None]
        n_surface = -nout_all[sel]  # the fluid side's normal into the body
        keep = torch.sum(n_hat * n_surface, dim=1) > 0.0
        if not bool(keep.any()):
            continue
        k = torch.nonzero(keep).flatten()
        p_b = _boundary_values(p, B0, sel)[k]
        U_b = _boundary_values(U, B0, sel)[k]
        U_c = U.data[cells[k]]
        ns = n_surface[k]
        wall = mag_d[k] * torch.sum(n_hat[k] * ns, dim=1).clamp_min(1.0e-12)
        dUdn = (U_b - U_c) / wall[:, None]
        gradU = dUdn[:, :, None] * ns[:, None, :]
        stress = rho * nu[cells[k]][:, None, None] * (gradU + gradU.transpose(1, 2))
        fp = (rho * p_b)[:, None] * Sf[k]
        fv = -(stress * Sf[k][:, None, :]).sum(dim=-1)
        face_force = fp + fv
        points = grid.cell_centers[cells[k]] + mag_d[k][:, None] * n_hat[k]
        force = force + face_force.sum(dim=0)
        moment = moment + torch.linalg.cross(points - center_of_rotation, face_force, dim=1).sum(dim=0)
        if collect is not None:
            collect.append(SideForces(patch, upper, sel[k], fp, fv, points))
    return force, moment


class ForceEvaluator:
    """post/forces/evaluator.py:22-180: force and moment coefficients of the
    configured immersed patches (``post_processing.forceCoeff``)."""

    def __init__(self, grid, phase: Optional[str] = None, config: Optional[ForceCoeffConfig] = None):
        from .fields import CellField

        self.grid = grid
        if config is None:
            pp = grid.sim_config.post_processing if grid.sim_config is not None else None
            if pp is None or pp.forceCoeff is None:
                raise ValueError("post_processing.forceCoeff is not configured")
            config = pp.forceCoeff
        self.config = config
        self.patches = list(config.patches)
        self.rho, self.magU_ref, self.A_ref, self.L_ref = config.rho, config.magU_ref, config.A_ref, config.L_ref
        self.local_coord = OrthonormalCoord.from_local_coord(config.local_coord)
        self.CofR = torch.tensor(config.local_coord.center_of_rotation, dtype=torch.float64)
        suffix = f".{phase}" if phase else ""
        self.U = grid.get_cellfield("U" + suffix) or CellField(grid, "U" + suffix, (3,))
        self.p = grid.get_cellfield("p" + suffix) or CellField(grid, "p" + suffix, ())
        if tuple(self.U.component_shape) != (3,) or tuple(self.p.component_shape) != ():
            raise ValueError("Unexpected component_shape for U or p")
        self.history: List[ForceCoeffs] = []
        self.side_forces: List[SideForces] = []

    def _scales(self) -> Tuple[float, float]:
        q = 0.5 * self.rho * self.magU_ref ** 2
        fs = q * self.A_ref
        ms = fs * self.L_ref
        if fs == 0.0 or ms == 0.0:
            raise ValueError("force coefficient reference scale is zero")
        return fs, ms

    def evaluate(self, grid=None, time: float = 0.0, turbulence=None) -> ForceCoeffs:
        """Integrate the surface forces and return the coefficients at
        ``time``; ``turbulence`` supplies nu_eff (an object with nu_eff(), a
        [C] tensor or a number; the case's laminar nu when None)."""
        grid = grid if grid is not None else self.grid
        if turbulence is None:
            nu = float(grid.sim_config.properties.transport.nu)
        elif hasattr(turbulence, "nu_eff"):
            nu = turbulence.nu_eff()
        else:
            nu = turbulence
        fs, ms = self._scales()
        dev = grid.device
        force = torch.zeros(3, dtype=torch.float64, device=dev)
        moment = torch.zeros(3, dtype=torch.float64, device=dev)
        self.side_forces = []
        cofr = self.CofR.to(dev)
        for patch in self.patches:
            f, m = integrate_patch(grid, p=self.p, U=self.U, nu_eff=nu, patch=patch, rho=self.rho,
                                   center_of_rotation=cofr, collect=self.side_forces)
            force, moment = force + f, moment + m
        co = ForceCoeffs.from_force_moment(time=time, force=force, moment=moment, coord=self.local_coord.to(dev),
                                           force_scale=fs, moment_scale=ms)
        self.history.append(co)
        return co

    def get_history(self) -> List[ForceCoeffs]:
        return self.history

    def write_csv(self, path) -> None:
        with pathlib.Path(path).open("w", newline="") as f:
            w = csv.writer(f)
            w.writerow(CSV_HEADER)
            for c in self.history:
                w.writerow([c.time] + [float(getattr(c, k)) for k in CSV_HEADER[1:]])
