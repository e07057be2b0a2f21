"""fvm / fvc operators on the HIP library (src/gridfoam/fv/).

``fvm.laplacian/div/ddt`` each assemble an FvMatrix with ``gf_assemble``;
``fvm.momentum`` assembles ``s_ddt*ddt + s_div*div - laplacian`` in one
pass (the sum the reference builds term by term, fvmatrix.py:181-201).
``fvc.*`` are the explicit operators.  Scheme lookup follows
fv/schemes/selection.py:36-113 (specific key, then ``default``, then the
built-in default).
"""

from __future__ import annotations

import ctypes
import math
from typing import Optional, Tuple, Union

import torch

from . import _native as nat
from .config import DdtScheme, DivScheme, GradScheme, LaplacianScheme, SnGradScheme
from .fields import CellField, FaceField
from .fvmatrix import FvMatrix

_LAP_ALIASES = {
    LaplacianScheme.LINEAR: LaplacianScheme.CORRECTED,
    LaplacianScheme.GAUSS_LINEAR_CORRECTED: LaplacianScheme.CORRECTED,
    LaplacianScheme.GAUSS_LINEAR_UNCORRECTED: LaplacianScheme.UNCORRECTED,
}
_LAP_HARMONIC = {  # schemes/laplacian.py: harmonic coefficient interpolation -> correction
    LaplacianScheme.GAUSS_HARMONIC_CORRECTED: LaplacianScheme.CORRECTED,
    LaplacianScheme.GAUSS_HARMONIC_UNCORRECTED: LaplacianScheme.UNCORRECTED,
}


def _lookup(table, key, default):
    if table is None:
        return default
    v = table.get(key)
    if v is None:
        v = table.get("default")
    return default if v is None else v


def div_scheme(grid, phi_name: str, field_name: str) -> DivScheme:
    s = _lookup(grid.sim_config.fvSchemes.divSchemes if grid.sim_config else None,
                f"div({phi_name}, {field_name})", DivScheme.UPWIND)
    if s not in (DivScheme.UPWIND, DivScheme.LINEAR):
        raise ValueError(f"div scheme {s} is not supported by this build")
    return s


def sn_grad_corrected(grid, field_name: str) -> bool:
    """fv/schemes/selection.py:94-102 (default: corrected)."""
    s = _lookup(grid.sim_config.fvSchemes.snGradSchemes if grid.sim_config else None,
                f"snGrad({field_name})", SnGradScheme.CORRECTED)
    return SnGradScheme(s) == SnGradScheme.CORRECTED


def laplacian_scheme(grid, field_name: str) -> Tuple[bool, bool]:
    """(corrected, harmonic) of the laplacian scheme configured for the field."""
    s = _lookup(grid.sim_config.fvSchemes.laplacianSchemes if grid.sim_config else None,
                f"laplacian({field_name})", LaplacianScheme.CORRECTED)
    harmonic = s in _LAP_HARMONIC
    s = _LAP_HARMONIC.get(s, _LAP_ALIASES.get(s, s))
    if s not in (LaplacianScheme.CORRECTED, LaplacianScheme.UNCORRECTED):
        raise ValueError(f"laplacian scheme {s} is not supported by this build")
    return s == LaplacianScheme.CORRECTED, harmonic


def laplacian_corrected(grid, field_name: str) -> bool:
    return laplacian_scheme(grid, field_name)[0]


def grad_scheme(grid, field_name: str) -> GradScheme:
    s = _lookup(grid.sim_config.fvSchemes.gradSchemes if grid.sim_config else None,
                f"grad({field_name})", GradScheme.LEASTSQUARE)
    if s != GradScheme.LEASTSQUARE:
        raise ValueError(f"grad scheme {s} is not supported by this build")
    return s


def ddt_scheme(grid, field_name: str) -> DdtScheme:
    return _lookup(grid.sim_config.fvSchemes.ddtSchemes if grid.sim_config else None,
                   f"ddt({field_name})", DdtScheme.EULER)


def ddt_coefficients(field: CellField) -> Tuple[float, float, float]:
    """fv/schemes/ddt.py:13-58 -- (a, b, c) in (a x - b x_old + c x_older)/dt."""
    dt = field.grid.dt
    if not math.isfinite(dt) or dt <= 0:
        raise ValueError("Time step must be finite and positive")
    if ddt_scheme(field.grid, field.name) == DdtScheme.EULER or field.older_data is None:
        return 1.0, 1.0, 0.0
    dt0 = field.previous_dt
    if dt0 is None or not math.isfinite(dt0) or dt0 <= 0:
        raise ValueError("backward ddt requires a finite positive previous time step")
    a = 1.0 + dt / (dt + dt0)
    c = (dt / dt0) * dt / (dt + dt0)
    return a, a + c, c


class fvc:
    @staticmethod
    def grad(field: CellField, *, hang_only: bool = False, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """least_squares.py:29-101 -> [C, *cs, 3] (hang_only: valid on hang cells)."""
        g = field.grid
        grad_scheme(g, field.name)
        st = field.boundary_state()
        shape = (g.num_cells, *field.component_shape, 3)
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=g.device)
        nat.call("gf_lsq_grad", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(field.data), nat.ptr(st.psi_b),
                 nat.ptr(st.covered), ctypes.c_int32(int(hang_only)), nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def interpolate(field: CellField, name: Optional[str] = None) -> FaceField:
        """fv/schemes/interpolate.py:20-63."""
        g = field.grid
        out = FaceField(g, name or f"{field.name}_f", field.component_shape)
        st = field.boundary_state()
        grad = fvc.grad(field, hang_only=True) if g.num_hanging_faces else None
        nat.call("gf_interpolate", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(field.data), nat.ptr(grad),
                 nat.ptr(out.single_buffer), nat.stream())
        out.bnd_buffer.copy_(st.psi_b.view(out.bnd_buffer.shape))
        return out

    @staticmethod
    def div(phi: FaceField) -> torch.Tensor:
        """fv/fvc/div.py:15-64 (face sum / V)."""
        g = phi.grid
        nc = 3 if phi.component_shape == (3,) else 1
        out = torch.empty((g.num_cells, *phi.component_shape), dtype=torch.float64, device=g.device)
        nat.call("gf_fvc_div", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer),
                 nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def sn_grad(field: CellField) -> torch.Tensor:
        """fv/schemes/sn_grad.py:109-181 (corrected)."""
        g = field.grid
        F = g.num_internal_faces
        diff = field.data[g.neighbour] - field.data[g.owner]
        dc = g.delta_coeffs[:F] if field.nc == 1 else g.delta_coeffs[:F, None]
        base = dc * diff
        if g.num_hanging_faces:
            grad = fvc.grad(field, hang_only=True)
            corr = torch.empty((g.num_hanging_faces, *field.component_shape), dtype=torch.float64, device=g.device)
            nat.call("gf_sn_grad_hang", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(grad), nat.ptr(corr),
                     nat.stream())
            base = base.clone()
            base.index_add_(0, g.hang_idx.long(), corr)
        return base


def _assembly(field: CellField, st) -> nat.GfAssembly:
    a = nat.GfAssembly()
    a.nc = field.nc
    a.psi = nat.ptr(field.data)
    a.frac, a.ref_v, a.ref_g = nat.ptr(st.frac), nat.ptr(st.ref_v), nat.ptr(st.ref_g)
    a.psi_b, a.covered = nat.ptr(st.psi_b), nat.ptr(st.covered)
    return a


class fvm:
    @staticmethod
    def assemble(field: CellField, *, ddt: float = 0.0, div: float = 0.0, lap: float = 0.0,
                 phi: Optional[FaceField] = None, gamma: Union[float, torch.Tensor, None] = None,
                 scheme: Optional[DivScheme] = None, corrected: Optional[bool] = None,
                 abc: Optional[Tuple[float, float, float]] = None,
                 grad: Optional[torch.Tensor] = None, harmonic: Optional[bool] = None) -> FvMatrix:
        """Fused ``ddt*fvm.ddt + div*fvm.div(phi) + lap*fvm.laplacian(gamma)``.

        The assembly is native and records no autograd graph: inputs that
        require gradients (diffusivity, flux, field values) raise instead of
        silently losing them.  Gradients reach the assembled LDU arrays and
        the source through the solve's implicit adjoint."""
        if torch.is_grad_enabled():
            ins = [gamma if isinstance(gamma, torch.Tensor) else None, grad, field.data, field.old_data,
                   field.older_data]
            if phi is not None:
                ins += [phi.single_buffer, phi.bnd_buffer]
            if any(t is not None and t.requires_grad for t in ins):
                raise NotImplementedError("gradients through the native FV assembly are not supported; "
                                          "differentiate with respect to the assembled matrix and source")
        g = field.grid
        mat = FvMatrix(field)
        st = field.boundary_state()
        a = _assembly(field, st)
        a.s_ddt, a.s_div, a.s_lap = float(ddt), float(div), float(lap)
        if ddt:
            aa, bb, cc = abc if abc is not None else ddt_coefficients(field)
            a.ddt_a, a.ddt_b, a.ddt_c, a.dt = aa, bb, cc, g.dt
            a.psi_old = nat.ptr(field.old_data)
            a.psi_older = nat.ptr(field.older_data) if cc != 0.0 else None
        need_grad = False
        if div:
            sch = scheme if scheme is not None else div_scheme(g, phi.name, field.name)
            a.div_scheme = 1 if sch == DivScheme.LINEAR else 0
            a.phi_s, a.phi_b = nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer)
            need_grad |= a.div_scheme == 1
        ffc = None
        if lap:
            corr_cfg, harm_cfg = laplacian_scheme(g, field.name)
            corr = corr_cfg if corrected is None else corrected
            harm = harm_cfg if harmonic is None else harmonic
            a.lap_corrected = int(corr)
            a.lap_harmonic = int(harm)
            if isinstance(gamma, torch.Tensor):
                if tuple(gamma.shape) != (g.num_cells,):
                    raise ValueError("diffusivity must have shape [C]")
                gamma_c = gamma.contiguous()  # keep alive until the launch below
                a.gamma_cell = nat.ptr(gamma_c)
            else:
                a.gamma = float(gamma)
            if harm:  # laplacian.py:45-54
                ok = (bool(torch.all(torch.isfinite(gamma_c) & (gamma_c >= 0))) if isinstance(gamma, torch.Tensor)
                      else math.isfinite(float(gamma)) and float(gamma) >= 0)
                if not ok:
                    raise ValueError("Harmonic diffusivity must be finite and nonnegative")
            need_grad |= corr
            if corr and g.num_hanging_faces:
                ffc = torch.zeros((max(g.num_internal_faces, 1), *field.component_shape), dtype=torch.float64,
                                  device=g.device)
        if need_grad and g.num_hanging_faces:
            if grad is None:
                grad = fvc.grad(field, hang_only=True)
            a.grad = nat.ptr(grad)
        nat.call("gf_assemble", g.mesh_ref, ctypes.byref(a), nat.ptr(mat._diag), nat.ptr(mat._upper),
                 nat.ptr(mat._lower), nat.ptr(mat._source), nat.ptr(ffc), nat.stream())
        if ffc is not None:
            mat.face_flux_correction = ffc[: g.num_internal_faces]
        return mat

    @staticmethod
    def laplacian(gamma, field: CellField) -> FvMatrix:
        """fv/fvm/laplacian.py:101-201."""
        return fvm.assemble(field, lap=1.0, gamma=gamma)

    @staticmethod
    def div(phi: FaceField, field: CellField) -> FvMatrix:
        """fv/fvm/div.py:15-73."""
        if phi.component_shape != ():
            raise ValueError("convection requires scalar face flux")
        return fvm.assemble(field, div=1.0, phi=phi)

    @staticmethod
    def ddt(field: CellField) -> FvMatrix:
        """fv/fvm/ddt.py:7-40."""
        return fvm.assemble(field, ddt=1.0)
