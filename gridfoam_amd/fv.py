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


# schemes/div.py:449-456 -- the scheme codes of gf_assembly.div_scheme
DIV_SCHEME_CODES = {
    DivScheme.UPWIND: 0, DivScheme.LINEAR: 1, DivScheme.VANLEER: 2, DivScheme.MINMOD: 3, DivScheme.SUPERBEE: 4,
    DivScheme.MONOTONIZED_CENTRAL: 5, DivScheme.LIMITEDLINEAR: 6, DivScheme.KOREN: 7,
}


def div_scheme(grid, phi_name: str, field_name: str) -> DivScheme:
    """schemes/selection.py: the scheme for div(phi, field), default upwind."""
    s = _lookup(grid.sim_config.fvSchemes.divSchemes if grid.sim_config else None,
                f"div({phi_name}, {field_name})", DivScheme.UPWIND)
    return DivScheme(s)


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
    """schemes/selection.py: the cell gradient of grad(field), default least squares."""
    s = _lookup(grid.sim_config.fvSchemes.gradSchemes if grid.sim_config else None,
                f"grad({field_name})", GradScheme.LEASTSQUARE)
    return GradScheme(s)


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


def _grad_pullback(field: CellField, st, g_grad: Optional[torch.Tensor], hang_only: bool, g_rv: torch.Tensor,
                   g_rg: torch.Tensor) -> torch.Tensor:
    """Carry a cotangent of the least-squares gradient (or None) and of the
    boundary reference values back to the field values: gf_lsq_grad_vjp,
    then gf_bc_state_vjp (which also moves slip faces' ref_v cotangent onto
    the cell values).  g_rv / g_rg are added to in place; returns d/d psi."""
    g = field.grid
    C, B, nc = g.num_cells, g.num_bnd_faces, field.nc
    f64 = dict(dtype=torch.float64, device=g.device)
    g_psi = torch.zeros((C, nc), **f64)
    g_psi_b = torch.zeros((max(B, 1), nc), **f64)
    if g_grad is not None:
        work = torch.zeros((C, nc, 3), **f64)
        nat.call("gf_lsq_grad_vjp", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(g_grad.contiguous()),
                 nat.ptr(st.covered), ctypes.c_int32(int(hang_only)), nat.ptr(work), nat.ptr(g_psi),
                 nat.ptr(g_psi_b), nat.stream())
    nat.call("gf_bc_state_vjp", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(st.frac), nat.ptr(st.covered),
             nat.ptr(st.slip), nat.ptr(g_psi_b), nat.ptr(g_psi), nat.ptr(g_rv), nat.ptr(g_rg), nat.stream())
    return g_psi.reshape(C, *field.component_shape)


class _LsqGradFn(torch.autograd.Function):
    """The least-squares gradient as an autograd node (boundary state
    included): the backward is gf_lsq_grad_vjp + gf_bc_state_vjp."""

    @staticmethod
    def forward(ctx, field, st, hang_only, psi, ref_v, ref_g):  # noqa: D401
        ctx.field, ctx.st, ctx.hang_only = field, st, hang_only
        g = field.grid
        out = torch.zeros((g.num_cells, *field.component_shape, 3), dtype=torch.float64, device=g.device)
        nat.call("gf_lsq_grad", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(field.data), nat.ptr(st.psi_b),
                 nat.ptr(st.covered), ctypes.c_int32(int(hang_only)), nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def backward(ctx, g_out):
        st, field = ctx.st, ctx.field
        g_rv, g_rg = torch.zeros_like(st.ref_v), torch.zeros_like(st.ref_g)
        g_psi = _grad_pullback(field, st, g_out, ctx.hang_only, g_rv, g_rg)
        need = ctx.needs_input_grad
        return (None, None, None, g_psi if need[3] else None, g_rv if need[4] else None,
                g_rg if need[5] else None)


class _BndValueFn(torch.autograd.Function):
    """A field's boundary face values (the value-fraction state,
    boundaries/base.py:71-106) as an autograd node: the backward is
    gf_bc_state_vjp (slip faces carry their cotangent on to the cell
    values)."""

    @staticmethod
    def forward(ctx, field, st, psi, ref_v, ref_g):  # noqa: D401
        ctx.field, ctx.st = field, st
        return st.psi_b.clone()

    @staticmethod
    def backward(ctx, g_out):
        st, field = ctx.st, ctx.field
        g = field.grid
        g_rv, g_rg = torch.zeros_like(st.ref_v), torch.zeros_like(st.ref_g)
        g_psi = torch.zeros((g.num_cells, field.nc), dtype=torch.float64, device=g.device)
        nat.call("gf_bc_state_vjp", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(st.frac), nat.ptr(st.covered),
                 nat.ptr(st.slip), nat.ptr(g_out.contiguous()), nat.ptr(g_psi), nat.ptr(g_rv), nat.ptr(g_rg),
                 nat.stream())
        need = ctx.needs_input_grad
        return (None, None, g_psi.reshape(g.num_cells, *field.component_shape) if need[2] else None,
                g_rv if need[3] else None, g_rg if need[4] else None)


def boundary_values(field: CellField) -> torch.Tensor:
    """The field's boundary face values [B, nc] (psi_b), carrying the graph
    of the field values and condition values when they require grad."""
    st = field.boundary_state()
    rv = st.ref_v_ad if st.ref_v_ad is not None else st.ref_v
    rg = st.ref_g_ad if st.ref_g_ad is not None else st.ref_g
    if not (torch.is_grad_enabled() and (field.data.requires_grad or rv.requires_grad or rg.requires_grad)):
        return st.psi_b
    if field.has_dynamic_state:
        st = st.frozen()
    return _BndValueFn.apply(field, st, field.data, rv, rg)


class _InterpFn(torch.autograd.Function):
    """Linear face values of the internal faces (with the hanging-face
    gradient corrections) as an autograd node: the backward is
    gf_interpolate_vjp; the gradient's own inputs follow through its node."""

    @staticmethod
    def forward(ctx, field, psi, grad):  # noqa: D401
        ctx.field, ctx.has_grad = field, grad is not None
        g = field.grid
        out = torch.zeros((max(g.num_internal_faces, 1), *field.component_shape), dtype=torch.float64,
                          device=g.device)
        nat.call("gf_interpolate", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(psi.contiguous()),
                 nat.ptr(None if grad is None else grad.contiguous()), nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def backward(ctx, g_out):
        field = ctx.field
        g = field.grid
        f64 = dict(dtype=torch.float64, device=g.device)
        g_psi = torch.zeros((g.num_cells, *field.component_shape), **f64)
        g_grad = torch.zeros((g.num_cells, *field.component_shape, 3), **f64) if ctx.has_grad else None
        nat.call("gf_interpolate_vjp", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(g_out.contiguous()),
                 nat.ptr(g_psi), nat.ptr(g_grad), nat.stream())
        need = ctx.needs_input_grad
        return None, g_psi if need[1] else None, g_grad if need[2] else None


class _DivFn(torch.autograd.Function):
    """fvc.div of a face flux as an autograd node: the backward is
    gf_fvc_div_vjp."""

    @staticmethod
    def forward(ctx, grid, nc, cshape, phi_s, phi_b):  # noqa: D401
        ctx.grid, ctx.nc = grid, nc
        out = torch.empty((grid.num_cells, *cshape), dtype=torch.float64, device=grid.device)
        nat.call("gf_fvc_div", grid.mesh_ref, ctypes.c_int32(nc), nat.ptr(phi_s.contiguous()),
                 nat.ptr(phi_b.contiguous()), nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def backward(ctx, g_out):
        return (None, None, None) + _div_pullback(ctx.grid, ctx.nc, g_out, ctx.needs_input_grad[3:])


class _GaussGradFn(torch.autograd.Function):
    """The Green-Gauss cell gradient of given face values as an autograd
    node: the backward is gf_gauss_grad_vjp."""

    @staticmethod
    def forward(ctx, grid, nc, cshape, psi_s, psi_b):  # noqa: D401
        ctx.grid, ctx.nc = grid, nc
        out = torch.empty((grid.num_cells, *cshape, 3), dtype=torch.float64, device=grid.device)
        nat.call("gf_gauss_grad", grid.mesh_ref, ctypes.c_int32(nc), nat.ptr(psi_s.contiguous()),
                 nat.ptr(psi_b.contiguous()), nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def backward(ctx, g_out):
        g, need = ctx.grid, ctx.needs_input_grad
        cs = tuple(g_out.shape[1:-1])
        f64 = dict(dtype=torch.float64, device=g.device)
        g_s = torch.zeros((max(g.num_internal_faces, 1), *cs), **f64)
        g_b = torch.zeros((max(g.num_bnd_faces, 1), *cs), **f64)
        nat.call("gf_gauss_grad_vjp", g.mesh_ref, ctypes.c_int32(ctx.nc), nat.ptr(g_out.contiguous()),
                 nat.ptr(g_s), nat.ptr(g_b), nat.stream())
        return None, None, None, (g_s if need[3] else None), (g_b if need[4] else None)


class _ReconstructFn(torch.autograd.Function):
    """fvc.reconstruct of a face flux as an autograd node: the backward is
    gf_reconstruct_vjp."""

    @staticmethod
    def forward(ctx, grid, phi_s, phi_b):  # noqa: D401
        ctx.grid = grid
        out = torch.empty((grid.num_cells, 3), dtype=torch.float64, device=grid.device)
        nat.call("gf_reconstruct", grid.mesh_ref, nat.ptr(phi_s.contiguous()), nat.ptr(phi_b.contiguous()),
                 nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def backward(ctx, g_out):
        g, need = ctx.grid, ctx.needs_input_grad
        f64 = dict(dtype=torch.float64, device=g.device)
        work = torch.empty((g.num_cells, 3), **f64)
        g_s = torch.zeros(max(g.num_internal_faces, 1), **f64)
        g_b = torch.zeros(max(g.num_bnd_faces, 1), **f64)
        nat.call("gf_reconstruct_vjp", g.mesh_ref, nat.ptr(g_out.contiguous()), nat.ptr(work), nat.ptr(g_s),
                 nat.ptr(g_b), nat.stream())
        return None, (g_s if need[1] else None), (g_b if need[2] else None)


def _div_pullback(g, nc: int, g_out: torch.Tensor, need) -> Tuple[Optional[torch.Tensor], Optional[torch.Tensor]]:
    cs = tuple(g_out.shape[1:])
    f64 = dict(dtype=torch.float64, device=g.device)
    g_s = torch.zeros((max(g.num_internal_faces, 1), *cs), **f64)
    g_b = torch.zeros((max(g.num_bnd_faces, 1), *cs), **f64)
    nat.call("gf_fvc_div_vjp", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(g_out.contiguous()), nat.ptr(g_s),
             nat.ptr(g_b), nat.stream())
    return (g_s if need[0] else None), (g_b if need[1] else None)


class fvc:
    @staticmethod
    def grad(field: CellField, *, hang_only: bool = False, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """The configured cell gradient [C, *cs, 3] (schemes/grad.py:82-101):
        least squares (least_squares.py:29-101) or Green-Gauss from the
        linear face values (grad.py:21-44).  hang_only: the least-squares
        gradient valid on the cells next to hanging faces, whatever the
        configuration (the hanging-face corrections always use it)."""
        g = field.grid
        scheme = grad_scheme(g, field.name)
        shape = (g.num_cells, *field.component_shape, 3)
        if out is None:
            out = torch.zeros(shape, dtype=torch.float64, device=g.device)
        st = field.boundary_state()
        rv = st.ref_v_ad if st.ref_v_ad is not None else st.ref_v
        rg = st.ref_g_ad if st.ref_g_ad is not None else st.ref_g
        track = torch.is_grad_enabled() and (field.data.requires_grad or rv.requires_grad or rg.requires_grad)
        if track and field.has_dynamic_state:
            st = st.frozen()
        if not hang_only and scheme == GradScheme.LINEAR:
            fi = fvc.interpolate(field)
            if track:  # through the linear face values (gf_gauss_grad_vjp, then the interpolation's node)
                return _GaussGradFn.apply(g, field.nc, field.component_shape, fi.single_buffer, fi.bnd_buffer)
            nat.call("gf_gauss_grad", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(fi.single_buffer),
                     nat.ptr(fi.bnd_buffer), nat.ptr(out), nat.stream())
            return out
        if track:  # a new tensor carrying the graph (``out`` is not written)
            out = _LsqGradFn.apply(field, st, hang_only, field.data, rv, rg)
        else:
            nat.call("gf_lsq_grad", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(field.data), nat.ptr(st.psi_b),
                     nat.ptr(st.covered), ctypes.c_int32(int(hang_only)), nat.ptr(out), nat.stream())
        if g.shape is not None and not hang_only and torch.is_grad_enabled():
            # |d| of the body's sides depends on its outline (a new tensor; ``out`` keeps the values)
            out = g.shape.add_lsq_terms(field, st, out)
        return out

    @staticmethod
    def interpolate(field: CellField, name: Optional[str] = None) -> FaceField:
        """fv/schemes/interpolate.py:20-63.  Differentiable with respect to
        the field values and its boundary values when either requires
        grad."""
        g = field.grid
        out = FaceField(g, name or f"{field.name}_f", field.component_shape)
        st = field.boundary_state()
        rv = st.ref_v_ad if st.ref_v_ad is not None else st.ref_v
        rg = st.ref_g_ad if st.ref_g_ad is not None else st.ref_g
        track = torch.is_grad_enabled() and (field.data.requires_grad or rv.requires_grad or rg.requires_grad)
        grad = fvc.grad(field, hang_only=True) if g.num_hanging_faces else None
        if track and field.has_dynamic_state:
            st = st.frozen()
        if track:  # face values carrying the graph (reverse mode: gf_interpolate_vjp, gf_bc_state_vjp)
            F, B = g.num_internal_faces, g.num_bnd_faces
            out.single_data = _InterpFn.apply(field, field.data, grad)[:F]
            out.bnd_data = _BndValueFn.apply(field, st, field.data, rv, rg).view(out.bnd_buffer.shape)[:B]
            return out
        nat.call("gf_interpolate", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(field.data), nat.ptr(grad),
                 nat.ptr(out.single_buffer), nat.stream())
        out.bnd_buffer.copy_(st.psi_b.view(out.bnd_buffer.shape))
        return out

    @staticmethod
    def div(phi: FaceField) -> torch.Tensor:
        """fv/fvc/div.py:15-64 (face sum / V); differentiable with respect to
        the face values when they carry a graph."""
        g = phi.grid
        nc = 3 if phi.component_shape == (3,) else 1
        if torch.is_grad_enabled() and (phi.single_buffer.requires_grad or phi.bnd_buffer.requires_grad):
            return _DivFn.apply(g, nc, phi.component_shape, phi.single_buffer, phi.bnd_buffer)
        out = torch.empty((g.num_cells, *phi.component_shape), dtype=torch.float64, device=g.device)
        nat.call("gf_fvc_div", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer),
                 nat.ptr(out), nat.stream())
        return out

    @staticmethod
    def reconstruct(phi: FaceField, U: CellField) -> torch.Tensor:
        """fvc/reconstruct.py:27-107: the cell velocity of a face flux (per
        axis, the area-weighted average of the face-normal velocities); U
        receives the data and it is returned."""
        if phi.component_shape != ():
            raise ValueError("phi must be a scalar face flux field.")
        if U.component_shape != (3,):
            raise ValueError("U must be a 3-component velocity field.")
        if phi.grid is not U.grid:
            raise ValueError("phi and U must share the same grid.")
        g = phi.grid
        if torch.is_grad_enabled() and (phi.single_buffer.requires_grad or phi.bnd_buffer.requires_grad):
            out = _ReconstructFn.apply(g, phi.single_buffer, phi.bnd_buffer)
            U.data = out
            return out
        out = torch.empty((g.num_cells, 3), dtype=torch.float64, device=g.device)
        nat.call("gf_reconstruct", g.mesh_ref, nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer), nat.ptr(out),
                 nat.stream())
        U.data = out
        return out

    @staticmethod
    def sn_grad_boundary(field: CellField) -> torch.Tensor:
        """fv/fvc/sn_grad.py:46-52: the boundary blocks of snGrad, per
        boundary face (domain, immersed upper, immersed lower): the mixed
        normal gradient f/|d| (ref_v - psi_P) + (1 - f) ref_g of
        boundary_ops.py:425-437 on faces with a condition, with the value
        term zero on snapped Dirichlet faces (a placeholder there, not an
        imposed gradient) and zero on faces without a condition."""
        g = field.grid
        B, nc = g.num_bnd_faces, field.nc
        st = field.boundary_state()
        f = st.frac[:B]
        active = (f != 0) & ~(g.bnd_snap[:B].bool() & (f == 1))
        dist = torch.where(active, g.bnd_distances()[:B], torch.ones_like(f))
        coef = torch.where(active, f / dist, torch.zeros_like(f))
        psi = field.data.reshape(g.num_cells, nc)[g.bnd_owner[:B]]
        out = coef[:, None] * (st.ref_v[:B] - psi) + (1.0 - f)[:, None] * st.ref_g[:B]
        out = torch.where(st.covered[:B].bool()[:, None], out, torch.zeros_like(out))
        return out.reshape((B, *field.component_shape))

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


class _Assembly:
    """The launch description of one gf_assemble call, kept for its VJP."""

    def __init__(self, field: CellField, st, ddt, div, lap, phi, gamma, scheme, corrected, abc, grad, harmonic):
        g = field.grid
        self.field, self.st, self.grid = field, st, g
        a = _assembly(field, st)
        a.s_ddt, a.s_div, a.s_lap = float(ddt), float(div), float(lap)
        self.keep = [field.data]
        if ddt:
            aa, bb, cc = abc if abc is not None else ddt_coefficients(field)
            a.ddt_a, a.ddt_b, a.ddt_c, a.dt = aa, bb, cc, g.dt
            a.psi_old = nat.ptr(field.old_data)
            a.psi_older = nat.ptr(field.older_data) if cc != 0.0 else None
            self.keep += [field.old_data, field.older_data]
        self.older_used = bool(ddt) and a.ddt_c != 0.0
        need_grad = False
        self.tvd = False
        if div:
            if phi is None or phi.component_shape != ():
                raise ValueError("convection requires scalar face flux")  # fvm/div.py:41-42
            sch = DivScheme(scheme) if scheme is not None else div_scheme(g, phi.name, field.name)
            a.div_scheme = DIV_SCHEME_CODES[sch]
            a.phi_s, a.phi_b = nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer)
            self.keep += [phi.single_buffer, phi.bnd_buffer]
            need_grad |= a.div_scheme == 1
            self.tvd = a.div_scheme >= 2
        self.gamma_t = None
        self.has_ffc = False
        if lap:
            corr_cfg, harm_cfg = laplacian_scheme(g, field.name)
            corr = corr_cfg if corrected is None else corrected
            harm = harm_cfg if harmonic is None else harmonic
            a.lap_corrected = int(corr)
            a.lap_harmonic = int(harm)
            if isinstance(gamma, torch.Tensor) and gamma.dim() > 0:
                if tuple(gamma.shape) != (g.num_cells,):
                    raise ValueError("diffusivity must have shape [C]")
                gamma_c = gamma.detach().contiguous()  # kept alive with the description
                self.gamma_t = gamma_c
                a.gamma_cell = nat.ptr(gamma_c)
            else:
                a.gamma = float(gamma.detach()) if isinstance(gamma, torch.Tensor) else float(gamma)
            if harm:  # laplacian.py:45-54
                ok = (bool(torch.all(torch.isfinite(self.gamma_t) & (self.gamma_t >= 0))) if self.gamma_t is not None
                      else math.isfinite(a.gamma) and a.gamma >= 0)
                if not ok:
                    raise ValueError("Harmonic diffusivity must be finite and nonnegative")
            need_grad |= corr
            self.has_ffc = bool(corr and g.num_hanging_faces)
        self.hang_terms = bool((need_grad or self.tvd) and g.num_hanging_faces)
        if self.tvd:  # r needs the gradient of every upwind cell (schemes/div.py:270-275)
            with torch.no_grad():  # the assembly's backward chains through it itself
                grad = fvc.grad(field)
            a.grad = nat.ptr(grad)
            self.keep.append(grad)
        elif self.hang_terms:
            if grad is None:
                with torch.no_grad():
                    grad = fvc.grad(field, hang_only=True)
            a.grad = nat.ptr(grad)
            self.keep.append(grad)
        # the field values enter the arrays (hanging corrections, TVD sources)
        self.psi_terms = self.hang_terms or self.tvd
        self.a = a

    def run(self):
        g, field = self.grid, self.field
        C, F = g.num_cells, g.num_internal_faces
        dev, cs = g.device, field.component_shape
        diag = torch.empty(C, dtype=torch.float64, device=dev)
        upper = torch.empty(max(F, 1), dtype=torch.float64, device=dev)
        lower = torch.empty(max(F, 1), dtype=torch.float64, device=dev)
        source = torch.empty((C, *cs), dtype=torch.float64, device=dev)
        if F == 0:
            upper.zero_()
            lower.zero_()
        ffc = torch.zeros((max(F, 1), *cs), dtype=torch.float64, device=dev) if self.has_ffc else None
        nat.call("gf_assemble", g.mesh_ref, ctypes.byref(self.a), nat.ptr(diag), nat.ptr(upper), nat.ptr(lower),
                 nat.ptr(source), nat.ptr(ffc), nat.stream())
        return diag, upper, lower, source, ffc


class _AssembleFn(torch.autograd.Function):
    """gf_assemble as an autograd node; the backward is gf_assemble_vjp."""

    @staticmethod
    def forward(ctx, spec, gamma, phi_s, phi_b, psi_old, psi_older, ref_v, ref_g, psi):  # noqa: D401
        ctx.spec = spec
        ctx.gamma_shape = None if gamma is None else tuple(gamma.shape)
        ctx.like = (phi_s, phi_b, psi_old, psi_older, ref_v, ref_g)
        diag, upper, lower, source, ffc = spec.run()
        if ffc is None:
            ffc = torch.zeros(0, dtype=torch.float64, device=diag.device)
            ctx.mark_non_differentiable(ffc)
        return diag, upper, lower, source, ffc

    @staticmethod
    def backward(ctx, g_diag, g_upper, g_lower, g_source, g_ffc):
        spec: _Assembly = ctx.spec
        g = spec.grid
        need = ctx.needs_input_grad
        F, C = g.num_internal_faces, g.num_cells
        dev = g.device
        f64 = dict(dtype=torch.float64, device=dev)

        def upstream(t, shape):
            return torch.zeros(shape, **f64) if t is None else t.contiguous()

        gd = upstream(g_diag, (C,))
        gu = upstream(g_upper, (max(F, 1),))
        gl = upstream(g_lower, (max(F, 1),))
        gs = upstream(g_source, (C, *spec.field.component_shape))
        gf = g_ffc.contiguous() if (g_ffc is not None and spec.has_ffc) else None
        work = torch.zeros(max(F, 1), **f64) if need[1] else None
        g_gamma = torch.zeros(C, **f64) if need[1] else None
        outs = [torch.zeros_like(t) if (need[k + 2] and t is not None) else None for k, t in enumerate(ctx.like)]
        g_phi_s, g_phi_b, g_old, g_older, g_rv, g_rg = outs
        # the field values and the reference values also reach the arrays through
        # the hang-cell gradient (hanging corrections) and the slip condition
        need_state = need[6] or need[7] or need[8]
        st = spec.st
        g_grad = None
        if need_state:
            g_rv = torch.zeros_like(st.ref_v) if g_rv is None else g_rv
            g_rg = torch.zeros_like(st.ref_g) if g_rg is None else g_rg
            if spec.hang_terms:
                g_grad = torch.zeros((C, spec.field.nc, 3), **f64)
        nat.call("gf_assemble_vjp", g.mesh_ref, ctypes.byref(spec.a), nat.ptr(gd), nat.ptr(gu), nat.ptr(gl),
                 nat.ptr(gs), nat.ptr(gf), nat.ptr(work), nat.ptr(g_gamma), nat.ptr(g_phi_s), nat.ptr(g_phi_b),
                 nat.ptr(g_old), nat.ptr(g_older), nat.ptr(g_rv), nat.ptr(g_rg), nat.ptr(g_grad), nat.stream())
        g_psi = _grad_pullback(spec.field, st, g_grad, True, g_rv, g_rg) if need_state else None
        if g_gamma is not None and ctx.gamma_shape != (C,):
            g_gamma = g_gamma.sum().reshape(ctx.gamma_shape)
        return (None, g_gamma, g_phi_s, g_phi_b, g_old, g_older, g_rv if need[6] else None,
                g_rg if need[7] else None, g_psi if need[8] else None)


class fvm:
    @staticmethod
    def assemble(field: CellField, *, ddt: float = 0.0, div: float = 0.0, lap: float = 0.0,
                 phi: Optional[FaceField] = None, gamma: Union[float, torch.Tensor, None] = None,
                 scheme: Optional[DivScheme] = None, corrected: Optional[bool] = None,
                 abc: Optional[Tuple[float, float, float]] = None,
                 grad: Optional[torch.Tensor] = None, harmonic: Optional[bool] = None) -> FvMatrix:
        """Fused ``ddt*fvm.ddt + div*fvm.div(phi) + lap*fvm.laplacian(gamma)``.

        Differentiable (when autograd is on and an input requires grad) with
        respect to the diffusivity (a [C] tensor or a 0-d tensor), the face
        flux, the old and older field levels and the Dirichlet values and
        Neumann gradients of the field's conditions: the backward is the
        native VJP ``gf_assemble_vjp``.  The current field values (and the
        boundary values a second time) enter through the hanging-face
        corrections, via the hang-cell least-squares gradient, and through
        slip conditions; both are differentiated (gf_lsq_grad_vjp,
        gf_bc_state_vjp).  Through TVD sources, whose ratio is nonlinear in
        them, field and boundary values that require grad raise, as does an
        explicit ``grad`` argument that requires grad."""
        g = field.grid
        st = field.boundary_state()
        # the autograd-connected reference values of this call (later
        # boundary_state() calls, e.g. for the cell gradient, rebuild them)
        rv = st.ref_v_ad if st.ref_v_ad is not None else st.ref_v
        rg = st.ref_g_ad if st.ref_g_ad is not None else st.ref_g
        if torch.is_grad_enabled() and field.has_dynamic_state:
            st = st.frozen()  # the VJP reads these arrays through the launch description's pointers
        spec = _Assembly(field, st, ddt, div, lap, phi, gamma, scheme, corrected, abc, grad, harmonic)
        mat = FvMatrix.__new__(FvMatrix)
        mat.field, mat.grid, mat.face_flux_correction = field, g, None
        track = torch.is_grad_enabled()
        gamma_in = gamma if (lap and isinstance(gamma, torch.Tensor)) else None
        phi_s = phi.single_buffer if (div and phi is not None) else None
        phi_b = phi.bnd_buffer if (div and phi is not None) else None
        old = field.old_data if ddt else None
        older = field.older_data if spec.older_used else None
        ins = (gamma_in, phi_s, phi_b, old, older, rv, rg, field.data)
        if track and any(t is not None and t.requires_grad for t in ins):
            if spec.tvd and (field.data.requires_grad or rv.requires_grad or rg.requires_grad):
                # the TVD ratio depends on the field values and its gradient nonlinearly
                raise NotImplementedError("gradients with respect to the field values or boundary values "
                                          "through TVD sources are not supported")
            if grad is not None and grad.requires_grad:
                raise NotImplementedError("gradients through an explicit cell gradient are not supported")
            diag, upper, lower, source, ffc = _AssembleFn.apply(spec, *ins)
        else:
            diag, upper, lower, source, ffc = spec.run()
        mat._diag, mat._upper, mat._lower, mat._source = diag, upper, lower, source
        if spec.has_ffc:
            mat.face_flux_correction = ffc[: g.num_internal_faces]
        if lap and g.shape is not None and track:  # |d| of the body's sides depends on its outline
            g.shape.add_laplacian_terms(mat, field, st, rv, rg, lap, gamma)
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
