"""LDU matrix container on the device (src/gridfoam/core/fvmatrix.py).

Storage is the reference's: ``diag [C]``, ``upper/lower [F]`` (owner row:
upper, neighbour row: lower), ``source [C, *cs]`` and an optional explicit
face-flux correction ``[F_single, *cs]``.  Products, H, flux and the algebra
run on the HIP library.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import torch

from . import _native as nat
from .fields import CellField


def _require(t: torch.Tensor, shape, what: str) -> None:  # core/shapes.py:29-36
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what}: expected shape {tuple(shape)}, got {tuple(t.shape)}")


class FvMatrix:
    def __init__(self, field: CellField):
        g = field.grid
        self.field = field
        self.grid = g
        C, F = g.num_cells, g.num_internal_faces
        dev = g.device
        self._diag = torch.zeros(C, dtype=torch.float64, device=dev)
        self._upper = torch.zeros(max(F, 1), dtype=torch.float64, device=dev)
        self._lower = torch.zeros(max(F, 1), dtype=torch.float64, device=dev)
        self._source = torch.zeros((C, *field.component_shape), dtype=torch.float64, device=dev)
        self.face_flux_correction: Optional[torch.Tensor] = None

    # ---- storage ---------------------------------------------------------
    @property
    def num_components(self) -> int:
        return self.field.nc

    @property
    def diag(self) -> torch.Tensor:
        return self._diag

    @diag.setter
    def diag(self, v: torch.Tensor) -> None:
        _require(v, (self.grid.num_cells,), "matrix diag")
        self._diag = v.contiguous()

    @property
    def upper(self) -> torch.Tensor:
        return self._upper[: self.grid.num_internal_faces]

    @upper.setter
    def upper(self, v: torch.Tensor) -> None:
        _require(v, (self.grid.num_internal_faces,), "matrix upper")
        self._upper[: self.grid.num_internal_faces] = v

    @property
    def lower(self) -> torch.Tensor:
        return self._lower[: self.grid.num_internal_faces]

    @lower.setter
    def lower(self, v: torch.Tensor) -> None:
        _require(v, (self.grid.num_internal_faces,), "matrix lower")
        self._lower[: self.grid.num_internal_faces] = v

    @property
    def source(self) -> torch.Tensor:
        return self._source

    @source.setter
    def source(self, v: torch.Tensor) -> None:
        _require(v, (self.grid.num_cells, *self.field.component_shape), "matrix source")
        self._source = v.contiguous()

    # ---- algebra (fvmatrix.py:181-240) -------------------------------------
    def _combine(self, other: "FvMatrix", sign: float) -> "FvMatrix":
        res = FvMatrix(self.field)
        if torch.is_grad_enabled() and any(t.requires_grad for m in (self, other) for t in (
                m._diag, m._upper, m._lower, m._source)):
            # differentiable term sum (the values equal the native axpby's)
            res._diag = self._diag + sign * other._diag
            res._upper = self._upper + sign * other._upper
            res._lower = self._lower + sign * other._lower
            res._source = self._source + sign * other._source
            fa, fb = self.face_flux_correction, other.face_flux_correction
            if fa is not None or fb is not None:
                res.face_flux_correction = (fa if fa is not None else 0.0) + sign * (fb if fb is not None else 0.0)
            return res
        st = nat.stream()
        for a, b, out in ((self._diag, other._diag, res._diag), (self._upper, other._upper, res._upper),
                          (self._lower, other._lower, res._lower), (self._source, other._source, res._source)):
            nat.call("gf_axpby", ctypes.c_int64(a.numel()), ctypes.c_double(1.0), nat.ptr(a),
                     ctypes.c_double(sign), nat.ptr(b), nat.ptr(out), st)
        fa, fb = self.face_flux_correction, other.face_flux_correction
        if fa is None and fb is None:
            res.face_flux_correction = None
        else:
            za = fa if fa is not None else torch.zeros_like(fb)
            zb = fb if fb is not None else torch.zeros_like(fa)
            out = torch.empty_like(za)
            nat.call("gf_axpby", ctypes.c_int64(za.numel()), ctypes.c_double(1.0), nat.ptr(za),
                     ctypes.c_double(sign), nat.ptr(zb), nat.ptr(out), st)
            res.face_flux_correction = out
        return res

    def __add__(self, other: "FvMatrix") -> "FvMatrix":
        return self._combine(other, 1.0)

    def __sub__(self, other: "FvMatrix") -> "FvMatrix":
        return self._combine(other, -1.0)

    def __neg__(self) -> "FvMatrix":
        zero = FvMatrix(self.field)
        return zero._combine(self, -1.0)

    def with_source(self, source: torch.Tensor) -> "FvMatrix":
        res = FvMatrix(self.field)
        res.source = source.clone()
        res._diag = self._diag.clone()
        res._upper = self._upper.clone()
        res._lower = self._lower.clone()
        if self.face_flux_correction is not None:
            res.face_flux_correction = self.face_flux_correction.clone()
        return res

    def as_transpose(self) -> "FvMatrix":
        """fvmatrix.py:396-412 -- shares storage, swaps upper and lower."""
        view = FvMatrix.__new__(FvMatrix)
        view.field, view.grid = self.field, self.grid
        view._diag, view._upper, view._lower = self._diag, self._lower, self._upper
        view._source = self._source
        view.face_flux_correction = self.face_flux_correction
        return view

    # ---- operators ---------------------------------------------------------
    def multiply(self, x: torch.Tensor) -> torch.Tensor:
        """fvmatrix.py:292-323 -- deterministic gather SpMV; differentiable
        with respect to x and the LDU entries when any of them requires grad
        (then also H)."""
        if torch.is_grad_enabled() and any(t.requires_grad for t in (x, self._diag, self._upper, self._lower)):
            return _MultiplyFn.apply(self.grid, self._diag, self._upper, self._lower, x)
        nc = 1 if x.dim() == 1 else x.shape[1]
        xc = x.contiguous()
        y = torch.empty_like(xc)
        nat.call("gf_ldu_multiply", self.grid.mesh_ref, ctypes.c_int32(nc), nat.ptr(self._diag),
                 nat.ptr(self._upper), nat.ptr(self._lower), nat.ptr(xc), nat.ptr(y), nat.stream())
        return y

    def A(self) -> torch.Tensor:
        return self._diag / self.grid.cell_volumes

    def H(self, x: torch.Tensor) -> torch.Tensor:
        d = self._diag if x.dim() == 1 else self._diag[:, None]
        V = self.grid.cell_volumes if x.dim() == 1 else self.grid.cell_volumes[:, None]
        return (self._source - self.multiply(x) + d * x) / V

    def H1(self) -> torch.Tensor:
        """fvmatrix.py:250-264 -- deterministic per-cell gather."""
        h1 = torch.empty_like(self._diag)
        nat.call("gf_ldu_h1", self.grid.mesh_ref, nat.ptr(self._upper), nat.ptr(self._lower), nat.ptr(h1),
                 nat.stream())
        return h1

    def flux(self, psi: torch.Tensor) -> torch.Tensor:
        """fvmatrix.py:424-468."""
        nc = 1 if psi.dim() == 1 else psi.shape[1]
        F = self.grid.num_internal_faces
        out = torch.empty((max(F, 1),) + tuple(psi.shape[1:]), dtype=torch.float64, device=psi.device)
        ffc = self.face_flux_correction
        pc = psi.contiguous()
        nat.call("gf_ldu_flux", self.grid.mesh_ref, ctypes.c_int32(nc), ctypes.c_double(1.0), nat.ptr(self._upper),
                 nat.ptr(self._lower), nat.ptr(ffc), nat.ptr(pc), nat.ptr(out), nat.stream())
        return out[:F]

    def with_fixed_values(self, cells: torch.Tensor, values: torch.Tensor, *, unique: bool = False) -> "FvMatrix":
        """fvmatrix.py:325-394: symmetric elimination of prescribed cell values
        into a separate solve matrix (no face-flux correction); ``self`` is
        unchanged.  Gradients flow to ``values`` and to the input source, and,
        when the input LDU entries carry a graph, to them as well (the
        elimination is then written as differentiable tensor operations)."""
        _require(values, (cells.numel(), *self.field.component_shape), "fixed values")
        if cells.numel() == 0:
            return self
        if torch.is_grad_enabled() and any(t.requires_grad for t in (self._diag, self._upper, self._lower)):
            return self._with_fixed_values_graph(cells, values)
        g = self.grid
        nc = self.field.nc
        C = g.num_cells
        cells32 = cells.to(device=g.device, dtype=torch.int32).contiguous()
        if not unique and torch.unique(cells32).numel() != cells32.numel():  # unique=True: the caller vouches
            raise ValueError("fixed-value cells must be unique")
        vals = values.detach().to(torch.float64).reshape(cells.numel(), nc).contiguous()
        src = self._source.detach().contiguous()
        fixed = torch.empty(C, dtype=torch.uint8, device=g.device)
        presc = torch.empty(C * nc, dtype=torch.float64, device=g.device)
        dsum = torch.empty(1, dtype=torch.float64, device=g.device)
        out = FvMatrix(self.field)
        nat.call("gf_fixed_values", g.mesh_ref, ctypes.c_int32(nc), ctypes.c_int32(cells.numel()), nat.ptr(cells32),
                 nat.ptr(vals), nat.ptr(self._diag), nat.ptr(self._upper), nat.ptr(self._lower), nat.ptr(src),
                 nat.ptr(fixed), nat.ptr(presc), nat.ptr(dsum), nat.ptr(out._diag), nat.ptr(out._upper),
                 nat.ptr(out._lower), nat.ptr(out._source), nat.stream())
        if values.requires_grad or self._source.requires_grad:
            out._source = _FixedSource.apply(values, self._source, out._source, self, out, cells32, fixed)
        return out


    def _with_fixed_values_graph(self, cells: torch.Tensor, values: torch.Tensor) -> "FvMatrix":
        """The elimination of fvmatrix.py:365-394 as tensor operations, so
        that autograd carries gradients to the LDU entries too."""
        g = self.grid
        C, F = g.num_cells, g.num_internal_faces
        cells = cells.to(device=g.device, dtype=torch.long)
        fixed = torch.zeros(C, dtype=torch.bool, device=g.device)
        fixed[cells] = True
        presc = torch.zeros_like(self._source).index_copy(0, cells, values.to(torch.float64))
        o, n = g.owner, g.neighbour
        diag, upper, lower = self._diag, self._upper[:F], self._lower[:F]
        unit = torch.where(diag.detach().sum() < 0, -torch.ones_like(diag), torch.ones_like(diag))
        cdiag = torch.where(diag != 0, diag, unit)
        conn = fixed[o] | fixed[n]
        out = FvMatrix(self.field)
        out._diag = torch.where(fixed, cdiag, diag)
        pad = self._upper.shape[0] - F
        zero = torch.zeros(pad, dtype=torch.float64, device=g.device)
        out._upper = torch.cat([torch.where(conn, torch.zeros_like(upper), upper), zero])
        out._lower = torch.cat([torch.where(conn, torch.zeros_like(lower), lower), zero])

        def bc(c, like):
            return c if like.dim() == 1 else c[:, None]
        from .immersed import det_index_add  # noqa: PLC0415

        fo = torch.nonzero(fixed[n]).flatten()  # faces whose neighbour is prescribed
        fn = torch.nonzero(fixed[o]).flatten()
        src = det_index_add(self._source, o[fo], -bc(upper[fo], presc[n[fo]]) * presc[n[fo]])
        src = det_index_add(src, n[fn], -bc(lower[fn], presc[o[fn]]) * presc[o[fn]])
        out._source = src.index_copy(0, cells, bc(cdiag[cells], values) * values.to(torch.float64))
        return out


class _FixedSource(torch.autograd.Function):
    """Source of the eliminated system as a function of the prescribed values
    and the input source: fixed rows cd*value, free rows source - A_off presc."""

    @staticmethod
    def forward(ctx, values, source, out_source, base, out, cells32, fixed):  # noqa: D401
        ctx.base, ctx.out, ctx.cells = base, out, cells32
        ctx.save_for_backward(fixed)
        ctx.vshape = values.shape
        return out_source.clone()

    @staticmethod
    def backward(ctx, grad):
        (fixed,) = ctx.saved_tensors
        base, out, cells = ctx.base, ctx.out, ctx.cells.long()
        free = (fixed == 0).to(grad.dtype)
        g_free = grad * (free if grad.dim() == 1 else free[:, None])
        # off-diagonal part of A^T applied to the free rows' gradient
        offT = FvMatrix.__new__(FvMatrix)
        offT.field, offT.grid = base.field, base.grid
        offT._diag = torch.zeros_like(base._diag)
        offT._upper, offT._lower = base._lower, base._upper
        offT._source = base._source
        offT.face_flux_correction = None
        at = offT.multiply(g_free.contiguous())
        cd = out._diag[cells]
        gv = (cd if grad.dim() == 1 else cd[:, None]) * grad[cells] - at[cells]
        return gv.reshape(ctx.vshape), g_free, None, None, None, None, None


class Equation:
    """core/equation.py:6-68: the target field and its solve matrix, which
    carries the immersed Dirichlet cell constraints (Gibou snapping) through
    ``with_fixed_values``; the input matrix itself when there are none."""

    def __init__(self, target: CellField, fv_matrix: FvMatrix):
        self.name = target.name
        self.target = target
        if target.grid.num_immersed_faces:
            from .immersed import immersed_dirichlet_constraints  # noqa: PLC0415 (immersed imports fields)

            cells, values = immersed_dirichlet_constraints(target)  # one cell at most once
            fv_matrix = fv_matrix.with_fixed_values(cells, values, unique=True)
        self.fv_matrix = fv_matrix

    def _init_plain(self, target: CellField, fv_matrix: "FvMatrix") -> "Equation":
        """An equation whose matrix already carries its constraints."""
        self.name, self.target, self.fv_matrix = target.name, target, fv_matrix
        return self


def equation(target: CellField, fv_matrix: FvMatrix) -> Equation:
    return Equation(target, fv_matrix)


class _MultiplyFn(torch.autograd.Function):
    """y = A x as an autograd node.  Backward: x gets A^T g (the same SpMV
    with upper and lower exchanged); the LDU entries get g_diag = g x,
    g_upper[f] = g_O x_N, g_lower[f] = g_N x_O (gf_ldu_grads with lambda =
    -g, summed over components)."""

    @staticmethod
    def forward(ctx, grid, diag, upper, lower, x):  # noqa: D401
        # upper/lower are the matrix's full storage ([max(F, 1)]); they are
        # saved, so an in-place change before backward raises in autograd
        # instead of giving a silently wrong A^T g
        ctx.grid = grid
        xc = x.contiguous()
        ctx.save_for_backward(diag, upper, lower, xc)
        nc = 1 if x.dim() == 1 else x.shape[1]
        y = torch.empty_like(xc)
        nat.call("gf_ldu_multiply", grid.mesh_ref, ctypes.c_int32(nc), nat.ptr(diag), nat.ptr(upper),
                 nat.ptr(lower), nat.ptr(xc), nat.ptr(y), nat.stream())
        return y

    @staticmethod
    def backward(ctx, g_y):
        diag, upper, lower, x = ctx.saved_tensors
        need, g = ctx.needs_input_grad, ctx.grid
        nc = 1 if x.dim() == 1 else x.shape[1]
        gy = g_y.contiguous()
        g_x = None
        if need[4]:
            g_x = torch.empty_like(gy)
            nat.call("gf_ldu_multiply", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(diag), nat.ptr(lower),
                     nat.ptr(upper), nat.ptr(gy), nat.ptr(g_x), nat.stream())
        gd = gu = gl = None
        if need[1] or need[2] or need[3]:
            gd, gu, gl = torch.empty_like(diag), torch.zeros_like(upper), torch.zeros_like(lower)
            lam = (-gy).contiguous()
            nat.call("gf_ldu_grads", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(lam), nat.ptr(x),
                     nat.ptr(gd), nat.ptr(gu), nat.ptr(gl), nat.stream())
        return None, gd, gu, gl, g_x
