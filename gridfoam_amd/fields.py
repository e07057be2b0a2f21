"""Cell/face fields and boundary conditions in value-fraction form.

Restates the data contracts of src/gridfoam/core/field.py (CellField
``[C, *component_shape]`` with old/older history, FaceField blocks
``[single | domain_bnd]``) and the BC reduction of
src/gridfoam/boundaries/base.py:71-106: every condition becomes, per
boundary face, ``(fraction, ref_value, ref_grad)``.  Here those are dense
per-face device arrays over all domain faces plus a ``covered`` mask (a
face is covered when a non-empty BC owns it), which is what the HIP
kernels read.  The boundary face value is evaluated by ``gf_bc_state``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, Optional, Tuple, Union

import torch

from . import _native as nat
from .config import BoundaryConditionType, ConditionConfig, DomainBoundaryPatch
from .grid import OctreeGrid


class BoundaryCondition:
    """boundaries/base.py:19-106."""

    type: BoundaryConditionType
    assignable: bool = True

    def fill(self, grid: OctreeGrid, mask: torch.Tensor, frac, ref_v, ref_g) -> None:
        raise NotImplementedError


class DirichletBC(BoundaryCondition):  # boundaries/basic/dirichlet.py:17-73
    type = BoundaryConditionType.DIRICHLET
    assignable = False

    def __init__(self, value):
        self.value = torch.as_tensor(value, dtype=torch.float64)

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 1.0
        ref_v[mask] = self.value.to(ref_v.device)
        ref_g[mask] = 0.0


class NeumannBC(BoundaryCondition):  # boundaries/basic/neumann.py:17-73
    type = BoundaryConditionType.NEUMANN

    def __init__(self, grad_value):
        self.grad_value = torch.as_tensor(grad_value, dtype=torch.float64)

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 0.0
        ref_v[mask] = 0.0
        ref_g[mask] = self.grad_value.to(ref_g.device)


class EmptyBC(BoundaryCondition):  # boundaries/basic/empty.py:13-51 (excluded from batches)
    type = BoundaryConditionType.EMPTY

    def fill(self, grid, mask, frac, ref_v, ref_g):
        pass


class SlipBC(BoundaryCondition):  # boundaries/basic/slip.py:17-101
    type = BoundaryConditionType.SLIP
    assignable = False

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 1.0
        ref_v[mask] = 0.0
        ref_g[mask] = 0.0


def create_boundary_condition(kind: BoundaryConditionType, value, cshape) -> BoundaryCondition:
    """boundaries/factory.py:25-69 (subset)."""
    if kind == BoundaryConditionType.DIRICHLET:
        return DirichletBC(value if value is not None else 0.0)
    if kind == BoundaryConditionType.NEUMANN:
        return NeumannBC(value if value is not None else 0.0)
    if kind == BoundaryConditionType.EMPTY:
        return EmptyBC()
    if kind == BoundaryConditionType.SLIP:
        return SlipBC()
    raise ValueError(f"boundary condition {kind} is not supported by this build")


@dataclass
class BoundaryState:
    frac: torch.Tensor  # [B]
    ref_v: torch.Tensor  # [B, nc]
    ref_g: torch.Tensor  # [B, nc]
    covered: torch.Tensor  # [B] uint8
    psi_b: torch.Tensor  # [B, nc]
    nonassignable: torch.Tensor  # [B] uint8 (constrainHbyA faces)
    adjust_kind: torch.Tensor  # [B] int8: 0 none, 1 fixed value, 2 adjustable
    slip: Optional[torch.Tensor]  # [B] uint8 or None


class CellField:
    """core/field.py:120-380 (cell-centred field with history)."""

    def __init__(self, grid: OctreeGrid, name: str, component_shape: Tuple[int, ...] = (),
                 transient: bool = False, condition: Optional[ConditionConfig] = None):
        self.grid = grid
        self.name = name
        self.component_shape = tuple(component_shape)
        if self.component_shape not in ((), (3,)):
            raise ValueError("component_shape must be () or (3,)")
        self.transient = transient
        self.data = torch.zeros((grid.num_cells, *self.component_shape), dtype=torch.float64, device=grid.device)
        self.bcs: Dict[DomainBoundaryPatch, BoundaryCondition] = {}
        self.export = condition.export if condition is not None else True  # meta/config.py:757-761
        if condition is not None:
            self.data[:] = torch.as_tensor(condition.internal, dtype=torch.float64)
            for bc_cfg in condition.boundary.values():
                bc = create_boundary_condition(bc_cfg.type, bc_cfg.value, self.component_shape)
                for patch in bc_cfg.patches:
                    self.bcs[DomainBoundaryPatch(patch)] = bc
        self._state: Optional[BoundaryState] = None
        self.old_data: Optional[torch.Tensor] = None
        self.older_data: Optional[torch.Tensor] = None
        self.previous_dt: Optional[float] = None
        self.update_history(reset=True)

    @property
    def nc(self) -> int:
        return 3 if self.component_shape == (3,) else 1

    def add_boundary_conditions(self, bcs: Dict[Union[str, DomainBoundaryPatch], BoundaryCondition]) -> None:
        for k, v in bcs.items():
            self.bcs[DomainBoundaryPatch(k)] = v
        self._state = None

    def update_history(self, *, reset: bool = False) -> None:
        """core/field.py:241-255."""
        if self.transient:
            self.older_data = None if reset else self.old_data
            self.previous_dt = None if reset else self.grid.dt
            self.old_data = self.data.clone()
        else:
            self.older_data = None
            self.previous_dt = None
            self.old_data = self.data

    def boundary_state(self) -> BoundaryState:
        """Per-face value-fraction arrays, psi_b refreshed from current data."""
        g = self.grid
        B, nc, dev = g.num_domain_bnd_faces, self.nc, g.device
        if self._state is None:
            frac = torch.zeros(max(B, 1), dtype=torch.float64, device=dev)
            ref_v = torch.zeros((max(B, 1), nc), dtype=torch.float64, device=dev)
            ref_g = torch.zeros((max(B, 1), nc), dtype=torch.float64, device=dev)
            covered = torch.zeros(max(B, 1), dtype=torch.uint8, device=dev)
            nonassign = torch.zeros(max(B, 1), dtype=torch.uint8, device=dev)
            kind = torch.zeros(max(B, 1), dtype=torch.int8, device=dev)
            slip = torch.zeros(max(B, 1), dtype=torch.uint8, device=dev)
            dirs = g.domain_bnd_dir_id
            for patch, bc in self.bcs.items():
                mask = torch.zeros(max(B, 1), dtype=torch.bool, device=dev)
                mask[:B] = dirs == patch.to_direction()
                kind[mask] = 1 if isinstance(bc, DirichletBC) else 2
                if bc.type == BoundaryConditionType.EMPTY:
                    continue
                bc.fill(g, mask, frac, ref_v, ref_g)
                covered[mask] = 1
                if not bc.assignable:
                    nonassign[mask] = 1
                if isinstance(bc, SlipBC):
                    slip[mask] = 1
            self._state = BoundaryState(frac, ref_v, ref_g, covered, None, nonassign, kind,
                                        slip if bool(slip.any()) else None)
        st = self._state
        if st.psi_b is None:
            st.psi_b = torch.zeros((max(B, 1), nc), dtype=torch.float64, device=dev)
        if st.slip is not None and nc == 3:
            nat.call("gf_bc_slip", g.mesh_ref, nat.ptr(self.data), nat.ptr(st.slip), nat.ptr(st.ref_v), nat.stream())
        nat.call("gf_bc_state", g.mesh_ref, ctypes.c_int32(nc), nat.ptr(self.data), nat.ptr(st.frac),
                 nat.ptr(st.ref_v), nat.ptr(st.ref_g), nat.ptr(st.covered), nat.ptr(st.psi_b), nat.stream())
        return st


class FaceField:
    """core/field.py:387-560: [single | domain_bnd] blocks, old single level."""

    def __init__(self, grid: OctreeGrid, name: str, component_shape: Tuple[int, ...] = ()):
        self.grid = grid
        self.name = name
        self.component_shape = tuple(component_shape)
        nc = 3 if self.component_shape == (3,) else 1
        F, B = grid.num_internal_faces, grid.num_domain_bnd_faces
        shp = (lambda n: (max(n, 1), 3)) if nc == 3 else (lambda n: (max(n, 1),))
        self._single = torch.zeros(shp(F), dtype=torch.float64, device=grid.device)
        self._bnd = torch.zeros(shp(B), dtype=torch.float64, device=grid.device)
        self.old_single_data: Optional[torch.Tensor] = None
        self.older_single_data: Optional[torch.Tensor] = None
        self.previous_dt: Optional[float] = None
        self.update_history(reset=True)

    @property
    def num_single_sided(self) -> int:
        return self.grid.num_internal_faces

    @property
    def single_data(self) -> torch.Tensor:
        return self._single[: self.grid.num_internal_faces]

    @single_data.setter
    def single_data(self, v: torch.Tensor) -> None:
        self._single[: self.grid.num_internal_faces] = v

    @property
    def domain_bnd_data(self) -> torch.Tensor:
        return self._bnd[: self.grid.num_domain_bnd_faces]

    @domain_bnd_data.setter
    def domain_bnd_data(self, v: torch.Tensor) -> None:
        self._bnd[: self.grid.num_domain_bnd_faces] = v

    @property
    def single_buffer(self) -> torch.Tensor:
        return self._single

    @property
    def bnd_buffer(self) -> torch.Tensor:
        return self._bnd

    def update_history(self, *, reset: bool = False) -> None:
        """core/field.py:483-495 -- always an independent copy."""
        self.older_single_data = None if reset else self.old_single_data
        self.previous_dt = None if reset else self.grid.dt
        self.old_single_data = self._single.clone()
