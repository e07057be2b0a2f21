"""Cell/face fields and boundary conditions in value-fraction form.

Restates the data contracts of src/gridfoam/core/field.py (CellField
``[C, *component_shape]`` with old/older history, FaceField blocks
``[single | domain_bnd]``) and the BC reduction of
src/gridfoam/boundaries/base.py:71-106: every condition becomes, per
boundary face, ``(fraction, ref_value, ref_grad)``.  Here those are dense
per-face device arrays over all domain faces plus a ``covered`` mask (a
face is covered when a non-empty BC owns it), which is what the HIP
kernels read.  The boundary face value is evaluated by ``gf_bc_state``.
Conditions that depend on other registered fields (inletOutlet reads the
flux ``phi``, fixedFluxPressure the predicted flux and the pressure
coefficient) refresh their faces on every ``boundary_state()`` call, as the
reference re-evaluates them at each use.
"""

from __future__ import annotations

import ctypes
import dataclasses
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
    dynamic: bool = False  # True: refresh() runs on every boundary_state() call

    def fill(self, grid: OctreeGrid, mask: torch.Tensor, frac, ref_v, ref_g) -> None:
        raise NotImplementedError

    def refresh(self, field: "CellField", patch: DomainBoundaryPatch, idx: torch.Tensor,
                state: "BoundaryState") -> None:
        """Rewrite (frac, ref_v, ref_g) on the faces ``idx`` from current data."""

    def graph_ref_g(self, field: "CellField", patch, idx: torch.Tensor) -> Optional[torch.Tensor]:
        """The refreshed reference gradient on ``idx`` as a graph-carrying
        tensor when its inputs require grad (else None)."""
        return None


class DirichletBC(BoundaryCondition):  # boundaries/basic/dirichlet.py:17-73
    type = BoundaryConditionType.DIRICHLET
    assignable = False

    def __init__(self, value):
        self.value = torch.as_tensor(value, dtype=torch.float64)

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 1.0
        ref_v[mask] = self.value.detach().to(ref_v.device)
        ref_g[mask] = 0.0


class NeumannBC(BoundaryCondition):  # boundaries/basic/neumann.py:17-73
    type = BoundaryConditionType.NEUMANN

    def __init__(self, grad_value):
        self.grad_value = torch.as_tensor(grad_value, dtype=torch.float64)

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 0.0
        ref_v[mask] = 0.0
        ref_g[mask] = self.grad_value.detach().to(ref_g.device)


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


class InletOutletBC(BoundaryCondition):
    """boundaries/derived/inlet_outlet.py:27-123: zero gradient where the
    registered flux ``phi`` leaves the domain or is zero (phi >= 0), the
    inlet value where it enters (phi < 0).  Assignable, and adjustable by
    adjustPhi (fv/adjust_phi.py:14-41)."""

    type = BoundaryConditionType.INLET_OUTLET
    dynamic = True

    def __init__(self, inlet_value, phi_name: str = "phi"):
        self.inlet_value = torch.as_tensor(inlet_value, dtype=torch.float64)
        self.phi_name = phi_name

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 0.0
        ref_v[mask] = 0.0
        ref_g[mask] = 0.0

    def refresh(self, field, patch, idx, state):
        if tuple(self.inlet_value.shape) != field.component_shape:
            raise ValueError(f"inlet value shape {tuple(self.inlet_value.shape)} does not match "
                             f"field shape {field.component_shape}")
        phi = field.grid.get_facefield(self.phi_name)
        if phi is None:
            raise ValueError(f"Field {self.phi_name} is not found.")
        inflow = phi.bnd_data.index_select(0, idx) < 0.0
        state.frac.index_copy_(0, idx, inflow.to(torch.float64))
        val = self.inlet_value.detach().to(state.ref_v.device).reshape(1, -1)
        state.ref_v.index_copy_(0, idx, torch.where(inflow[:, None], val, torch.zeros_like(val)))


class FixedFluxPressureBC(BoundaryCondition):
    """boundaries/derived/fixed_flux_pressure.py:124-252: a normal gradient
    that makes the corrected boundary flux equal the velocity boundary flux,
    ``snGrad(p) = (phiHbyA_b - U_b . Sf) / (rAtU_P |Sf|)`` (zero where the
    denominator is not positive).  The coefficient is ``rAtU``, else
    ``rAU``; without ``phiHbyA`` the predicted flux is ``HbyA_P . Sf``;
    without a velocity condition on the patch the target is the boundary
    ``phi``.  Until the coupling fields are registered the gradient is zero.
    Scalar fields only."""

    type = BoundaryConditionType.FIXED_FLUX_PRESSURE
    dynamic = True

    def __init__(self, phase_suffix: str = ""):
        sfx = phase_suffix
        self.phi_name, self.phi_hbya_name, self.HbyA_name = "phi" + sfx, "phiHbyA" + sfx, "HbyA" + sfx
        self.rAU_name, self.rAtU_name, self.U_name = "rAU" + sfx, "rAtU" + sfx, "U" + sfx

    def fill(self, grid, mask, frac, ref_v, ref_g):
        frac[mask] = 0.0
        ref_v[mask] = 0.0
        ref_g[mask] = 0.0

    def _gradient(self, field, patch, idx, graph: bool) -> Optional[torch.Tensor]:
        """(phiHbyA_b - U_b . Sf) / (rAtU_P |Sf|) on ``idx``; None: zero gradient."""
        if field.component_shape != ():
            raise ValueError("FixedFluxPressure requires a scalar field")
        g = field.grid
        phi, phi_hbya = g.get_facefield(self.phi_name), g.get_facefield(self.phi_hbya_name)
        HbyA, U = g.get_cellfield(self.HbyA_name), g.get_cellfield(self.U_name)
        coeff = g.get_cellfield(self.rAtU_name)
        if coeff is None:
            coeff = g.get_cellfield(self.rAU_name)
        if coeff is None or (phi_hbya is None and HbyA is None):
            return None
        cells = g.bnd_owner.index_select(0, idx)
        Sf = g.bnd_Sf.index_select(0, idx)
        if phi_hbya is not None:
            pred = phi_hbya.bnd_data.index_select(0, idx)
        else:
            pred = (HbyA.data.index_select(0, cells) * Sf).sum(dim=1)
        if U is not None and patch in U.bcs:
            if graph:
                from .fv import boundary_values  # noqa: PLC0415 (fv imports this module)

                ub = boundary_values(U)
            else:
                ub = U.boundary_state().psi_b
            target = (ub.index_select(0, idx) * Sf).sum(dim=1)
        elif phi is not None:
            target = phi.bnd_data.index_select(0, idx)
        else:
            return None
        denom = coeff.data.index_select(0, cells) * torch.linalg.vector_norm(Sf, dim=1)
        pos = denom > 0.0
        return torch.where(pos, (pred - target) / torch.where(pos, denom, torch.ones_like(denom)),
                           torch.zeros_like(denom))

    def refresh(self, field, patch, idx, state):
        grad = self._gradient(field, patch, idx, graph=False)
        if grad is None:
            state.ref_g.index_fill_(0, idx, 0.0)
        else:
            state.ref_g.index_copy_(0, idx, grad[:, None])

    def graph_ref_g(self, field, patch, idx):
        g = field.grid
        ins = [g.get_facefield(self.phi_name), g.get_facefield(self.phi_hbya_name)]
        cells = [g.get_cellfield(n) for n in (self.HbyA_name, self.U_name, self.rAtU_name, self.rAU_name)]
        if not any(f is not None and f.bnd_buffer.requires_grad for f in ins) and not any(
                c is not None and (c.data.requires_grad or any(
                    getattr(b, "value", torch.zeros(())).requires_grad for b in c.bcs.values())) for c in cells):
            return None
        grad = self._gradient(field, patch, idx, graph=True)
        return grad if grad is not None and grad.requires_grad else None


def patch_key(patch) -> Union[DomainBoundaryPatch, str]:
    """A condition's patch: a domain patch, or the name of an immersed body."""
    if isinstance(patch, DomainBoundaryPatch):
        return patch
    name = str(patch)
    return DomainBoundaryPatch(name) if name in {p.value for p in DomainBoundaryPatch} else name


def create_boundary_condition(kind: BoundaryConditionType, value, cshape) -> BoundaryCondition:
    """boundaries/factory.py:25-69."""
    if kind == BoundaryConditionType.DIRICHLET:
        return DirichletBC(value if value is not None else 0.0)
    if kind == BoundaryConditionType.NEUMANN:
        return NeumannBC(value if value is not None else 0.0)
    if kind == BoundaryConditionType.EMPTY:
        return EmptyBC()
    if kind == BoundaryConditionType.SLIP:
        return SlipBC()
    if kind == BoundaryConditionType.INLET_OUTLET:
        if value is None:
            raise ValueError("inletOutlet requires an inlet value")
        return InletOutletBC(torch.broadcast_to(torch.as_tensor(value, dtype=torch.float64), cshape))
    if kind == BoundaryConditionType.FIXED_FLUX_PRESSURE:
        return FixedFluxPressureBC()
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
    # faces whose flux U_f . Sf is kept by flux_from_face_velocity: covered
    # domain faces and every immersed side (fv/flux.py:59-71)
    flux_keep: Optional[torch.Tensor] = None
    # autograd-connected copies of ref_v / ref_g, rebuilt on every call when a
    # Dirichlet value or Neumann gradient requires grad (else None)
    ref_v_ad: Optional[torch.Tensor] = None
    ref_g_ad: Optional[torch.Tensor] = None

    def frozen(self) -> "BoundaryState":
        """A copy of the per-face values that later boundary_state() calls do
        not rewrite (they refresh inletOutlet, fixedFluxPressure and slip
        faces in place): what an autograd node keeps for its backward."""
        return dataclasses.replace(self, frac=self.frac.clone(), ref_v=self.ref_v.clone(),
                                   ref_g=self.ref_g.clone(), psi_b=self.psi_b.clone())


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
        self.bcs: Dict[Union[DomainBoundaryPatch, str], BoundaryCondition] = {}
        self.export = condition.export if condition is not None else True  # meta/config.py:757-761
        if condition is not None:
            self.data[:] = torch.as_tensor(condition.internal, dtype=torch.float64)
            for bc_cfg in condition.boundary.values():
                bc = create_boundary_condition(bc_cfg.type, bc_cfg.value, self.component_shape)
                for patch in bc_cfg.patches:
                    self.bcs[patch_key(patch)] = bc
        self._state: Optional[BoundaryState] = None
        self._dynamic: list = []
        self._valued: list = []
        grid.register_cellfield(self)
        self.old_data: Optional[torch.Tensor] = None
        self.older_data: Optional[torch.Tensor] = None
        self.previous_dt: Optional[float] = None
        self.update_history(reset=True)

    @property
    def nc(self) -> int:
        return 3 if self.component_shape == (3,) else 1

    def add_boundary_conditions(self, bcs: Dict[Union[str, DomainBoundaryPatch], BoundaryCondition]) -> None:
        for k, v in bcs.items():
            self.bcs[patch_key(k)] = v
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

    @property
    def has_dynamic_state(self) -> bool:
        """Whether boundary_state() rewrites per-face values on each call."""
        if self._state is None:
            self.boundary_state()
        return bool(self._dynamic) or self._state.slip is not None

    def boundary_state(self) -> BoundaryState:
        """Per-face value-fraction arrays, psi_b refreshed from current data."""
        g = self.grid
        B, nc, dev = g.num_bnd_faces, self.nc, g.device
        if self._state is None:
            self._valued = []
            frac = torch.zeros(max(B, 1), dtype=torch.float64, device=dev)
            ref_v = torch.zeros((max(B, 1), nc), dtype=torch.float64, device=dev)
            ref_g = torch.zeros((max(B, 1), nc), dtype=torch.float64, device=dev)
            covered = torch.zeros(max(B, 1), dtype=torch.uint8, device=dev)
            nonassign = torch.zeros(max(B, 1), dtype=torch.uint8, device=dev)
            kind = torch.zeros(max(B, 1), dtype=torch.int8, device=dev)
            slip = torch.zeros(max(B, 1), dtype=torch.uint8, device=dev)
            dynamic = []
            for patch, bc in self.bcs.items():
                mask = torch.zeros(max(B, 1), dtype=torch.bool, device=dev)
                mask[:B] = g.bnd_patch_mask(patch)
                if isinstance(patch, DomainBoundaryPatch):  # adjustPhi balances domain patches only
                    kind[mask] = 1 if isinstance(bc, DirichletBC) else 2
                if bc.type == BoundaryConditionType.EMPTY:
                    continue
                bc.fill(g, mask, frac, ref_v, ref_g)
                covered[mask] = 1
                if not bc.assignable:
                    nonassign[mask] = 1
                if isinstance(bc, SlipBC):
                    slip[mask] = 1
                if bc.dynamic:
                    idx = torch.nonzero(mask[:B]).flatten()
                    if idx.numel():
                        dynamic.append((patch, bc, idx))
                if isinstance(bc, (DirichletBC, NeumannBC)):
                    self._valued.append((bc, mask))
            flux_keep = covered.clone()
            flux_keep[g.num_domain_bnd_faces:B] = 1
            self._state = BoundaryState(frac, ref_v, ref_g, covered, None, nonassign, kind,
                                        slip if bool(slip.any()) else None, flux_keep)
            self._dynamic = dynamic
        st = self._state
        if st.psi_b is None:
            st.psi_b = torch.zeros((max(B, 1), nc), dtype=torch.float64, device=dev)
        with torch.no_grad():  # the state arrays are device buffers; gradients do not run through these refreshes
            for patch, bc, idx in self._dynamic:
                bc.refresh(self, patch, idx, st)
        if st.slip is not None and nc == 3:
            nat.call("gf_bc_slip", g.mesh_ref, nat.ptr(self.data), nat.ptr(st.slip), nat.ptr(st.ref_v), nat.stream())
        st.ref_v_ad = st.ref_g_ad = None
        if torch.is_grad_enabled() and any(
                (bc.value if isinstance(bc, DirichletBC) else bc.grad_value).requires_grad for bc, _ in self._valued):
            # values that require grad are re-read on every call: the device
            # arrays get their current values, the copies carry the graph
            rv, rg = st.ref_v.detach().clone(), st.ref_g.detach().clone()
            for bc, mask in self._valued:
                if isinstance(bc, DirichletBC):
                    rv = torch.where(mask[:, None], bc.value.to(dev).reshape(1, -1), rv)
                else:
                    rg = torch.where(mask[:, None], bc.grad_value.to(dev).reshape(1, -1), rg)
            st.ref_v.copy_(rv.detach())
            st.ref_g.copy_(rg.detach())
            st.ref_v_ad, st.ref_g_ad = rv, rg
        if torch.is_grad_enabled():  # conditions whose refreshed gradient depends on fields with a graph
            for patch, bc, idx in self._dynamic:
                gv = bc.graph_ref_g(self, patch, idx)
                if gv is not None:
                    base = st.ref_g_ad if st.ref_g_ad is not None else st.ref_g.detach().clone()
                    st.ref_g_ad = base.index_copy(0, idx, gv.reshape(-1, *base.shape[1:]))
        if g.bnd_zero_dist.numel() and not torch.cuda.is_current_stream_capturing():
            from .immersed import check_zero_distance  # noqa: PLC0415 (immersed imports this module)

            check_zero_distance(self, st)
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
        F, B = grid.num_internal_faces, grid.num_bnd_faces
        shp = (lambda n: (max(n, 1), 3)) if nc == 3 else (lambda n: (max(n, 1),))
        self._single = torch.zeros(shp(F), dtype=torch.float64, device=grid.device)
        self._bnd = torch.zeros(shp(B), dtype=torch.float64, device=grid.device)
        grid.register_facefield(self)
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
    def bnd_data(self) -> torch.Tensor:
        """Every boundary face: [domain | immersed upper | immersed lower]."""
        return self._bnd[: self.grid.num_bnd_faces]

    @bnd_data.setter
    def bnd_data(self, v: torch.Tensor) -> None:
        self._bnd[: self.grid.num_bnd_faces] = v

    @property
    def immersed_upper(self) -> torch.Tensor:
        """core/field.py:455-470: owner sides of the immersed faces (along +Sf)."""
        B, Fi = self.grid.num_domain_bnd_faces, self.grid.num_immersed_faces
        return self._bnd[B: B + Fi]

    @immersed_upper.setter
    def immersed_upper(self, v: torch.Tensor) -> None:
        B, Fi = self.grid.num_domain_bnd_faces, self.grid.num_immersed_faces
        self._bnd[B: B + Fi] = v

    @property
    def immersed_lower(self) -> torch.Tensor:
        """Neighbour sides of the immersed faces (along -Sf)."""
        B, Fi = self.grid.num_domain_bnd_faces, self.grid.num_immersed_faces
        return self._bnd[B + Fi: B + 2 * Fi]

    @immersed_lower.setter
    def immersed_lower(self, v: torch.Tensor) -> None:
        B, Fi = self.grid.num_domain_bnd_faces, self.grid.num_immersed_faces
        self._bnd[B + Fi: B + 2 * Fi] = v

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
