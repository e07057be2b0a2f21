"""Batched steady SIMPLE: several cases advanced by one launch sequence.

Cases of a serving stream are small enough that a SIMPLE iteration of one
case leaves the GPU mostly idle: its pressure CG is bound by the latency of
its two launches per iteration, not by its size, up to a few 1e5 cells.
Here the admitted cases are concatenated into one mesh (cells, faces and
boundary faces of case k follow those of case k-1; no face couples two
cases), so assembly, gradients, fluxes and corrections run as one launch
each for all of them, and the pressure equation is solved by the segmented
PCG (``gf_seg_pcg``), whose per-case scalars and stopping tests make every
case follow its own CG iteration; the momentum equations likewise go
through the segmented BiCGSTAB (``gf_seg_bicgstab``).  Per-case
quantities -- viscosity (Reynolds number), pressure reference cell,
adjustPhi balance, initial residuals and residualControl -- are kept per
segment, so a case's fields do not depend on which other cases share its
batch.  The SIMPLE iteration follows algorithms/simple.py:287-408 (the
SIMPLE driver of this package), without SIMPLEC.
"""

from __future__ import annotations

import copy
import ctypes
import weakref
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple

import torch

from . import _native as nat
from .algorithms import (compute_phi_hbya, correct_flux, correct_phi, correct_velocity, normalize_residual_control,
                         residual_satisfied)
from .config import GridfoamConfig, NormType, PreconditionerType, SIMPLEAlgorithm, SolverType
from .fields import CellField, DirichletBC, FaceField
from .fv import fvc, fvm
from .fvmatrix import Equation, FvMatrix
from .grid import OctreeGrid
from .immersed import N_DOMAIN_PATCHES, apply_immersed_dirichlet_values

SEG_MIN_ROWS = 256  # gf_seg_pcg: every case needs at least this many cells


# --------------------------------------------------------------------- mesh
@dataclass
class CaseSlices:
    """Where case k's cells, faces and boundary faces sit in the combined mesh."""

    cells: Tuple[int, int]
    faces: Tuple[int, int]
    bnd: torch.Tensor  # [B_k] combined position of each of the case's boundary faces (its own order)


def concat_grids(grids: Sequence[OctreeGrid], sim_config) -> Tuple[OctreeGrid, List[CaseSlices]]:
    """One OctreeGrid holding the given grids side by side.  Boundary faces
    keep the [domain | immersed upper | immersed lower] order over all
    cases; immersed patches are unified by body name.  The solver layout is
    built on integer coordinates shifted apart along x, so no brick mixes two
    cases."""
    if not grids:
        raise ValueError("no grids to combine")
    dev = grids[0].device
    g = OctreeGrid.__new__(OctreeGrid)
    g.device, g.dtype, g.dt = dev, torch.float64, 1.0
    g.sim_config = sim_config
    g._cellfields = weakref.WeakValueDictionary()
    g._facefields = weakref.WeakValueDictionary()
    g.face_order = g.bnd_face_order = None
    C = [k.num_cells for k in grids]
    F = [k.num_internal_faces for k in grids]
    Bd = [k.num_domain_bnd_faces for k in grids]
    Fi = [k.num_immersed_faces for k in grids]
    Fa = [k.num_all_internal_faces for k in grids]
    coff = [sum(C[:k]) for k in range(len(grids))]
    g.num_cells, g.num_internal_faces = sum(C), sum(F)
    g.num_domain_bnd_faces, g.num_immersed_faces = sum(Bd), sum(Fi)
    g.num_bnd_faces = g.num_domain_bnd_faces + 2 * g.num_immersed_faces
    g.num_all_internal_faces = sum(Fa)
    g.lmax = max(k.lmax for k in grids)

    def pad(t, n):
        if t.shape[0] >= max(n, 1):
            return t.contiguous()
        return torch.cat([t, torch.zeros((max(n, 1) - t.shape[0],) + tuple(t.shape[1:]), dtype=t.dtype,
                                         device=dev)]).contiguous()

    def cells(name):
        return torch.cat([getattr(k, name)[: C[i]] for i, k in enumerate(grids)])

    def faces(name, shift=False):
        return pad(torch.cat([getattr(k, name)[: F[i]] + (coff[i] if shift else 0) for i, k in enumerate(grids)]),
                   g.num_internal_faces)

    g.level = cells("level")
    # integer coordinates shifted apart along x (multiples of 4 root blocks)
    shifts, x0 = [], 0
    for k in grids:
        shifts.append(x0)
        x0 += (k.fluxel.root_resolution[0] + 7) // 4 * 4
    ijk = []
    for i, k in enumerate(grids):
        t = k.ijk[: C[i]].clone()
        lv = k.level[: C[i]].long()
        t[:, 0] += ((torch.ones_like(lv) * shifts[i]) << lv).to(t.dtype)
        ijk.append(t)
    g.ijk = torch.cat(ijk)
    root = [x0, max(k.fluxel.root_resolution[1] for k in grids), max(k.fluxel.root_resolution[2] for k in grids)]
    g.fluxel = grids[0].fluxel.model_copy(update={"root_resolution": root, "immersed_bodies": []})
    g.owner_i32 = faces("owner_i32", True)
    g.neighbour_i32 = faces("neighbour_i32", True)
    g.axis_i32 = faces("axis_i32")
    g.face_centers = faces("face_centers")
    g.Sf = faces("Sf")
    g.cell_centers, g.cell_sizes, g.cell_volumes = cells("cell_centers"), cells("cell_sizes"), cells("cell_volumes")
    # immersed bodies unified by name
    names: Dict[str, int] = {}
    bodies = []
    for k in grids:
        for b in k.immersed_bodies:
            if b.name not in names:
                names[b.name] = N_DOMAIN_PATCHES + len(bodies)
                bodies.append(b)
    g.immersed_bodies = bodies
    g.patch_name_to_id = names

    def remap(k):  # local patch id -> combined patch id
        m = torch.arange(N_DOMAIN_PATCHES + max(len(k.immersed_bodies), 1), device=dev)
        for j, b in enumerate(k.immersed_bodies):
            m[N_DOMAIN_PATCHES + j] = names[b.name]
        return m

    # boundary faces: [domain of all | upper of all | lower of all]
    parts = {"dom": [], "up": [], "lo": []}
    for i, k in enumerate(grids):
        parts["dom"].append((i, 0, Bd[i]))
        parts["up"].append((i, Bd[i], Bd[i] + Fi[i]))
        parts["lo"].append((i, Bd[i] + Fi[i], Bd[i] + 2 * Fi[i]))
    order = parts["dom"] + parts["up"] + parts["lo"]

    def bnd(name, shift=False, fn=None):
        out = []
        for i, a, b in order:
            t = getattr(grids[i], name)[a:b]
            if shift:
                t = t + coff[i]
            if fn is not None:
                t = fn(i, t)
            out.append(t)
        return pad(torch.cat(out), g.num_bnd_faces)

    g.bnd_owner_i32 = bnd("bnd_owner_i32", True)
    g.bnd_dir_i32 = bnd("bnd_dir_i32")
    g.bnd_patch_id = bnd("bnd_patch_id", fn=lambda i, t: remap(grids[i])[t])
    g.bnd_d = bnd("bnd_d")
    g.bnd_mag_d = bnd("bnd_mag_d")
    g.bnd_mag_sf = bnd("bnd_mag_sf")
    g.bnd_sign = bnd("bnd_sign")
    g.bnd_Sf = bnd("bnd_Sf")
    g.bnd_snap = bnd("bnd_snap")
    g.domain_bnd_face_centers = pad(torch.cat([k.domain_bnd_face_centers[: Bd[i]] for i, k in enumerate(grids)]),
                                    g.num_domain_bnd_faces)
    g.domain_bnd_Sf = pad(torch.cat([k.domain_bnd_Sf[: Bd[i]] for i, k in enumerate(grids)]), g.num_domain_bnd_faces)
    g.bnd_zero_dist = torch.nonzero(g.bnd_mag_d[: g.num_bnd_faces] == 0).flatten()
    # axis-projected payload, concatenated in case order
    g.ap_is_immersed_faces = torch.cat([k.ap_is_immersed_faces for k in grids])
    for nm in ("ap_dist_owner_to_bnd", "ap_dist_neighbour_to_bnd", "ap_owner_near_boundary",
               "ap_neighbour_near_boundary", "ap_owner_normal", "ap_neighbour_normal"):
        setattr(g, nm, torch.cat([getattr(k, nm) for k in grids]))
    for nm in ("ap_owner_bnd_patch_id", "ap_neighbour_bnd_patch_id"):
        setattr(g, nm, torch.cat([remap(k)[N_DOMAIN_PATCHES + getattr(k, nm)] - N_DOMAIN_PATCHES
                                  if getattr(k, nm).numel() else getattr(k, nm) for k in grids]))
    g._build_operator()
    # per-case positions
    slices = []
    dom0 = [sum(Bd[:i]) for i in range(len(grids))]
    up0 = [g.num_domain_bnd_faces + sum(Fi[:i]) for i in range(len(grids))]
    lo0 = [g.num_domain_bnd_faces + g.num_immersed_faces + sum(Fi[:i]) for i in range(len(grids))]
    for i in range(len(grids)):
        pos = torch.cat([torch.arange(dom0[i], dom0[i] + Bd[i], device=dev),
                         torch.arange(up0[i], up0[i] + Fi[i], device=dev),
                         torch.arange(lo0[i], lo0[i] + Fi[i], device=dev)])
        slices.append(CaseSlices((coff[i], coff[i] + C[i]), (sum(F[:i]), sum(F[:i]) + F[i]), pos))
    return g, slices


# ------------------------------------------------------------ segmented PCG
@dataclass
class SegStats:
    iterations: List[int]
    initial_residual: List[float]
    final_residual: List[float]
    converged: List[bool]


class SegmentedCG:
    """Jacobi PCG over the segments of a block-diagonal operator
    (``gf_seg_pcg``): each segment has its own scalars, threshold
    max(atol, rtol |r0|) and stop."""

    def __init__(self, atol: float, rtol: float, max_iter: int, precond: bool = True, check_every: int = 16):
        self.atol, self.rtol, self.max_iter = float(atol), float(rtol), int(max_iter)
        self.precond, self.check_every = bool(precond), int(check_every)
        self._ws: Optional[torch.Tensor] = None
        self._flag = torch.zeros(2, dtype=torch.int32).pin_memory()  # [2][n_seg] active flags

    def solve(self, A: FvMatrix, b: torch.Tensor, x0: torch.Tensor, seg_row: torch.Tensor,
              seg_row_host: List[int], row_seg: torch.Tensor) -> Tuple[torch.Tensor, SegStats]:
        g = A.grid
        n_seg = len(seg_row_host) - 1
        nbytes = ctypes.c_int64(0)
        nat.call("gf_seg_pcg_workspace", ctypes.c_int32(g.num_cells), ctypes.c_int32(n_seg),
                 ctypes.c_int64(int(g.mesh.nnz_pad)), ctypes.byref(nbytes))
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.zeros(nbytes.value, dtype=torch.uint8, device=g.device)
        if self._flag.numel() < 2 * n_seg:
            self._flag = torch.zeros(2 * n_seg, dtype=torch.int32).pin_memory()
        x = x0.detach().clone().contiguous()
        host_rows = (ctypes.c_int32 * (n_seg + 1))(*seg_row_host)
        iters = (ctypes.c_int32 * n_seg)()
        res = (ctypes.c_double * (2 * n_seg))()
        conv = (ctypes.c_uint8 * n_seg)()
        nat.call("gf_seg_pcg", g.mesh_ref, ctypes.c_int32(n_seg), nat.ptr(seg_row), host_rows, nat.ptr(row_seg),
                 nat.ptr(A._diag), nat.ptr(A._upper), nat.ptr(A._lower), nat.ptr(b.detach().contiguous()),
                 nat.ptr(x), ctypes.c_int32(int(self.precond)), ctypes.c_double(self.atol),
                 ctypes.c_double(self.rtol), ctypes.c_int32(self.max_iter), ctypes.c_int32(self.check_every),
                 nat.ptr(self._ws), ctypes.c_int64(self._ws.numel()), iters, res, conv,
                 ctypes.c_void_p(self._flag.data_ptr()), nat.stream())
        return x, SegStats(list(iters), [res[2 * s] for s in range(n_seg)], [res[2 * s + 1] for s in range(n_seg)],
                           [bool(c) for c in conv])


class SegmentedBiCGSTAB:
    """Jacobi right-preconditioned BiCGSTAB of a [C, 3] system over the
    segments of a block-diagonal operator (``gf_seg_bicgstab``): every
    (segment, component) has its own scalars, threshold, restarts and stop."""

    def __init__(self, atol: float, rtol: float, max_iter: int, max_restart: int, precond: bool = True,
                 check_every: int = 8):
        self.atol, self.rtol, self.max_iter = float(atol), float(rtol), int(max_iter)
        self.max_restart, self.precond, self.check_every = int(max_restart), bool(precond), int(check_every)
        self._ws: Optional[torch.Tensor] = None
        self._flag = torch.zeros(2, dtype=torch.int32).pin_memory()

    def solve(self, A: FvMatrix, b: torch.Tensor, x0: torch.Tensor,
              seg_row_host: List[int]) -> Tuple[torch.Tensor, SegStats]:
        g = A.grid
        n_seg = len(seg_row_host) - 1
        if tuple(b.shape) != (g.num_cells, 3) or tuple(x0.shape) != (g.num_cells, 3):
            raise ValueError("the segmented BiCGSTAB solves [C, 3] systems")
        nbytes = ctypes.c_int64(0)
        nat.call("gf_seg_bicgstab_workspace", ctypes.c_int32(g.num_cells), ctypes.c_int32(n_seg),
                 ctypes.c_int64(int(g.mesh.nnz_pad)), ctypes.byref(nbytes))
        if self._ws is None or self._ws.numel() < nbytes.value:
            self._ws = torch.zeros(nbytes.value, dtype=torch.uint8, device=g.device)
        if self._flag.numel() < 2 * n_seg:
            self._flag = torch.zeros(2 * n_seg, dtype=torch.int32).pin_memory()
        x = x0.detach().clone().contiguous()
        host_rows = (ctypes.c_int32 * (n_seg + 1))(*seg_row_host)
        n = 3 * n_seg
        iters = (ctypes.c_int32 * n)()
        res = (ctypes.c_double * (2 * n))()
        conv = (ctypes.c_uint8 * n)()
        nat.call("gf_seg_bicgstab", g.mesh_ref, ctypes.c_int32(n_seg), host_rows, nat.ptr(A._diag),
                 nat.ptr(A._upper), nat.ptr(A._lower), nat.ptr(b.detach().contiguous()), nat.ptr(x),
                 ctypes.c_int32(int(self.precond)), ctypes.c_double(self.atol), ctypes.c_double(self.rtol),
                 ctypes.c_int32(self.max_iter), ctypes.c_int32(self.max_restart), ctypes.c_int32(self.check_every),
                 nat.ptr(self._ws), ctypes.c_int64(self._ws.numel()), iters, res, conv,
                 ctypes.c_void_p(self._flag.data_ptr()), nat.stream())
        return x, SegStats(list(iters), [res[2 * k] for k in range(n)], [res[2 * k + 1] for k in range(n)],
                           [bool(c) for c in conv])


# ---------------------------------------------------------- batched SIMPLE
@dataclass
class CaseState:
    """One case's SIMPLE state between batches (its own cell/face order)."""

    rid: int
    config: dict
    grid: OctreeGrid
    nu: float
    ref: Optional[Tuple[int, float]]
    U: torch.Tensor
    p: torch.Tensor
    phi_s: torch.Tensor
    phi_b: torch.Tensor
    iterations: int = 0
    p_iterations: int = 0
    initial: Dict[str, float] = field(default_factory=dict)
    current: Dict[str, float] = field(default_factory=dict)
    done: bool = False


def new_case(rid: int, config: GridfoamConfig, raw: dict) -> CaseState:
    """Build a request's grid and its initial fields (the SIMPLE driver's
    construction, simple.py:155-190)."""
    from .grid import create_grid

    g = create_grid(config)
    sim = config.simulator
    cond = sim.conditions
    U = CellField(g, "U", (3,), condition=cond.get("U"))
    p = CellField(g, "p", (), condition=cond.get("p"))
    phi = FaceField(g, "phi")
    correct_flux(phi, U, update_internal=True)
    alg = sim.fvSolution.algorithm
    if not isinstance(alg, SIMPLEAlgorithm) or alg.consistent:
        raise ValueError("the batched server runs plain SIMPLE cases")
    needs_ref = not any(isinstance(bc, DirichletBC) for bc in p.bcs.values())
    ref = (int(alg.pRefCell), float(alg.pRefValue)) if needs_ref else None
    if g.num_cells < SEG_MIN_ROWS:
        raise ValueError(f"a batched case needs at least {SEG_MIN_ROWS} cells")
    return CaseState(rid, raw, g, float(sim.properties.transport.nu), ref, U.data.clone(), p.data.clone(),
                     phi.single_data.clone(), phi.bnd_data.clone())


def merged_config(cases: Sequence[CaseState]) -> GridfoamConfig:
    """The first case's configuration with every condition's patch list
    extended by the other cases' patches of the same entry (their obstacle
    names); solvers, schemes and algorithm settings must agree."""
    from .config import load_config

    raw = copy.deepcopy(cases[0].config)
    keys = ("fvSolution", "fvSchemes")
    for c in cases[1:]:
        for k in keys:
            if c.config["simulator"].get(k) != raw["simulator"].get(k):
                raise ValueError(f"case {c.rid}: {k} differs from the batch")
        for fname, fc in c.config["simulator"]["conditions"].items():
            mine = raw["simulator"]["conditions"][fname]["boundary"]
            for ename, entry in fc["boundary"].items():
                if ename not in mine:
                    mine[ename] = copy.deepcopy(entry)
                    continue
                if entry["type"] != mine[ename]["type"] or entry.get("value") != mine[ename].get("value"):
                    raise ValueError(f"case {c.rid}: condition {fname}.{ename} differs from the batch")
                for patch in entry["patches"]:
                    if patch not in mine[ename]["patches"]:
                        mine[ename]["patches"].append(patch)
    return load_config(raw)


class Batch:
    """The combined mesh, fields and solvers of a set of admitted cases."""

    def __init__(self, cases: Sequence[CaseState]):
        self.cases = list(cases)
        template = merged_config(self.cases)
        sim = template.simulator
        self.grid, self.slices = concat_grids([c.grid for c in self.cases], sim)
        g = self.grid
        dev = g.device
        alg = sim.fvSolution.algorithm
        self.alpha_U = alg.relaxationFactors.equations["U"]
        self.alpha_p = alg.relaxationFactors.equations["p"]
        self.rc = normalize_residual_control(alg.residualControl)
        # the fields, with every case's conditions (shared types; obstacle names unified)
        cond = copy.deepcopy(sim.conditions)
        self.U = CellField(g, "U", (3,), condition=cond.get("U"))
        self.p = CellField(g, "p", (), condition=cond.get("p"))
        self.phi = FaceField(g, "phi")
        self.phi_hbya = FaceField(g, "phiHbyA")
        self.rAU = CellField(g, "rAU")
        self.HbyA = CellField(g, "HbyA", (3,))
        self.U.data = torch.cat([c.U for c in self.cases]).contiguous()
        self.p.data = torch.cat([c.p for c in self.cases]).contiguous()
        phi_s = torch.cat([c.phi_s for c in self.cases])
        phi_b = torch.empty(g.num_bnd_faces, dtype=torch.float64, device=dev)
        for c, sl in zip(self.cases, self.slices):
            phi_b[sl.bnd] = c.phi_b
        self.phi.single_data = phi_s
        self.phi.bnd_data = phi_b
        self.nu = torch.cat([torch.full((c.grid.num_cells,), c.nu, dtype=torch.float64, device=dev)
                             for c in self.cases])
        rows = [0]
        for c in self.cases:
            rows.append(rows[-1] + c.grid.num_cells)
        self.seg_row_host = rows
        self.seg_row = torch.tensor(rows, dtype=torch.int32, device=dev)
        self.row_seg = torch.repeat_interleave(torch.arange(len(self.cases), dtype=torch.int32, device=dev),
                                               torch.tensor([c.grid.num_cells for c in self.cases], device=dev))
        # adjustPhi per case: the case of every boundary face
        bseg = torch.empty(g.num_bnd_faces, dtype=torch.long, device=dev)
        for k, sl in enumerate(self.slices):
            bseg[sl.bnd] = k
        self.bnd_seg = bseg
        self.dom_rows = [0]
        for c in self.cases:
            self.dom_rows.append(self.dom_rows[-1] + c.grid.num_domain_bnd_faces)
        self.adjust = sim.fvSolution.adjustPhi and not any(isinstance(bc, DirichletBC) for bc in self.p.bcs.values())
        self.ref_cells = [(sl.cells[0] + c.ref[0], c.ref[1]) for c, sl in zip(self.cases, self.slices)
                          if c.ref is not None]
        sol = sim.fvSolution.solvers
        us = sol["U"]
        if us.method != SolverType.BiCGSTAB or us.norm_type != NormType.L_2:
            raise ValueError("the batched server solves momentum with the 2-norm BiCGSTAB")
        self.solver_U = SegmentedBiCGSTAB(us.tolerance, us.rel_tolerance, us.max_iter, us.max_restart,
                                          precond=us.preconditioner == PreconditionerType.JACOBI)
        ps = sol["p"]
        self.solver_p = SegmentedCG(ps.tolerance, ps.rel_tolerance, ps.max_iter,
                                    precond=ps.preconditioner == PreconditionerType.JACOBI)

    # ---- per-segment reductions -----------------------------------------------------
    def _seg_sum(self, v: torch.Tensor, bounds: Optional[List[int]] = None) -> torch.Tensor:
        """Sums of v over each case's contiguous range (cells by default):
        one reduction per case; atomics into a few slots would serialise."""
        b = self.seg_row_host if bounds is None else bounds
        return torch.stack([v[b[k]:b[k + 1]].sum() for k in range(len(b) - 1)])

    def initial_residuals(self, mat: FvMatrix, field: CellField) -> torch.Tensor:
        """residual.py:12-36 per case: |A psi - b|_2 / (|b|_2 + eps)."""
        r = mat.multiply(field.data) - mat.source
        C = self.grid.num_cells
        rr = (r * r).reshape(C, -1).sum(dim=1)
        bb = (mat.source * mat.source).reshape(C, -1).sum(dim=1)
        return torch.sqrt(self._seg_sum(rr)) / (torch.sqrt(self._seg_sum(bb)) + torch.finfo(torch.float64).eps)

    def _adjust_phi(self) -> None:
        """fv/adjust_phi.py:44-127 for every case at once (its own domain faces)."""
        g = self.grid
        B = g.num_domain_bnd_faces
        st = self.U.boundary_state()
        kind = st.adjust_kind[:B].long()
        ph = self.phi_hbya.bnd_data[:B]
        seg = self.bnd_seg[:B]
        z = torch.zeros_like(ph)

        def seg_sum(v):  # the domain faces of case k are contiguous
            return self._seg_sum(v, self.dom_rows)
        mass_in = seg_sum(torch.where((kind > 0) & (ph < 0), -ph, z))
        fixed_out = seg_sum(torch.where((kind == 1) & (ph > 0), ph, z))
        adj_out = seg_sum(torch.where((kind == 2) & (ph > 0), ph, z))
        ok = adj_out > 0
        corr = torch.where(ok, (mass_in - fixed_out) / torch.where(ok, adj_out, torch.ones_like(adj_out)),
                           torch.ones_like(adj_out))
        scale = torch.where((kind == 2) & (ph > 0) & ok[seg], corr[seg], torch.ones_like(ph))
        bnd = self.phi_hbya.bnd_data.clone()
        bnd[:B] = ph * scale
        self.phi_hbya.bnd_data = bnd

    def step(self) -> List[int]:
        """One SIMPLE iteration of every case; returns the indices of cases
        that met their residualControl in it (simple.py:241-285)."""
        g, U, p = self.grid, self.U, self.p
        C = g.num_cells
        UEqn = fvm.assemble(U, div=1.0, lap=-1.0, phi=self.phi, gamma=self.nu)
        nat.call("gf_relax_momentum", ctypes.c_int32(C), ctypes.c_double(self.alpha_U), nat.ptr(UEqn._diag),
                 nat.ptr(UEqn._source), nat.ptr(U.data), nat.stream())
        gp = fvc.grad(p)
        pred = FvMatrix.__new__(FvMatrix)
        pred.field, pred.grid = U, g
        pred._diag, pred._upper, pred._lower = UEqn._diag, UEqn._upper, UEqn._lower
        pred.face_flux_correction = None
        pred._source = torch.empty_like(UEqn._source)
        nat.call("gf_predictor_source", ctypes.c_int32(C), ctypes.c_int32(3), nat.ptr(UEqn._source),
                 nat.ptr(g.cell_volumes), nat.ptr(gp), nat.ptr(pred._source), nat.stream())
        u_eq = Equation(U, pred)
        res_U = self.initial_residuals(u_eq.fv_matrix, U) if "U" in self.rc else None
        U.data = self._solve_momentum(u_eq.fv_matrix)
        UEqn = Equation(U, UEqn).fv_matrix
        p_old = p.data.clone()
        nat.call("gf_rau_hbya", g.mesh_ref, nat.ptr(UEqn._diag), nat.ptr(UEqn._upper), nat.ptr(UEqn._lower),
                 nat.ptr(UEqn._source), nat.ptr(U.data), nat.ptr(self.rAU.data), nat.ptr(self.HbyA.data),
                 nat.stream())
        compute_phi_hbya(self.phi_hbya, self.HbyA, U)
        if self.adjust:
            self._adjust_phi()
        div = fvc.div(self.phi_hbya)
        mat = fvm.assemble(p, lap=-1.0, gamma=self.rAU.data)
        nat.call("gf_predictor_source", ctypes.c_int32(C), ctypes.c_int32(1), nat.ptr(mat._source),
                 nat.ptr(g.cell_volumes), nat.ptr(div), nat.ptr(mat._source), nat.stream())
        for cell, val in self.ref_cells:
            nat.call("gf_set_reference", nat.ptr(mat._diag), nat.ptr(mat._source), ctypes.c_int32(cell),
                     ctypes.c_double(val), nat.stream())
        eq = Equation(p, mat)
        res_p = self.initial_residuals(eq.fv_matrix, p) if "p" in self.rc else None
        A = eq.fv_matrix
        x, stats = self.solver_p.solve(A, A.source, p.data, self.seg_row, self.seg_row_host, self.row_seg)
        p.data = x
        correct_phi(self.phi, self.phi_hbya, mat, p, self.rAU)
        nat.call("gf_relax_field", ctypes.c_int64(C), ctypes.c_double(self.alpha_p), nat.ptr(p_old),
                 nat.ptr(p.data), nat.stream())
        apply_immersed_dirichlet_values(p)
        correct_velocity(U, self.HbyA, self.rAU, p)
        res_phi = None
        if "phi" in self.rc:
            d = fvc.div(self.phi)
            res_phi = torch.sqrt(self._seg_sum(d * d))
        # per-case bookkeeping (one host copy of the residuals)
        vals = {}
        for name, t in (("U", res_U), ("p", res_p), ("phi", res_phi)):
            if t is not None:
                vals[name] = t.tolist()
        converged = []
        for k, c in enumerate(self.cases):
            if c.done:
                continue
            c.iterations += 1
            c.p_iterations += stats.iterations[k]
            for name, v in vals.items():
                c.initial.setdefault(name, v[k])
                c.current[name] = v[k]
            if self._case_converged(c):
                converged.append(k)
        return converged

    def _solve_momentum(self, A: FvMatrix) -> torch.Tensor:
        """The segmented BiCGSTAB of every case's momentum rows."""
        x, _ = self.solver_U.solve(A, A.source, self.U.data, self.seg_row_host)
        return x

    def _case_converged(self, c: CaseState) -> bool:
        """simple.py:241-271: every residualControl entry below its absolute tolerance."""
        if not self.rc:
            return False
        for name, (tol, _rtol) in self.rc.items():
            if name not in c.current or name not in c.initial:
                return False
            if not residual_satisfied(c.current[name], tol, 0.0, c.initial[name]):
                return False
        return True

    def extract(self, k: int) -> None:
        """Copy case k's fields back into its own state."""
        c, sl = self.cases[k], self.slices[k]
        a, b = sl.cells
        c.U = self.U.data[a:b].clone()
        c.p = self.p.data[a:b].clone()
        f0, f1 = sl.faces
        c.phi_s = self.phi.single_data[f0:f1].clone()
        c.phi_b = self.phi.bnd_data[sl.bnd].clone()


# ------------------------------------------------------------------ server
@dataclass
class ServedResult:
    rid: int
    cells: int
    U: torch.Tensor
    p: torch.Tensor
    iterations: int
    p_iterations: int
    converged: bool
    arrival: float
    start: float
    end: float

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def base_cells(raw: dict) -> int:
    """Root cells of a request's grid (its size class before refinement)."""
    r = raw["fluxel"]["root_resolution"]
    return int(r[0]) * int(r[1]) * int(r[2])


class BatchServer:
    """Continuous batching of steady SIMPLE requests on one GPU.

    Requests arrive over time.  Two kinds of lanes share the device, each on
    its own stream:

    * the batch lane (the calling thread): between chunks of ``chunk``
      batched SIMPLE iterations it retires the cases that met their
      residualControl (or their iteration budget, ``endTime / deltaT``),
      admits waiting requests while the batch holds fewer than
      ``max_cases`` cases and its measured device footprint stays within
      ``max_bytes``, and rebuilds the combined mesh when its membership
      changed.  A case that converges mid-chunk is copied out at that
      iteration, so its result is the state at which it would have stopped
      running alone;
    * ``large_lanes`` alone lanes (threads): requests whose root grid has at
      least ``large_cells`` cells are bandwidth-bound on their own, so they
      run one at a time per lane through the SIMPLE driver with their own
      solver settings, concurrently with the batch.

    The footprint of a batch is its cells times the bytes per cell measured
    by ``calibrate`` (peak device allocation while one request is built,
    batched and iterated), done before the clock starts when ``max_bytes``
    is set."""

    def __init__(self, max_cases: int = 8, max_cells: int = 4_000_000, chunk: int = 10,
                 large_cells: Optional[int] = None, large_lanes: int = 2, max_bytes: Optional[int] = None):
        if max_cases < 1 or chunk < 1 or large_lanes < 0:
            raise ValueError("max_cases and chunk must be >= 1, large_lanes >= 0")
        if large_cells is not None and large_lanes == 0:
            raise ValueError("large requests need at least one alone lane")
        self.max_cases, self.max_cells, self.chunk = max_cases, max_cells, chunk
        self.large_cells, self.large_lanes = large_cells, large_lanes
        self.max_bytes = max_bytes
        self.bytes_per_cell: Optional[float] = None
        self._lane_error: Optional[BaseException] = None
        # called with each ServedResult as it completes (under the results lock)
        self.on_result: Optional[Callable[["ServedResult"], None]] = None

    def _is_large(self, raw: dict) -> bool:
        return self.large_cells is not None and base_cells(raw) >= self.large_cells

    def _alone_lane(self, queue, lock, results, t0) -> None:
        import time

        from .serving import solve_case

        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                while True:
                    with lock:
                        if not queue or self._lane_error is not None:
                            return
                        rid, raw, arr = queue.popleft()
                    wait = arr - (time.perf_counter() - t0)
                    if wait > 0:
                        time.sleep(wait)
                    start = time.perf_counter() - t0
                    cells, U, p, its, conv, p_its = solve_case(raw)
                    res = ServedResult(rid, cells, U, p, its, p_its, conv, arr, start, time.perf_counter() - t0)
                    with lock:
                        results.append(res)
                        if self.on_result is not None:
                            self.on_result(res)
        except BaseException as exc:  # surfaced by run()
            with lock:
                self._lane_error = exc

    def run(self, requests: Sequence[Tuple[int, dict, float]]) -> List[ServedResult]:
        """``requests``: (rid, raw case configuration, arrival offset in s)."""
        import collections
        import threading
        import time

        order = sorted(requests, key=lambda r: (r[2], r[0]))
        large = collections.deque(r for r in order if self._is_large(r[1]))
        small = [r for r in order if not self._is_large(r[1])]
        if self.max_bytes is not None and self.bytes_per_cell is None and small:
            self.calibrate(small[0][1])  # before the clock starts
        results: List[ServedResult] = []
        lock = threading.Lock()
        self._lane_error = None
        t0 = time.perf_counter()
        lanes = [threading.Thread(target=self._alone_lane, args=(large, lock, results, t0), daemon=True)
                 for _ in range(min(self.large_lanes, len(large)))]
        for th in lanes:
            th.start()
        try:
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                self._batch_lane(small, t0, results, lock)
            stream.synchronize()
        finally:
            for th in lanes:
                th.join()
        if self._lane_error is not None:
            raise RuntimeError(f"alone lane failed: {self._lane_error!r}") from self._lane_error
        return results

    def calibrate(self, raw: dict) -> float:
        """Device bytes per cell of a batch: the peak allocation while one
        request is built, batched and iterated once, over its cells."""
        from .config import load_config

        torch.cuda.synchronize()
        base = torch.cuda.memory_allocated()
        torch.cuda.reset_peak_memory_stats()
        case = new_case(-1, load_config(raw), raw)
        b = Batch([case])
        b.step()
        torch.cuda.synchronize()
        self.bytes_per_cell = (torch.cuda.max_memory_allocated() - base) / case.grid.num_cells
        return self.bytes_per_cell

    def _fits(self, cells: int) -> bool:
        if cells > self.max_cells:
            return False
        if self.max_bytes is not None and self.bytes_per_cell is not None:
            return cells * self.bytes_per_cell <= self.max_bytes
        return True

    def _batch_lane(self, order, t0, results, lock) -> None:
        import time

        from .config import load_config

        queue: List[Tuple[int, dict, float]] = []
        active: List[CaseState] = []
        meta: Dict[int, Tuple[float, float, int]] = {}  # rid -> (arrival, start, budget)
        nxt, done = 0, 0
        batch: Optional[Batch] = None
        while done < len(order):
            if self._lane_error is not None:
                return
            now = time.perf_counter() - t0
            while nxt < len(order) and order[nxt][2] <= now:
                queue.append(order[nxt])
                nxt += 1
            changed = False
            cells = sum(c.grid.num_cells for c in active)
            while queue and len(active) < self.max_cases:
                rid, raw, arr = queue[0]
                cfg = load_config(raw)
                case = new_case(rid, cfg, raw)
                if active and not self._fits(cells + case.grid.num_cells):
                    break
                queue.pop(0)
                ctl = cfg.simulator.control
                meta[rid] = (arr, time.perf_counter() - t0, max(1, int(round(ctl.endTime / ctl.deltaT))))
                active.append(case)
                cells += case.grid.num_cells
                changed = True
            if not active:
                if nxt < len(order):
                    time.sleep(max(0.0, min(0.05, order[nxt][2] - (time.perf_counter() - t0))))
                continue
            if changed or batch is None:
                if batch is not None:
                    for k, c in enumerate(batch.cases):
                        if not c.done:
                            batch.extract(k)
                batch = Batch(active)
            retired = False
            for _ in range(self.chunk):
                conv = set(batch.step())
                for k, c in enumerate(batch.cases):
                    if c.done:
                        continue
                    if k in conv or c.iterations >= meta[c.rid][2]:
                        batch.extract(k)
                        c.done = True
                        retired = True
                        arr, start, _ = meta[c.rid]
                        U, p = c.U.cpu(), c.p.cpu()  # synchronises this lane's stream
                        res = ServedResult(c.rid, c.grid.num_cells, U, p, c.iterations, c.p_iterations,
                                           k in conv, arr, start, time.perf_counter() - t0)
                        with lock:
                            results.append(res)
                            if self.on_result is not None:
                                self.on_result(res)
                        done += 1
                if all(c.done for c in batch.cases):
                    break
            if retired:
                active = [c for c in active if not c.done]
                batch = None if not active else batch
                if active:
                    for k, c in enumerate(batch.cases):
                        if not c.done:
                            batch.extract(k)
                    batch = Batch(active)
