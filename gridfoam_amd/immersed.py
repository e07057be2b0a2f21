"""Axis-projected immersed boundaries on the octree grid.

The reference treats an immersed face as two boundary faces and no internal
coupling (fv/fvm/laplacian.py:149-153 cuts its coefficient;
fv/boundary_ops.py:150-186 yields an "upper" batch on the owner side, with
outward area +Sf and distance ap_dist_owner_to_bnd, and a "lower" batch on
the neighbour side with -Sf and ap_dist_neighbour_to_bnd).  Here that is a
transformation of the mesh before the operator is built: the internal face
list keeps the single-sided faces in ascending order (the reference's
``single_idx``, fv/kernels/face_geometry.py:128-131) and the boundary list
becomes [domain faces | upper sides | lower sides], each side with its ray
distance, so every kernel that handles boundary faces handles immersed ones.

The payload (cut flags, ray distances, near-boundary flags, patch ids) comes
from ``gf_ap_cut`` for analytic bodies (config ``fluxel.immersed_bodies``).
Snapped Dirichlet cells (Gibou et al., data_contracts.rst:181-231) are
handled by ``immersed_dirichlet_constraints`` below and the elimination of
``FvMatrix.with_fixed_values``.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import List, Sequence

import torch

from . import _native as nat
from .config import ImmersedBodyConfig

N_DOMAIN_PATCHES = 6  # boundary patch ids 0..5 are the domain directions; bodies follow


def shapes_from_config(bodies: Sequence[ImmersedBodyConfig]):
    """gf_shape array for gf_ap_cut; body k has patch id k."""
    arr = (nat.GfShape * max(len(bodies), 1))()
    for k, b in enumerate(bodies):
        s = arr[k]
        s.patch = k
        if b.kind == "plane":
            s.kind, s.axis = 0, int(b.axis)
            s.p[0] = float(b.offset)
        elif b.kind == "box":
            s.kind = 1
            for j in range(3):
                s.p[j], s.p[3 + j] = float(b.min[j]), float(b.max[j])
        elif b.kind == "bluff":
            s.kind, s.axis = 3, len(b.struts)
            for j in range(3):
                s.p[j], s.p[3 + j] = float(b.min[j]), float(b.max[j])
            s.p[6] = float(b.radius)
            s.p[7], s.p[8] = math.sin(math.radians(b.slant_angle)), math.cos(math.radians(b.slant_angle))
            s.p[9], s.p[10], s.p[11] = float(b.slant_length), float(b.strut_radius), float(b.strut_height)
            for k1, (x, y) in enumerate(b.struts):
                s.p[12 + 2 * k1], s.p[13 + 2 * k1] = float(x), float(y)
        else:
            s.kind = 2
            s.p[0], s.p[1], s.p[2] = float(b.center[0]), float(b.center[1]), float(b.radius)
            for k1, v in enumerate(b.a):
                s.p[3 + k1] = float(v)
            for k1, v in enumerate(b.b):
                s.p[11 + k1] = float(v)
    return arr


@dataclass
class APCut:
    is_immersed: torch.Tensor  # [F] bool
    dist_owner: torch.Tensor   # [F_i]
    dist_neighbour: torch.Tensor
    owner_near: torch.Tensor   # [F_i] bool
    neighbour_near: torch.Tensor
    owner_patch: torch.Tensor  # [F_i] body index
    neighbour_patch: torch.Tensor
    owner_normal: torch.Tensor  # [F_i, 3] body-outward unit normal at the owner-side hit
    neighbour_normal: torch.Tensor


def ap_cut(owner, neighbour, axis, centers, sizes, face_centers, bodies: Sequence[ImmersedBodyConfig],
           ref_length: float) -> APCut:
    """Run gf_ap_cut over every internal face."""
    F = int(owner.shape[0])
    dev = owner.device
    imm = torch.zeros(max(F, 1), dtype=torch.uint8, device=dev)
    do = torch.zeros(max(F, 1), dtype=torch.float64, device=dev)
    dn = torch.zeros_like(do)
    no = torch.zeros_like(imm)
    nn = torch.zeros_like(imm)
    po = torch.zeros(max(F, 1), dtype=torch.int32, device=dev)
    pn = torch.zeros_like(po)
    nrm_o = torch.zeros((max(F, 1), 3), dtype=torch.float64, device=dev)
    nrm_n = torch.zeros_like(nrm_o)
    shapes = shapes_from_config(bodies)
    nat.call("gf_ap_cut", ctypes.c_int32(F), *map(nat.ptr, (owner, neighbour, axis, centers, sizes, face_centers)),
             ctypes.c_int32(len(bodies)), shapes, ctypes.c_double(ref_length),
             *map(nat.ptr, (imm, do, dn, no, nn, po, pn, nrm_o, nrm_n)), nat.stream())
    cut = imm[:F].bool()
    return APCut(cut, do[:F][cut].contiguous(), dn[:F][cut].contiguous(), no[:F][cut].bool(), nn[:F][cut].bool(),
                 po[:F][cut].long(), pn[:F][cut].long(), nrm_o[:F][cut].contiguous(), nrm_n[:F][cut].contiguous())


def _snap_candidates(field, st) -> torch.Tensor:
    """Boundary faces that may carry a cell constraint: snapped immersed
    sides, covered by a condition other than Neumann, with fraction 1
    (boundary_ops.py:388-402)."""
    g = field.grid
    B = g.num_bnd_faces
    from .fields import NeumannBC  # noqa: PLC0415 (fields imports this module)

    neumann = torch.zeros(B, dtype=torch.bool, device=g.device)
    for patch, bc in field.bcs.items():
        if type(bc) is NeumannBC:
            neumann |= g.bnd_patch_mask(patch)
    return g.bnd_snap[:B].bool() & st.covered[:B].bool() & (st.frac[:B] == 1.0) & ~neumann


def _batch_order_key(field, idx: torch.Tensor) -> torch.Tensor:
    """Position of boundary faces ``idx`` (immersed sides) in the reference's
    batch concatenation: conditions in the field's order, a patch's upper
    sides before its lower sides, face order within a side."""
    g = field.grid
    B_d, F_i = g.num_domain_bnd_faces, g.num_immersed_faces
    n_pid = N_DOMAIN_PATCHES + len(g.immersed_bodies)
    pos = torch.full((n_pid,), n_pid, dtype=torch.long, device=g.device)
    for i, patch in enumerate(field.bcs):
        pid = g.patch_id(patch)
        if pid >= 0:
            pos[pid] = i
    side = (idx - B_d) // F_i
    within = (idx - B_d) % F_i
    return (pos[g.bnd_patch_id[idx]] * 2 + side) * F_i + within


def _select_constraints(field, st) -> tuple[torch.Tensor, torch.Tensor]:
    """(cells, chosen boundary face per cell): the closest snapped face, ties
    in batch order (boundary_ops.py:480-496)."""
    g = field.grid
    idx = torch.nonzero(_snap_candidates(field, st)).flatten()
    if idx.numel() == 0:
        empty = torch.empty(0, dtype=torch.long, device=g.device)
        return empty, empty
    cell = g.bnd_owner[idx]
    dist = g.bnd_mag_d[idx]
    closest = torch.full((g.num_cells,), float("inf"), dtype=torch.float64, device=g.device)
    closest.scatter_reduce_(0, cell, dist, reduce="amin")
    key = _batch_order_key(field, idx)
    none = torch.iinfo(torch.long).max
    cand = torch.where(dist == closest[cell], key, torch.full_like(key, none))
    sel = torch.full((g.num_cells,), none, dtype=torch.long, device=g.device)
    sel.scatter_reduce_(0, cell, cand, reduce="amin")
    cells = torch.nonzero(sel < none).flatten()
    order = torch.argsort(key)
    pick = idx[order[torch.searchsorted(key[order], sel[cells])]]
    return cells, pick


def immersed_dirichlet_constraints(field) -> tuple[torch.Tensor, torch.Tensor]:
    """fv/boundary_ops.py:440-535: cells fixed by Gibou snapping and their
    values.  A snapped face is an immersed side within the Gibou distance
    whose condition is not Neumann and has fraction 1 there; several faces
    on one cell: the closest one wins, ties in the reference's batch order.
    The selection is kept while every immersed condition is a fixed value
    (it depends on geometry only, as the reference caches it); values are
    re-read on every call (autograd-connected when they require grad)."""
    g = field.grid
    if g.num_immersed_faces == 0:
        return torch.empty(0, dtype=torch.long, device=g.device), field.data[:0]
    st = field.boundary_state()
    from .fields import DirichletBC, NeumannBC  # noqa: PLC0415

    key = tuple((patch, id(bc)) for patch, bc in field.bcs.items())
    static = all(g.patch_id(patch) < N_DOMAIN_PATCHES or type(bc) in (DirichletBC, NeumannBC)
                 for patch, bc in field.bcs.items())
    cache = getattr(field, "_ib_constraints", None)
    if static and cache is not None and cache[0] == key:
        cells, pick = cache[1], cache[2]
    else:
        cells, pick = _select_constraints(field, st)
        field._ib_constraints = (key, cells, pick) if static else None
    if cells.numel() == 0:
        return cells, field.data[:0]
    rv = st.ref_v_ad if st.ref_v_ad is not None else st.ref_v
    vals = rv[pick].reshape((cells.numel(), *field.component_shape))
    return cells, vals


def apply_immersed_dirichlet_values(field) -> None:
    """fv/boundary_ops.py:537-541: restore the prescribed values of snapped
    cells after an explicit update."""
    if field.grid.num_immersed_faces == 0:
        return
    cells, values = immersed_dirichlet_constraints(field)
    if cells.numel():
        field.data = field.data.index_copy(0, cells, values.to(field.data.dtype))


def det_index_add(base: torch.Tensor, idx: torch.Tensor, vals: torch.Tensor) -> torch.Tensor:
    """``base.index_add(0, idx, vals)`` with a fixed summation order: the
    entries aimed at one row are added in their order in ``idx``, one per
    launch, so no two threads of a launch add into the same row (float
    atomics would make the sum depend on scheduling, e.g. on work running
    concurrently on another stream).  Meant for sparse contributions: one
    launch per multiplicity level."""
    if idx.numel() == 0:
        return base
    order = torch.argsort(idx, stable=True)
    si = idx[order]
    rank = torch.arange(si.numel(), device=si.device) - torch.searchsorted(si, si)
    out = base
    for r in range(int(rank.max()) + 1):
        sel = order[rank == r]
        out = out.index_add(0, idx[sel], vals[sel])
    return out


def close_fixed_cell_flux(phi, p) -> None:
    """pressure_correction.py:308-335: on cells fixed by snapping the
    Poisson row no longer balances the flux; the cell's imbalance is taken
    off its snapped Dirichlet faces in proportion to their area."""
    g = p.grid
    if g.num_immersed_faces == 0:
        return
    cells, _ = immersed_dirichlet_constraints(p)
    if cells.numel() == 0:
        return
    st = p.boundary_state()
    B = g.num_bnd_faces
    from .fv import fvc  # noqa: PLC0415

    fixed = _snap_candidates(p, st)
    area = torch.where(fixed, g.bnd_mag_sf[:B], torch.zeros_like(g.bnd_mag_sf[:B]))
    owner = g.bnd_owner[:B]
    snapped = torch.nonzero(fixed).flatten()
    area_sum = det_index_add(torch.zeros(g.num_cells, dtype=torch.float64, device=g.device), owner[snapped],
                             area[snapped])
    imbalance = fvc.div(phi) * g.cell_volumes[: g.num_cells]
    total = area_sum[owner]
    frac = area / torch.where(total > 0, total, torch.ones_like(total))
    phi.bnd_data = phi.bnd_data - frac * imbalance[owner]


def check_zero_distance(field, st) -> None:
    """boundary_ops.py:414-420: a mixed condition (not a snapped fixed
    value, not pure gradient) on a boundary face at zero distance cannot be
    evaluated."""
    g = field.grid
    idx = g.bnd_zero_dist
    if idx.numel() == 0:
        return
    f = st.frac[idx]
    active = st.covered[idx].bool() & (f != 0) & ~(g.bnd_snap[idx].bool() & (f == 1))
    if bool(active.any()):
        raise ValueError("A mixed boundary at zero distance needs a cell constraint")


def patch_ids(bodies: Sequence[ImmersedBodyConfig]) -> dict:
    return {b.name: N_DOMAIN_PATCHES + k for k, b in enumerate(bodies)}


def names(bodies: Sequence[ImmersedBodyConfig]) -> List[str]:
    return [b.name for b in bodies]
