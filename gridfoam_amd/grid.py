"""Octree grid on the device: topology, geometry and the face-neighbour operator.

``OctreeGrid`` exposes the GridBase contract of the reference
(src/gridfoam/core/grid/base.py:19-275, axis_projected.py:521-625) --
owner/neighbour/axis, domain faces, cell centres/sizes/volumes, face
centres, Sf -- built entirely by the HIP library:

1. ``gf_octree_build`` (Morton-ordered leaves, surface and region refinement, 2:1 face
   balance, face enumeration) replaces fluxel-cfd (core/grid/factory.py:32-65);
2. ``gf_grid_geometry`` restates axis_projected.py:130-161, :715-811;
3. ``gf_face_geometry`` / ``gf_lsq_ata_inv`` / ``gf_hang_geometry`` restate
   fv/kernels/face_geometry.py:107-262;
4. ``gf_ap_cut`` cuts the internal faces by analytic immersed bodies
   (immersed.py): a cut face becomes two boundary faces, so the internal
   list holds the single-sided faces and the boundary list is
   [domain | immersed upper sides | immersed lower sides];
5. the per-cell CSR face lists, the slice-of-64 ELL operator and the PCG
   solver layout are the device layout the assembly and Krylov kernels read.
``num_internal_faces`` counts the single-sided faces (the LDU arrays cover
them); ``num_all_internal_faces`` and ``ap_is_immersed_faces`` describe the
faces before the cut.
"""

from __future__ import annotations

import ctypes
import weakref
from functools import cached_property

import torch

from . import _native as nat
from .config import FluxelConfig, GridfoamConfig
from .solver_layout import build_solver_layout


def _i32(n: int, dev) -> torch.Tensor:
    return torch.zeros(max(n, 1), dtype=torch.int32, device=dev)


def _f64(shape, dev) -> torch.Tensor:
    if isinstance(shape, int):
        shape = (max(shape, 1),)
    return torch.zeros(shape, dtype=torch.float64, device=dev)


def _scan(cnt: torch.Tensor, n: int) -> torch.Tensor:
    out = torch.zeros_like(cnt)
    nat.call("gf_exclusive_scan_i32", nat.ptr(cnt), nat.ptr(out), ctypes.c_int32(n), nat.stream())
    return out


class OctreeGrid:
    """Device-resident octree grid (no immersed boundary)."""

    # the active FourierShape (gridfoam_amd.shape) while shape derivatives are taken
    shape = None

    def __init__(self, fluxel: FluxelConfig, *, dt: float = 1.0, device: str | torch.device = "cuda",
                 topology: dict | None = None):
        """Build the octree on the GPU from the fluxel block, or adopt a
        given integer topology (``topology``: level [C], ijk [C,3], owner,
        neighbour, axis [F], bnd_owner, bnd_dir [B]; e.g. a fluxel mesh),
        whose faces may come in any order: they are brought into the
        face-order contract by a stable sort by owner, and ``face_order`` /
        ``bnd_face_order`` record the given index of every face."""
        nat.require_gpu()
        self.device = torch.device(device)
        self.dtype = torch.float64
        self.dt = float(dt)
        self.fluxel = fluxel
        self.sim_config = None
        # field registry by name (axis_projected.py:78-79, :370-386): fields
        # register on construction; conditions look up the fields they read
        self._cellfields: "weakref.WeakValueDictionary" = weakref.WeakValueDictionary()
        self._facefields: "weakref.WeakValueDictionary" = weakref.WeakValueDictionary()
        lo = (ctypes.c_double * 3)(*fluxel.domain.lower)
        hi = (ctypes.c_double * 3)(*fluxel.domain.upper)
        root = (ctypes.c_int32 * 3)(*fluxel.root_resolution)
        if topology is None:
            self._build_topology(fluxel, lo, hi, root)
        else:
            self._adopt_topology(topology)
        self._build_geometry(lo, hi, root)
        self._apply_immersed(list(fluxel.immersed_bodies))
        self._build_operator()

    def _build_topology(self, fluxel: FluxelConfig, lo, hi, root) -> None:
        dev = self.device
        regs = fluxel.refinement_regions
        nr = len(regs)
        rmin = (ctypes.c_double * (3 * nr + 3))(*[v for r in regs for v in r.min])
        rmax = (ctypes.c_double * (3 * nr + 3))(*[v for r in regs for v in r.max])
        rlev = (ctypes.c_int32 * (nr + 1))(*[r.level for r in regs])
        handle = ctypes.c_void_p()
        counts = (ctypes.c_int64 * 4)()
        bodies = list(fluxel.immersed_bodies)
        if fluxel.target_level > 0 and bodies:  # refine to target_level around the bodies' surfaces
            from .immersed import shapes_from_config

            nat.call("gf_octree_build_surface", lo, hi, root, ctypes.c_int32(nr), rmin, rmax, rlev,
                     ctypes.c_int32(len(bodies)), shapes_from_config(bodies), ctypes.c_int32(fluxel.target_level),
                     ctypes.c_int32(fluxel.n_leaf_refinement), ctypes.byref(handle), counts, nat.stream())
        else:
            nat.call("gf_octree_build", lo, hi, root, ctypes.c_int32(nr), rmin, rmax, rlev,
                     ctypes.c_int32(fluxel.n_leaf_refinement), ctypes.byref(handle), counts, nat.stream())
        C, F, B, lmax = (int(v) for v in counts)
        self.num_cells, self.num_internal_faces, self.num_domain_bnd_faces, self.lmax = C, F, B, lmax
        self.level = _i32(C, dev)
        self.ijk = torch.zeros((max(C, 1), 3), dtype=torch.int32, device=dev)
        self.owner_i32 = _i32(F, dev)
        self.neighbour_i32 = _i32(F, dev)
        self.axis_i32 = _i32(F, dev)
        self.bnd_owner_i32 = _i32(B, dev)
        self.bnd_dir_i32 = _i32(B, dev)
        try:
            nat.call("gf_octree_fetch", handle, *map(nat.ptr, (
                self.level, self.ijk, self.owner_i32, self.neighbour_i32, self.axis_i32,
                self.bnd_owner_i32, self.bnd_dir_i32)), nat.stream())
        finally:
            nat.lib().gf_octree_free(handle)
        self.face_order = None
        self.bnd_face_order = None

    def _adopt_topology(self, t: dict) -> None:
        dev = self.device

        def i32(x, n, shape=None):
            x = torch.as_tensor(x).to(device=dev, dtype=torch.int32)
            out = _i32(n, dev) if shape is None else torch.zeros(shape, dtype=torch.int32, device=dev)
            out[: x.shape[0]] = x
            return out

        level = torch.as_tensor(t["level"])
        C, F, B = int(level.shape[0]), int(len(t["owner"])), int(len(t["bnd_owner"]))
        self.num_cells, self.num_internal_faces, self.num_domain_bnd_faces = C, F, B
        self.lmax = int(level.max()) if C else 0
        order, bnd_order = canonical_face_order(torch.as_tensor(t["owner"]), torch.as_tensor(t["bnd_owner"]))
        self.face_order, self.bnd_face_order = order.to(dev), bnd_order.to(dev)
        self.level = i32(level, C)
        self.ijk = i32(t["ijk"], C, (max(C, 1), 3))
        self.owner_i32 = i32(torch.as_tensor(t["owner"])[order], F)
        self.neighbour_i32 = i32(torch.as_tensor(t["neighbour"])[order], F)
        self.axis_i32 = i32(torch.as_tensor(t["axis"])[order], F)
        self.bnd_owner_i32 = i32(torch.as_tensor(t["bnd_owner"])[bnd_order], B)
        self.bnd_dir_i32 = i32(torch.as_tensor(t["bnd_dir"])[bnd_order], B)

    # ------------------------------------------------------------------ build
    def _build_geometry(self, lo, hi, root) -> None:
        C, F, B, dev = self.num_cells, self.num_internal_faces, self.num_domain_bnd_faces, self.device
        self.cell_centers = _f64((C, 3), dev)
        self.cell_sizes = _f64((C, 3), dev)
        self.cell_volumes = _f64(C, dev)
        self.face_centers = _f64((max(F, 1), 3), dev)
        self.Sf = _f64((max(F, 1), 3), dev)
        self.domain_bnd_face_centers = _f64((max(B, 1), 3), dev)
        self.domain_bnd_Sf = _f64((max(B, 1), 3), dev)
        nat.call("gf_grid_geometry", ctypes.c_int32(C), ctypes.c_int32(F), ctypes.c_int32(B), lo, hi, root,
                 *map(nat.ptr, (self.level, self.ijk, self.owner_i32, self.neighbour_i32, self.axis_i32,
                                self.bnd_owner_i32, self.bnd_dir_i32, self.cell_centers, self.cell_sizes,
                                self.cell_volumes, self.face_centers, self.Sf, self.domain_bnd_face_centers,
                                self.domain_bnd_Sf)), nat.stream())
        # boundary vectors from the owner centre (boundary_ops.py:128-146)
        cb = self.cell_centers[self.bnd_owner_i32.long()]
        self.bnd_d = (self.domain_bnd_face_centers[:B] - cb[:B]).contiguous()
        if B and bool((torch.linalg.vector_norm(self.bnd_d, dim=1) == 0).any()):
            raise ValueError("zero cell-to-boundary distance")

    def _apply_immersed(self, bodies: list) -> None:
        """Cut the internal faces by the immersed bodies (gf_ap_cut) and turn
        every cut face into two boundary faces: the internal list keeps the
        single-sided faces in ascending order, the boundary list becomes
        [domain | upper sides (owner, +Sf) | lower sides (neighbour, -Sf)]
        (fv/boundary_ops.py:150-186, fv/kernels/face_geometry.py:128-131,
        boundary_lsq_vectors :198-243)."""
        from .immersed import N_DOMAIN_PATCHES, ap_cut, patch_ids

        dev = self.device
        F_all, B_d = self.num_internal_faces, self.num_domain_bnd_faces
        self.num_all_internal_faces = F_all
        self.immersed_bodies = bodies
        self.patch_name_to_id = patch_ids(bodies)
        dom_dir = self.bnd_dir_i32[:B_d].long()
        bnd_sf = self.domain_bnd_Sf[:B_d]
        self.ap_is_immersed_faces = torch.zeros(F_all, dtype=torch.bool, device=dev)
        if bodies and F_all:
            lo, hi = self.fluxel.domain.lower, self.fluxel.domain.upper
            L = max(h - l for l, h in zip(lo, hi))
            cut = ap_cut(self.owner_i32[:F_all], self.neighbour_i32[:F_all], self.axis_i32[:F_all],
                         self.cell_centers, self.cell_sizes, self.face_centers[:F_all], bodies, L)
            self.ap_is_immersed_faces = cut.is_immersed
        F_i = int(self.ap_is_immersed_faces.sum())
        self.num_immersed_faces = F_i
        if F_i == 0:
            self.bnd_patch_id = dom_dir.contiguous()
            self.bnd_mag_d = torch.linalg.vector_norm(self.bnd_d, dim=1).contiguous()
            self.bnd_mag_sf = torch.linalg.vector_norm(bnd_sf, dim=1).contiguous()
            self.bnd_sign = (2.0 * (dom_dir % 2).double() - 1.0).contiguous()
            self.bnd_Sf = bnd_sf.contiguous()
            self.bnd_snap = torch.zeros(max(B_d, 1), dtype=torch.uint8, device=dev)
            self.bnd_zero_dist = torch.zeros(0, dtype=torch.long, device=dev)
            self.num_bnd_faces = B_d
            self.ap_dist_owner_to_bnd = self.ap_dist_neighbour_to_bnd = torch.zeros(0, dtype=torch.float64,
                                                                                    device=dev)
            self.ap_owner_near_boundary = self.ap_neighbour_near_boundary = torch.zeros(0, dtype=torch.bool,
                                                                                      device=dev)
            self.ap_owner_bnd_patch_id = self.ap_neighbour_bnd_patch_id = torch.zeros(0, dtype=torch.long,
                                                                                    device=dev)
            self.ap_owner_normal = self.ap_neighbour_normal = torch.zeros((0, 3), dtype=torch.float64, device=dev)
            self.ap_face_centers = torch.zeros((0, 3), dtype=torch.float64, device=dev)
            return
        imm = self.ap_is_immersed_faces
        single = ~imm
        self.ap_dist_owner_to_bnd, self.ap_dist_neighbour_to_bnd = cut.dist_owner, cut.dist_neighbour
        self.ap_owner_near_boundary, self.ap_neighbour_near_boundary = cut.owner_near, cut.neighbour_near
        self.ap_owner_bnd_patch_id, self.ap_neighbour_bnd_patch_id = cut.owner_patch, cut.neighbour_patch
        self.ap_owner_normal, self.ap_neighbour_normal = cut.owner_normal, cut.neighbour_normal
        o_all, n_all, a_all = (t[:F_all].long() for t in (self.owner_i32, self.neighbour_i32, self.axis_i32))
        io, inb, ia = o_all[imm], n_all[imm], a_all[imm]
        self.ap_face_centers = self.face_centers[:F_all][imm].contiguous()  # the axis rays' tangential position
        Sf_i = self.Sf[:F_all][imm]
        mag_sf_i = torch.linalg.vector_norm(Sf_i, dim=1)
        e = torch.zeros((F_i, 3), dtype=torch.float64, device=dev)
        e[torch.arange(F_i, device=dev), ia] = 1.0
        # single-sided internal faces (order kept: still grouped by owner)
        F_s = F_all - F_i
        self.num_internal_faces = F_s

        def keep_single(t, shape):
            out = torch.zeros(shape, dtype=t.dtype, device=dev)
            out[:F_s] = t[:F_all][single]
            return out

        self.owner_i32 = keep_single(self.owner_i32, max(F_s, 1))
        self.neighbour_i32 = keep_single(self.neighbour_i32, max(F_s, 1))
        self.axis_i32 = keep_single(self.axis_i32, max(F_s, 1))
        self.face_centers = keep_single(self.face_centers, (max(F_s, 1), 3))
        self.Sf = keep_single(self.Sf, (max(F_s, 1), 3))
        # boundary list [domain | upper | lower]
        d_up = cut.dist_owner[:, None] * e
        d_lo = -cut.dist_neighbour[:, None] * e
        B = B_d + 2 * F_i
        self.num_bnd_faces = B
        self.bnd_owner_i32 = torch.cat([self.bnd_owner_i32[:B_d], io.int(), inb.int()]).contiguous()
        self.bnd_dir_i32 = torch.cat([self.bnd_dir_i32[:B_d], (2 * ia + 1).int(), (2 * ia).int()]).contiguous()
        self.bnd_patch_id = torch.cat([dom_dir, N_DOMAIN_PATCHES + cut.owner_patch,
                                       N_DOMAIN_PATCHES + cut.neighbour_patch]).contiguous()
        self.bnd_d = torch.cat([self.bnd_d[:B_d], d_up, d_lo]).contiguous()
        self.bnd_mag_d = torch.cat([torch.linalg.vector_norm(self.bnd_d[:B_d], dim=1), cut.dist_owner,
                                    cut.dist_neighbour]).contiguous()
        self.bnd_mag_sf = torch.cat([torch.linalg.vector_norm(bnd_sf, dim=1), mag_sf_i, mag_sf_i]).contiguous()
        self.bnd_sign = torch.cat([2.0 * (dom_dir % 2).double() - 1.0,
                                   torch.ones(F_i, dtype=torch.float64, device=dev),
                                   -torch.ones(F_i, dtype=torch.float64, device=dev)]).contiguous()
        self.bnd_Sf = torch.cat([bnd_sf, Sf_i, -Sf_i]).contiguous()
        self.bnd_snap = torch.cat([torch.zeros(B_d, dtype=torch.uint8, device=dev),
                                   cut.owner_near.to(torch.uint8), cut.neighbour_near.to(torch.uint8)]).contiguous()
        self.bnd_zero_dist = torch.nonzero(self.bnd_mag_d == 0).flatten()

    def _build_operator(self) -> None:
        C, F, B, dev = self.num_cells, self.num_internal_faces, self.num_bnd_faces, self.device
        m = nat.GfMesh()
        self._mesh = m
        self._keep: list[torch.Tensor] = []
        m.n_cells, m.n_faces, m.n_bnd = C, F, B
        m.owner, m.neighbour, m.axis = (nat.ptr(t) for t in (self.owner_i32, self.neighbour_i32, self.axis_i32))
        m.bnd_owner, m.bnd_dir = nat.ptr(self.bnd_owner_i32), nat.ptr(self.bnd_dir_i32)
        m.centers, m.sizes, m.volumes = (nat.ptr(t) for t in (self.cell_centers, self.cell_sizes, self.cell_volumes))
        m.bnd_d, m.bnd_mag_d = nat.ptr(self.bnd_d), nat.ptr(self.bnd_mag_d)
        m.bnd_mag_sf, m.bnd_sign = nat.ptr(self.bnd_mag_sf), nat.ptr(self.bnd_sign)
        mp = ctypes.byref(m)
        st = nat.stream()
        # per-cell CSR of incident faces
        own_cnt, nbr_cnt, bnd_cnt = _i32(C + 1, dev), _i32(C + 1, dev), _i32(C + 1, dev)
        nat.call("gf_row_counts", mp, nat.ptr(own_cnt), nat.ptr(nbr_cnt), nat.ptr(bnd_cnt), st)
        self.own_row = _scan(own_cnt, C + 1)
        self.nbr_row = _scan(nbr_cnt, C + 1)
        self.bnd_row = _scan(bnd_cnt, C + 1)
        m.own_row, m.nbr_row, m.bnd_row = (nat.ptr(t) for t in (self.own_row, self.nbr_row, self.bnd_row))
        self.bnd_face = _i32(B, dev)
        m.bnd_face = nat.ptr(self.bnd_face)
        nat.call("gf_fill_bnd_lists", mp, nat.ptr(self.bnd_face), st)
        m.bnd_snap = nat.ptr(self.bnd_snap)
        flags = ctypes.c_int32(0)
        nat.call("gf_mesh_validate", mp, ctypes.byref(flags), st)  # raises on a face-order violation
        self.nbr_face = _i32(F, dev)
        m.nbr_face = nat.ptr(self.nbr_face)
        nat.call("gf_fill_nbr_lists", mp, nat.ptr(self.nbr_face), st)
        # face weights, delta coefficients, hanging flags
        self.w = _f64(F, dev)
        self.mag_Sf = _f64(F, dev)
        self.delta_coeffs = _f64(F, dev)
        is_hang = torch.zeros(max(F, 1), dtype=torch.uint8, device=dev)
        h_min = float(self.cell_sizes.min())
        nat.call("gf_face_geometry", mp, nat.ptr(self.face_centers), nat.ptr(self.w), nat.ptr(self.mag_Sf),
                 nat.ptr(self.delta_coeffs), nat.ptr(is_hang), ctypes.c_double(h_min), st)
        m.w, m.mag_sf, m.delta = nat.ptr(self.w), nat.ptr(self.mag_Sf), nat.ptr(self.delta_coeffs)
        self.hang_idx = torch.nonzero(is_hang[:F]).flatten().to(torch.int32).contiguous()
        H = int(self.hang_idx.numel())
        m.n_hang = H
        self.hang_face = self.hang_idx if H else _i32(1, dev)
        m.hang_face = nat.ptr(self.hang_face)
        self.hang_offset = _f64((max(H, 1), 3), dev)
        self.hang_move_O = _f64((max(H, 1), 3), dev)
        self.hang_move_N = _f64((max(H, 1), 3), dev)
        nat.call("gf_hang_geometry", mp, nat.ptr(self.face_centers), nat.ptr(self.hang_offset),
                 nat.ptr(self.hang_move_O), nat.ptr(self.hang_move_N), st)
        m.hang_offset, m.hang_move_o, m.hang_move_n = (
            nat.ptr(t) for t in (self.hang_offset, self.hang_move_O, self.hang_move_N))
        # cells next to hanging faces and their hanging-face lists
        mark = torch.zeros(C, dtype=torch.bool, device=dev)
        if H:
            hf = self.hang_idx.long()
            mark[self.owner_i32.long()[hf]] = True
            mark[self.neighbour_i32.long()[hf]] = True
        self.hang_cell_mask = mark
        hang_cell = torch.nonzero(mark).flatten().to(torch.int32).contiguous()
        HC = int(hang_cell.numel())
        self.hang_cell = hang_cell if HC else _i32(1, dev)
        cell_to_hc = torch.full((C,), -1, dtype=torch.int32, device=dev)
        if HC:
            cell_to_hc[hang_cell.long()] = torch.arange(HC, dtype=torch.int32, device=dev)
        hc_cnt = _i32(HC + 1, dev)
        if H:
            hf = self.hang_idx.long()
            ends = torch.cat([cell_to_hc[self.owner_i32.long()[hf]], cell_to_hc[self.neighbour_i32.long()[hf]]])
            hc_cnt[:HC] = torch.bincount(ends.long(), minlength=HC).to(torch.int32)
        self.hc_row = _scan(hc_cnt, HC + 1)
        self.hc_code = _i32(2 * H, dev)
        m.n_hang_cells = HC
        m.hang_cell, m.hc_row, m.hc_code = (nat.ptr(t) for t in (self.hang_cell, self.hc_row, self.hc_code))
        nat.call("gf_hang_cell_lists", mp, nat.ptr(cell_to_hc), nat.ptr(self.hc_code), st)
        # geometric least-squares normal-matrix inverses
        self.lsq_ata_inv = _f64((C, 9), dev)
        nat.call("gf_lsq_ata_inv", mp, nat.ptr(self.lsq_ata_inv), st)
        m.ata_inv = nat.ptr(self.lsq_ata_inv)
        # slice-of-64 ELL operator
        ns = (C + 63) // 64
        m.n_slices = ns
        slice_len = _i32(ns + 1, dev)
        slice_len[ns:] = 0
        nat.call("gf_sell_widths", mp, nat.ptr(slice_len), st)
        self.sell_ptr = _scan(slice_len, ns + 1)
        nnz = int(self.sell_ptr[ns].item())
        m.nnz_pad = nnz
        self.sell_col = _i32(nnz, dev)
        self.sell_code = _i32(nnz, dev)
        m.sell_ptr, m.sell_col, m.sell_code = (nat.ptr(t) for t in (self.sell_ptr, self.sell_col, self.sell_code))
        nat.call("gf_sell_fill", mp, nat.ptr(self.sell_col), nat.ptr(self.sell_code), st)
        self._build_solver_layout()
        torch.cuda.current_stream().synchronize()

    def _build_solver_layout(self) -> None:
        """Solver row order and brick layout of the symmetric PCG operator
        (solver_layout.py), kept on the device with the rest of the mesh."""
        F = self.num_internal_faces
        lay = build_solver_layout(self.level[: self.num_cells], self.ijk[: self.num_cells], self.owner_i32[:F],
                                  self.neighbour_i32[:F], self.axis_i32[:F], list(self.fluxel.root_resolution))
        self.solver_layout = lay
        m = self._mesh
        m.n_bricks, m.n_expl, m.pnnz_pad = lay.n_bricks, lay.n_expl, int(lay.psell_col.numel())
        keep = lambda t: t if t.numel() else torch.zeros(1, dtype=torch.int32, device=self.device)  # noqa: E731
        self._layout_keep = [keep(t) for t in (lay.row_perm, lay.row_iperm, lay.brick_nbr, lay.brick_vp, lay.brick_vm,
                                               lay.expl_col, lay.expl_code, lay.psell_ptr, lay.psell_col,
                                               lay.psell_code)]
        (m.row_perm, m.row_iperm, m.brick_nbr, m.brick_vp, m.brick_vm, m.expl_col, m.expl_code, m.psell_ptr,
         m.psell_col, m.psell_code) = (nat.ptr(t) for t in self._layout_keep)

    # ------------------------------------------------------- reference contract
    @property
    def mesh(self) -> nat.GfMesh:
        return self._mesh

    @property
    def mesh_ref(self):
        return ctypes.byref(self._mesh)

    @cached_property
    def owner(self) -> torch.Tensor:
        return self.owner_i32[: self.num_internal_faces].long()

    @cached_property
    def neighbour(self) -> torch.Tensor:
        return self.neighbour_i32[: self.num_internal_faces].long()

    @cached_property
    def axis(self) -> torch.Tensor:
        return self.axis_i32[: self.num_internal_faces].long()

    @cached_property
    def domain_bnd_owner(self) -> torch.Tensor:
        return self.bnd_owner_i32[: self.num_domain_bnd_faces].long()

    @cached_property
    def domain_bnd_dir_id(self) -> torch.Tensor:
        return self.bnd_dir_i32[: self.num_domain_bnd_faces].long()

    def get_domain_bnd_mask(self, patch) -> torch.Tensor:
        from .config import DomainBoundaryPatch

        return self.domain_bnd_dir_id == DomainBoundaryPatch(patch).to_direction()

    def patch_id(self, patch) -> int:
        """Boundary patch id: domain patches 0..5 (their direction), immersed
        bodies 6 + their index; -1 for an unknown name."""
        from .config import DomainBoundaryPatch

        key = patch.value if isinstance(patch, DomainBoundaryPatch) else str(patch)
        if key in {p.value for p in DomainBoundaryPatch}:
            return DomainBoundaryPatch(key).to_direction()
        return self.patch_name_to_id.get(key, -1)

    def bnd_distances(self) -> torch.Tensor:
        """|d| of every boundary face; while a FourierShape is active, the
        same values carrying their derivative with respect to its
        coefficients."""
        return self.bnd_mag_d if self.shape is None else self.shape.bnd_mag_d

    def bnd_patch_mask(self, patch) -> torch.Tensor:
        """Mask over all boundary faces (domain and immersed sides) of a patch."""
        return self.bnd_patch_id[: self.num_bnd_faces] == self.patch_id(patch)

    def ap_get_patch_mask(self, patch_name: str):
        """axis_projected.py:394-410: (upper, lower) masks over the immersed faces."""
        pid = self.patch_name_to_id.get(str(patch_name), -1) - 6
        return self.ap_owner_bnd_patch_id == pid, self.ap_neighbour_bnd_patch_id == pid

    @property
    def bnd_owner(self) -> torch.Tensor:
        """Owner cell of every boundary face (domain and immersed sides)."""
        return self.bnd_owner_i32[: self.num_bnd_faces].long()

    def register_cellfield(self, field) -> None:
        self._cellfields[field.name] = field

    def register_facefield(self, field) -> None:
        self._facefields[field.name] = field

    def get_cellfield(self, name: str):
        return self._cellfields.get(name)

    def get_facefield(self, name: str):
        return self._facefields.get(name)

    @property
    def num_hanging_faces(self) -> int:
        return int(self._mesh.n_hang)

    def topology_host(self) -> dict:
        """Integer build on the host (for parity checks)."""
        F, B = self.num_internal_faces, self.num_domain_bnd_faces
        return {
            "level": self.level.cpu().numpy(),
            "ijk": self.ijk.cpu().numpy(),
            "owner": self.owner_i32[:F].cpu().numpy(),
            "neighbour": self.neighbour_i32[:F].cpu().numpy(),
            "axis": self.axis_i32[:F].cpu().numpy(),
            "bnd_owner": self.bnd_owner_i32[:B].cpu().numpy(),
            "bnd_dir": self.bnd_dir_i32[:B].cpu().numpy(),
        }


def canonical_face_order(owner: torch.Tensor, bnd_owner: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
    """Permutations (new index -> given index) that group internal and domain
    faces by owner in ascending order, keeping the given order inside a
    group (stable sort): the face-order contract of gf_mesh."""
    o = torch.as_tensor(owner).long().cpu()
    b = torch.as_tensor(bnd_owner).long().cpu()
    return torch.sort(o, stable=True).indices, torch.sort(b, stable=True).indices


def create_grid(config: GridfoamConfig, device: str | torch.device = "cuda") -> OctreeGrid:
    """core/grid/factory.py:11-73 equivalent (native octree build)."""
    grid = OctreeGrid(config.fluxel, dt=config.simulator.control.deltaT, device=device)
    grid.sim_config = config.simulator
    return grid
