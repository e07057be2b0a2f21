"""ORACLE -- test infrastructure only.

FP64 CPU (numpy) restatement of gridfoam's pressure-velocity path, used by
``tests/`` as the expected answer, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py``.  Nothing in ``gridfoam_amd`` imports
it; the product runs only on the HIP library.

Scatters use ``np.bincount`` (sequential, deterministic summation order),
never atomics.  Each function cites the reference file:line it restates.
Every function is pinned in ``tests/test_oracle.py`` by the reference's own
known-answer tests, closed forms and invariants.
"""

from __future__ import annotations

import ctypes
import math
import pathlib
from dataclasses import dataclass, field
from typing import Any

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB: ctypes.CDLL | None = None

PATCHES = ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")


# ---------------------------------------------------------------------------
# Octree build (C restatement, oracle/octree_oracle.c)
# ---------------------------------------------------------------------------
def _lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        path = _HERE / "liboctree_oracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library missing: {path} (run make -C oracle)")
        lib = ctypes.CDLL(str(path))
        lib.oct_oracle_build.restype = ctypes.c_void_p
        lib.oct_oracle_build_surface.restype = ctypes.c_void_p
        lib.oct_oracle_fetch.restype = None
        lib.oct_oracle_free.restype = None
        lib.oct_oracle_free.argtypes = [ctypes.c_void_p]
        _LIB = lib
    return _LIB


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


@dataclass
class OctreeTopology:
    lmax: int
    level: np.ndarray  # [C] int32
    ijk: np.ndarray  # [C,3] int32 (coordinates at the cell's own level)
    owner: np.ndarray  # [F] int32
    neighbour: np.ndarray  # [F] int32
    axis: np.ndarray  # [F] int32
    bnd_owner: np.ndarray  # [F_b] int32
    bnd_dir: np.ndarray  # [F_b] int32


def _shape_params(body: dict):
    """(kind code, axis, p[20]) of a body for the C build: plane 0 (axis,
    offset), box 1 (min, max), bluff body 3 (min, max, nose radius, slant
    sin / cos, slant length, strut radius and height, strut count in axis,
    strut centres)."""
    p = np.zeros(20)
    kind = body["kind"]
    if kind == "plane":
        p[0] = float(body.get("offset", 0.0))
        return 0, int(body.get("axis", 0)), p
    if kind in ("box", "bluff"):
        p[0:3], p[3:6] = body["min"], body["max"]
        if kind == "box":
            return 1, 0, p
        ang = math.radians(float(body.get("slant_angle", 0.0)))
        p[6], p[7], p[8] = float(body.get("radius", 0.0)), math.sin(ang), math.cos(ang)
        p[9], p[10], p[11] = (float(body.get(k, 0.0)) for k in ("slant_length", "strut_radius", "strut_height"))
        struts = body.get("struts", []) or []
        for k, (x, y) in enumerate(struts):
            p[12 + 2 * k], p[13 + 2 * k] = x, y
        return 3, len(struts), p
    raise ValueError(f"surface refinement is not defined for {kind!r} bodies")


def build_octree(lower, upper, root, regions=(), n_leaf: int = 0, bodies=(), target_level: int = 0) -> OctreeTopology:
    """Octree leaves and face lists (contract: core/grid/factory.py:32-65);
    with bodies and target_level > 0, leaves are refined to target_level
    where a body's surface may pass (meta/config.py:100-106)."""
    lib = _lib()
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    hi = np.ascontiguousarray(upper, dtype=np.float64)
    rt = np.ascontiguousarray(root, dtype=np.int32)
    nr = len(regions)
    rmin = np.ascontiguousarray([r[0] for r in regions] or [[0, 0, 0]], dtype=np.float64)
    rmax = np.ascontiguousarray([r[1] for r in regions] or [[0, 0, 0]], dtype=np.float64)
    rlev = np.ascontiguousarray([r[2] for r in regions] or [0], dtype=np.int32)
    counts = np.zeros(4, dtype=np.int64)
    if target_level > 0 and len(bodies):
        packed = [_shape_params(b) for b in bodies]
        kinds = np.ascontiguousarray([k for k, _, _ in packed], dtype=np.int32)
        axes = np.ascontiguousarray([a for _, a, _ in packed], dtype=np.int32)
        params = np.ascontiguousarray(np.stack([q for _, _, q in packed]), dtype=np.float64)
        h = lib.oct_oracle_build_surface(
            _ptr(lo), _ptr(hi), _ptr(rt), ctypes.c_int32(nr), _ptr(rmin), _ptr(rmax), _ptr(rlev),
            ctypes.c_int32(len(bodies)), _ptr(kinds), _ptr(axes), _ptr(params), ctypes.c_int32(target_level),
            ctypes.c_int32(n_leaf), _ptr(counts),
        )
    else:
        h = lib.oct_oracle_build(
            _ptr(lo), _ptr(hi), _ptr(rt), ctypes.c_int32(nr), _ptr(rmin), _ptr(rmax),
            _ptr(rlev), ctypes.c_int32(n_leaf), _ptr(counts),
        )
    try:
        C, F, B, lmax = (int(v) for v in counts)
        level = np.zeros(C, np.int32)
        ijk = np.zeros((C, 3), np.int32)
        fo = np.zeros(F, np.int32)
        fn = np.zeros(F, np.int32)
        fa = np.zeros(F, np.int32)
        bo = np.zeros(B, np.int32)
        bd = np.zeros(B, np.int32)
        lib.oct_oracle_fetch(ctypes.c_void_p(h), _ptr(level), _ptr(ijk), _ptr(fo), _ptr(fn),
                             _ptr(fa), _ptr(bo), _ptr(bd))
    finally:
        lib.oct_oracle_free(ctypes.c_void_p(h))
    return OctreeTopology(lmax, level, ijk, fo, fn, fa, bo, bd)


# ---------------------------------------------------------------------------
# Grid geometry  (core/grid/axis_projected.py:130-161, :715-811)
# ---------------------------------------------------------------------------
def _scatter(idx: np.ndarray, vals: np.ndarray, n: int) -> np.ndarray:
    """Deterministic index_add over axis 0 (any trailing shape)."""
    if vals.ndim == 1:
        return np.bincount(idx, weights=vals, minlength=n)
    flat = vals.reshape(vals.shape[0], -1)
    out = np.stack([np.bincount(idx, weights=flat[:, k], minlength=n)
                    for k in range(flat.shape[1])], axis=1)
    return out.reshape((n,) + vals.shape[1:])


# ---------------------------------------------------------------------------
# Axis-projected immersed-boundary payload for analytic bodies
# ---------------------------------------------------------------------------
N_DOMAIN_PATCHES = 6


def _fourier_outline(body: dict, x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """|P - c| - r(theta), r = radius (1 + sum_k a_k cos k theta + b_k sin k theta):
    negative inside the outline."""
    a = list(body.get("a", [])) + [0.0] * (8 - len(body.get("a", [])))
    b = list(body.get("b", [])) + [0.0] * (8 - len(body.get("b", [])))
    cx, cy = body["center"][0], body["center"][1]
    dx, dy = x - cx, y - cy
    rho = np.sqrt(dx * dx + dy * dy)
    th = np.arctan2(dy, dx)
    r = np.ones_like(th)
    for k in range(1, 9):
        r = r + (a[k - 1] * np.cos(k * th) + b[k - 1] * np.sin(k * th))
    return rho - body["radius"] * r


def _bluff_outline(body: dict, x: np.ndarray, y: np.ndarray, z: np.ndarray) -> np.ndarray:
    """Implicit function of a bluff body (negative inside): the box min - max
    with its nose (low-x) edges rounded by `radius` (for x < x0 + r:
    max(|(a, b_y)|, |(a, b_z)|) - r, a = x0 + r - x, b_y and b_z the
    distances past the rounding's inner lines), the top rear edge cut by the
    plane of outward normal (sin, 0, cos) of `slant_angle` through the top
    `slant_length` ahead of the rear, and vertical struts of `strut_radius`
    from `strut_height` below the body into it.  Every piece is 1-Lipschitz,
    so |G| bounds the distance to the surface from below."""
    _, n_st, p = _shape_params(body)
    x0, y0, z0, x1, y1, z1, r = p[0], p[1], p[2], p[3], p[4], p[5], p[6]
    g = np.maximum(np.maximum(np.maximum(x0 - x, x - x1), np.maximum(y0 - y, y - y1)),
                   np.maximum(z0 - z, z - z1))
    a = (x0 + r) - x
    by = np.maximum(np.maximum((y0 + r) - y, y - (y1 - r)), 0.0)
    bz = np.maximum(np.maximum((z0 + r) - z, z - (z1 - r)), 0.0)
    nose = np.maximum(np.sqrt(a * a + by * by) - r, np.sqrt(a * a + bz * bz) - r)
    g = np.where(x < x0 + r, np.maximum(g, nose), g)
    g = np.maximum(g, p[7] * (x - (x1 - p[9])) + p[8] * (z - z1))
    for k in range(n_st):
        dx, dy = x - p[12 + 2 * k], y - p[13 + 2 * k]
        gs = np.maximum(np.sqrt(dx * dx + dy * dy) - p[10], np.maximum((z0 - p[11]) - z, z - (z0 + p[10])))
        g = np.minimum(g, gs)
    return g


def _body_hits(body: dict, ax: np.ndarray, P: np.ndarray, s0: np.ndarray, s1: np.ndarray):
    """Hit coordinate along each face's axis ray [s0, s1] (through the face
    centre P, endpoints included) seen from the s0 end and from the s1 end;
    NaN where the ray misses the body."""
    nan = np.full(len(ax), np.nan)
    kind = body["kind"]
    if kind == "plane":
        s = float(body.get("offset", 0.0))
        hit = (ax == int(body.get("axis", 0))) & (s0 <= s) & (s <= s1)
        h = np.where(hit, s, nan)
        return h, h.copy()
    if kind == "box":
        lo3, hi3 = np.asarray(body["min"], np.float64), np.asarray(body["max"], np.float64)
        inside = np.ones(len(ax), bool)
        for j in range(3):
            inside &= (ax == j) | ((P[:, j] >= lo3[j]) & (P[:, j] <= hi3[j]))
        lo, hi = lo3[ax], hi3[ax]
        in_lo = inside & (lo >= s0) & (lo <= s1)
        in_hi = inside & (hi >= s0) & (hi <= s1)
        from_lo = np.where(in_lo, lo, np.where(in_hi, hi, nan))
        from_hi = np.where(in_hi, hi, np.where(in_lo, lo, nan))
        return from_lo, from_hi
    if kind in ("fourier", "bluff"):
        def outline(s):
            x, y = np.where(ax == 0, s, P[:, 0]), np.where(ax == 1, s, P[:, 1])
            if kind == "bluff":
                return _bluff_outline(body, x, y, np.where(ax == 2, s, P[:, 2]))
            return _fourier_outline(body, x, y)
        g0, g1 = outline(s0), outline(s1)
        ok = ((ax != 2) | (kind == "bluff")) & ~(g0 * g1 > 0.0)
        h = np.where(ok & (g0 == 0.0), s0, np.where(ok & (g1 == 0.0), s1, nan))
        need = ok & (g0 != 0.0) & (g1 != 0.0)
        lo, hi, glo = s0.copy(), s1.copy(), g0.copy()
        active = need.copy()
        for _ in range(64):  # bisection on the sign change
            mid = 0.5 * (lo + hi)
            active &= ~((mid <= lo) | (mid >= hi))
            gm = outline(mid)
            same = (gm <= 0.0) == (glo <= 0.0)
            lo = np.where(active & same, mid, lo)
            glo = np.where(active & same, gm, glo)
            hi = np.where(active & ~same, mid, hi)
        h = np.where(need, 0.5 * (lo + hi), h)
        return h, h.copy()
    raise ValueError(f"unknown immersed body kind {kind!r}")


@dataclass
class APCut:
    is_immersed: np.ndarray  # [F] bool
    dist_owner: np.ndarray   # [F_i]
    dist_neighbour: np.ndarray
    owner_near: np.ndarray   # [F_i] bool
    neighbour_near: np.ndarray
    owner_patch: np.ndarray  # [F_i] body index
    neighbour_patch: np.ndarray
    owner_normal: np.ndarray  # [F_i, 3] body-outward unit normal at the owner-side hit
    neighbour_normal: np.ndarray


def _body_normal(body: dict, ax: np.ndarray, P: np.ndarray, s: np.ndarray) -> np.ndarray:
    """Body-outward unit normal at the hit s on each ray: a plane's own +axis
    normal; a box's face normal (-axis on its lower face); the gradient of
    |P - c| - R r(theta) for the Fourier outline."""
    n = np.zeros((len(ax), 3))
    kind = body["kind"]
    if kind == "plane":
        n[:, int(body.get("axis", 0))] = 1.0
    elif kind == "box":
        lo = np.asarray(body["min"], np.float64)[ax]
        n[np.arange(len(ax)), ax] = np.where(s == lo, -1.0, 1.0)
    elif kind == "bluff":  # central differences of the implicit function, step 1e-7
        q = P.copy()
        q[np.arange(len(ax)), ax] = s
        e = 1e-7
        gr = np.zeros((len(ax), 3))
        for j in range(3):
            qp, qm = q.copy(), q.copy()
            qp[:, j] = q[:, j] + e
            qm[:, j] = q[:, j] - e
            gr[:, j] = (_bluff_outline(body, *qp.T) - _bluff_outline(body, *qm.T)) / (2.0 * e)
        gn = np.sqrt(gr[:, 0] * gr[:, 0] + gr[:, 1] * gr[:, 1] + gr[:, 2] * gr[:, 2])
        n = gr / gn[:, None]
    else:
        x = np.where(ax == 0, s, P[:, 0])
        y = np.where(ax == 1, s, P[:, 1])
        dx, dy = x - body["center"][0], y - body["center"][1]
        rho2 = dx * dx + dy * dy
        rho = np.sqrt(rho2)
        th = np.arctan2(dy, dx)
        a = list(body.get("a", [])) + [0.0] * (8 - len(body.get("a", [])))
        b = list(body.get("b", [])) + [0.0] * (8 - len(body.get("b", [])))
        dr = np.zeros_like(th)
        for k in range(1, 9):
            dr = dr + k * (b[k - 1] * np.cos(k * th) - a[k - 1] * np.sin(k * th))
        Rdr = body["radius"] * dr
        gx, gy = dx / rho + Rdr * dy / rho2, dy / rho - Rdr * dx / rho2
        gn = np.sqrt(gx * gx + gy * gy)
        n[:, 0], n[:, 1] = gx / gn, gy / gn
    return n


def ap_cut(owner, neighbour, axis, centers, sizes, face_centers, bodies, ref_length: float) -> APCut:
    """The AP-IBM payload fluxel hands to AxisProjectedGrid
    (core/grid/axis_projected.py:163-211), defined by
    docs/source/contributor_guide/architecture/data_contracts.rst:181-231, for
    analytic bodies: a face is immersed when its axis ray from the owner's to
    the neighbour's centre meets a body; the distances are ray distances from
    each centre to the first hit on its side, the near flags Gibou's
    distance / dx <= dx / L (dx the cell's width on the axis, L the largest
    domain extent), and each side carries the patch of the body it hits (the
    first body listed wins a tie)."""
    F = len(owner)
    s0 = centers[owner, axis]
    s1 = centers[neighbour, axis]
    best_o, best_n = np.full(F, np.inf), np.full(F, np.inf)
    hit_o, hit_n = np.zeros(F), np.zeros(F)
    po, pn = np.full(F, -1), np.full(F, -1)
    for k, body in enumerate(bodies):
        h_lo, h_hi = _body_hits(body, axis, face_centers, s0, s1)
        do, dn = h_lo - s0, s1 - h_hi
        up_o = ~np.isnan(h_lo) & (do < best_o)
        up_n = ~np.isnan(h_hi) & (dn < best_n)
        best_o, po, hit_o = np.where(up_o, do, best_o), np.where(up_o, k, po), np.where(up_o, h_lo, hit_o)
        best_n, pn, hit_n = np.where(up_n, dn, best_n), np.where(up_n, k, pn), np.where(up_n, h_hi, hit_n)
    cut = po >= 0
    dxo, dxn = sizes[owner, axis], sizes[neighbour, axis]
    near_o = cut & (best_o / dxo <= dxo / ref_length)
    near_n = cut & (best_n / dxn <= dxn / ref_length)
    nrm_o, nrm_n = np.zeros((F, 3)), np.zeros((F, 3))
    for k, body in enumerate(bodies):
        for pk, hit, nrm in ((po, hit_o, nrm_o), (pn, hit_n, nrm_n)):
            sel = pk == k
            if sel.any():
                nrm[sel] = _body_normal(body, axis[sel], face_centers[sel], hit[sel])
    return APCut(cut, best_o[cut], best_n[cut], near_o[cut], near_n[cut], po[cut], pn[cut], nrm_o[cut], nrm_n[cut])


class Grid:
    """Restates AxisProjectedGrid's topology + geometry; with ``bodies`` the
    internal faces cut by them become two boundary faces each (upper: owner
    side, +Sf, distance to the body from the owner; lower: neighbour side,
    -Sf), fv/boundary_ops.py:150-186.  The boundary arrays then hold
    [domain faces | upper sides | lower sides]; the internal arrays the
    single-sided faces in ascending order (face_geometry.py:128-131)."""

    def __init__(self, lower, upper, root, regions=(), n_leaf=0, dt=1.0, bodies=(), target_level=0):
        topo = build_octree(lower, upper, root, regions, n_leaf, bodies, target_level)
        self.topo = topo
        self.lower = np.asarray(lower, np.float64)
        self.upper = np.asarray(upper, np.float64)
        self.root = np.asarray(root, np.int64)
        self.dt = float(dt)
        lev = topo.level.astype(np.int64)
        ext = self.upper - self.lower
        self.cell_sizes = ext[None, :] / (self.root[None, :] * (2.0 ** lev)[:, None])
        self.cell_centers = self.lower[None, :] + (topo.ijk + 0.5) * self.cell_sizes
        s = self.cell_sizes
        self.cell_volumes = s[:, 0] * s[:, 1] * s[:, 2]
        self.owner = topo.owner.astype(np.int64)
        self.neighbour = topo.neighbour.astype(np.int64)
        self.axis = topo.axis.astype(np.int64)
        self.bnd_owner = topo.bnd_owner.astype(np.int64)
        self.bnd_dir = topo.bnd_dir.astype(np.int64)
        self.num_cells = len(lev)
        self.num_faces = len(self.owner)
        self.num_bnd = len(self.bnd_owner)
        # axis_projected.py:715-737 -- overlap-box centre
        cO, cN = self.cell_centers[self.owner], self.cell_centers[self.neighbour]
        sO, sN = s[self.owner], s[self.neighbour]
        fmin = np.maximum(cO - sO / 2.0, cN - sN / 2.0)
        fmax = np.minimum(cO + sO / 2.0, cN + sN / 2.0)
        self.face_centers = (fmin + fmax) / 2.0
        # axis_projected.py:740-767 -- finer-cell cross section on the axis
        mins = np.minimum(sO, sN)
        fi = np.arange(self.num_faces)
        area = (mins[:, 0] * mins[:, 1] * mins[:, 2]) / mins[fi, self.axis]
        self.Sf = np.zeros((self.num_faces, 3))
        self.Sf[fi, self.axis] = area
        # axis_projected.py:770-811
        bi = np.arange(self.num_bnd)
        baxis = self.bnd_dir // 2
        sign = 2.0 * (self.bnd_dir % 2) - 1.0
        normals = np.zeros((self.num_bnd, 3))
        normals[bi, baxis] = sign
        cb, sb = self.cell_centers[self.bnd_owner], s[self.bnd_owner]
        self.bnd_face_centers = cb + 0.5 * sb * normals
        barea = sign * (sb[:, 0] * sb[:, 1] * sb[:, 2]) / sb[bi, baxis]
        self.bnd_Sf = np.zeros((self.num_bnd, 3))
        self.bnd_Sf[bi, baxis] = barea
        # cell-to-boundary vectors (boundary_ops.py:128-146) and patch ids
        self.bnd_d = self.bnd_face_centers - cb
        self.bnd_mag_d = np.sqrt(np.sum(self.bnd_d * self.bnd_d, axis=1))
        self.bnd_patch = self.bnd_dir.copy()
        self.bnd_snap = np.zeros(self.num_bnd, bool)
        self.num_domain_bnd = self.num_bnd
        self.num_all_faces = self.num_faces
        self.bodies = [dict(b) for b in bodies]
        self.patch_ids = {b["name"]: N_DOMAIN_PATCHES + k for k, b in enumerate(self.bodies)}
        self.is_immersed = np.zeros(self.num_faces, bool)
        self.num_immersed = 0
        if self.bodies:
            self._cut(float(np.max(ext)))
        self.geo = FaceGeometry(self)
        # core/grid/axis_projected.py:370-386 -- registered fields by name;
        # each entry is a zero-argument callable returning the current value
        # (U, p: Field; phi, phiHbyA: FaceValues; HbyA, rAU, rAtU: arrays)
        self.registry: dict[str, Any] = {}

    def _cut(self, L: float) -> None:
        cut = ap_cut(self.owner, self.neighbour, self.axis, self.cell_centers, self.cell_sizes, self.face_centers,
                     self.bodies, L)
        self.ap = cut
        imm = cut.is_immersed
        Fi = int(imm.sum())
        self.is_immersed, self.num_immersed = imm, Fi
        if Fi == 0:
            return
        Sf_i = self.Sf[imm]
        mag = np.sqrt(np.sum(Sf_i * Sf_i, axis=1))
        n_hat = Sf_i / mag[:, None]
        ax_i = self.axis[imm]
        d_up = cut.dist_owner[:, None] * n_hat
        d_lo = cut.dist_neighbour[:, None] * (-n_hat)
        own_i, nbr_i = self.owner[imm], self.neighbour[imm]
        self.bnd_owner = np.concatenate([self.bnd_owner, own_i, nbr_i])
        self.bnd_dir = np.concatenate([self.bnd_dir, 2 * ax_i + 1, 2 * ax_i])
        self.bnd_patch = np.concatenate([self.bnd_patch, N_DOMAIN_PATCHES + cut.owner_patch,
                                         N_DOMAIN_PATCHES + cut.neighbour_patch])
        self.bnd_d = np.concatenate([self.bnd_d, d_up, d_lo])
        self.bnd_mag_d = np.concatenate([self.bnd_mag_d, cut.dist_owner, cut.dist_neighbour])
        self.bnd_Sf = np.concatenate([self.bnd_Sf, Sf_i, -Sf_i])
        self.bnd_face_centers = np.concatenate([self.bnd_face_centers, self.cell_centers[own_i] + d_up,
                                                self.cell_centers[nbr_i] + d_lo])
        self.bnd_snap = np.concatenate([self.bnd_snap, cut.owner_near, cut.neighbour_near])
        self.num_bnd = len(self.bnd_owner)
        single = ~imm
        self.owner, self.neighbour, self.axis = self.owner[single], self.neighbour[single], self.axis[single]
        self.face_centers, self.Sf = self.face_centers[single], self.Sf[single]
        self.num_faces = len(self.owner)

    def patch_id(self, patch: str) -> int:
        if patch in PATCHES:
            return PATCHES.index(patch)
        return self.patch_ids.get(patch, -1)

    def patch_mask(self, patch: str) -> np.ndarray:
        return self.bnd_patch == self.patch_id(patch)

    @property
    def domain_mask(self) -> np.ndarray:
        return np.arange(self.num_bnd) < self.num_domain_bnd


class FaceGeometry:
    """fv/kernels/face_geometry.py:107-262 over the single-sided faces; the
    least-squares normal matrix takes every boundary face's cell-to-boundary
    vector (boundary_lsq_vectors, :198-243)."""

    def __init__(self, g: Grid):
        C = g.cell_centers
        o, n, ax = g.owner, g.neighbour, g.axis
        fi = np.arange(g.num_faces)
        d_all = C[n] - C[o]
        d_fN = C[n] - g.face_centers
        d_n = np.abs(d_all[fi, ax])
        self.w = np.abs(d_fN[fi, ax]) / d_n
        self.mag_Sf = np.sqrt(np.sum(g.Sf * g.Sf, axis=1))
        self.delta = 1.0 / d_n
        centroid = self.w[:, None] * C[o] + (1.0 - self.w)[:, None] * C[n]
        offset = g.face_centers - centroid
        tol = 1.0e-8 * float(g.cell_sizes.min())
        hang = np.sqrt(np.sum(offset * offset, axis=1)) > tol
        self.hang_idx = np.nonzero(hang)[0]
        self.hang_offset = offset[self.hang_idx]
        h = self.hang_idx
        mO = g.face_centers[h] - C[o[h]]
        mN = g.face_centers[h] - C[n[h]]
        mO[np.arange(len(h)), ax[h]] = 0.0
        mN[np.arange(len(h)), ax[h]] = 0.0
        self.hang_move_O, self.hang_move_N = mO, mN
        self.hang_cell_mask = np.zeros(g.num_cells, bool)
        self.hang_cell_mask[o[h]] = True
        self.hang_cell_mask[n[h]] = True
        # face_geometry.py:191-262 -- geometric LSQ normal matrix and pinv
        ata = np.zeros((g.num_cells, 3, 3))
        outer = _lsq_outer(d_all)
        ata += _scatter(o, outer, g.num_cells)
        ata += _scatter(n, outer, g.num_cells)
        ata += _scatter(g.bnd_owner, _lsq_outer(g.bnd_d), g.num_cells)
        self.lsq_ata_inv = np.linalg.pinv(ata, hermitian=True)


def _lsq_outer(d: np.ndarray) -> np.ndarray:
    d2 = np.sum(d * d, axis=1)
    w2 = 1.0 / np.where(d2 > 0, d2, 1.0)
    return w2[:, None, None] * d[:, :, None] * d[:, None, :]


# ---------------------------------------------------------------------------
# Fields and boundary conditions  (boundaries/*.py, fv/boundary_ops.py)
# ---------------------------------------------------------------------------
@dataclass
class BC:
    kind: str  # dirichlet | neumann | empty | slip | inlet_outlet | fixed_flux_pressure
    value: np.ndarray | float = 0.0

    @property
    def assignable(self) -> bool:  # boundaries/base.py:38-48; dirichlet/slip False
        return self.kind not in ("dirichlet", "slip")


@dataclass
class Field:
    name: str
    data: np.ndarray
    bcs: dict[str, BC] = field(default_factory=dict)
    old: np.ndarray | None = None
    older: np.ndarray | None = None

    @property
    def cshape(self) -> tuple[int, ...]:
        return tuple(self.data.shape[1:])


@dataclass
class Batch:
    patch: str
    mask: np.ndarray
    cells: np.ndarray
    d_vec: np.ndarray
    mag_d: np.ndarray
    Sf: np.ndarray
    mag_Sf: np.ndarray
    snap: np.ndarray  # immersed side within the Gibou distance
    immersed: bool = False


def batches(g: Grid, fld: Field) -> list[Batch]:
    """boundary_ops.py:121-186 in the order of the field's conditions, empty
    skipped; an immersed patch is one batch of its upper sides followed by
    its lower sides (the reference's upper batch, then its lower batch)."""
    out = []
    for patch, bc in fld.bcs.items():
        if bc.kind == "empty":
            continue
        m = g.patch_mask(patch)
        if not m.any():
            continue
        cells = g.bnd_owner[m]
        Sf = g.bnd_Sf[m]
        out.append(Batch(patch, m, cells, g.bnd_d[m], g.bnd_mag_d[m], Sf, np.sqrt(np.sum(Sf * Sf, axis=1)),
                         g.bnd_snap[m], patch not in PATCHES))
    return out


def bc_evaluate(g: Grid, fld: Field, b: Batch):
    """(fraction, ref_v, ref_g): dirichlet.py:46-73, neumann.py:42-73, slip.py:46-101."""
    bc = fld.bcs[b.patch]
    n = len(b.cells)
    cs = fld.cshape
    if bc.kind == "dirichlet":
        v = np.broadcast_to(np.asarray(bc.value, np.float64), cs)
        return np.ones(n), np.ones((n,) + cs) * v, np.zeros((n,) + cs)
    if bc.kind == "neumann":
        v = np.broadcast_to(np.asarray(bc.value, np.float64), cs)
        return np.zeros(n), np.zeros((n,) + cs), np.ones((n,) + cs) * v
    if bc.kind == "slip":
        if cs == ():
            return np.zeros(n), np.zeros(n), np.zeros(n)
        nvec = np.sign(b.Sf)
        psi = fld.data[b.cells]
        un = np.sum(psi * nvec, axis=-1)
        return np.ones(n), psi - un[:, None] * nvec, np.zeros((n,) + cs)
    if bc.kind == "inlet_outlet":
        return _inlet_outlet(g, fld, b, bc)
    if bc.kind == "fixed_flux_pressure":
        return _fixed_flux_pressure(g, fld, b)
    raise ValueError(f"unsupported bc {bc.kind}")


def _inlet_outlet(g: Grid, fld: Field, b: Batch, bc: BC):
    """boundaries/derived/inlet_outlet.py:61-123: the inlet value where the
    registered boundary flux phi is negative (inflow), zero gradient where it
    is zero or positive."""
    cs = fld.cshape
    v = np.asarray(bc.value, np.float64)
    if v.shape != cs:
        raise ValueError(f"inlet value shape {v.shape} does not match field shape {cs}")
    if "phi" not in g.registry:
        raise ValueError("Field phi is not found.")
    inflow = g.registry["phi"]().bnd[b.mask] < 0.0
    n = len(b.cells)
    rv = np.zeros((n,) + cs)
    rv[inflow] = v
    return inflow.astype(np.float64), rv, np.zeros((n,) + cs)


def _fixed_flux_pressure(g: Grid, fld: Field, b: Batch):
    """boundaries/derived/fixed_flux_pressure.py:185-252: fraction 0 and the
    normal gradient (phiHbyA_b - U_b . Sf) / (rAtU_P |Sf|), zero where that
    denominator is not positive; rAU stands in for an unregistered rAtU,
    HbyA_P . Sf for an unregistered phiHbyA, the boundary phi for a velocity
    without a condition on the patch; zero gradient while the coefficient or
    both predicted-flux fields are unregistered."""
    if fld.cshape != ():
        raise ValueError("FixedFluxPressure requires a scalar field")
    reg = g.registry
    n = len(b.cells)
    zero = np.zeros(n)
    coeff = reg.get("rAtU", reg.get("rAU"))
    if coeff is None or ("phiHbyA" not in reg and "HbyA" not in reg):
        return zero, zero.copy(), zero.copy()
    if "phiHbyA" in reg:
        pred = reg["phiHbyA"]().bnd[b.mask]
    else:
        pred = np.sum(reg["HbyA"]()[b.cells] * b.Sf, axis=1)
    U = reg["U"]() if "U" in reg else None
    if U is not None and b.patch in U.bcs:
        if U.bcs[b.patch].kind == "empty":  # empty.py:31-51: zero-gradient face value
            ub = U.data[b.cells]
        else:
            ub = bc_state(g, U, b)[3]
        target = np.sum(ub * b.Sf, axis=1)
    elif "phi" in reg:
        target = reg["phi"]().bnd[b.mask]
    else:
        return zero, zero.copy(), zero.copy()
    denom = coeff()[b.cells] * b.mag_Sf
    pos = denom > 0.0
    grad = np.where(pos, (pred - target) / np.where(pos, denom, 1.0), 0.0)
    return zero, zero.copy(), grad


def bc_state(g: Grid, fld: Field, b: Batch):
    """boundary_ops.py:311-326 -> (f, ref_v, ref_g, psi_b)."""
    f, rv, rg = bc_evaluate(g, fld, b)
    psi = fld.data[b.cells]
    fb = f.reshape((-1,) + (1,) * len(fld.cshape))
    db = b.mag_d.reshape(fb.shape)
    psi_b = fb * rv + (1.0 - fb) * (psi + db * rg)
    return f, rv, rg, psi_b


def fixed_value_mask(b: Batch, f: np.ndarray) -> np.ndarray:
    """boundary_ops.py:388-402: snapped immersed sides with fraction 1."""
    return b.snap & (f == 1)


def value_grad_coeff(b: Batch, f: np.ndarray) -> np.ndarray:
    """boundary_ops.py:405-422: f / |d|, zero on pure Neumann and snapped
    Dirichlet faces."""
    active = (f != 0) & ~fixed_value_mask(b, f)
    if np.any(active & (b.mag_d == 0)):
        raise ValueError("A mixed boundary at zero distance needs a cell constraint")
    dist = np.where(active, b.mag_d, 1.0)
    return np.where(active, f / dist, 0.0)


def immersed_constraints(g: Grid, fld: Field) -> tuple[np.ndarray, np.ndarray]:
    """boundary_ops.py:440-496: cells fixed by Gibou snapping and their
    values.  Candidates are the snapped sides with fraction 1 of immersed
    patches whose condition is not Neumann; a cell with several takes the
    closest, ties in batch/face order."""
    cells, dist, vals = [], [], []
    for b in batches(g, fld):
        if not b.immersed or fld.bcs[b.patch].kind == "neumann":
            continue
        f, rv, _ = bc_evaluate(g, fld, b)
        fixed = fixed_value_mask(b, f)
        if fixed.any():
            cells.append(b.cells[fixed])
            dist.append(b.mag_d[fixed])
            vals.append(rv[fixed])
    if not cells:
        return np.zeros(0, np.int64), fld.data[:0].copy()
    cell, d, v = np.concatenate(cells), np.concatenate(dist), np.concatenate(vals)
    n = len(cell)
    closest = np.full(g.num_cells, np.inf)
    np.minimum.at(closest, cell, d)
    cand = np.where(d == closest[cell], np.arange(n), n)
    sel = np.full(g.num_cells, n)
    np.minimum.at(sel, cell, cand)
    ucells = np.nonzero(sel < n)[0]
    return ucells, v[sel[ucells]]


def apply_immersed_values(g: Grid, fld: Field) -> None:
    """boundary_ops.py:538-542: restore prescribed values after an explicit update."""
    cells, values = immersed_constraints(g, fld)
    if len(cells):
        fld.data = fld.data.copy()
        fld.data[cells] = values


def close_fixed_cell_flux(g: Grid, phi: FaceValues, p: Field) -> None:
    """pressure_correction.py:308-335: on cells fixed by snapping, the cell's
    flux imbalance is taken off its snapped Dirichlet sides in proportion
    to their area."""
    selected = []
    area_sum = np.zeros(g.num_cells)
    for b in batches(g, p):
        f = bc_evaluate(g, p, b)[0]
        fixed = fixed_value_mask(b, f)
        if not fixed.any():
            continue
        area = np.where(fixed, b.mag_Sf, 0.0)
        area_sum += _scatter(b.cells, area, g.num_cells)
        selected.append((b, area))
    if not selected:
        return
    imbalance = fvc_div(g, phi) * g.cell_volumes
    for b, area in selected:
        total = area_sum[b.cells]
        frac = area / np.where(total > 0, total, 1.0)
        phi.bnd[b.mask] = phi.bnd[b.mask] - frac * imbalance[b.cells]


def sn_grad_boundary(g: Grid, fld: Field) -> np.ndarray:
    """fv/fvc/sn_grad.py:40-60 boundary blocks: the mixed normal gradient
    coeff (ref_v - psi_P) + (1 - f) ref_g of boundary_ops.py:425-437 on
    covered faces (zero on snapped Dirichlet faces), zero elsewhere."""
    out = np.zeros((g.num_bnd,) + fld.cshape)
    for b in batches(g, fld):
        f, rv, rg, _ = bc_state(g, fld, b)
        c = value_grad_coeff(b, f)
        out[b.mask] = _bcast(c, rv) * (rv - fld.data[b.cells]) + _bcast(1.0 - f, rg) * rg
    return out


def _bcast(v: np.ndarray, like: np.ndarray) -> np.ndarray:
    return v.reshape(v.shape + (1,) * (like.ndim - v.ndim))


# ---------------------------------------------------------------------------
# LDU matrix  (core/fvmatrix.py)
# ---------------------------------------------------------------------------
class Matrix:
    def __init__(self, g: Grid, cshape=()):
        self.g = g
        self.cshape = tuple(cshape)
        self.diag = np.zeros(g.num_cells)
        self.upper = np.zeros(g.num_faces)
        self.lower = np.zeros(g.num_faces)
        self.source = np.zeros((g.num_cells,) + self.cshape)
        self.ffc: np.ndarray | None = None

    def copy(self) -> "Matrix":
        m = Matrix(self.g, self.cshape)
        m.diag, m.upper, m.lower = self.diag.copy(), self.upper.copy(), self.lower.copy()
        m.source = self.source.copy()
        m.ffc = None if self.ffc is None else self.ffc.copy()
        return m

    def _combine(self, other: "Matrix", sign: float) -> "Matrix":  # fvmatrix.py:166-201
        m = Matrix(self.g, self.cshape)
        m.diag = self.diag + sign * other.diag
        m.upper = self.upper + sign * other.upper
        m.lower = self.lower + sign * other.lower
        m.source = self.source + sign * other.source
        a, b = self.ffc, other.ffc
        if a is None:
            m.ffc = None if b is None else sign * b
        else:
            m.ffc = a.copy() if b is None else a + sign * b
        return m

    def __add__(self, o):
        return self._combine(o, 1.0)

    def __sub__(self, o):
        return self._combine(o, -1.0)

    def __neg__(self):  # fvmatrix.py:203-211
        m = Matrix(self.g, self.cshape)
        m.diag, m.upper, m.lower, m.source = -self.diag, -self.upper, -self.lower, -self.source
        m.ffc = None if self.ffc is None else -self.ffc
        return m

    def multiply(self, x: np.ndarray) -> np.ndarray:  # fvmatrix.py:292-323
        g = self.g
        res = _bcast(self.diag, x) * x
        res = res + _scatter(g.owner, _bcast(self.upper, x[g.neighbour]) * x[g.neighbour], g.num_cells)
        res = res + _scatter(g.neighbour, _bcast(self.lower, x[g.owner]) * x[g.owner], g.num_cells)
        return res

    def A(self) -> np.ndarray:  # fvmatrix.py:242-248
        return self.diag / self.g.cell_volumes

    def H(self, x: np.ndarray) -> np.ndarray:  # fvmatrix.py:266-290
        V = _bcast(self.g.cell_volumes, x)
        return (self.source - self.multiply(x) + _bcast(self.diag, x) * x) / V

    def H1(self) -> np.ndarray:  # fvmatrix.py:250-264
        g = self.g
        h1 = _scatter(g.owner, -self.upper, g.num_cells) + _scatter(g.neighbour, -self.lower, g.num_cells)
        return h1 / g.cell_volumes

    def flux(self, psi: np.ndarray) -> np.ndarray:  # fvmatrix.py:424-468
        g = self.g
        pN, pO = psi[g.neighbour], psi[g.owner]
        out = _bcast(self.upper, pN) * pN - _bcast(self.lower, pO) * pO
        if self.ffc is not None:
            out = out + self.ffc
        return out

    def with_fixed_values(self, cells: np.ndarray, values: np.ndarray) -> "Matrix":
        """fvmatrix.py:325-394: symmetric elimination of prescribed cell
        values.  Prescribed columns move to the source, row and column
        couplings of a fixed cell are zeroed, the fixed row keeps its
        diagonal (a zero one becomes -1 when the diagonal sum is negative,
        else +1) with source diag * value; no face-flux correction."""
        if len(cells) == 0:
            return self
        g = self.g
        fixed = np.zeros(g.num_cells, bool)
        fixed[cells] = True
        presc = np.zeros_like(self.source)
        presc[cells] = values
        unit = -1.0 if self.diag.sum() < 0 else 1.0
        cdiag = np.where(self.diag != 0, self.diag, unit)
        m = Matrix(g, self.cshape)
        m.diag = np.where(fixed, cdiag, self.diag)
        conn = fixed[g.owner] | fixed[g.neighbour]
        m.upper = np.where(conn, 0.0, self.upper)
        m.lower = np.where(conn, 0.0, self.lower)
        pn, po = presc[g.neighbour], presc[g.owner]
        src = self.source + _scatter(g.owner, -_bcast(self.upper, pn) * pn, g.num_cells)
        src = src + _scatter(g.neighbour, -_bcast(self.lower, po) * po, g.num_cells)
        src[cells] = _bcast(cdiag[cells], values) * values
        m.source = src
        m.ffc = None
        return m

    def transpose(self) -> "Matrix":  # fvmatrix.py:396-412
        m = Matrix(self.g, self.cshape)
        m.diag, m.upper, m.lower, m.source, m.ffc = self.diag, self.lower, self.upper, self.source, self.ffc
        return m


# ---------------------------------------------------------------------------
# Explicit kernels  (fv/kernels/*.py, fv/fvc/*.py)
# ---------------------------------------------------------------------------
def lsq_grad(g: Grid, fld: Field) -> np.ndarray:
    """fv/kernels/least_squares.py:29-101 (full-stencil form)."""
    C = g.cell_centers
    psi = fld.data
    o, n = g.owner, g.neighbour
    d = C[n] - C[o]
    d2 = np.sum(d * d, axis=1)
    wd = (1.0 / np.where(d2 > 0, d2, 1.0))[:, None] * d
    dpsi = psi[n] - psi[o]
    atb_face = dpsi[..., None] * wd.reshape((len(d),) + (1,) * len(fld.cshape) + (3,))
    atb = _scatter(o, atb_face, g.num_cells) + _scatter(n, atb_face, g.num_cells)
    for b in batches(g, fld):
        _, _, _, psi_b = bc_state(g, fld, b)
        db = b.d_vec
        dd = np.sum(db * db, axis=1)
        wdb = (1.0 / np.where(dd > 0, dd, 1.0))[:, None] * db
        dp = psi_b - psi[b.cells]
        atb = atb + _scatter(b.cells, dp[..., None] * wdb.reshape((len(db),) + (1,) * len(fld.cshape) + (3,)),
                             g.num_cells)
    return np.einsum("nij,n...j->n...i", g.geo.lsq_ata_inv, atb)


def gauss_grad(g: Grid, fld: Field) -> np.ndarray:
    """fv/schemes/grad.py:21-44 + fv/kernels/gauss_gradient.py:15-73:
    Green-Gauss, sum over faces of psi_f Sf / V, with the linear face values
    (hanging faces corrected by the least-squares gradient) and the
    boundary condition's face values (zero gradient where none)."""
    fv = interpolate(g, fld)
    flux = fv.single[..., None] * g.Sf.reshape((g.num_faces,) + (1,) * len(fld.cshape) + (3,))
    bflux = fv.bnd[..., None] * g.bnd_Sf.reshape((g.num_bnd,) + (1,) * len(fld.cshape) + (3,))
    out = _scatter(g.owner, flux, g.num_cells) - _scatter(g.neighbour, flux, g.num_cells)
    out = out + _scatter(g.bnd_owner, bflux, g.num_cells)
    return out / g.cell_volumes.reshape((g.num_cells,) + (1,) * (out.ndim - 1))


def cell_grad(g: Grid, fld: Field, scheme: str = "leastsquare") -> np.ndarray:
    """fv/schemes/grad.py:82-101: the configured cell gradient."""
    if scheme == "linear":
        return gauss_grad(g, fld)
    if scheme != "leastsquare":
        raise ValueError(f"unsupported grad scheme {scheme}")
    return lsq_grad(g, fld)


def linear_face_values(g: Grid, psi: np.ndarray) -> np.ndarray:
    """face_interpolation.py:114-138."""
    w = _bcast(g.geo.w, psi[g.owner])
    return w * psi[g.owner] + (1.0 - w) * psi[g.neighbour]


def correct_face_values(g: Grid, base: np.ndarray, grad: np.ndarray) -> np.ndarray:
    """face_interpolation.py:141-197 -- hanging face-centre offset."""
    geo = g.geo
    h = geo.hang_idx
    if len(h) == 0:
        return base
    gO, gN = grad[g.owner[h]], grad[g.neighbour[h]]
    w = _bcast(geo.w[h], gO)
    gf = w * gO + (1.0 - w) * gN
    corr = np.einsum("n...j,nj->n...", gf, geo.hang_offset)
    out = base.copy()
    out[h] += corr
    return out


def sn_grad_hang_correction(g: Grid, grad: np.ndarray) -> np.ndarray:
    """face_interpolation.py:200-246."""
    geo = g.geo
    h = geo.hang_idx
    cN = np.einsum("n...j,nj->n...", grad[g.neighbour[h]], geo.hang_move_N)
    cO = np.einsum("n...j,nj->n...", grad[g.owner[h]], geo.hang_move_O)
    diff = cN - cO
    return _bcast(geo.delta[h], diff) * diff


def sn_grad(g: Grid, fld: Field, corrected: bool = True) -> np.ndarray:
    """fv/schemes/sn_grad.py:109-181."""
    psi = fld.data
    diff = psi[g.neighbour] - psi[g.owner]
    base = _bcast(g.geo.delta, diff) * diff
    if not corrected or len(g.geo.hang_idx) == 0:
        return base
    out = base.copy()
    out[g.geo.hang_idx] += sn_grad_hang_correction(g, lsq_grad(g, fld))
    return out


@dataclass
class FaceValues:
    single: np.ndarray
    bnd: np.ndarray


def interpolate(g: Grid, fld: Field) -> FaceValues:
    """fv/schemes/interpolate.py:20-63 + boundary_ops.py:545-593."""
    base = linear_face_values(g, fld.data)
    bnd = fld.data[g.bnd_owner].copy()  # zero-gradient extrapolation by default
    for b in batches(g, fld):
        bnd[b.mask] = bc_state(g, fld, b)[3]
    if len(g.geo.hang_idx):
        single = correct_face_values(g, base, lsq_grad(g, fld))
    else:
        single = base
    return FaceValues(single, bnd)


def fvc_div(g: Grid, phi: FaceValues) -> np.ndarray:
    """fv/fvc/div.py:15-64."""
    d = _scatter(g.owner, phi.single, g.num_cells)
    d = d + _scatter(g.neighbour, -phi.single, g.num_cells)
    d = d + _scatter(g.bnd_owner, phi.bnd, g.num_cells)
    return d / _bcast(g.cell_volumes, d)


def reconstruct(g: Grid, phi: FaceValues) -> np.ndarray:
    """fv/fvc/reconstruct.py:10-107: U_P = (sum_f |Sf| n n^T)^-1 sum_f n phi_f
    over every face of the cell (owner and neighbour with the same sign);
    the tensor is diagonal on axis-aligned faces."""
    def acc(cells, Sf, flux):
        mag = np.sqrt(np.sum(Sf * Sf, axis=1))
        n = Sf / mag[:, None]
        return _scatter(cells, flux[:, None] * n, g.num_cells), _scatter(cells, mag[:, None] * n * n, g.num_cells)
    num = np.zeros((g.num_cells, 3))
    den = np.zeros((g.num_cells, 3))
    for cells, Sf, flux in ((g.owner, g.Sf, phi.single), (g.neighbour, g.Sf, phi.single),
                            (g.bnd_owner, g.bnd_Sf, phi.bnd)):
        a, b = acc(cells, Sf, flux)
        num, den = num + a, den + b
    return num / den


def covered_domain(g: Grid, fld: Field) -> np.ndarray:
    cov = np.zeros(g.num_bnd, bool)
    for b in batches(g, fld):
        cov |= b.mask
    return cov


def flux_from_face_velocity(g: Grid, Uf: FaceValues, U: Field) -> FaceValues:
    """fv/flux.py:32-71: domain faces without a velocity condition carry no
    flux; immersed sides always take U_f . Sf."""
    single = np.sum(Uf.single * g.Sf, axis=1)
    dom = np.sum(Uf.bnd * g.bnd_Sf, axis=1)
    dom = np.where(~covered_domain(g, U) & g.domain_mask, 0.0, dom)
    return FaceValues(single, dom)


def correct_flux(g: Grid, U: Field, phi: FaceValues | None = None) -> FaceValues:
    """fv/flux.py:132-162 (update_internal=True when phi is None)."""
    if phi is None:
        phi = flux_from_face_velocity(g, interpolate(g, U), U)
    for b in batches(g, U):
        phi.bnd[b.mask] = np.sum(bc_state(g, U, b)[3] * b.Sf, axis=1)
    return phi


def ddt_coefficients(scheme: str, have_older: bool, dt: float, dt0: float | None):
    """fv/schemes/ddt.py:13-58."""
    if scheme == "euler" or not have_older:
        return 1.0, 1.0, 0.0
    a = 1.0 + dt / (dt + dt0)
    c = (dt / dt0) * dt / (dt + dt0)
    return a, a + c, c


def ddt_corr(g: Grid, U: Field, phi_old: np.ndarray, abc, phi_older=None) -> np.ndarray:
    """fv/fvc/ddt_corr.py:13-78."""
    w = g.geo.w[:, None]
    U0f = w * U.old[g.owner] + (1.0 - w) * U.old[g.neighbour]
    corr = phi_old - np.sum(U0f * g.Sf, axis=1)
    _, b, c = abc
    hist = b * corr
    if c != 0.0:
        U00f = w * U.older[g.owner] + (1.0 - w) * U.older[g.neighbour]
        hist = hist - c * (phi_older - np.sum(U00f * g.Sf, axis=1))
    eps = np.finfo(np.float64).tiny
    coupling = 1.0 - np.minimum(np.abs(corr) / (np.abs(phi_old) + eps), 1.0)
    return coupling * hist / g.dt


# ---------------------------------------------------------------------------
# Implicit operators  (fv/fvm/*.py, fv/schemes/div.py)
# ---------------------------------------------------------------------------
def interpolate_gamma(g: Grid, gamma, harmonic: bool = False):
    """fv/fvm/laplacian.py:38-73: linear, or distance-weighted harmonic
    (series resistances 1 / ((1-w)/gamma_O + w/gamma_N), evaluated scaled by
    max(gamma_O, gamma_N); a zero diffusivity blocks the face)."""
    if harmonic:
        ok = np.all(np.isfinite(gamma)) and np.all(np.asarray(gamma) >= 0)
        if not ok:
            raise ValueError("Harmonic diffusivity must be finite and nonnegative")
    if np.ndim(gamma) == 0:
        return float(gamma)
    go, gn, w = gamma[g.owner], gamma[g.neighbour], g.geo.w
    if not harmonic:
        return w * go + (1.0 - w) * gn
    scale = np.maximum(go, gn)
    safe = np.where(scale > 0, scale, 1.0)
    den = w * (go / safe) + (1.0 - w) * (gn / safe)
    return np.minimum(go, gn) / np.where(den > 0, den, 1.0)


def fvm_laplacian(g: Grid, gamma, fld: Field, corrected: bool = True, harmonic: bool = False) -> Matrix:
    """fv/fvm/laplacian.py:101-201."""
    geo = g.geo
    m = Matrix(g, fld.cshape)
    gamma_f = interpolate_gamma(g, gamma, harmonic)
    coeff = gamma_f * geo.mag_Sf * geo.delta
    coeff = np.broadcast_to(coeff, (g.num_faces,)).copy()
    m.upper = coeff.copy()
    m.lower = coeff.copy()
    m.diag = m.diag + _scatter(g.owner, -coeff, g.num_cells) + _scatter(g.neighbour, -coeff, g.num_cells)
    h = geo.hang_idx
    if corrected and len(h):
        corr = sn_grad_hang_correction(g, lsq_grad(g, fld))
        gh = gamma_f if np.ndim(gamma_f) == 0 else gamma_f[h]
        src = _bcast(gh * geo.mag_Sf[h], corr) * corr
        m.source = m.source + _scatter(g.owner[h], -src, g.num_cells) + _scatter(g.neighbour[h], src, g.num_cells)
        ffc = np.zeros((g.num_faces,) + fld.cshape)
        ffc[h] = src
        m.ffc = ffc
    for b in batches(g, fld):
        f, rv, rg, _ = bc_state(g, fld, b)
        gb = gamma if np.ndim(gamma) == 0 else gamma[b.cells]
        cond = gb * b.mag_Sf
        c_value = cond * value_grad_coeff(b, f)
        c_grad = cond * (1.0 - f)
        src = _bcast(c_value, rv) * rv + _bcast(c_grad, rg) * rg
        m.diag = m.diag + _scatter(b.cells, -c_value, g.num_cells)
        m.source = m.source + _scatter(b.cells, -src, g.num_cells)
    return m


# schemes/div.py:145-196 -- TVD limiters psi(r)
TVD_LIMITERS = {
    "vanleer": lambda r: (r + np.abs(r)) / (1.0 + np.abs(r)),
    "minmod": lambda r: np.maximum(np.minimum(1.0, r), 0.0),
    "superbee": lambda r: np.maximum(np.maximum(np.minimum(2.0 * r, 1.0), np.minimum(r, 2.0)), 0.0),
    "monotonized_central": lambda r: np.maximum(np.minimum(np.minimum(2.0 * r, 0.5 * r + 0.5), 2.0), 0.0),
    "limitedlinear": lambda r: np.maximum(np.minimum(2.0 * r, 1.0), 0.0),
    "koren": lambda r: np.maximum(np.minimum(np.minimum(2.0 * r, (2.0 + r) / 3.0), 2.0), 0.0),
}


def tvd_r(g: Grid, fld: Field, grad: np.ndarray, upwind_owner: np.ndarray) -> np.ndarray:
    """schemes/div.py:199-340: r = 2 (grad_U . d) / (psi_N - psi_O) - 1 (for
    vectors the component sums of (psi_N - psi_O) . (grad_U d) over
    |psi_N - psi_O|^2), with r = 2000 sign(num) sign(den) - 1 where
    |num| >= 1000 |den| (steep)."""
    d = g.cell_centers[g.neighbour] - g.cell_centers[g.owner]
    gU = np.where(_bcast(upwind_owner, grad[g.owner]), grad[g.owner], grad[g.neighbour])
    dpsi = fld.data[g.neighbour] - fld.data[g.owner]
    gd = np.einsum("n...j,nj->n...", gU, d)
    if fld.cshape == ():
        den, num = dpsi, gd
    else:
        den, num = np.sum(dpsi * dpsi, axis=1), np.sum(dpsi * gd, axis=1)
    steep = np.abs(num) >= 1000.0 * np.abs(den)
    safe = np.where(steep, 1.0, den)
    return np.where(steep, 2.0 * 1000.0 * np.sign(num) * np.sign(den) - 1.0, 2.0 * num / safe - 1.0)


def fvm_div(g: Grid, phi: FaceValues, fld: Field, scheme: str, grad_scheme: str = "leastsquare") -> Matrix:
    """fv/fvm/div.py:15-73 with schemes/div.py:36-461 (upwind, linear and the
    TVD schemes: upwind matrix plus the deferred-correction source
    -phi psi(r) (psi_linear - psi_upwind) per face)."""
    geo = g.geo
    m = Matrix(g, fld.cshape)
    ph = phi.single
    if scheme == "upwind" or scheme in TVD_LIMITERS:
        pos, neg = np.maximum(ph, 0.0), np.minimum(ph, 0.0)
        upper, lower, dO, dN = neg, -pos, pos, -neg
        sface = np.zeros((g.num_faces,) + fld.cshape)
        if scheme in TVD_LIMITERS:
            grad = cell_grad(g, fld, grad_scheme)  # also corrects the hanging faces' linear values
            base = linear_face_values(g, fld.data)
            lin = correct_face_values(g, base, grad) if len(geo.hang_idx) else base
            up_own = ph > 0
            psi_up = np.where(_bcast(up_own, base), fld.data[g.owner], fld.data[g.neighbour])
            lim = TVD_LIMITERS[scheme](tvd_r(g, fld, grad, up_own))
            sface = -_bcast(ph * lim, lin) * (lin - psi_up)
    elif scheme == "linear":
        w = geo.w
        dO, upper = ph * w, ph * (1.0 - w)
        lower, dN = -ph * w, -ph * (1.0 - w)
        sface = np.zeros((g.num_faces,) + fld.cshape)
        if len(geo.hang_idx):
            base = linear_face_values(g, fld.data)
            corrected = correct_face_values(g, base, lsq_grad(g, fld))
            sface = -_bcast(ph, base) * (corrected - base)
    else:
        raise ValueError(f"unsupported div scheme {scheme}")
    m.upper = upper.copy()
    m.lower = lower.copy()
    m.diag = m.diag + _scatter(g.owner, dO, g.num_cells) + _scatter(g.neighbour, dN, g.num_cells)
    m.source = m.source + _scatter(g.owner, sface, g.num_cells) + _scatter(g.neighbour, -sface, g.num_cells)
    for b in batches(g, fld):
        f, rv, rg, _ = bc_state(g, fld, b)
        fv = _bcast(f, rv)
        bsrc = fv * rv + (1.0 - fv) * _bcast(b.mag_d, rg) * rg
        flux = phi.bnd[b.mask]
        m.diag = m.diag + _scatter(b.cells, flux * (1.0 - f), g.num_cells)
        m.source = m.source + _scatter(b.cells, -_bcast(flux, bsrc) * bsrc, g.num_cells)
    return m


def fvm_ddt(g: Grid, fld: Field, abc) -> Matrix:
    """fv/fvm/ddt.py:7-40."""
    a, b, c = abc
    m = Matrix(g, fld.cshape)
    vdt = g.cell_volumes / g.dt
    m.diag = m.diag + a * vdt
    hist = b * fld.old
    if c != 0.0:
        hist = hist - c * fld.older
    m.source = m.source + _bcast(vdt, hist) * hist
    return m


# ---------------------------------------------------------------------------
# Krylov solvers  (solvers/*.py)
# ---------------------------------------------------------------------------
@dataclass
class Stats:
    solver: str
    initial_residual: float
    final_residual: float
    iterations: int
    converged: bool


@dataclass
class SolverCfg:
    method: str = "cg"
    preconditioner: str = "none"
    tolerance: float = 1e-6
    rel_tolerance: float = 1e-6
    max_iter: int = 1000
    norm_type: str = "l_2"
    max_restart: int = 5


def _norm(v: np.ndarray, kind: str) -> float:
    return float(np.max(np.abs(v))) if kind == "l_inf" else float(np.sqrt(np.dot(v, v)))


def _inv_diag(m: Matrix, kind: str) -> np.ndarray:
    """solvers/preconditioners.py:76-96."""
    if kind == "none":
        return np.ones_like(m.diag)
    eps = np.finfo(np.float64).eps
    safe = np.where(m.diag == 0, 1.0, m.diag)
    return np.where(np.abs(m.diag) < eps, 1.0, 1.0 / safe)


# Host-timing aid: with threads > 0 the Krylov solves multiply through a CSR
# copy of the LDU matrix (built once per solve) on that many torch CPU
# threads.  Same products; each row is summed in column order instead of
# face order, so results agree with the LDU multiply to rounding.  Off (0)
# for every parity use: the expected answers use the LDU multiply.
_SPMV_THREADS = 0


def use_threaded_spmv(threads: int) -> None:
    global _SPMV_THREADS
    _SPMV_THREADS = int(threads)


def _operator(m: Matrix):
    """The multiply of a Krylov solve: the LDU multiply, or its CSR copy."""
    if _SPMV_THREADS <= 0:
        return m.multiply
    import torch

    torch.set_num_threads(_SPMV_THREADS)
    g, C = m.g, m.g.num_cells
    rows = np.concatenate([np.arange(C), g.owner, g.neighbour])
    cols = np.concatenate([np.arange(C), g.neighbour, g.owner])
    vals = np.concatenate([m.diag, m.upper, m.lower])
    order = np.lexsort((cols, rows))
    indptr = np.zeros(C + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=C), out=indptr[1:])
    A = torch.sparse_csr_tensor(torch.from_numpy(indptr), torch.from_numpy(cols[order].astype(np.int64)),
                                torch.from_numpy(vals[order]), size=(C, C))
    return lambda x: (A @ torch.from_numpy(np.ascontiguousarray(x))).numpy()


def cg_single(m: Matrix, b: np.ndarray, x: np.ndarray, cfg: SolverCfg, inv_d: np.ndarray, op=None):
    """solvers/cg.py:75-174."""
    mul = op or m.multiply
    r = b - mul(x)
    nr0 = _norm(r, cfg.norm_type)
    thresh = max(cfg.tolerance, cfg.rel_tolerance * nr0)
    if nr0 < thresh:
        return x, Stats("cg", nr0, nr0, 0, True)
    z = inv_d * r
    p = z.copy()
    rz_old = np.dot(r, z)
    final, conv, it = nr0, False, 0
    tmp = np.empty_like(p)
    # in-place updates: the same IEEE operations as x + alpha p, r - alpha Ap,
    # z + p beta, without a temporary allocation per operation
    for it in range(1, cfg.max_iter + 1):
        Ap = mul(p)
        alpha = rz_old / np.dot(p, Ap)
        np.add(x, np.multiply(p, alpha, out=tmp), out=x)
        np.subtract(r, np.multiply(Ap, alpha, out=tmp), out=r)
        final = _norm(r, cfg.norm_type)
        if final < thresh:
            conv = True
            break
        np.multiply(inv_d, r, out=z)
        rz_new = np.dot(r, z)
        np.add(z, np.multiply(p, rz_new / rz_old, out=p), out=p)
        rz_old = rz_new
    return x, Stats("cg", nr0, final, it, conv)


def bicgstab_single(m: Matrix, b: np.ndarray, x: np.ndarray, cfg: SolverCfg, inv_d: np.ndarray, op=None):
    """solvers/bicgstab.py:76-272."""
    mul = op or m.multiply
    r = b - mul(x)
    nr0 = _norm(r, cfg.norm_type)
    thresh = max(cfg.tolerance, cfg.rel_tolerance * nr0)
    if nr0 < thresh:
        return x, Stats("bicgstab", nr0, nr0, 0, True)

    def l2(v):
        return float(np.sqrt(np.dot(v, v)))

    def restart(r):
        return r.copy(), r.copy(), np.dot(r, r)

    r_hat, p, rho_old = restart(r)
    restarts, final, conv, it = 0, nr0, False, 0
    for it in range(1, cfg.max_iter + 1):
        y = inv_d * p
        v = mul(y)
        den = np.dot(r_hat, v)
        if abs(den) <= 1e-6 * l2(r_hat) * l2(v):
            if restarts >= cfg.max_restart:
                break
            restarts += 1
            r_hat, p, rho_old = restart(r)
            continue
        alpha = rho_old / den
        s = r - alpha * v
        final = _norm(s, cfg.norm_type)
        if final < thresh:
            x = x + alpha * y
            conv = True
            break
        z = inv_d * s
        t = mul(z)
        dom = np.dot(t, t)
        if abs(dom) <= 1e-6 * l2(t) * l2(t):
            if restarts >= cfg.max_restart:
                break
            restarts += 1
            r_hat, p, rho_old = restart(r)
            continue
        omega = np.dot(t, s) / dom
        if abs(omega) <= 1e-6:
            if restarts >= cfg.max_restart:
                break
            restarts += 1
            r_hat, p, rho_old = restart(r)
            continue
        x = x + alpha * y + omega * z
        r = s - omega * t
        final = _norm(r, cfg.norm_type)
        if final < thresh:
            conv = True
            break
        rho_new = np.dot(r_hat, r)
        if abs(rho_new) <= 1e-6 * l2(r_hat) * l2(r):
            if restarts >= cfg.max_restart:
                break
            restarts += 1
            r_hat, p, rho_old = restart(r)
            continue
        beta = (rho_new / rho_old) * (alpha / omega)
        p = r + beta * (p - omega * v)
        rho_old = rho_new
    return x, Stats("bicgstab", nr0, final, it, conv)


def solve(m: Matrix, x0: np.ndarray, cfg: SolverCfg):
    """solvers/krylov.py:34-82 -- one shared preconditioner, per component."""
    inv_d = _inv_diag(m, cfg.preconditioner)
    single = cg_single if cfg.method == "cg" else bicgstab_single
    op = _operator(m)
    if x0.ndim == 1:
        x, st = single(m, m.source, x0.copy(), cfg, inv_d, op)
        return x, (st,)
    xs, sts = [], []
    for k in range(x0.shape[1]):
        x, st = single(m, m.source[:, k], x0[:, k].copy(), cfg, inv_d, op)
        xs.append(x)
        sts.append(st)
    return np.stack(xs, axis=1), tuple(sts)


def initial_residual(m: Matrix, psi: np.ndarray) -> float:
    """algorithms/utils/residual.py:12-36."""
    res = m.multiply(psi) - m.source
    return float(np.linalg.norm(res.ravel()) / (np.linalg.norm(m.source.ravel()) + np.finfo(np.float64).eps))


# ---------------------------------------------------------------------------
# Pressure-velocity coupling  (algorithms/utils/pressure_correction.py)
# ---------------------------------------------------------------------------
def adjust_phi(g: Grid, phi: FaceValues, U: Field, p: Field) -> bool:
    """fv/adjust_phi.py:44-127."""
    if any(bc.kind == "dirichlet" for bc in p.bcs.values()):
        return False
    mass_in = fixed_out = adj_out = 0.0
    for patch, bc in U.bcs.items():
        if patch not in PATCHES:  # immersed patches are not balanced
            continue
        m = g.patch_mask(patch)
        if not m.any():
            continue
        ph = phi.bnd[m]
        mass_in -= ph[ph < 0].sum()
        if bc.kind == "dirichlet":
            fixed_out += ph[ph > 0].sum()
        else:
            adj_out += ph[ph > 0].sum()
    if adj_out <= 0.0:
        return False
    corr = (mass_in - fixed_out) / adj_out
    for patch, bc in U.bcs.items():
        if bc.kind == "dirichlet" or patch not in PATCHES:
            continue
        m = g.patch_mask(patch)
        ph = phi.bnd[m]
        phi.bnd[m] = np.where(ph > 0, ph * corr, ph)
    return True


def needs_reference(p: Field) -> bool:  # algorithms/utils/reference_value.py:14-30
    return not any(bc.kind == "dirichlet" for bc in p.bcs.values())


def solve_pressure(g, p, rAtU, div_hbya, cfg_p, cfg_final, n_nonorth, ref, corrected=True, last=None,
                   harmonic=False):
    """pressure_correction.py:68-160 + reference_value.py:33-51.  ``last``
    (a list) receives the initial residual of the last pass."""
    div_src = div_hbya * g.cell_volumes
    first = None
    for corr in range(n_nonorth + 1):
        mat = -fvm_laplacian(g, rAtU, p, corrected, harmonic)
        mat.source = mat.source - div_src
        if ref is not None:
            cell, val = ref
            rd = mat.diag[cell]
            mat.diag[cell] += rd
            mat.source[cell] += rd * val
        smat = mat.with_fixed_values(*immersed_constraints(g, p))  # core/equation.py:42-47
        res0 = initial_residual(smat, p.data)
        if first is None:
            first = res0
        if last is not None:
            last[:] = [res0]
        cfg = cfg_final if (cfg_final is not None and corr == n_nonorth) else cfg_p
        p.data, st = solve(smat, p.data, cfg)
    return mat, st, first


def correct_phi(g, phi_hbya: FaceValues, pmat: Matrix, p: Field, rAtU: np.ndarray) -> FaceValues:
    """pressure_correction.py:197-217, :270-335 (snapped cells closed by
    their mass balance)."""
    flux_p = (-pmat).flux(p.data)
    out = FaceValues(phi_hbya.single - flux_p, phi_hbya.bnd.copy())
    for b in batches(g, p):
        f, rv, rg, _ = bc_state(g, p, b)
        coeff = value_grad_coeff(b, f)
        sng = coeff * (rv - p.data[b.cells]) + (1.0 - f) * rg
        out.bnd[b.mask] = out.bnd[b.mask] - rAtU[b.cells] * b.mag_Sf * sng
    close_fixed_cell_flux(g, out, p)
    return out


def simplec_rAtU(UEqn: Matrix, rAU: np.ndarray, bounded: bool = True) -> np.ndarray:
    """pressure_correction.py:163-193."""
    inv = 1.0 / rAU - UEqn.H1()
    if bounded:
        inv = np.maximum(inv, 0.1 / rAU)
    return 1.0 / inv


def apply_simplec(g: Grid, phi_h: FaceValues, HbyA: np.ndarray, p: Field, rAU: np.ndarray, rAtU: np.ndarray,
                  corrected: bool = True, publish=None, grad_scheme: str = "leastsquare"):
    """pressure_correction.py:222-267: phiHbyA += (rAtU-rAU)_f |Sf| snGrad(p)
    (boundary faces: the condition's normal gradient); HbyA -= (rAU-rAtU) grad p.
    The reference updates the registered phiHbyA in place before grad(p), so
    conditions that read it see the updated flux there: ``publish(out)``
    registers the new flux at that point."""
    d_r = rAtU - rAU
    w = g.geo.w
    d_r_f = w * d_r[g.owner] + (1.0 - w) * d_r[g.neighbour]
    out = FaceValues(phi_h.single + d_r_f * g.geo.mag_Sf * sn_grad(g, p, corrected), phi_h.bnd.copy())
    for b in batches(g, p):
        f, rv, rg, _ = bc_state(g, p, b)
        coeff = value_grad_coeff(b, f)
        sng = coeff * (rv - p.data[b.cells]) + (1.0 - f) * rg
        out.bnd[b.mask] = out.bnd[b.mask] + d_r[b.cells] * b.mag_Sf * sng
    if publish is not None:
        publish(out)
    return out, HbyA - (rAU - rAtU)[:, None] * cell_grad(g, p, grad_scheme)


def _residual_control(rc: dict) -> dict:
    """meta/config.py:364-388: a float is an absolute tolerance."""
    out = {}
    for k, v in (rc or {}).items():
        if isinstance(v, (int, float)):
            out[k] = (float(v), 0.0)
        else:
            out[k] = (float(v["tolerance"]), float(v.get("rel_tolerance", 0.0)))
    return out


def _residual_satisfied(res: float, tol: float, rtol: float, init: float) -> bool:
    """algorithms/utils/residual.py:56-87."""
    return res < tol or (rtol > 0.0 and res < rtol * init)


# ---------------------------------------------------------------------------
# Case description and drivers  (algorithms/simple.py, piso.py, pimple.py)
# ---------------------------------------------------------------------------
def _bc_from_cfg(entry: dict, cshape) -> BC:
    """boundaries/factory.py:25-69."""
    kind = entry["type"]
    if kind in ("dirichlet", "neumann"):
        v = entry.get("value", 0.0)
        return BC(kind, np.asarray(v, np.float64))
    if kind == "inlet_outlet":
        if entry.get("value") is None:
            raise ValueError("inletOutlet requires an inlet value")
        return BC(kind, np.broadcast_to(np.asarray(entry["value"], np.float64), tuple(cshape)).copy())
    if kind in ("empty", "slip", "fixed_flux_pressure"):
        return BC(kind)
    raise ValueError(f"unsupported boundary type {kind}")


def field_from_cfg(g: Grid, name: str, cond: dict, cshape) -> Field:
    data = np.zeros((g.num_cells,) + tuple(cshape))
    data[:] = np.asarray(cond.get("internal", 0.0), np.float64)
    bcs: dict[str, BC] = {}
    for entry in cond.get("boundary", {}).values():
        bc = _bc_from_cfg(entry, cshape)
        for patch in entry["patches"]:
            bcs[patch] = bc
    return Field(name, data, bcs)


def grid_from_cfg(cfg: dict) -> Grid:
    fx = cfg["fluxel"]
    regions = [(r["min"], r["max"], r["level"]) for r in fx.get("refinement_regions", []) or []]
    return Grid(fx["domain"]["lower"], fx["domain"]["upper"], fx["root_resolution"], regions,
                fx.get("n_leaf_refinement", 0), cfg["simulator"]["control"]["deltaT"],
                fx.get("immersed_bodies", []) or [], int(fx.get("target_level", 0) or 0))


def _solver_cfg(d: dict) -> SolverCfg:
    keys = ("method", "preconditioner", "tolerance", "rel_tolerance", "max_iter", "norm_type", "max_restart")
    return SolverCfg(**{k: d[k] for k in keys if k in d})


class Case:
    """SIMPLE / PISO drivers over the restated operators."""

    def __init__(self, cfg: dict, grid: Grid | None = None):
        sim = cfg["simulator"]
        self.g = g = grid if grid is not None else grid_from_cfg(cfg)
        cond = sim["conditions"]
        self.U = field_from_cfg(g, "U", cond["U"], (3,))
        self.p = field_from_cfg(g, "p", cond["p"], ())
        self.nu = float(sim["properties"]["transport"]["nu"])
        fvs = sim["fvSolution"]
        self.alg = fvs["algorithm"]
        self.kind = self.alg["type"]
        self.adjust = fvs.get("adjustPhi", True)
        self.solvers = {k: _solver_cfg(v) for k, v in fvs["solvers"].items()}
        sch = sim.get("fvSchemes", {}) or {}

        def look(group, key, default):
            d = sch.get(group) or {}
            return d.get(key, d.get("default", default))

        self.div_scheme = look("divSchemes", "div(phi, U)", "upwind")
        self.grad_p = look("gradSchemes", "grad(p)", "leastsquare")
        self.grad_U = look("gradSchemes", "grad(U)", "leastsquare")
        lap_u = look("laplacianSchemes", "laplacian(U)", "corrected")
        lap_p = look("laplacianSchemes", "laplacian(p)", "corrected")
        uncorr = ("uncorrected", "Gauss linear uncorrected", "Gauss harmonic uncorrected")
        self.lap_u_corr = lap_u not in uncorr
        self.lap_p_corr = lap_p not in uncorr
        self.lap_u_harm = lap_u.startswith("Gauss harmonic")
        self.lap_p_harm = lap_p.startswith("Gauss harmonic")
        self.ddt_scheme = look("ddtSchemes", "ddt(U)", "euler")
        self.consistent = bool(self.alg.get("consistent", False)) and self.kind in ("simple", "pimple")
        self.sn_p_corr = look("snGradSchemes", "snGrad(p)", "corrected") != "uncorrected"
        self.rc = _residual_control(self.alg.get("residualControl", {}))
        self.outer_converged = False
        self.simple_init: dict[str, float] = {}  # simple.py:277-285: first recorded value is kept
        self.simple_cur: dict[str, float] = {}
        self.ref = None
        if needs_reference(self.p):
            self.ref = (int(self.alg["pRefCell"]), float(self.alg["pRefValue"]))
        # the algorithms register U, p, phi, phiHbyA, rAU, rAtU and HbyA at
        # construction, zero-initialised (piso.py:143-176, simple.py:155-190)
        C = g.num_cells
        self.phi = FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
        self.phi_h = FaceValues(np.zeros(g.num_faces), np.zeros(g.num_bnd))
        self.HbyA_d, self.rAU_d, self.rAtU_d = np.zeros((C, 3)), np.zeros(C), np.zeros(C)
        g.registry.update(U=lambda: self.U, p=lambda: self.p, phi=lambda: self.phi, phiHbyA=lambda: self.phi_h,
                          HbyA=lambda: self.HbyA_d, rAU=lambda: self.rAU_d, rAtU=lambda: self.rAtU_d)
        self.phi = correct_flux(g, self.U)
        self.U.old = self.U.data.copy()
        self.U.older = None
        self.phi_old = self.phi.single.copy()
        self.phi_older = None
        self.prev_dt = None
        self.stats: dict[str, Any] = {}

    def _momentum(self, transient: bool):
        g, U = self.g, self.U
        m = (fvm_div(g, self.phi, U, self.div_scheme, self.grad_U)
             - fvm_laplacian(g, self.nu, U, self.lap_u_corr, self.lap_u_harm))
        if transient:
            abc = ddt_coefficients(self.ddt_scheme, U.older is not None, g.dt, self.prev_dt)
            m = fvm_ddt(g, U, abc) + m
        return m

    def _pressure_pass(self, UEqn: Matrix, transient: bool, final: bool, abc=None):
        g, U, p = self.g, self.U, self.p
        rAU = 1.0 / UEqn.A()
        HbyA = Field("HbyA", rAU[:, None] * UEqn.H(U.data))
        self.rAU_d, self.HbyA_d = rAU, HbyA.data
        Hf = interpolate(g, HbyA)
        phi_h = flux_from_face_velocity(g, Hf, U)
        if transient:
            rAUf = linear_face_values(g, rAU)
            phi_h.single = phi_h.single + rAUf * ddt_corr(g, U, self.phi_old, abc, self.phi_older)
        for b in batches(g, U):  # flux.py:74-87 constrainHbyA
            if not U.bcs[b.patch].assignable:
                phi_h.bnd[b.mask] = np.sum(bc_state(g, U, b)[3] * b.Sf, axis=1)
        if self.adjust:
            adjust_phi(g, phi_h, U, p)
        self.phi_h = phi_h
        if self.consistent:
            rAtU = simplec_rAtU(UEqn, rAU)
            self.rAtU_d = rAtU

            def publish(out):
                self.phi_h = out
            phi_h, HbyA_d = apply_simplec(g, phi_h, HbyA.data, p, rAU, rAtU, self.sn_p_corr, publish, self.grad_p)
            HbyA = Field("HbyA", HbyA_d)
            self.HbyA_d = HbyA_d
            rAU = rAtU
        else:
            self.rAtU_d = rAU
        div_h = fvc_div(g, phi_h)
        cfg_final = None
        if final:
            cfg_final = self.solvers.get("pFinal", self.solvers.get("p_final"))
        last = []
        pmat, st, res0 = solve_pressure(g, p, rAU, div_h, self.solvers["p"], cfg_final,
                                        int(self.alg.get("nNonOrthogonalCorrectors", 0)), self.ref,
                                        self.lap_p_corr, last, self.lap_p_harm)
        self._last_res0 = last[0]
        self.stats["p"] = st
        self.stats["p_initial_residual"] = res0
        self.stats["p_last_initial_residual"] = self._last_res0
        return HbyA, rAU, phi_h, pmat

    def step_simple(self):
        """algorithms/simple.py:287-408."""
        g, U, p = self.g, self.U, self.p
        eq = self.alg.get("relaxationFactors", {}).get("equations", {})
        aU, ap = float(eq["U"]), float(eq["p"])
        UEqn = self._momentum(False)
        A_old = UEqn.diag.copy()
        UEqn.diag = A_old / aU
        UEqn.source = UEqn.source + (((1.0 - aU) / aU) * A_old)[:, None] * U.data
        gp = cell_grad(g, p, self.grad_p)
        pred = UEqn.copy()
        pred.source = UEqn.source - g.cell_volumes[:, None] * gp
        pred = pred.with_fixed_values(*immersed_constraints(g, U))
        if "U" in self.rc:  # simple.py:319-323
            self._simple_record("U", initial_residual(pred, U.data))
        U.data, self.stats["U"] = solve(pred, U.data, self.solvers["U"])
        UEqn = UEqn.with_fixed_values(*immersed_constraints(g, U))  # simple.py:335
        p_old = p.data.copy()
        HbyA, rAU, phi_h, pmat = self._pressure_pass(UEqn, False, False)
        if "p" in self.rc:  # simple.py:379-380
            self._simple_record("p", self.stats["p_initial_residual"])
        self.phi = correct_phi(g, phi_h, pmat, p, rAU)
        p.data = p_old + ap * (p.data - p_old)
        apply_immersed_values(g, p)
        U.data = HbyA.data - rAU[:, None] * cell_grad(g, p, self.grad_p)
        apply_immersed_values(g, U)
        if "phi" in self.rc:  # simple.py:394-395
            self._simple_record("phi", float(np.linalg.norm(fvc_div(g, self.phi))))

    def _simple_record(self, name: str, r: float) -> None:
        self.simple_init.setdefault(name, float(r))
        self.simple_cur[name] = float(r)

    def simple_converged(self) -> bool:
        """simple.py:241-271: every residualControl entry below its absolute
        tolerance (relative tolerance ignored); False when none is set."""
        if not self.rc:
            return False
        for name, (tol, _rtol) in self.rc.items():
            if name not in self.simple_cur or name not in self.simple_init:
                return False
            if not _residual_satisfied(self.simple_cur[name], tol, 0.0, self.simple_init[name]):
                return False
        return True

    def step_piso(self):
        """algorithms/piso.py:227-326."""
        g, U, p = self.g, self.U, self.p
        abc = ddt_coefficients(self.ddt_scheme, U.older is not None, g.dt, self.prev_dt)
        UEqn = self._momentum(True)
        gp = cell_grad(g, p, self.grad_p)
        pred = UEqn.copy()
        pred.source = UEqn.source - g.cell_volumes[:, None] * gp
        pred = pred.with_fixed_values(*immersed_constraints(g, U))
        U.data, self.stats["U"] = solve(pred, U.data, self.solvers["U"])
        UEqn = UEqn.with_fixed_values(*immersed_constraints(g, U))  # piso.py:257
        n_corr = int(self.alg.get("nCorrectors", 2))
        for i in range(n_corr):
            HbyA, rAU, phi_h, pmat = self._pressure_pass(UEqn, True, i == n_corr - 1, abc)
            self.phi = correct_phi(g, phi_h, pmat, p, rAU)
            U.data = HbyA.data - rAU[:, None] * cell_grad(g, p, self.grad_p)
            apply_immersed_values(g, U)
        # field.py:241-255 -- history advance once per step
        U.older, U.old = U.old, U.data.copy()
        self.phi_older, self.phi_old = self.phi_old, self.phi.single.copy()
        self.prev_dt = g.dt

    def step_pimple(self):
        """algorithms/pimple.py:290-454 (laminar): outer correctors around the
        PISO loop, residualControl exit after one more (final) outer pass."""
        g, U, p = self.g, self.U, self.p
        abc = ddt_coefficients(self.ddt_scheme, U.older is not None, g.dt, self.prev_dt)
        n_outer = int(self.alg.get("nOuterCorrectors", 1))
        n_corr = int(self.alg.get("nCorrectors", 2))
        init, cur = {}, {}
        self.outer_converged = False
        self.stats["outer"] = 0

        def record(name, r):
            init.setdefault(name, r)
            cur[name] = r

        def converged(allow_relative):
            if not self.rc:
                return False
            for name, (tol, rtol) in self.rc.items():
                if name not in cur or name not in init:
                    return False
                if not _residual_satisfied(cur[name], tol, rtol if allow_relative else 0.0, init[name]):
                    return False
            return True

        for outer in range(n_outer):
            final_after = 0 < outer < n_outer - 1 and converged(outer > 1)
            if final_after:
                self.outer_converged = True
            UEqn = self._momentum(True)
            pred = UEqn.copy()
            pred.source = UEqn.source - g.cell_volumes[:, None] * cell_grad(g, p, self.grad_p)
            pred = pred.with_fixed_values(*immersed_constraints(g, U))
            if "U" in self.rc:
                record("U", initial_residual(pred, U.data))
            U.data, self.stats["U"] = solve(pred, U.data, self.solvers["U"])
            UEqn = UEqn.with_fixed_values(*immersed_constraints(g, U))  # pimple.py:346
            for i in range(n_corr):
                HbyA, rAtU, phi_h, pmat = self._pressure_pass(UEqn, True, i == n_corr - 1, abc)
                if "p" in self.rc:
                    if i == 0:
                        record("p", self.stats["p_initial_residual"])
                    cur["p"] = self.stats["p_last_initial_residual"]
                self.phi = correct_phi(g, phi_h, pmat, p, rAtU)
                U.data = HbyA.data - rAtU[:, None] * cell_grad(g, p, self.grad_p)
                apply_immersed_values(g, U)
            if "phi" in self.rc:
                record("phi", float(np.linalg.norm(fvc_div(g, self.phi))))
            self.stats["outer"] = outer + 1
            if final_after:
                break
        U.older, U.old = U.old, U.data.copy()
        self.phi_older, self.phi_old = self.phi_old, self.phi.single.copy()
        self.prev_dt = g.dt

    def step(self):
        if self.kind == "simple":
            self.step_simple()
        elif self.kind == "piso":
            self.step_piso()
        elif self.kind == "pimple":
            self.step_pimple()
        else:
            raise ValueError(f"unsupported algorithm {self.kind}")


# ---------------------------------------------------------------------------
# Force coefficients  (post/forces/*.py)
# ---------------------------------------------------------------------------
def _unit(v: np.ndarray) -> np.ndarray:
    n = np.linalg.norm(v)
    if n == 0:
        raise ValueError("vector is zero vector")
    return v / n


def orthonormal_coord(drag_dir, lift_dir=None, pitch_axis=None):
    """post/forces/coord.py:44-64 -> (e1, e2, e3)."""
    e1 = _unit(np.asarray(drag_dir, np.float64))
    if lift_dir is not None:
        lift = np.asarray(lift_dir, np.float64)
        e3 = _unit(lift - e1 * (lift @ e1))
        return e1, np.cross(e3, e1), e3
    pitch = np.asarray(pitch_axis, np.float64)
    e2 = _unit(pitch - e1 * (pitch @ e1))
    return e1, e2, np.cross(e1, e2)


def force_coeffs(force, moment, coord, force_scale, moment_scale) -> dict:
    """post/forces/coeffs.py:46-74."""
    e1, e2, e3 = coord
    Cd, Cs, Cl = (float(force @ e) / force_scale for e in (e1, e2, e3))
    Cr, Cp, Cy = (float(moment @ e) / moment_scale for e in (e1, e2, e3))
    return dict(Cd=Cd, Cd_f=0.5 * Cd + Cr, Cd_r=0.5 * Cd - Cr, Cl=Cl, Cl_f=0.5 * Cl + Cp, Cl_r=0.5 * Cl - Cp,
                CmPitch=Cp, CmRoll=Cr, CmYaw=Cy, Cs=Cs, Cs_f=0.5 * Cs + Cy, Cs_r=0.5 * Cs - Cy)


def _side_values(g: Grid, fld: Field, idx: np.ndarray) -> np.ndarray:
    """sampling.py:28-65: the condition's boundary value on boundary faces
    idx, the cell value where the patch has no condition."""
    out = fld.data[g.bnd_owner[idx]].copy()
    pos = {int(k): i for i, k in enumerate(idx)}
    for b in batches(g, fld):
        faces = np.nonzero(b.mask)[0]
        hit = np.array([k in pos for k in faces], bool)
        if hit.any():
            psi_b = bc_state(g, fld, b)[3]
            out[[pos[int(k)] for k in faces[hit]]] = psi_b[hit]
    return out


def patch_force(g: Grid, U: Field, p: Field, nu, patch: str, rho: float, cofr) -> tuple[np.ndarray, np.ndarray]:
    """integration.py:61-187 + sampling.py:68-168: (force, moment) of one
    immersed patch, summed over its kept axis-projected side samples."""
    force, moment = np.zeros(3), np.zeros(3)
    if g.num_immersed == 0:
        return force, moment
    Bd, Fi = g.num_domain_bnd, g.num_immersed
    k_body = g.patch_ids[patch] - N_DOMAIN_PATCHES
    nu_c = np.full(g.num_cells, float(nu)) if np.ndim(nu) == 0 else np.asarray(nu)
    for side in (0, 1):
        pid = g.ap.owner_patch if side == 0 else g.ap.neighbour_patch
        sel = np.nonzero(pid == k_body)[0]
        if len(sel) == 0:
            continue
        idx = Bd + side * Fi + sel
        cells, Sf, mag_d = g.bnd_owner[idx], g.bnd_Sf[idx], g.bnd_mag_d[idx]
        n_hat = Sf / np.linalg.norm(Sf, axis=1)[:, None]
        n_out = (g.ap.owner_normal if side == 0 else g.ap.neighbour_normal)[sel]
        n_s = -n_out
        keep = np.sum(n_hat * n_s, axis=1) > 0.0
        if not keep.any():
            continue
        p_b = _side_values(g, p, idx[keep])
        U_b = _side_values(g, U, idx[keep])
        U_c = U.data[cells[keep]]
        ns = n_s[keep]
        wall = mag_d[keep] * np.maximum(np.sum(n_hat[keep] * ns, axis=1), 1.0e-12)
        dUdn = (U_b - U_c) / wall[:, None]
        gradU = dUdn[:, :, None] * ns[:, None, :]
        stress = rho * nu_c[cells[keep]][:, None, None] * (gradU + np.transpose(gradU, (0, 2, 1)))
        fp = (rho * p_b)[:, None] * Sf[keep]
        fv = -np.sum(stress * Sf[keep][:, None, :], axis=-1)
        ff = fp + fv
        pts = g.cell_centers[cells[keep]] + mag_d[keep][:, None] * n_hat[keep]
        force += ff.sum(axis=0)
        moment += np.cross(pts - np.asarray(cofr, np.float64), ff).sum(axis=0)
    return force, moment


def continuity_error(g: Grid, phi: FaceValues, dt: float) -> tuple[float, float]:
    """post/diagnostics/continuity.py:12-49: dt * volume-weighted mean of
    |div phi| (local) and of div phi (global)."""
    d = fvc_div(g, phi)
    V = g.cell_volumes
    return float(dt * np.sum(np.abs(d) * V) / np.sum(V)), float(dt * np.sum(d * V) / np.sum(V))


def continuity_max(g: Grid, phi: FaceValues) -> float:
    return float(np.max(np.abs(fvc_div(g, phi))))


def _unused(*_a):  # keeps linters quiet about math import on some paths
    return math.inf
