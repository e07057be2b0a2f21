"""Solver row order and the brick operator layout of the PCG direction kernel.

The Krylov solve of a symmetric operator runs on its own row order:
  * first every "brick" -- 64 leaves of one level that fill an aligned
    4x4x4 block of that level's lattice -- as one slice of 64 rows (one
    wavefront), lane = x + 4 y + 16 z for the cell's in-block coordinates
    (so its +-x, +-y, +-z neighbours in the brick are lanes +-1, +-4, +-16);
  * then every other cell, in the grid's order.
Bricks are ordered by their first cell in the grid's order, so Morton
locality is kept.

Inside a brick every face between two of its cells joins lanes whose
indices follow from the lane alone, so the direction kernel reads the
neighbour's value with a cross-lane permute instead of a memory gather, and
(for symmetric operators) keeps one value per face: the +axis face of each
lane.  A brick's side toward another brick of the same level is implicit
too (the neighbour is lane "mirror(lane)" of that brick), and only its
-axis face values are stored (16 per side).  A side toward anything else (a
2:1 level change) is an explicit block of 64 slots: 16 boundary cells x up
to 4 faces, slot = 4 q + k.  Rows outside bricks keep the sliced-ELL layout
(on solver rows), which also serves every non-direction kernel.

Everything here is built once per grid with tensor operations on the grid's
device (or on the CPU, for tests).  Direction d of a brick is 2 axis + side
(0: -axis, 1: +axis); q of a boundary cell on an axis-a side is
c_b + 4 c_c with b = (a+1) % 3, c = (a+2) % 3.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

INT32_MIN = -(2 ** 31)


def brick_lane(c: torch.Tensor) -> torch.Tensor:
    """Lane of a cell with in-brick coordinates c [n,3] in 0..3."""
    return c[:, 0] + 4 * c[:, 1] + 16 * c[:, 2]


def face_position(c: torch.Tensor, a: torch.Tensor) -> torch.Tensor:
    """q of a cell on an axis-a side of its brick."""
    b = (a + 1) % 3
    cc = (a + 2) % 3
    ar = torch.arange(c.shape[0], device=c.device)
    return c[ar, b] + 4 * c[ar, cc]


@dataclass
class SolverLayout:
    n_bricks: int
    n_expl: int
    row_perm: torch.Tensor    # [C] solver row -> cell
    row_iperm: torch.Tensor   # [C] cell -> solver row
    brick_nbr: torch.Tensor   # [nb*6] >= 0 brick, -1 no faces (domain boundary), -2-e explicit block e
    brick_vp: torch.Tensor    # [nb*3*64] +axis face f of each lane (value upper[f]); INT32_MIN: none or explicit
    brick_vm: torch.Tensor    # [nb*3*16] -axis face toward a neighbour brick as -f-1 (value lower[f]); INT32_MIN: none
    expl_col: torch.Tensor    # [ne*64] solver-row column of each slot (the slot's own boundary row when empty)
    expl_code: torch.Tensor   # [ne*64] f: upper[f], -f-1: lower[f], INT32_MIN: empty
    psell_ptr: torch.Tensor   # [ceil(C/64)+1] sliced ELL over solver rows
    psell_col: torch.Tensor
    psell_code: torch.Tensor


def build_solver_layout(level, ijk, owner, neighbour, axis, root) -> SolverLayout:
    """level [C], ijk [C,3] (cell index in its level's lattice), internal
    faces owner/neighbour/axis [F] with the neighbour on the +axis side,
    root = root lattice size [3]."""
    dev = level.device
    L = level.long()
    I = ijk.long()
    C = int(L.shape[0])
    o, n, a = owner.long(), neighbour.long(), axis.long()
    F = int(o.shape[0])
    root = torch.as_tensor(root, dtype=torch.long, device=dev)
    cell = torch.arange(C, device=dev)

    # ---- bricks: 64 leaves of one level in one aligned 4x4x4 block
    B = I >> 2
    key = (L << 57) | (B[:, 0] << 38) | (B[:, 1] << 19) | B[:, 2]
    uniq, inv, cnt = torch.unique(key, return_inverse=True, return_counts=True)
    full_group = cnt == 64
    first = torch.full((uniq.shape[0],), C, dtype=torch.long, device=dev)
    first.scatter_reduce_(0, inv, cell, reduce="amin")
    gid = torch.nonzero(full_group).flatten()
    order = torch.argsort(first[gid])  # bricks by their first cell
    gid = gid[order]
    nb = int(gid.shape[0])
    brick_of_group = torch.full((uniq.shape[0],), -1, dtype=torch.long, device=dev)
    brick_of_group[gid] = torch.arange(nb, device=dev)
    brick = brick_of_group[inv]                       # [C] brick rank or -1
    inb = brick >= 0
    c4 = I & 3
    lane = brick_lane(c4)

    # ---- solver rows
    iperm = torch.empty(C, dtype=torch.long, device=dev)
    iperm[inb] = brick[inb] * 64 + lane[inb]
    rest = torch.nonzero(~inb).flatten()
    iperm[rest] = nb * 64 + torch.arange(rest.shape[0], device=dev)
    perm = torch.empty(C, dtype=torch.long, device=dev)
    perm[iperm] = cell

    # ---- brick neighbours: the same-level brick across each side, if any
    bkey = uniq[gid]                                  # keys of the bricks, rank order
    sk, sidx = torch.sort(bkey)
    bl = bkey >> 57
    bB = torch.stack([(bkey >> 38) & ((1 << 19) - 1), (bkey >> 19) & ((1 << 19) - 1), bkey & ((1 << 19) - 1)], 1)
    lat = root[None, :] << bl[:, None]                 # cell lattice extent at the brick's level
    nbr = torch.full((nb, 6), -1, dtype=torch.long, device=dev)
    for d in range(6):
        ax, side = d // 2, d % 2
        nB = bB.clone()
        nB[:, ax] += 1 if side else -1
        inside = (nB[:, ax] >= 0) & (4 * nB[:, ax] < lat[:, ax])  # cells exist across this side
        nk = (bl << 57) | (nB[:, 0].clamp(min=0) << 38) | (nB[:, 1].clamp(min=0) << 19) | nB[:, 2].clamp(min=0)
        pos = torch.searchsorted(sk, nk).clamp(max=max(nb - 1, 0))
        found = inside & (sk[pos] == nk) if nb else inside & False
        nbr[:, d] = torch.where(found, sidx[pos], torch.where(inside, torch.full_like(pos, -2), torch.full_like(pos, -1)))
    # explicit blocks get consecutive ids
    is_ex = nbr == -2
    ex_id = torch.cumsum(is_ex.flatten().long(), 0).reshape(nb, 6) - 1
    ne = int(is_ex.sum())
    nbr = torch.where(is_ex, -2 - ex_id, nbr)

    # ---- face codes
    vp = torch.full((nb * 3 * 64,), INT32_MIN, dtype=torch.long, device=dev)
    vm = torch.full((nb * 3 * 16,), INT32_MIN, dtype=torch.long, device=dev)
    expl_col = torch.zeros(ne * 64, dtype=torch.long, device=dev)
    expl_code = torch.full((ne * 64,), INT32_MIN, dtype=torch.long, device=dev)
    fidx = torch.arange(F, device=dev)
    # (cell, side) records of every face: owner side +a, neighbour side -a
    rec_cell = torch.cat([o, n])
    rec_dir = torch.cat([2 * a + 1, 2 * a])
    rec_code = torch.cat([fidx, -fidx - 1])
    rec_col = torch.cat([iperm[n], iperm[o]])
    rb = brick[rec_cell]
    keep = rb >= 0
    rec_cell, rec_dir, rec_code, rec_col, rb = rec_cell[keep], rec_dir[keep], rec_code[keep], rec_col[keep], rb[keep]
    ra = rec_dir // 2
    rplus = (rec_dir % 2) == 1
    coord = c4[rec_cell, ra]
    on_side = torch.where(rplus, coord == 3, coord == 0)
    side_mode = nbr[rb, rec_dir]
    rlane = lane[rec_cell]
    # +axis values of lanes: inside the brick, or toward a neighbour brick
    m = rplus & ((~on_side) | (side_mode >= 0))
    vp[(rb[m] * 3 + ra[m]) * 64 + rlane[m]] = rec_code[m]
    # -axis values toward a neighbour brick
    m = (~rplus) & on_side & (side_mode >= 0)
    q = face_position(c4[rec_cell], ra)
    vm[(rb[m] * 3 + ra[m]) * 16 + q[m]] = rec_code[m]
    # explicit blocks: slot 4 q + k, k = rank of the face among the cell's faces on that side
    m = on_side & (side_mode <= -2)
    e_id = -2 - side_mode[m]
    e_q = q[m]
    grp = e_id * 16 + e_q
    srt = torch.argsort(grp * (2 * F + 2) + (rec_code[m] + F + 1), stable=True)
    g_sorted = grp[srt]
    start = torch.zeros_like(g_sorted)
    if g_sorted.numel():
        newg = torch.ones_like(g_sorted, dtype=torch.bool)
        newg[1:] = g_sorted[1:] != g_sorted[:-1]
        pos_all = torch.arange(g_sorted.numel(), device=dev)
        start = torch.cummax(torch.where(newg, pos_all, torch.zeros_like(pos_all)), 0).values
        k = pos_all - start
        if int(k.max()) > 3:
            raise ValueError("a cell side has more than 4 faces: the grid is not 2:1 balanced")
        slot = e_id[srt] * 64 + e_q[srt] * 4 + k
        expl_code[slot] = rec_code[m][srt]
        expl_col[slot] = rec_col[m][srt]
    # empty slots point at their own boundary row (a valid, cached address)
    if ne:
        eb = torch.nonzero(is_ex.flatten()).flatten()  # brick*6 + d, in block-id order
        e_brick, e_dir = eb // 6, eb % 6
        slots = torch.arange(64, device=dev)
        qq = slots // 4
        e_ax = e_dir // 2
        cside = torch.where((e_dir % 2) == 1, 3, 0)
        cq = torch.zeros((ne, 64, 3), dtype=torch.long, device=dev)
        b_ax = (e_ax + 1) % 3
        c_ax = (e_ax + 2) % 3
        ar = torch.arange(ne, device=dev)
        cq[ar, :, e_ax] = cside[:, None]
        cq[ar[:, None], slots[None, :], b_ax[:, None]] = (qq % 4)[None, :]
        cq[ar[:, None], slots[None, :], c_ax[:, None]] = (qq // 4)[None, :]
        own_row = e_brick[:, None] * 64 + brick_lane(cq.reshape(-1, 3)).reshape(ne, 64)
        empty = (expl_code == INT32_MIN).reshape(ne, 64)
        ec = expl_col.reshape(ne, 64)
        ec[empty] = own_row[empty]
        expl_col = ec.reshape(-1)

    # ---- sliced ELL over solver rows (entries ascending by solver column)
    rows = torch.cat([iperm[o], iperm[n]])
    cols = torch.cat([iperm[n], iperm[o]])
    codes = torch.cat([fidx, -fidx - 1])
    srt = torch.argsort(rows * C + cols)
    rows, cols, codes = rows[srt], cols[srt], codes[srt]
    cnt_r = torch.bincount(rows, minlength=C)
    rstart = torch.cumsum(cnt_r, 0) - cnt_r
    kk = torch.arange(rows.shape[0], device=dev) - rstart[rows]
    ns = (C + 63) // 64
    wpad = torch.zeros(ns * 64, dtype=torch.long, device=dev)
    wpad[:C] = cnt_r
    W = wpad.reshape(ns, 64).max(1).values
    ptr = torch.zeros(ns + 1, dtype=torch.long, device=dev)
    ptr[1:] = torch.cumsum(64 * W, 0)
    nnz = int(ptr[-1])
    pcol = torch.empty(nnz, dtype=torch.long, device=dev)
    pcode = torch.full((nnz,), INT32_MIN, dtype=torch.long, device=dev)
    # padding points at the row itself
    srow = torch.repeat_interleave(torch.arange(ns, device=dev), 64 * W)
    within = torch.arange(nnz, device=dev) - ptr[srow]
    prow = srow * 64 + within % 64
    pcol[:] = torch.clamp(prow, max=C - 1)
    e = ptr[rows >> 6] + kk * 64 + (rows & 63)
    pcol[e] = cols
    pcode[e] = codes

    i32 = lambda t: t.to(torch.int32).contiguous()  # noqa: E731
    return SolverLayout(nb, ne, i32(perm), i32(iperm), i32(nbr.flatten()), i32(vp), i32(vm), i32(expl_col),
                        i32(expl_code), i32(ptr), i32(pcol), i32(pcode))


# ---------------------------------------------------------------- emulation
def emulate_direction_spmv(lay: SolverLayout, diag, upper, lower, x_rows) -> torch.Tensor:
    """y = A x on solver rows, computed the way the direction kernel does it
    (brick slices through lanes, permutes and explicit blocks; the rest
    through the sliced ELL); for checking the layout on the CPU.  x_rows and
    the result are in solver-row order; diag is in cell order."""
    dev = x_rows.device
    C = x_rows.shape[0]
    nb = lay.n_bricks
    perm = lay.row_perm.long()
    y = diag[perm] * x_rows

    def val(code):
        code = code.long()
        v = torch.zeros(code.shape, dtype=upper.dtype, device=dev)
        up = code >= 0
        lo = (code < 0) & (code != INT32_MIN)
        v[up] = upper[code[up]]
        v[lo] = lower[-code[lo] - 1]
        return v

    if nb:
        lanes = torch.arange(64, device=dev)
        cx = torch.stack([lanes & 3, (lanes >> 2) & 3, lanes >> 4], 1)
        xb = x_rows[: nb * 64].reshape(nb, 64)
        acc = y[: nb * 64].reshape(nb, 64).clone()
        nbr = lay.brick_nbr.long().reshape(nb, 6)
        vp = val(lay.brick_vp).reshape(nb, 3, 64)
        vm = val(lay.brick_vm).reshape(nb, 3, 16)
        for ax in range(3):
            up_c = cx.clone()
            up_c[:, ax] = (cx[:, ax] + 1).clamp(max=3)
            lp = brick_lane(up_c)
            dn_c = cx.clone()
            dn_c[:, ax] = (cx[:, ax] - 1).clamp(min=0)
            lm = brick_lane(dn_c)
            mirror_p = cx.clone()
            mirror_p[:, ax] = 0
            mirror_m = cx.clone()
            mirror_m[:, ax] = 3
            q = face_position(cx, torch.full((64,), ax, device=dev))
            inner_p = cx[:, ax] < 3
            inner_m = cx[:, ax] > 0
            # +axis neighbour
            t = torch.where(inner_p[None, :], vp[:, ax, :] * xb[:, lp], torch.zeros_like(xb))
            nbp = nbr[:, 2 * ax + 1]
            ext_p = nbp[:, None].clamp(min=0) * 64 + brick_lane(mirror_p)[None, :]
            t = t + torch.where((~inner_p)[None, :] & (nbp >= 0)[:, None], vp[:, ax, :] * x_rows[ext_p],
                                torch.zeros_like(xb))
            acc = acc + t
            # -axis neighbour (symmetric: the -neighbour's +axis value)
            t = torch.where(inner_m[None, :], vp[:, ax, lm] * xb[:, lm], torch.zeros_like(xb))
            nbm = nbr[:, 2 * ax]
            ext_m = nbm[:, None].clamp(min=0) * 64 + brick_lane(mirror_m)[None, :]
            t = t + torch.where((~inner_m)[None, :] & (nbm >= 0)[:, None], vm[:, ax, q] * x_rows[ext_m],
                                torch.zeros_like(xb))
            acc = acc + t
        # explicit blocks
        ecol = lay.expl_col.long().reshape(-1, 64)
        ev = val(lay.expl_code).reshape(-1, 64)
        for b in range(nb):
            for d in range(6):
                s = int(nbr[b, d])
                if s > -2:
                    continue
                e = -2 - s
                c = ev[e] * x_rows[ecol[e]]
                tot = c.reshape(16, 4).sum(1)
                ax, side = d // 2, d % 2
                on = cx[:, ax] == (3 if side else 0)
                qq = face_position(cx, torch.full((64,), ax, device=dev))
                acc[b] = acc[b] + torch.where(on, tot[qq], torch.zeros_like(tot[qq]))
        y = torch.cat([acc.reshape(-1), y[nb * 64:]])
    # general rows through the sliced ELL
    ptr, col, code = lay.psell_ptr.long(), lay.psell_col.long(), lay.psell_code.long()
    ns = ptr.shape[0] - 1
    for s in range(nb, ns):
        W = int(ptr[s + 1] - ptr[s]) // 64
        r0 = s * 64
        nr = min(64, C - r0)
        for k in range(W):
            e = ptr[s] + k * 64 + torch.arange(nr, device=dev)
            y[r0:r0 + nr] = y[r0:r0 + nr] + val(code[e]) * x_rows[col[e]]
    return y
