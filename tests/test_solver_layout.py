"""CPU checks of the PCG solver row order and brick layout
(gridfoam_amd/solver_layout.py) on the oracle's octree topologies: the
layout, applied the way the direction kernel applies it (lanes, permutes,
neighbour bricks, explicit 2:1 blocks, sliced ELL for the rest), must give
the oracle's LDU multiply (core/fvmatrix.py:292-323) of a symmetric
operator, row for row, and every index must lie in its buffer."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gridfoam_amd.solver_layout import INT32_MIN, build_solver_layout, emulate_direction_spmv
from oracle import fv_oracle as fo
from tests.cases import FLUXEL

WALLS16 = dict(lower=[0, 0, 0], upper=[1, 1, 1], root=[16, 16, 16],
               regions=[([0, 0, 0], [0.1, 1, 1], 2), ([0.9, 0, 0], [1, 1, 1], 2), ([0, 0, 0], [1, 0.1, 1], 2),
                        ([0, 0.9, 0], [1, 1, 1], 2), ([0, 0, 0], [1, 1, 0.1], 2), ([0, 0, 0.9], [1, 1, 1], 2)],
               n_leaf=0)
GRIDS = dict(FLUXEL, walls16=WALLS16, uniform8=dict(lower=[0, 0, 0], upper=[1, 1, 1], root=[8, 8, 8], regions=[],
                                                    n_leaf=0),
             slab=dict(lower=[0, 0, 0], upper=[2, 1, 0.5], root=[8, 4, 2],
                       regions=[([0.0, 0.0, 0.0], [0.6, 1.0, 0.5], 1)], n_leaf=0))


def _layout(name):
    f = GRIDS[name]
    t = fo.build_octree(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"])
    T = {k: torch.from_numpy(np.asarray(getattr(t, k)).astype(np.int64)) for k in
         ("level", "ijk", "owner", "neighbour", "axis")}
    lay = build_solver_layout(T["level"], T["ijk"], T["owner"], T["neighbour"], T["axis"], f["root"])
    return t, lay


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_layout_reproduces_the_ldu_multiply(name):
    t, lay = _layout(name)
    C, F = len(t.level), len(t.owner)
    rng = np.random.default_rng(0)
    upper = torch.from_numpy(rng.standard_normal(F))
    lower = upper.clone()  # symmetric operator
    diag = torch.from_numpy(rng.standard_normal(C) + 7.0)
    x = torch.from_numpy(rng.standard_normal(C))
    perm = lay.row_perm.long()
    y_rows = emulate_direction_spmv(lay, diag, upper, lower, x[perm])
    y = torch.empty_like(y_rows)
    y[perm] = y_rows
    o, n = torch.from_numpy(t.owner.astype(np.int64)), torch.from_numpy(t.neighbour.astype(np.int64))
    ref = diag * x
    ref = ref.index_add(0, o, upper * x[n]).index_add(0, n, lower * x[o])
    torch.testing.assert_close(y, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("name", sorted(GRIDS))
def test_layout_indices_stay_in_their_buffers(name):
    t, lay = _layout(name)
    C, F = len(t.level), len(t.owner)
    nb, ne = lay.n_bricks, lay.n_expl
    assert sorted(lay.row_perm.tolist()) == list(range(C))
    assert torch.equal(lay.row_iperm.long()[lay.row_perm.long()], torch.arange(C))
    assert lay.brick_nbr.numel() == 6 * nb and lay.brick_vp.numel() == 192 * nb and lay.brick_vm.numel() == 48 * nb
    nbr = lay.brick_nbr.long()
    assert bool(((nbr >= -1) & (nbr < nb) | (nbr <= -2) & (-2 - nbr < ne)).all())
    assert sorted((-2 - nbr[nbr <= -2]).tolist()) == list(range(ne))
    vp, vm = lay.brick_vp.long(), lay.brick_vm.long()
    assert bool(((vp == INT32_MIN) | ((vp >= 0) & (vp < F))).all())
    assert bool(((vm == INT32_MIN) | ((vm >= -F) & (vm < 0))).all())
    col = lay.expl_col.long()
    assert bool(((col >= 0) & (col < C)).all())
    code = lay.expl_code.long()
    assert bool(((code == INT32_MIN) | ((code >= -F) & (code < F))).all())
    ptr = lay.psell_ptr.long()
    assert int(ptr[0]) == 0 and ptr.numel() == (C + 63) // 64 + 1
    assert bool((torch.diff(ptr) % 64 == 0).all()) and int(ptr[-1]) == lay.psell_col.numel()
    assert bool(((lay.psell_col.long() >= 0) & (lay.psell_col.long() < C)).all())
    # every internal face appears twice in the sliced ELL (once per row)
    pc = lay.psell_code.long()
    assert int((pc != INT32_MIN).sum()) == 2 * F


def test_cavity_pattern_is_mostly_bricks():
    """The wall-refined cavity pattern: most cells lie in bricks, and the
    explicit blocks sit only on level changes."""
    t, lay = _layout("walls16")
    C = len(t.level)
    assert lay.n_bricks * 64 > 0.5 * C
    assert lay.n_expl > 0
