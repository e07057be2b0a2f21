"""The fixed-order scatter-add (immersed.det_index_add) used where several
contributions land on one cell: it equals a sequential accumulation in
index order exactly, and index_add up to rounding."""

from __future__ import annotations

import torch

from gridfoam_amd.immersed import det_index_add


def test_fixed_order_scatter_add_equals_sequential_sums():
    gen = torch.Generator().manual_seed(4)
    base = torch.randn(50, 3, dtype=torch.float64, generator=gen)
    idx = torch.randint(0, 50, (400,), generator=gen)
    vals = torch.randn(400, 3, dtype=torch.float64, generator=gen) * torch.logspace(-8, 8, 400)[:, None].double()
    got = det_index_add(base, idx, vals)
    ref = base.clone()
    for k in range(idx.numel()):  # in index order, one at a time
        ref[idx[k]] += vals[k]
    assert torch.equal(got, ref)
    assert torch.allclose(got, base.index_add(0, idx, vals), rtol=1e-12, atol=1e-6)
    assert torch.equal(det_index_add(base, idx[:0], vals[:0]), base)
