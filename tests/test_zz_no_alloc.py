"""GPU: gf_initial_residual_noalloc, the allocation-free variant of
gf_initial_residual (residual.py:12-36), against the existing entry point
through the C ABI: the same kernels, so the results must be bit-identical.
Two calls in a row, the second on a smaller grid (the static partials array
keeps the first call's contents), for a scalar and a vector field; the
oracle's residual is checked as well, and sizes beyond the array are
rejected before anything is launched."""

from __future__ import annotations

import ctypes
import sys

import numpy as np
import pytest
import torch

from oracle import fv_oracle as fo
from tests.cases import FLUXEL, fluxel_block

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _case(name: str, nc: int, seed: int):
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.fields import CellField, DirichletBC, NeumannBC
    from gridfoam_amd.fv import fvm
    from gridfoam_amd.grid import OctreeGrid

    g = OctreeGrid(FluxelConfig.model_validate(fluxel_block(name)), dt=0.01)
    f = FLUXEL[name]
    o = fo.Grid(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"], dt=0.01)
    rng = np.random.default_rng(seed)
    cs = () if nc == 1 else (3,)
    psi = rng.standard_normal((g.num_cells,) + cs)
    gamma = 0.5 + rng.random(g.num_cells)
    val = 0.3 if nc == 1 else [0.3, -0.1, 0.2]
    fld = CellField(g, f"q{nc}", cs)
    fld.data = torch.tensor(psi, device=g.device)
    fld.add_boundary_conditions({p: (DirichletBC(val) if p == "x_minus" else NeumannBC(0.0 if nc == 1 else [0.0] * 3))
                                 for p in fo.PATCHES})
    m = fvm.assemble(fld, lap=-1.0, gamma=torch.tensor(gamma, device=g.device), corrected=True)
    ofld = fo.Field("q", psi.copy(), {p: (fo.BC("dirichlet", np.asarray(val)) if p == "x_minus"
                                          else fo.BC("neumann", np.zeros(cs))) for p in fo.PATCHES})
    om = -fo.fvm_laplacian(o, gamma, ofld, True)
    return g, m, fld, fo.initial_residual(om, psi)


def _both(g, m, fld):
    from gridfoam_amd import _native as nat

    nb = (g.num_cells + 255) // 256
    scratch = torch.zeros(2 * nb + 8, dtype=torch.float64, device=g.device)
    a = torch.full((1,), -1.0, dtype=torch.float64, device=g.device)
    b = torch.full((1,), -1.0, dtype=torch.float64, device=g.device)
    args = (g.mesh_ref, ctypes.c_int32(fld.nc), nat.ptr(m._diag), nat.ptr(m._upper), nat.ptr(m._lower),
            nat.ptr(m._source), nat.ptr(fld.data))
    nat.call("gf_initial_residual", *args, nat.ptr(scratch), nat.ptr(a), nat.stream())
    nat.call("gf_initial_residual_noalloc", *args, nat.ptr(b), nat.stream())
    torch.cuda.synchronize()
    return float(a[0]), float(b[0])


@pytest.mark.parametrize("nc", [1, 3])
def test_noalloc_initial_residual_equals_the_scratch_entry_on_two_calls(nc):
    for name, seed in (("walls8", 1), ("refined3d", 2)):  # the second call on the smaller grid
        _say(f"noalloc initial residual {name} nc={nc}")
        g, m, fld, ref = _case(name, nc, seed)
        a, b = _both(g, m, fld)
        assert a == b, (name, a, b)
        assert abs(b - ref) <= 1e-12 * abs(ref), (name, b, ref)


def test_noalloc_initial_residual_rejects_grids_beyond_its_scratch():
    from gridfoam_amd import _native as nat

    _say("noalloc size limit")
    g, m, fld, _ = _case("refined3d", 1, 3)
    big = nat.GfMesh()
    ctypes.pointer(big)[0] = g.mesh
    big.n_cells = 40_000_000  # rejected on the host before any launch
    out = torch.zeros(1, dtype=torch.float64, device=g.device)
    with pytest.raises(RuntimeError, match="static scratch capacity"):
        nat.call("gf_initial_residual_noalloc", ctypes.byref(big), ctypes.c_int32(1), nat.ptr(m._diag),
                 nat.ptr(m._upper), nat.ptr(m._lower), nat.ptr(m._source), nat.ptr(fld.data), nat.ptr(out),
                 nat.stream())
    # the library still works afterwards
    a, b = _both(g, m, fld)
    assert a == b
