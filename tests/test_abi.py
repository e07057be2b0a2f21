"""CPU checks of the C ABI: the library loads, exports every symbol that
include/gridfoam_hip.h declares, and the ctypes signatures match the header
(no compute calls: there is no GPU here)."""

from __future__ import annotations

import ctypes
import re

import pytest

from tests.conftest import ROOT

HEADER = ROOT / "include" / "gridfoam_hip.h"
LIB = ROOT / "gridfoam_amd" / "libgridfoam_hip.so"


def _declarations() -> dict[str, int]:
    """name -> parameter count, parsed from the header."""
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"\b(?:int|int64_t|void|const char \*)\s*\*?\s*(gf_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        name, params = m.group(1), m.group(2).strip()
        n = 0 if params in ("", "void") else params.count(",") + 1
        decls[name] = n
    return decls


def test_header_declares_functions():
    d = _declarations()
    assert "gf_krylov_solve" in d and "gf_octree_build" in d and "gf_assemble" in d
    assert len(d) >= 40


@pytest.mark.skipif(not LIB.exists(), reason="HIP library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    for name in _declarations():
        assert hasattr(lib, name), name


def test_ctypes_signatures_match_header():
    from gridfoam_amd import _native

    d = _declarations()
    assert set(d) == set(_native._SIGS), set(d) ^ set(_native._SIGS)
    for name, n in d.items():
        assert len(_native._SIGS[name][1]) == n, f"{name}: header {n} vs ctypes {len(_native._SIGS[name][1])}"


@pytest.mark.skipif(not LIB.exists(), reason="HIP library not built")
def test_struct_layouts_match_library():
    from gridfoam_amd import _native

    sizes = (ctypes.c_int64 * 3)()
    assert _native.lib().gf_abi_sizes(sizes) == 0
    assert sizes[0] == ctypes.sizeof(_native.GfMesh)
    assert sizes[1] == ctypes.sizeof(_native.GfAssembly)
    assert sizes[2] == ctypes.sizeof(_native.GfKrylovCfg)
    # gf_mesh: six int32 counts + one int64, 30 pointers, the solver layout
    # (two int32 counts + one int64 and 10 pointers), then bnd_face, bnd_snap
    assert _native.GfMesh.owner.offset == 32 and _native.GfMesh.n_bricks.offset == 32 + 30 * 8
    assert ctypes.sizeof(_native.GfMesh) == 32 + 30 * 8 + 16 + 10 * 8 + 2 * 8
    assert ctypes.sizeof(_native.GfShape) == 16 + 20 * 8


def test_product_path_has_no_cpu_fallback():
    """Without a GPU every grid construction must fail loudly."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.grid import OctreeGrid

    fx = FluxelConfig.model_validate({"domain": {"lower": [0, 0, 0], "upper": [1, 1, 1]}, "root_resolution": [2, 2, 2]})
    with pytest.raises(RuntimeError):
        OctreeGrid(fx)
