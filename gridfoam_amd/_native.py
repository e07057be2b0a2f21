"""ctypes binding of libgridfoam_hip.so (declared in include/gridfoam_hip.h).

The product path has no CPU fallback: if the library is missing, or no GPU
is visible, every native call raises.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
from typing import Any

import torch

LIB_PATH = pathlib.Path(os.environ.get("GRIDFOAM_HIP_LIB", pathlib.Path(__file__).resolve().parent / "libgridfoam_hip.so"))

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double


class GfMesh(ctypes.Structure):
    _fields_ = [
        ("n_cells", _i32), ("n_faces", _i32), ("n_bnd", _i32), ("n_hang", _i32),
        ("n_hang_cells", _i32), ("n_slices", _i32), ("nnz_pad", _i64),
        ("owner", _vp), ("neighbour", _vp), ("axis", _vp), ("bnd_owner", _vp), ("bnd_dir", _vp),
        ("own_row", _vp), ("nbr_row", _vp), ("nbr_face", _vp), ("bnd_row", _vp),
        ("sell_ptr", _vp), ("sell_col", _vp), ("sell_code", _vp),
        ("centers", _vp), ("sizes", _vp), ("volumes", _vp), ("w", _vp), ("mag_sf", _vp), ("delta", _vp),
        ("bnd_d", _vp), ("bnd_mag_d", _vp), ("bnd_mag_sf", _vp), ("bnd_sign", _vp), ("ata_inv", _vp),
        ("hang_face", _vp), ("hang_offset", _vp), ("hang_move_o", _vp), ("hang_move_n", _vp),
        ("hang_cell", _vp), ("hc_row", _vp), ("hc_code", _vp),
        ("n_bricks", _i32), ("n_expl", _i32), ("pnnz_pad", _i64),
        ("row_perm", _vp), ("row_iperm", _vp), ("brick_nbr", _vp), ("brick_vp", _vp), ("brick_vm", _vp),
        ("expl_col", _vp), ("expl_code", _vp), ("psell_ptr", _vp), ("psell_col", _vp), ("psell_code", _vp),
        ("bnd_face", _vp), ("bnd_snap", _vp),
    ]


class GfShape(ctypes.Structure):
    _fields_ = [("kind", _i32), ("axis", _i32), ("patch", _i32), ("pad", _i32), ("p", _f64 * 20)]


class GfAssembly(ctypes.Structure):
    _fields_ = [
        ("nc", _i32), ("s_ddt", _f64), ("s_div", _f64), ("s_lap", _f64),
        ("ddt_a", _f64), ("ddt_b", _f64), ("ddt_c", _f64), ("dt", _f64),
        ("psi_old", _vp), ("psi_older", _vp),
        ("div_scheme", _i32), ("phi_s", _vp), ("phi_b", _vp),
        ("gamma", _f64), ("gamma_cell", _vp), ("lap_corrected", _i32), ("lap_harmonic", _i32),
        ("psi", _vp), ("grad", _vp),
        ("frac", _vp), ("ref_v", _vp), ("ref_g", _vp), ("psi_b", _vp), ("covered", _vp),
    ]


class GfKrylovCfg(ctypes.Structure):
    _fields_ = [
        ("method", _i32), ("precond", _i32), ("norm", _i32), ("max_iter", _i32), ("max_restart", _i32),
        ("atol", _f64), ("rtol", _f64), ("mixed", _i32), ("inner_rtol", _f64), ("check_every", _i32),
        ("symmetric", _i32),
    ]


_P = ctypes.POINTER
_SIGS: dict[str, tuple[Any, list[Any]]] = {
    "gf_last_error": (ctypes.c_char_p, []),
    "gf_device_count": (_i32, [_vp]),
    "gf_abi_sizes": (_i32, [_vp]),
    "gf_octree_build": (_i32, [_vp, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _vp, _vp, _vp]),
    "gf_octree_build_surface": (_i32, [_vp, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _P(GfShape), _i32, _i32, _vp,
                                       _vp, _vp]),
    "gf_octree_fetch": (_i32, [_vp] * 8 + [_vp]),
    "gf_octree_free": (None, [_vp]),
    "gf_grid_geometry": (_i32, [_i32, _i32, _i32] + [_vp] * 17 + [_vp]),
    "gf_face_geometry": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp, _vp, _f64, _vp]),
    "gf_row_counts": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp]),
    "gf_mesh_validate": (_i32, [_P(GfMesh), _P(ctypes.c_int32), _vp]),
    "gf_exclusive_scan_i32": (_i32, [_vp, _vp, _i32, _vp]),
    "gf_fill_nbr_lists": (_i32, [_P(GfMesh), _vp, _vp]),
    "gf_fill_bnd_lists": (_i32, [_P(GfMesh), _vp, _vp]),
    "gf_ap_cut": (_i32, [_i32] + [_vp] * 6 + [_i32, _P(GfShape), _f64] + [_vp] * 9 + [_vp]),
    "gf_sell_widths": (_i32, [_P(GfMesh), _vp, _vp]),
    "gf_sell_fill": (_i32, [_P(GfMesh), _vp, _vp, _vp]),
    "gf_lsq_ata_inv": (_i32, [_P(GfMesh), _vp, _vp]),
    "gf_hang_geometry": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp, _vp]),
    "gf_hang_cell_lists": (_i32, [_P(GfMesh), _vp, _vp, _vp]),
    "gf_bc_state": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_bc_slip": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp]),
    "gf_lsq_grad": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _i32, _vp, _vp]),
    "gf_interpolate": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp]),
    "gf_gauss_grad": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp]),
    "gf_reconstruct": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp]),
    "gf_fvc_div": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp]),
    "gf_fvc_div_vjp": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp]),
    "gf_reconstruct_vjp": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp, _vp]),
    "gf_gauss_grad_vjp": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp]),
    "gf_interpolate_vjp": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp]),
    "gf_continuity_error": (_i32, [_P(GfMesh), _vp, _vp, _f64, _vp, _vp, _vp]),
    "gf_sn_grad_hang": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp]),
    "gf_assemble": (_i32, [_P(GfMesh), _P(GfAssembly), _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_assemble_vjp": (_i32, [_P(GfMesh), _P(GfAssembly)] + [_vp] * 15),
    "gf_lsq_grad_vjp": (_i32, [_P(GfMesh), _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    "gf_bc_state_vjp": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_ldu_multiply": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_rau_hbya": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_ldu_flux": (_i32, [_P(GfMesh), _i32, _f64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_face_flux": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_ddt_corr_add": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp, _vp, _f64, _f64, _f64, _vp, _vp]),
    "gf_constrain_bnd_flux": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp]),
    "gf_adjust_phi": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp]),
    "gf_correct_phi": (_i32, [_P(GfMesh)] + [_vp] * 13 + [_vp]),
    "gf_correct_velocity": (_i32, [_i32, _vp, _vp, _vp, _vp, _vp]),
    "gf_set_reference": (_i32, [_vp, _vp, _i32, _f64, _vp]),
    "gf_axpby": (_i32, [_i64, _f64, _vp, _f64, _vp, _vp, _vp]),
    "gf_initial_residual": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_initial_residual_noalloc": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_krylov_workspace_bytes": (_i64, [_P(GfMesh), _i32, _P(GfKrylovCfg)]),
    "gf_krylov_solve": (_i32, [_P(GfMesh), _P(GfKrylovCfg), _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_krylov_use_graphs": (_i32, [_i32]),
    "gf_krylov_release": (_i32, [_vp]),
    "gf_ldu_grads": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_pcg_dir_bench_bytes": (_i64, [_P(GfMesh), _i32]),
    "gf_pcg_dir_bench": (_i32, [_P(GfMesh), _i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "gf_seg_pcg_workspace": (_i32, [_i32, _i32, _i64, _vp]),
    "gf_seg_bicgstab_workspace": (_i32, [_i32, _i32, _i64, _vp]),
    "gf_seg_bicgstab": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _f64, _f64, _i32, _i32, _i32,
                               _vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "gf_seg_pcg": (_i32, [_P(GfMesh), _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _f64, _f64, _i32, _i32,
                          _vp, _i64, _vp, _vp, _vp, _vp, _vp]),
    "gf_relax_momentum": (_i32, [_i32, _f64, _vp, _vp, _vp, _vp]),
    "gf_ldu_h1": (_i32, [_P(GfMesh), _vp, _vp, _vp, _vp]),
    "gf_fixed_values": (_i32, [_P(GfMesh), _i32, _i32] + [_vp] * 13 + [_vp]),
    "gf_simplec_rautu": (_i32, [_i32, _vp, _vp, _i32, _vp, _vp]),
    "gf_predictor_source": (_i32, [_i32, _i32, _vp, _vp, _vp, _vp, _vp]),
    "gf_relax_field": (_i32, [_i64, _f64, _vp, _vp, _vp]),
    "gf_stream_copy": (_i32, [_vp, _vp, _i64, _vp]),
}

_LIB: ctypes.CDLL | None = None


def exported_symbols() -> list[str]:
    return sorted(_SIGS)


def lib() -> ctypes.CDLL:
    """Load the HIP library once; raise loudly when it is missing."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m gridfoam_amd.build` "
                "(the pressure-velocity path has no CPU fallback)"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("gridfoam_amd needs a ROCm GPU (torch.cuda.is_available() is False)")


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p | None:
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def call(name: str, *args: Any) -> int:
    fn = getattr(lib(), name)
    rc = fn(*args)
    if isinstance(rc, int) and rc < 0 and fn.restype is _i32:
        msg = lib().gf_last_error()
        raise RuntimeError(f"{name} failed ({rc}): {msg.decode() if msg else 'unknown error'}")
    return rc
