"""Writes the golden fixtures in this folder from the FP64 CPU oracle.

    python tests/golden/make_golden.py

cavity2d_simple_500.npz   2-D lid-driven cavity (64x64x1, Re 100, SIMPLE,
                          upwind, relaxation U 0.7 / p 0.3): U [C,3] and p [C]
                          after the case's 500 outer iterations, plus the
                          per-iteration pressure-solve iteration counts.
octree_refined3d.npz      integer octree build of a small refined 3-D block
                          (levels, ijk, face lists).
cavity3d_octree_piso_2.npz  3-D lid-driven cavity on the octree (32^3 root,
                          two wall levels, 1,125,384 cells, Re 1000, PISO 2
                          correctors, linear convection, CFL 0.5) after 2
                          steps in FP64: U and p at 20,000 cells sampled with
                          numpy seed 0, the L2 norms of the whole U and p, and
                          the whole-field functionals sum(p r_p), sum(U r_U)
                          with r ~ N(0, 1) from numpy seed 1 (p first).
channel2d_fourier_pimple_2.npz  2-D channel around the Fourier body (40x16x1
                          root, four levels within one diameter, 205,600
                          cells, Re 40, PIMPLE 2 outer x 2 inner correctors)
                          after 2 steps in FP64: the same samples, norms and
                          functionals as above, the AP payload (ray distances
                          and near flags of the 2,144 immersed faces) and the
                          corrected flux on their upper and lower sides.

The reference itself cannot run in this environment (Python >= 3.12 source
and the absent fluxel-cfd mesher), so the fixtures freeze the oracle, whose
parts are pinned by tests/test_oracle.py against the reference's own
known-answer tests and closed forms.  They let the GPU path and any later
oracle change be checked against a fixed answer.
"""

from __future__ import annotations

import pathlib
import sys

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from oracle import fv_oracle as fo  # noqa: E402
from tests.cases import FLUXEL, case_dict  # noqa: E402


def cavity2d(iterations: int = 500) -> None:
    case = fo.Case(case_dict("cavity2d_simple.yaml"))
    p_iters = []
    for _ in range(iterations):
        case.step()
        p_iters.append(sum(st.iterations for st in case.stats["p"]))
    np.savez_compressed(HERE / "cavity2d_simple_500.npz", U=case.U.data, p=case.p.data,
                        p_iterations=np.asarray(p_iters, np.int32), iterations=np.int32(iterations))


def cavity3d_octree(steps: int = 2, n_sample: int = 20000) -> None:
    case = fo.Case(case_dict("cavity3d_octree_piso.yaml"))
    for _ in range(steps):
        case.step()
    C = case.g.num_cells
    idx = np.sort(np.random.default_rng(0).choice(C, size=n_sample, replace=False)).astype(np.int32)
    rng = np.random.default_rng(1)
    r_p = rng.standard_normal(C)
    r_U = rng.standard_normal((C, 3))
    U, p = case.U.data, case.p.data
    np.savez_compressed(HERE / f"cavity3d_octree_piso_{steps}.npz", idx=idx, U=U[idx], p=p[idx],
                        U_norm=np.linalg.norm(U), p_norm=np.linalg.norm(p),
                        f_p=float(np.dot(p, r_p)), f_U=float(np.sum(U * r_U)),
                        cells=np.int64(C), steps=np.int32(steps),
                        p_iterations=np.int32(sum(st.iterations for st in case.stats["p"])))


def channel2d_fourier(steps: int = 2, n_sample: int = 20000) -> None:
    fo.use_threaded_spmv(8)
    case = fo.Case(case_dict("channel2d_fourier_pimple.yaml"))
    for _ in range(steps):
        case.step()
    g = case.g
    C = g.num_cells
    idx = np.sort(np.random.default_rng(0).choice(C, size=n_sample, replace=False)).astype(np.int32)
    rng = np.random.default_rng(1)
    r_p = rng.standard_normal(C)
    r_U = rng.standard_normal((C, 3))
    U, p = case.U.data, case.p.data
    B = g.num_domain_bnd
    np.savez_compressed(HERE / f"channel2d_fourier_pimple_{steps}.npz", idx=idx, U=U[idx], p=p[idx],
                        U_norm=np.linalg.norm(U), p_norm=np.linalg.norm(p),
                        f_p=float(np.dot(p, r_p)), f_U=float(np.sum(U * r_U)),
                        cells=np.int64(C), steps=np.int32(steps), immersed=np.int64(g.num_immersed),
                        dist_owner=g.ap.dist_owner, dist_neighbour=g.ap.dist_neighbour,
                        owner_near=g.ap.owner_near, neighbour_near=g.ap.neighbour_near,
                        phi_immersed=case.phi.bnd[B:], div_max=fo.continuity_max(g, case.phi),
                        p_iterations=np.int32(sum(st.iterations for st in case.stats["p"])))


def octree() -> None:
    f = FLUXEL["refined3d"]
    t = fo.build_octree(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"])
    np.savez_compressed(HERE / "octree_refined3d.npz", level=t.level, ijk=t.ijk, owner=t.owner,
                        neighbour=t.neighbour, axis=t.axis, bnd_owner=t.bnd_owner, bnd_dir=t.bnd_dir)


if __name__ == "__main__":
    which = sys.argv[1:] or ["octree", "cavity2d", "cavity3d", "channel2d"]
    if "octree" in which:
        octree()
    if "cavity2d" in which:
        cavity2d()
    if "cavity3d" in which:
        cavity3d_octree()  # about 3 minutes per step on 8 host cores
    if "channel2d" in which:
        channel2d_fourier()  # about 80 s per step on 8 host cores
