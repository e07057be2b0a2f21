"""Writes the golden fixtures in this folder from the FP64 CPU oracle.

    python tests/golden/make_golden.py

cavity2d_simple_500.npz   2-D lid-driven cavity (64x64x1, Re 100, SIMPLE,
                          upwind, relaxation U 0.7 / p 0.3): U [C,3] and p [C]
                          after the case's 500 outer iterations, plus the
                          per-iteration pressure-solve iteration counts.
octree_refined3d.npz      integer octree build of a small refined 3-D block
                          (levels, ijk, face lists).

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


def octree() -> None:
    f = FLUXEL["refined3d"]
    t = fo.build_octree(f["lower"], f["upper"], f["root"], f["regions"], f["n_leaf"])
    np.savez_compressed(HERE / "octree_refined3d.npz", level=t.level, ijk=t.ijk, owner=t.owner,
                        neighbour=t.neighbour, axis=t.axis, bnd_owner=t.bnd_owner, bnd_dir=t.bnd_dir)


if __name__ == "__main__":
    octree()
    cavity2d()
