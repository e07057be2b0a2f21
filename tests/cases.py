"""Small case configurations shared by the CPU and GPU parity tests."""

from __future__ import annotations

import copy

import yaml

from gridfoam_amd.config import CASES

FLUXEL = {
    # tests/conftest.py:48-109 of the reference: 4x4x1, unit x 0.1
    "uniform2d": dict(lower=[0, 0, 0], upper=[1, 1, 0.1], root=[4, 4, 1], regions=[], n_leaf=0),
    # tests/helpers/configs.py:51-108: one centred level-1 region
    "refined2d": dict(lower=[0, 0, 0], upper=[1, 1, 0.1], root=[4, 4, 1],
                      regions=[([0.25, 0.25, 0.0], [0.75, 0.75, 0.1], 1)], n_leaf=0),
    # tests/helpers/configs.py:111-: 4^3 root, centred level-1 cube
    "refined3d": dict(lower=[0, 0, 0], upper=[1, 1, 1], root=[4, 4, 4],
                      regions=[([0.25, 0.25, 0.25], [0.75, 0.75, 0.75], 1)], n_leaf=0),
    # the 3-D octree cavity pattern on an 8^3 root grid
    "walls8": dict(lower=[0, 0, 0], upper=[1, 1, 1], root=[8, 8, 8],
                   regions=[([0, 0, 0], [0.1, 1, 1], 2), ([0.9, 0, 0], [1, 1, 1], 2),
                            ([0, 0, 0], [1, 0.1, 1], 2), ([0, 0.9, 0], [1, 1, 1], 2),
                            ([0, 0, 0], [1, 1, 0.1], 2), ([0, 0, 0.9], [1, 1, 1], 2)], n_leaf=0),
    # two cells (one internal face) and a 3^3 block: one partial slice of 64 rows
    "pair": dict(lower=[0, 0, 0], upper=[2, 1, 1], root=[2, 1, 1], regions=[], n_leaf=0),
    "cube3": dict(lower=[0, 0, 0], upper=[1, 1, 1], root=[3, 3, 3], regions=[], n_leaf=0),
    # odd root block counts, nested regions and a leaf refinement
    "odd": dict(lower=[-1, 0, 2], upper=[2, 1, 3.5], root=[3, 5, 2],
                regions=[([0.1, 0.1, 2.1], [0.7, 0.4, 2.6], 2), ([-0.5, 0.6, 3.0], [0.3, 0.9, 3.4], 1)], n_leaf=1),
}


def case_dict(name: str) -> dict:
    return yaml.safe_load((CASES / name).read_text())


def small_cavity3d(root=8, steps=None) -> dict:
    """The octree PISO cavity case on a smaller root grid."""
    cfg = copy.deepcopy(case_dict("cavity3d_octree_piso.yaml"))
    cfg["fluxel"]["root_resolution"] = [root, root, root]
    h_min = 1.0 / (root * 4)
    cfg["simulator"]["control"]["deltaT"] = 0.5 * h_min
    cfg["simulator"]["control"]["endTime"] = 0.5 * h_min * (steps or 10)
    return cfg


def fluxel_block(name: str) -> dict:
    f = FLUXEL[name]
    return {
        "domain": {"lower": f["lower"], "upper": f["upper"]},
        "root_resolution": f["root"],
        "target_level": 0,
        "n_leaf_refinement": f["n_leaf"],
        "refinement_regions": [{"min": list(r[0]), "max": list(r[1]), "level": r[2]} for r in f["regions"]],
    }


def small_cavity3d_pimple(root=8, steps=None, n_outer=2, residual_control=None, consistent=False,
                          cfl=2.0) -> dict:
    """The octree cavity with PIMPLE at a larger time step (outer correctors)."""
    cfg = small_cavity3d(root, steps)
    h_min = 1.0 / (root * 4)
    cfg["simulator"]["control"]["deltaT"] = cfl * h_min
    cfg["simulator"]["control"]["endTime"] = cfl * h_min * (steps or 10)
    alg = cfg["simulator"]["fvSolution"]["algorithm"]
    alg["type"] = "pimple"
    alg["nOuterCorrectors"] = n_outer
    alg["consistent"] = consistent
    if residual_control is not None:
        alg["residualControl"] = residual_control
    return cfg


def small_cavity2d(n=16, consistent=False, alpha_p=0.3) -> dict:
    """The 2-D SIMPLE cavity on an n x n grid."""
    cfg = copy.deepcopy(case_dict("cavity2d_simple.yaml"))
    cfg["fluxel"]["root_resolution"] = [n, n, 1]
    cfg["fluxel"]["domain"]["upper"] = [1.0, 1.0, 1.0 / n]
    alg = cfg["simulator"]["fvSolution"]["algorithm"]
    alg["consistent"] = consistent
    alg["relaxationFactors"]["equations"]["p"] = alpha_p
    return cfg


def channel_flow(algorithm: str = "piso", *, p_outlet: str = "dirichlet", walls: str = "slip",
                 refined: bool = True, consistent: bool = False, u_internal=(1.0, 0.0, 0.0),
                 ffp_walls: bool = False) -> dict:
    """Inlet/outlet channel of the reference's pressure-coupling tests
    (tests/helpers/configs.py:335-461): [0,2]x[0,0.5]x[0,0.1] on an 8x2x1
    root with an optional centred level-1 block; U Dirichlet (1,0,0) at
    x_minus, inletOutlet (inlet value 0) at x_plus, slip or Dirichlet-0
    walls at y; p fixedFluxPressure at x_minus, zero gradient at the walls,
    Dirichlet 0 or zero gradient (pRefCell 0) at x_plus; z faces carry no
    condition.  ``ffp_walls`` moves fixedFluxPressure onto the walls with a
    zero-gradient U there (the corrected wall flux then follows U), so that
    its normal gradient is not zero."""
    steady = algorithm == "simple"
    if algorithm == "simple":
        alg = {"type": "simple", "consistent": consistent, "pRefCell": 0, "pRefValue": 0.0,
               "relaxationFactors": {"equations": {"U": 0.9 if consistent else 0.7, "p": 0.3}}}
    elif algorithm == "piso":
        alg = {"type": "piso", "nCorrectors": 2, "pRefCell": 0, "pRefValue": 0.0}
    else:
        alg = {"type": "pimple", "nCorrectors": 2, "nOuterCorrectors": 2, "consistent": consistent,
               "pRefCell": 0, "pRefValue": 0.0}
    wall_patches = ["y_minus", "y_plus"]
    u_walls = ({"type": "neumann", "patches": wall_patches, "value": [0.0, 0.0, 0.0]} if ffp_walls else
               {"type": walls, "patches": wall_patches, "value": [0.0, 0.0, 0.0]})
    p_bcs = {
        "fixedFluxWalls": {"type": "fixed_flux_pressure", "patches": ["x_minus"]},
        "walls": {"type": "neumann", "patches": wall_patches, "value": 0.0},
        "outlet": {"type": p_outlet, "patches": ["x_plus"], "value": 0.0},
    }
    if ffp_walls:
        p_bcs["fixedFluxWalls"]["patches"] = ["x_minus", *wall_patches]
        del p_bcs["walls"]
    regions = [{"name": "center", "min": [0.75, 0.0, 0.0], "max": [1.25, 0.5, 0.1], "level": 1}] if refined else []
    solver = {"preconditioner": "jacobi", "rel_tolerance": 0.0, "tolerance": 1e-12, "max_iter": 1000,
              "norm_type": "l_2"}
    return {
        "version": "0.3.0",
        "fluxel": {"domain": {"lower": [0.0, 0.0, 0.0], "upper": [2.0, 0.5, 0.1]}, "root_resolution": [8, 2, 1],
                   "target_level": 0, "n_leaf_refinement": 0, "refinement_regions": regions},
        "simulator": {
            "control": {"deltaT": 1.0 if steady else 0.05, "endTime": 10.0 if steady else 1.0, "writeInterval": 1,
                        "precision": "float64"},
            "fvSolution": {"adjustPhi": True, "algorithm": alg,
                           "solvers": {"U": {"method": "bicgstab", **solver}, "p": {"method": "cg", **solver}}},
            "conditions": {
                "U": {"internal": list(u_internal), "boundary": {
                    "inlet": {"type": "dirichlet", "patches": ["x_minus"], "value": [1.0, 0.0, 0.0]},
                    "outlet": {"type": "inlet_outlet", "patches": ["x_plus"], "value": [0.0, 0.0, 0.0]},
                    "walls": u_walls}},
                "p": {"internal": 0.0, "boundary": p_bcs},
            },
            "properties": {"transport": {"type": "newtonian", "nu": 0.01}, "turbulence": {"type": "laminar"}},
        },
    }


def fourier_channel(root=(20, 8, 1), level: int = 1, algorithm: str = "pimple", dt: float = 0.05,
                    consistent: bool = False) -> dict:
    """The Fourier-body channel case on a coarser root grid and fewer levels
    (same body, conditions and solvers)."""
    cfg = copy.deepcopy(case_dict("channel2d_fourier_pimple.yaml"))
    fx = cfg["fluxel"]
    fx["root_resolution"] = list(root)
    fx["domain"]["upper"][2] = 20.0 / root[0]
    fx["refinement_regions"][0]["level"] = level
    fx["refinement_regions"][0]["max"][2] = 20.0 / root[0]
    sim = cfg["simulator"]
    sim["control"]["deltaT"] = dt
    sim["control"]["endTime"] = 10 * dt
    alg = sim["fvSolution"]["algorithm"]
    alg["type"] = algorithm
    alg["consistent"] = consistent
    alg.setdefault("relaxationFactors", {"equations": {"U": 0.7, "p": 0.3}})
    for s in sim["fvSolution"]["solvers"].values():
        s["tolerance"], s["rel_tolerance"] = 1e-11, 0.0
    return cfg


def box_cavity(n: int = 16, boxes=None, algorithm: str = "simple", depth: int = 1, consistent: bool = False) -> dict:
    """The lid-driven cavity (lid u = 1, Re 100) with axis-aligned box
    obstacles: no-slip, zero pressure gradient on the boxes; ``depth`` root
    blocks in z (1: a 2-D slab with empty z faces)."""
    cfg = small_cavity2d(n, consistent=consistent)
    boxes = boxes if boxes is not None else [dict(name="block", kind="box", min=[0.28125, 0.25, -1.0],
                                                  max=[0.53125, 0.5, 2.0])]
    fx = cfg["fluxel"]
    fx["immersed_bodies"] = boxes
    fx["root_resolution"] = [n, n, depth]
    fx["domain"]["upper"] = [1.0, 1.0, depth / n]
    sim = cfg["simulator"]
    sim["fvSolution"]["algorithm"]["type"] = algorithm
    sim["fvSolution"]["algorithm"].setdefault("nCorrectors", 2)
    sim["fvSolution"]["algorithm"].setdefault("nOuterCorrectors", 2)
    sim["control"]["deltaT"] = 1.0 if algorithm == "simple" else 0.02
    names = [b["name"] for b in boxes]
    cond = sim["conditions"]
    cond["U"]["boundary"]["obstacles"] = {"type": "dirichlet", "patches": names, "value": [0.0, 0.0, 0.0]}
    cond["p"]["boundary"]["obstacles"] = {"type": "neumann", "patches": names, "value": 0.0}
    if depth > 1:
        for fld in ("U", "p"):
            cond[fld]["boundary"].pop("empty")
        cond["U"]["boundary"]["wall"]["patches"] = ["x_minus", "x_plus", "y_minus", "z_minus", "z_plus"]
        cond["p"]["boundary"]["wall"]["patches"] = ["x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus"]
    for s in sim["fvSolution"]["solvers"].values():
        s["tolerance"], s["rel_tolerance"] = 1e-11, 0.0
        s["max_iter"] = 4000
    return cfg


def bluff_body(root=(20, 4, 3), target_level: int = 4, n_leaf: int = 0, wake_level=None, dt=None,
               mixed: bool = True) -> dict:
    """The 3-D bluff-body tunnel case with a given root grid, surface level,
    leaf refinement and wake-box level (None: no wake box); dt defaults to
    CFL 0.5 on the finest cells."""
    cfg = copy.deepcopy(case_dict("bluff3d_piso.yaml"))
    fx = cfg["fluxel"]
    fx["root_resolution"] = list(root)
    fx["target_level"] = target_level
    fx["n_leaf_refinement"] = n_leaf
    if wake_level is None:
        fx["refinement_regions"] = []
    else:
        fx["refinement_regions"][0]["level"] = wake_level
    h0 = 10.0 / root[0]
    levels = max(target_level, wake_level or 0) + n_leaf
    sim = cfg["simulator"]
    sim["control"]["deltaT"] = dt if dt is not None else 0.5 * h0 / 2 ** levels
    sim["control"]["endTime"] = 100 * sim["control"]["deltaT"]
    sim["fvSolution"]["solvers"]["p"]["mixed_precision"] = mixed
    return cfg
