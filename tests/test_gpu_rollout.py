"""GPU: gradients through a checkpointed transient rollout
(gridfoam_amd.adjoint): implicit scalar transport across a refinement
interface, every step a native assembly + BiCGSTAB solve.  The gradient of a
functional of the final field with respect to the diffusivity, the inlet
value, the flux and the initial field equals the one from the full graph
(checkpoint segments of 4 steps recompute the same deterministic steps) and
central differences (1e-6), for Euler and BDF2."""

from __future__ import annotations

import sys

import pytest
import torch

pytestmark = pytest.mark.gpu

PATCHES = ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _case():
    from gridfoam_amd.config import FluxelConfig
    from gridfoam_amd.fields import CellField, FaceField
    from gridfoam_amd.grid import OctreeGrid

    n, slab = 16, 0.1
    fx = {"domain": {"lower": [0, 0, 0], "upper": [1.0, slab, slab]}, "root_resolution": [n, 1, 1],
          "target_level": 0, "n_leaf_refinement": 0,
          "refinement_regions": [{"min": [0.5, 0.0, 0.0], "max": [1.0, slab, slab], "level": 1}]}
    g = OctreeGrid(FluxelConfig.model_validate(fx), dt=0.1 / n)
    q = CellField(g, "q", (), transient=True)
    phi = FaceField(g, "phi")
    return g, q, phi


@pytest.mark.parametrize("bdf2", [False, True])
def test_checkpointed_rollout_gradients(bdf2):
    from gridfoam_amd.adjoint import rollout_scalar_transport
    from gridfoam_amd.fields import DirichletBC, NeumannBC

    g, q, phi = _case()
    _say(f"checkpointed scalar transport rollout bdf2={bdf2}")
    x = g.cell_centers[:, 0]
    F, B = g.num_internal_faces, g.num_domain_bnd_faces
    gen = torch.Generator(g.device).manual_seed(5)
    base = {
        "nu": torch.tensor(2e-3, dtype=torch.float64, device=g.device),
        "inlet": torch.tensor(0.3, dtype=torch.float64, device=g.device),
        "phi_s": g.Sf[:F, 0].clone() * (1.0 + 0.1 * torch.rand(F, dtype=torch.float64, device=g.device,
                                                                generator=gen)),
        "q0": ((x >= 0.25) & (x < 0.45)).to(torch.float64),
    }
    w = torch.randn(g.num_cells, dtype=torch.float64, device=g.device, generator=gen)

    def J(vals, every):
        from gridfoam_amd.fields import FaceField

        q.add_boundary_conditions({p: (DirichletBC(vals["inlet"]) if p == "x_minus" else NeumannBC(0.0))
                                   for p in PATCHES})
        phi = FaceField(g, "phi")  # fresh buffers: each evaluation builds its own graph
        phi.single_data = vals["phi_s"]
        phi.domain_bnd_data = g.domain_bnd_Sf[:B, 0].clone()
        qn = rollout_scalar_transport(q, phi, vals["nu"], vals["q0"], 12, every=every, bdf2=bdf2)
        return (w * qn).sum()

    grads = {}
    for every in (None, 4):
        params = {k: v.clone().requires_grad_() for k, v in base.items()}
        Jv = J(params, every)
        grads[every] = dict(zip(params, torch.autograd.grad(Jv, list(params.values()))))
    for k in base:
        torch.testing.assert_close(grads[4][k], grads[None][k], rtol=1e-12, atol=1e-15)
    with torch.no_grad():
        for k, b in base.items():
            v = torch.randn(b.shape, dtype=torch.float64, device=g.device, generator=gen)
            if k == "phi_s":  # y and z faces carry no flux: upwind's kink at phi = 0 takes
                v = v * (b != 0)  # torch.clamp's one-sided rule there, not a central difference
            h = 1e-5 * max(float(b.abs().max()), 1e-3)
            fd = (float(J({**base, k: b + h * v}, None)) - float(J({**base, k: b - h * v}, None))) / (2 * h)
            an = float((grads[4][k] * v).sum())
            assert abs(an - fd) <= 1e-6 * max(abs(fd), 1e-6), (k, an, fd)


def test_rollout_arguments():
    from gridfoam_amd.adjoint import checkpointed_rollout

    step = lambda s: (s[0] * 2.0,)  # noqa: E731
    x = torch.ones(3, dtype=torch.float64)
    assert torch.equal(checkpointed_rollout(step, (x,), 5, every=2)[0], x * 32.0)
    with pytest.raises(ValueError):
        checkpointed_rollout(step, (x,), -1)
    with pytest.raises(ValueError):
        checkpointed_rollout(step, (x,), 3, every=0)
