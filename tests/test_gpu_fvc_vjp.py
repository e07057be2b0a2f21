"""GPU: reverse mode of the explicit face operators -- fvc.interpolate
(fv/schemes/interpolate.py:20-63 with the hanging-face corrections of
fv/kernels/face_interpolation.py:114-197, boundary values from the
conditions), fvc.div (fv/fvc/div.py:15-64), fvc.reconstruct
(fv/fvc/reconstruct.py:27-107) and the Green-Gauss gradient
(fv/schemes/grad.py:21-44) -- against central
differences.  Both maps are linear in the field values and the boundary
values, so the differences are exact to rounding; the transport step built
on an interpolated flux is differentiable away from phi = 0 (upwind), and
the flux there is kept well away from zero."""

from __future__ import annotations

import copy
import sys

import pytest
import torch

from tests.cases import fluxel_block, small_cavity2d

pytestmark = pytest.mark.gpu


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", file=sys.stderr, flush=True)


def _grid(name):
    from gridfoam_amd.config import load_config
    from gridfoam_amd.grid import create_grid

    cfg = copy.deepcopy(small_cavity2d(8))
    cfg["fluxel"] = {**cfg["fluxel"], **fluxel_block(name)}
    return create_grid(load_config(cfg))


def _flux(g, Uf, name):
    """phi = U_f . Sf on every face block (fv/flux.py:32-129 without the
    boundary-condition constraints)."""
    from gridfoam_amd.fields import FaceField

    F, B = g.num_internal_faces, g.num_domain_bnd_faces
    phi = FaceField(g, name)
    phi.single_data = (Uf.single_data * g.Sf[:F]).sum(1)
    phi.domain_bnd_data = (Uf.domain_bnd_data * g.domain_bnd_Sf[:B]).sum(1)
    return phi


def _check(J, base, gen, dev, tol=1e-7):
    params = {k: v.clone().requires_grad_() for k, v in base.items()}
    grads = dict(zip(params, torch.autograd.grad(J(params), list(params.values()))))
    with torch.no_grad():
        for k, b in base.items():
            v = torch.randn(b.shape, dtype=torch.float64, device=dev, generator=gen)
            h = 1e-4 * max(float(b.abs().max()), 1.0)
            fd = (float(J({**base, k: b + h * v})) - float(J({**base, k: b - h * v}))) / (2 * h)
            an = float((grads[k] * v).sum())
            assert abs(an - fd) <= tol * max(abs(fd), 1e-3), (k, an, fd)


@pytest.mark.parametrize("name", ["refined2d", "refined3d"])
@pytest.mark.parametrize("cs", [(), (3,)])
def test_interpolate_vjp_matches_central_differences(name, cs):
    """Face values (internal and boundary blocks) of fvc.interpolate with
    respect to the cell values, a Dirichlet value and a Neumann gradient
    (vector fields also with a slip wall)."""
    from gridfoam_amd.fields import CellField, DirichletBC, NeumannBC, SlipBC
    from gridfoam_amd.fv import fvc

    g = _grid(name)
    assert g.num_hanging_faces > 0
    dev = g.device
    gen = torch.Generator(dev).manual_seed(5)
    _say(f"interpolate vjp {name} cs={cs}")
    F, B = g.num_internal_faces, g.num_domain_bnd_faces
    P = CellField(g, f"ip{len(cs)}", cs)
    base = {"psi": torch.randn(g.num_cells, *cs, dtype=torch.float64, device=dev, generator=gen),
            "dir": torch.randn(cs, dtype=torch.float64, device=dev, generator=gen),
            "neu": torch.randn(cs, dtype=torch.float64, device=dev, generator=gen)}
    Ws = torch.randn(F, *cs, dtype=torch.float64, device=dev, generator=gen)
    Wb = torch.randn(B, *cs, dtype=torch.float64, device=dev, generator=gen)

    def J(vals):
        bcs = {"x_minus": DirichletBC(vals["dir"]), "y_plus": NeumannBC(vals["neu"])}
        if cs:
            bcs["y_minus"] = SlipBC()
        P.add_boundary_conditions(bcs)
        P.data = vals["psi"]
        f = fvc.interpolate(P)
        return (f.single_data * Ws).sum() + (f.domain_bnd_data * Wb).sum()

    _check(J, base, gen, dev)
    # the tracked path gives the same values as the untracked one
    with torch.no_grad():
        P.data = base["psi"]
        ref = fvc.interpolate(P)
    P.data = base["psi"].clone().requires_grad_()
    got = fvc.interpolate(P)
    assert torch.equal(got.single_data.detach(), ref.single_data)
    assert torch.equal(got.domain_bnd_data.detach(), ref.domain_bnd_data)


@pytest.mark.parametrize("name", ["refined2d", "refined3d"])
def test_div_of_interpolated_flux_vjp_matches_central_differences(name):
    """fvc.div of the flux U_f . Sf (scalar) and of the vector face values
    themselves, with respect to the cell velocity and a Dirichlet wall
    velocity."""
    from gridfoam_amd.fields import CellField, DirichletBC
    from gridfoam_amd.fv import fvc

    g = _grid(name)
    dev = g.device
    gen = torch.Generator(dev).manual_seed(8)
    _say(f"div vjp {name}")
    U = CellField(g, "Udiv", (3,))
    base = {"U": torch.randn(g.num_cells, 3, dtype=torch.float64, device=dev, generator=gen),
            "lid": torch.randn(3, dtype=torch.float64, device=dev, generator=gen)}
    w1 = torch.randn(g.num_cells, dtype=torch.float64, device=dev, generator=gen)
    w3 = torch.randn(g.num_cells, 3, dtype=torch.float64, device=dev, generator=gen)

    def J(vals):
        U.add_boundary_conditions({"y_plus": DirichletBC(vals["lid"]), "x_minus": DirichletBC([0.0, 0.0, 0.0])})
        U.data = vals["U"]
        Uf = fvc.interpolate(U)
        return (fvc.div(_flux(g, Uf, "phi_div")) * w1).sum() + (fvc.div(Uf) * w3).sum()

    _check(J, base, gen, dev)
    with torch.no_grad():
        U.data = base["U"]
        Uf = fvc.interpolate(U)
        ref = fvc.div(_flux(g, Uf, "phi_ref"))
    U.data = base["U"].clone().requires_grad_()
    got = fvc.div(_flux(g, fvc.interpolate(U), "phi_got"))
    assert torch.equal(got.detach(), ref)


@pytest.mark.parametrize("bdf2", [False, True])
def test_transport_by_an_interpolated_velocity_is_differentiable(bdf2):
    """Four implicit transport steps of a scalar by the flux of an
    interpolated velocity field, checkpointed in segments of two: the
    gradient of a functional of the final scalar with respect to the cell
    velocity and the inlet velocity equals central differences."""
    from gridfoam_amd.adjoint import rollout_scalar_transport
    from gridfoam_amd.fields import CellField, DirichletBC, NeumannBC
    from gridfoam_amd.fv import fvc

    g = _grid("refined3d")
    dev = g.device
    gen = torch.Generator(dev).manual_seed(12)
    _say(f"transport by interpolated velocity bdf2={bdf2}")
    q = CellField(g, "q_tr", ())
    q.add_boundary_conditions({"x_minus": DirichletBC(1.0), "x_plus": NeumannBC(0.0)})
    x = g.cell_centers[:, 0]
    q0 = torch.exp(-20.0 * (x - 0.3) ** 2)
    mean = torch.tensor([1.0, 0.4, 0.3], dtype=torch.float64, device=dev)
    base = {"U": mean + 0.05 * torch.randn(g.num_cells, 3, dtype=torch.float64, device=dev, generator=gen),
            "inlet": mean.clone()}
    W = torch.randn(g.num_cells, dtype=torch.float64, device=dev, generator=gen)
    U = CellField(g, "U_tr", (3,))

    def J(vals):
        U.add_boundary_conditions({"x_minus": DirichletBC(vals["inlet"])})
        U.data = vals["U"]
        phi = _flux(g, fvc.interpolate(U), "phi_tr")
        return (rollout_scalar_transport(q, phi, 1e-3, q0, 4, every=2, bdf2=bdf2) * W).sum()

    with torch.no_grad():
        U.add_boundary_conditions({"x_minus": DirichletBC(base["inlet"])})
        U.data = base["U"]
        phi = _flux(g, fvc.interpolate(U), "phi_chk")
        smallest = min(float(phi.single_data.abs().min()), float(phi.domain_bnd_data.abs().min()))
    assert smallest > 1e-6  # no face flux near upwind's kink at phi = 0 for the steps below
    _check(J, base, gen, dev, tol=1e-6)


@pytest.mark.parametrize("name", ["refined2d", "refined3d"])
@pytest.mark.parametrize("cs", [(), (3,)])
def test_green_gauss_gradient_vjp_matches_central_differences(name, cs):
    """The Green-Gauss gradient (gradSchemes linear, fv/schemes/grad.py:21-44)
    of the linear face values with respect to the cell values, a Dirichlet
    value and a Neumann gradient; the tracked gradient equals the untracked
    one."""
    from gridfoam_amd.config import load_config
    from gridfoam_amd.fields import CellField, DirichletBC, NeumannBC
    from gridfoam_amd.fv import fvc
    from gridfoam_amd.grid import create_grid

    cfg = copy.deepcopy(small_cavity2d(8))
    cfg["fluxel"] = {**cfg["fluxel"], **fluxel_block(name)}
    cfg["simulator"]["fvSchemes"]["gradSchemes"] = {"default": "linear"}
    g = create_grid(load_config(cfg))
    dev = g.device
    gen = torch.Generator(dev).manual_seed(17)
    _say(f"green-gauss vjp {name} cs={cs}")
    P = CellField(g, f"gg{len(cs)}", cs)
    base = {"psi": torch.randn(g.num_cells, *cs, dtype=torch.float64, device=dev, generator=gen),
            "dir": torch.randn(cs, dtype=torch.float64, device=dev, generator=gen),
            "neu": torch.randn(cs, dtype=torch.float64, device=dev, generator=gen)}
    W = torch.randn(g.num_cells, *cs, 3, dtype=torch.float64, device=dev, generator=gen)

    def J(vals):
        P.add_boundary_conditions({"x_plus": DirichletBC(vals["dir"]), "y_minus": NeumannBC(vals["neu"])})
        P.data = vals["psi"]
        return (fvc.grad(P) * W).sum()

    _check(J, base, gen, dev)
    with torch.no_grad():
        P.data = base["psi"]
        ref = fvc.grad(P).clone()
    P.data = base["psi"].clone().requires_grad_()
    assert torch.equal(fvc.grad(P).detach(), ref)


@pytest.mark.parametrize("name", ["refined2d", "refined3d"])
def test_reconstruct_vjp_matches_central_differences(name):
    """fvc.reconstruct (fvc/reconstruct.py:27-107) of a face flux with
    respect to its internal and boundary values; the tracked result equals
    the untracked one."""
    from gridfoam_amd.fields import CellField, FaceField
    from gridfoam_amd.fv import fvc

    g = _grid(name)
    dev = g.device
    gen = torch.Generator(dev).manual_seed(23)
    _say(f"reconstruct vjp {name}")
    F, B = g.num_internal_faces, g.num_domain_bnd_faces
    base = {"s": torch.randn(F, dtype=torch.float64, device=dev, generator=gen),
            "b": torch.randn(B, dtype=torch.float64, device=dev, generator=gen)}
    W = torch.randn(g.num_cells, 3, dtype=torch.float64, device=dev, generator=gen)
    U = CellField(g, "U_rec", (3,))

    def J(vals):
        phi = FaceField(g, "phi_rec")
        phi.single_data = vals["s"]
        phi.domain_bnd_data = vals["b"]
        return (fvc.reconstruct(phi, U) * W).sum()

    _check(J, base, gen, dev)
    with torch.no_grad():
        phi = FaceField(g, "phi_rec_ref")
        phi.single_data, phi.domain_bnd_data = base["s"], base["b"]
        ref = fvc.reconstruct(phi, U).clone()
    phi = FaceField(g, "phi_rec_got")
    phi.single_data = base["s"].clone().requires_grad_()
    phi.domain_bnd_data = base["b"]
    assert torch.equal(fvc.reconstruct(phi, U).detach(), ref)


@pytest.mark.parametrize("cs", [(), (3,)])
def test_matrix_h_and_a_vjp_match_central_differences(cs):
    """FvMatrix.H(x) and A() (core/fvmatrix.py:242-290) of an assembled
    ddt + div - laplacian matrix, with respect to the diffusivity (through
    the assembly's reverse mode) and x: the SpMV's reverse mode is the
    transposed SpMV and the LDU outer products."""
    from gridfoam_amd.fields import CellField, DirichletBC, FaceField, NeumannBC
    from gridfoam_amd.fv import fvm

    g = _grid("refined3d")
    dev = g.device
    gen = torch.Generator(dev).manual_seed(29)
    _say(f"matrix H/A vjp cs={cs}")
    F, B = g.num_internal_faces, g.num_domain_bnd_faces
    q = CellField(g, f"qH{len(cs)}", cs)
    q.add_boundary_conditions({"x_minus": DirichletBC(torch.ones(cs, dtype=torch.float64, device=dev)),
                               "x_plus": NeumannBC(torch.zeros(cs, dtype=torch.float64, device=dev))})
    q.data = torch.randn(g.num_cells, *cs, dtype=torch.float64, device=dev, generator=gen)
    q.update_history(reset=True)
    phi = FaceField(g, "phi_H")
    phi.single_data = torch.randn(F, dtype=torch.float64, device=dev, generator=gen) * 0.01
    phi.domain_bnd_data = torch.zeros(B, dtype=torch.float64, device=dev)
    base = {"gamma": torch.tensor(0.05, dtype=torch.float64, device=dev),
            "x": torch.randn(g.num_cells, *cs, dtype=torch.float64, device=dev, generator=gen)}
    W1 = torch.randn(g.num_cells, *cs, dtype=torch.float64, device=dev, generator=gen)
    W2 = torch.randn(g.num_cells, dtype=torch.float64, device=dev, generator=gen)

    def J(vals):
        m = fvm.assemble(q, ddt=1.0, div=1.0, lap=-1.0, phi=phi, gamma=vals["gamma"])
        return (m.H(vals["x"]) * W1).sum() + (m.A() * W2).sum()

    _check(J, base, gen, dev)
    with torch.no_grad():
        m = fvm.assemble(q, ddt=1.0, div=1.0, lap=-1.0, phi=phi, gamma=base["gamma"])
        ref = m.H(base["x"]).clone()
    got = m.H(base["x"].clone().requires_grad_())
    assert torch.equal(got.detach(), ref)
