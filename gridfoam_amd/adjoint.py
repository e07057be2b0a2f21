"""Checkpointed transient rollouts for gradients through many time steps.

A transient run whose step is differentiable -- native assembly
(``gf_assemble`` / ``gf_assemble_vjp``) plus a Krylov solve (transpose solve
as its backward) -- keeps the autograd graph of every step by default, which
grows with the step count.  ``checkpointed_rollout`` keeps only the state at
segment boundaries and recomputes each segment's steps in the backward pass
(the steps are deterministic: fixed reduction orders in every kernel), so the
memory held is one segment's graph plus the boundary states.

``ScalarTransport`` is such a step: implicit (Euler or BDF2) transport of a
scalar by a given face flux with a diffusivity, ``ddt(q) + div(phi, q) -
laplacian(gamma, q) = 0`` (fv/fvm/ddt.py, fv/fvm/div.py, fv/fvm/laplacian.py),
differentiable with respect to the initial field, the diffusivity, the flux
and the boundary values.
"""

from __future__ import annotations

from typing import Callable, Optional, Sequence, Tuple, Union

import torch
from torch.utils.checkpoint import checkpoint

from .config import DivScheme, SolverConfig
from .fields import CellField, DirichletBC, FaceField
from .fv import fvm
from .fvmatrix import Equation
from .solvers import LinearSolver, create_solver

State = Tuple[torch.Tensor, ...]


class _Segment(torch.autograd.Function):
    """``k`` calls of ``step`` run without a graph; the backward recomputes
    them with one from the saved input state and differentiates the
    recomputation with respect to that state and to ``params`` (tensors the
    step reads from its closure).  Only the segment's input state is held
    between the passes, whatever the step's own nodes keep."""

    @staticmethod
    def forward(ctx, step, k, n_state, *tensors):  # noqa: D401
        state, params = tensors[:n_state], tensors[n_state:]
        ctx.step, ctx.k, ctx.n_state, ctx.params = step, k, n_state, params
        ctx.save_for_backward(*state)
        out = tuple(state)
        for _ in range(k):
            out = step(out)
        return tuple(o.clone() if any(o is t for t in state) else o for o in out)

    @staticmethod
    def backward(ctx, *g_out):
        n = ctx.n_state
        need = ctx.needs_input_grad[3:]
        with torch.enable_grad():
            s = tuple(t.detach().requires_grad_(bool(need[i])) for i, t in enumerate(ctx.saved_tensors))
            out = s
            for _ in range(ctx.k):
                out = ctx.step(out)
        pairs = [(o, g) for o, g in zip(out, g_out) if g is not None and o.requires_grad]
        wanted = [i for i in range(len(need)) if need[i]]
        ins = [s[i] if i < n else ctx.params[i - n] for i in wanted]
        grads = [None] * len(need)
        if pairs and ins:
            got = torch.autograd.grad([o for o, _ in pairs], ins, [g for _, g in pairs], allow_unused=True)
            for i, g in zip(wanted, got):
                grads[i] = g
        return (None, None, None, *grads)


def checkpointed_rollout(step: Callable[[State], State], state: State, n_steps: int,
                         every: Optional[int] = None, params: Optional[Sequence[torch.Tensor]] = None) -> State:
    """Advance ``state`` by ``n_steps`` calls of ``step``.  With ``every``,
    the steps run in segments of that many and the backward pass recomputes
    a segment from its saved input state.

    With ``params`` (every tensor that the step reads from outside ``state``
    and whose gradient is wanted: a viscosity, boundary values, the
    immersed geometry of a ``FourierShape``), a segment's forward keeps no
    graph at all, so the memory held is the segment boundary states plus,
    in the backward pass, one segment's graph; gradients reach exactly the
    state and ``params``, which must not be computed from one another (a
    gradient reaching both a tensor and its ancestor would count twice).
    Without ``params``, torch's activation
    checkpointing runs the segments: gradients reach every tensor the step
    reads, but the nodes of the native operations keep their launch
    descriptions (and the tensors those hold) for every step."""
    if n_steps < 0:
        raise ValueError("n_steps must be >= 0")
    if every is not None and every < 1:
        raise ValueError("every must be >= 1")

    def segment(k: int):
        def run(*s):
            out = tuple(s)
            for _ in range(k):
                out = step(out)
            return out
        return run

    done = 0
    while done < n_steps:
        k = n_steps - done if every is None else min(every, n_steps - done)
        if every is None or not torch.is_grad_enabled():
            state = segment(k)(*state)
        elif params is not None:
            ps = tuple(t for t in params if torch.is_tensor(t))  # numbers carry no gradient
            state = _Segment.apply(step, k, len(state), *state, *ps)
        else:
            state = checkpoint(segment(k), *state, use_reentrant=False)
        done += k
    return state


class ScalarTransport:
    """One implicit transport step of a scalar field ``q`` (with its boundary
    conditions) by the face flux ``phi``; the state is ``(q^n,)`` for Euler
    and ``(q^n, q^{n-1})`` for BDF2 (constant time step).  The explicit parts
    (hanging-face corrections) use the previous level, as the reference's
    deferred corrections do.  The step sets ``q``'s data and history, so
    after a backward pass that recomputed segments ``q`` holds the last
    recomputed step's values."""

    def __init__(self, q: CellField, phi: FaceField, gamma: Union[float, torch.Tensor], *,
                 scheme: DivScheme = DivScheme.UPWIND, bdf2: bool = False,
                 solver: Optional[LinearSolver] = None):
        if q.component_shape != ():
            raise ValueError("ScalarTransport advances a scalar field")
        self.q, self.phi, self.gamma = q, phi, gamma
        self.scheme, self.bdf2 = scheme, bdf2
        self.solver = solver if solver is not None else create_solver(SolverConfig(
            method="bicgstab", preconditioner="jacobi", tolerance=0.0, rel_tolerance=1e-13, max_iter=2000))

    def __call__(self, state: State) -> State:
        q = self.q
        old = state[0]
        older = state[1] if self.bdf2 else None
        abc = (1.5, 2.0, 0.5) if self.bdf2 else (1.0, 1.0, 0.0)
        q.old_data, q.older_data = old, older
        q.data = old
        m = fvm.assemble(q, ddt=1.0, div=1.0, lap=-1.0, phi=self.phi, gamma=self.gamma, abc=abc,
                         scheme=self.scheme)
        new = self.solver.solve(Equation(q, m)).solution
        return (new, old) if self.bdf2 else (new,)


def rollout_scalar_transport(q: CellField, phi: FaceField, gamma, q0: torch.Tensor, n_steps: int, *,
                             every: Optional[int] = None, bdf2: bool = False,
                             scheme: DivScheme = DivScheme.UPWIND,
                             solver: Optional[LinearSolver] = None) -> torch.Tensor:
    """q after ``n_steps`` implicit steps from ``q0`` (BDF2 starts from
    ``q^{-1} = q0``), checkpointed in segments of ``every`` steps."""
    step = ScalarTransport(q, phi, gamma, scheme=scheme, bdf2=bdf2, solver=solver)
    state: Sequence[torch.Tensor] = (q0, q0) if bdf2 else (q0,)
    return checkpointed_rollout(step, tuple(state), n_steps, every)[0]


# --------------------------------------------------------------------------
# Differentiable PISO / PIMPLE step
# --------------------------------------------------------------------------
def _pad(t: torch.Tensor, rows: int) -> torch.Tensor:
    if t.shape[0] == rows:
        return t
    return torch.cat([t, torch.zeros((rows - t.shape[0],) + tuple(t.shape[1:]), dtype=t.dtype, device=t.device)])


def _set_face(ff: FaceField, single: torch.Tensor, bnd: torch.Tensor) -> None:
    """Point a face field's blocks at graph-carrying tensors."""
    ff._single = _pad(single, ff._single.shape[0])
    ff._bnd = _pad(bnd, ff._bnd.shape[0])


class TransientStep:
    """One PISO (algorithms/piso.py:227-326) or PIMPLE (pimple.py:290-454)
    step of an existing driver, written with operations that carry the
    autograd graph: native assembly with its VJP kernels (``gf_assemble`` /
    ``gf_assemble_vjp``), Krylov solves with the transpose-solve adjoint, the
    least-squares gradient, interpolation, boundary-value and divergence
    nodes (``gf_lsq_grad_vjp``, ``gf_interpolate_vjp``, ``gf_bc_state_vjp``,
    ``gf_fvc_div_vjp``), and tensor arithmetic for A, H, HbyA, the face
    fluxes, ddtCorr, constrainHbyA, adjustPhi, the reference cell,
    correct_phi and correct_velocity (pressure_correction.py:68-363,
    fv/flux.py:32-129, fv/fvc/ddt_corr.py:13-78, fv/adjust_phi.py:44-127).
    Gradients reach the initial state, the viscosity (a 0-d tensor given as
    ``nu``) and boundary-condition values that require grad.

    The state is ``(U, phi_single, phi_boundary, p, phiHbyA_single,
    phiHbyA_boundary, rAU)`` at the start of a step (Euler time derivative:
    U_old = U, phi_old = phi_single); the last three are the registered
    fields that conditions such as fixedFluxPressure read, carried so that a
    recomputed step sees what the original saw.  Those conditions' own
    values (inletOutlet's switch, fixedFluxPressure's gradient) are
    re-evaluated without a graph, so gradients do not run through them.
    PIMPLE runs every outer corrector (no residual-control exit).  Velocity
    cells fixed by immersed snapping (Gibou constraints, fv/boundary_ops.py:
    440-535) are eliminated from the momentum solve by the differentiable
    ``with_fixed_values`` and get their prescribed values back after the
    velocity correction (pressure_correction.py:338-363).  Not covered:
    SIMPLEC (``consistent``), backward (BDF2) time derivatives, TVD
    convection schemes and snapped pressure cells (a pressure condition that
    prescribes values on an immersed body); those raise."""

    def __init__(self, alg, *, nu: Union[float, torch.Tensor, None] = None):
        from .algorithms import PIMPLE, PISO
        from .config import DdtScheme
        from .fv import ddt_scheme, div_scheme

        if not isinstance(alg, (PISO, PIMPLE)):
            raise TypeError("TransientStep runs a PISO or PIMPLE driver")
        if alg.consistent:
            raise NotImplementedError("SIMPLEC (consistent) steps are not differentiable here")
        g = alg.grid
        if ddt_scheme(g, "U") != DdtScheme.EULER:
            raise NotImplementedError("only the Euler time derivative is differentiable here")
        if div_scheme(g, "phi", "U") not in (DivScheme.UPWIND, DivScheme.LINEAR):
            raise NotImplementedError("only upwind and linear convection are differentiable here")
        self.alg, self.grid = alg, g
        self.nu = alg.nu if nu is None else nu
        self.n_outer = getattr(alg, "n_outer_correctors", 1)
        self.n_corr = alg.n_correctors

    # ---- pieces -------------------------------------------------------------
    def _check_unconstrained(self, field) -> None:
        g = self.grid
        if g.num_immersed_faces:
            from .immersed import immersed_dirichlet_constraints

            if immersed_dirichlet_constraints(field)[0].numel():
                raise NotImplementedError("snapped immersed pressure cells are not differentiable here")

    def _snapped(self, Udata: torch.Tensor) -> torch.Tensor:
        """The prescribed values of snapped velocity cells written back
        (boundary_ops.py:537-541) with a graph-carrying index_copy."""
        if self.grid.num_immersed_faces == 0:
            return Udata
        from .immersed import immersed_dirichlet_constraints

        cells, vals = immersed_dirichlet_constraints(self.alg.U)
        if cells.numel() == 0:
            return Udata
        return Udata.index_copy(0, cells, vals.to(Udata.dtype))

    def _bnd_values(self, field) -> torch.Tensor:
        from .fv import boundary_values

        return boundary_values(field)[: self.grid.num_bnd_faces]

    def _phi_hbya(self, HbyA: torch.Tensor, rAU: torch.Tensor, U_old: torch.Tensor, phi_old: torch.Tensor):
        """compute_phi_hbya (fv/flux.py:74-129) with rAU_f ddtCorr."""
        from .fv import fvc

        g, alg = self.grid, self.alg
        F, B = g.num_internal_faces, g.num_bnd_faces
        U = alg.U
        Hf_field = alg.HbyA
        Hf_field.data = HbyA
        Hf = fvc.interpolate(Hf_field)
        Sf, bSf = g.Sf[:F], g.bnd_Sf[:B]
        st = U.boundary_state()
        keep = st.flux_keep[:B].to(torch.float64)
        phi_s = (Hf.single_data * Sf).sum(dim=1)
        phi_b = (Hf.bnd_data * bSf).sum(dim=1) * keep
        # ddtCorr (Euler: b = 1, c = 0)
        o, n = g.owner, g.neighbour
        w = g.w[:F]
        U0f = w[:, None] * U_old[o] + (1.0 - w)[:, None] * U_old[n]
        corr = phi_old - (U0f * Sf).sum(dim=1)
        # 1 - min(|corr| / (|phi_old| + tiny), 1), differentiated only where the
        # ratio is below 1 and phi_old is not zero (elsewhere its derivative is
        # 0 and the quotient rule would give 0 * inf)
        tiny = torch.finfo(torch.float64).tiny
        a, den = corr.abs(), phi_old.abs() + tiny
        smooth = (a < den) & (phi_old != 0)
        ratio = torch.where(smooth, a / torch.where(smooth, den, torch.ones_like(den)),
                            torch.where(a < den, a.detach() / den.detach(), torch.ones_like(den)))
        coupling = 1.0 - ratio
        rAUf = w * rAU[o] + (1.0 - w) * rAU[n]
        phi_s = phi_s + rAUf * (coupling * corr / g.dt)
        # constrainHbyA: the velocity condition's flux on non-assignable faces
        nonassign = st.nonassignable[:B].bool()
        Ub = self._bnd_values(U).reshape(B, 3)
        phi_b = torch.where(nonassign, (Ub * bSf).sum(dim=1), phi_b)
        if alg.adjust_phi_enabled and not any(isinstance(bc, DirichletBC) for bc in alg.p.bcs.values()):
            phi_b = self._adjust_phi(phi_b, st.adjust_kind[:B])
        return phi_s, phi_b

    @staticmethod
    def _adjust_phi(phi_b: torch.Tensor, kind: torch.Tensor) -> torch.Tensor:
        """fv/adjust_phi.py:44-127 on the domain faces (kind 1 fixed value, 2 adjustable)."""
        zero = torch.zeros_like(phi_b)
        on = kind > 0
        mass_in = -torch.where(on & (phi_b < 0), phi_b, zero).sum()
        fixed_out = torch.where((kind == 1) & (phi_b > 0), phi_b, zero).sum()
        adj_out = torch.where((kind == 2) & (phi_b > 0), phi_b, zero).sum()
        if float(adj_out.detach()) <= 0.0:
            return phi_b
        corr = (mass_in - fixed_out) / adj_out
        return torch.where((kind == 2) & (phi_b > 0), phi_b * corr, phi_b)

    def _pressure(self, rAU: torch.Tensor, phi_s: torch.Tensor, phi_b: torch.Tensor, final: bool):
        """solve_pressure_poisson (pressure_correction.py:68-160)."""
        from .fv import fvc, fvm
        from .solvers import resolve_solver

        alg, g = self.alg, self.grid
        _set_face(alg.phi_hbya, phi_s, phi_b)
        div = fvc.div(alg.phi_hbya)
        V = g.cell_volumes[: g.num_cells]
        p = alg.p
        self._check_unconstrained(p)
        rf = alg.rAU
        rf.data = rAU
        mat = None
        for corr in range(alg.n_non_orthogonal_correctors + 1):
            mat = fvm.assemble(p, lap=-1.0, gamma=rAU)
            src = mat.source - div * V
            if alg.p_needs_ref:
                c, v = alg.p_ref_cell, alg.p_ref_value
                rd = mat.diag[c]
                d = mat.diag.clone()
                d[c] = d[c] + rd
                src = src.clone()
                src[c] = src[c] + rd * v
                mat._diag = d
            mat._source = src
            last = corr == alg.n_non_orthogonal_correctors
            solver = resolve_solver(alg.solvers, p.name, is_final=final) if last else alg.solvers["p"]
            p.data = solver.solve(Equation(p, mat)).solution
        return mat

    def _correct_phi(self, phi_s, phi_b, pmat, rAU):
        """correct_phi (pressure_correction.py:270-305): phiHbyA minus the
        pressure flux; boundary faces with the condition's normal gradient."""
        alg, g = self.alg, self.grid
        F, B = g.num_internal_faces, g.num_bnd_faces
        p = alg.p.data
        o, n = g.owner, g.neighbour
        flux = pmat.upper * p[n] - pmat.lower * p[o]
        if pmat.face_flux_correction is not None:
            flux = flux + pmat.face_flux_correction[:F]
        phi_s = phi_s + flux  # pmat = -laplacian, so the laplacian's flux enters with a minus
        st = alg.p.boundary_state()
        f = st.frac[:B].clone()  # later refreshes rewrite the state arrays in place
        rv = (st.ref_v_ad if st.ref_v_ad is not None else st.ref_v.clone())[:B, 0]
        rg = (st.ref_g_ad if st.ref_g_ad is not None else st.ref_g.clone())[:B, 0]
        active = (f != 0) & ~(g.bnd_snap[:B].bool() & (f == 1))
        dist = torch.where(active, g.bnd_distances()[:B], torch.ones_like(f))
        coef = torch.where(active, f / dist, torch.zeros_like(f))
        cells = g.bnd_owner[:B]
        sng = coef * (rv - p[cells]) + (1.0 - f) * rg
        cov = st.covered[:B].bool()
        phi_b = torch.where(cov, phi_b - rAU[cells] * g.bnd_mag_sf[:B] * sng, phi_b)
        return phi_s, phi_b

    # ---- the step -------------------------------------------------------------
    def __call__(self, state: State) -> State:
        from .fv import fvc

        U0, phi_s, phi_b, p0, ph_s, ph_b, rAU = state
        alg, g = self.alg, self.grid
        U, p = alg.U, alg.p
        V = g.cell_volumes[: g.num_cells]
        U.old_data, U.older_data = U0, None
        U.data = U0
        p.data = p0
        _set_face(alg.phi_hbya, ph_s, ph_b)
        alg.rAU.data = rAU
        phi_old = phi_s
        for outer in range(self.n_outer):
            _set_face(alg.phi, phi_s, phi_b)
            UEqn = fvm.assemble(U, ddt=1.0, div=1.0, lap=-1.0, phi=alg.phi, gamma=self.nu, abc=(1.0, 1.0, 0.0))
            pred = UEqn.with_source(UEqn.source - V[:, None] * fvc.grad(p))
            U.data = alg.solvers["U"].solve(Equation(U, pred)).solution
            UEqn = Equation(U, UEqn).fv_matrix  # piso.py:257 / pimple.py:346: snapped cells eliminated
            for i in range(self.n_corr):
                final = outer == self.n_outer - 1 and i == self.n_corr - 1
                rAU = 1.0 / UEqn.A()
                HbyA = rAU[:, None] * UEqn.H(U.data)
                ph_s, ph_b = self._phi_hbya(HbyA, rAU, U0, phi_old)
                pmat = self._pressure(rAU, ph_s, ph_b, final)
                phi_s, phi_b = self._correct_phi(ph_s, ph_b, pmat, rAU)
                _set_face(alg.phi, phi_s, phi_b)
                U.data = self._snapped(HbyA - rAU[:, None] * fvc.grad(p))
        return U.data, phi_s, phi_b, p.data, ph_s, ph_b, rAU


def transient_state(alg) -> State:
    """The (U, phi_single, phi_boundary, p, phiHbyA_single,
    phiHbyA_boundary, rAU) state of a PISO / PIMPLE driver."""
    return tuple(t.detach().clone() for t in (alg.U.data, alg.phi.single_data, alg.phi.bnd_data, alg.p.data,
                                              alg.phi_hbya.single_data, alg.phi_hbya.bnd_data, alg.rAU.data))
