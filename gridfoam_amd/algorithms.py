"""SIMPLE / PISO / PIMPLE pressure-velocity drivers on the HIP path.

The drivers stay host Python, as in the reference (src/gridfoam/algorithms/
simple.py:287-408, piso.py:227-326, pimple.py:290-454,
utils/pressure_correction.py:68-363),
and call only native kernels for the per-step work: fused momentum
assembly, batched 3-component BiCGSTAB, rAU/HbyA, interpolation and flux
assembly, adjustPhi, pressure assembly with the hanging-face correction,
PCG (FP64 or FP32-in-FP64-refinement), flux and velocity correction.
"""

from __future__ import annotations

import ctypes
import logging
from dataclasses import dataclass
from typing import Dict, Optional, Tuple

import torch

from . import _native as nat
from .config import ManualAlgorithm, PIMPLEAlgorithm, PISOAlgorithm, PrecisionType, SIMPLEAlgorithm
from .fields import CellField, DirichletBC, FaceField
from .fv import ddt_coefficients, fvc, fvm, sn_grad_corrected
from .fvmatrix import Equation, FvMatrix
from .grid import OctreeGrid, create_grid
from .immersed import apply_immersed_dirichlet_values, close_fixed_cell_flux
from .solvers import LinearSolver, SolveStats, create_solver, resolve_solver

logger = logging.getLogger(__name__)


@dataclass(frozen=True)
class PressureSolveResult:  # pressure_correction.py:43-65
    matrix: FvMatrix
    stats: Tuple[SolveStats, ...]
    initial_residual: Optional[torch.Tensor]
    last_initial_residual: Optional[torch.Tensor]


def needs_reference_value(p: CellField) -> bool:
    """reference_value.py:14-30."""
    return not any(isinstance(bc, DirichletBC) for bc in p.bcs.values())


def set_reference_value(mat: FvMatrix, ref_cell: int, ref_value: float) -> None:
    """reference_value.py:33-51."""
    nat.call("gf_set_reference", nat.ptr(mat._diag), nat.ptr(mat._source), ctypes.c_int32(ref_cell),
             ctypes.c_double(ref_value), nat.stream())


def field_initial_residual(mat: FvMatrix, field: CellField) -> torch.Tensor:
    """residual.py:12-36 -> device scalar ``|A psi - b|_2 / (|b|_2 + eps)``."""
    g = mat.grid
    nb = (g.num_cells + 255) // 256
    scratch = torch.zeros(2 * nb + 8, dtype=torch.float64, device=g.device)
    out = torch.zeros(1, dtype=torch.float64, device=g.device)
    nat.call("gf_initial_residual", g.mesh_ref, ctypes.c_int32(field.nc), nat.ptr(mat._diag), nat.ptr(mat._upper),
             nat.ptr(mat._lower), nat.ptr(mat._source), nat.ptr(field.data), nat.ptr(scratch), nat.ptr(out),
             nat.stream())
    return out[0]


def correct_flux(phi: FaceField, U: CellField, update_internal: bool = False) -> None:
    """fv/flux.py:132-162."""
    g = U.grid
    st = U.boundary_state()
    if update_internal:
        Uf = fvc.interpolate(U)
        nat.call("gf_face_flux", g.mesh_ref, nat.ptr(Uf.single_buffer), nat.ptr(Uf.bnd_buffer), nat.ptr(st.flux_keep),
                 nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer), nat.stream())
    nat.call("gf_constrain_bnd_flux", g.mesh_ref, nat.ptr(st.psi_b), nat.ptr(st.covered), nat.ptr(phi.bnd_buffer),
             nat.stream())


def compute_phi_hbya(phi_hbya: FaceField, HbyA: CellField, U: CellField, *, rAU: Optional[CellField] = None,
                     phi: Optional[FaceField] = None, abc=None) -> FaceField:
    """fv/flux.py:74-129 (+ rAU_f * ddtCorr when rAU/phi are given)."""
    g = U.grid
    Hf = fvc.interpolate(HbyA)
    st = U.boundary_state()
    nat.call("gf_face_flux", g.mesh_ref, nat.ptr(Hf.single_buffer), nat.ptr(Hf.bnd_buffer), nat.ptr(st.flux_keep),
             nat.ptr(phi_hbya.single_buffer), nat.ptr(phi_hbya.bnd_buffer), nat.stream())
    if rAU is not None:
        _, b, c = abc
        older = U.older_data if c != 0.0 else None
        phi_older = phi.older_single_data if c != 0.0 else None
        nat.call("gf_ddt_corr_add", g.mesh_ref, nat.ptr(U.old_data), nat.ptr(older), nat.ptr(phi.old_single_data),
                 nat.ptr(phi_older), nat.ptr(rAU.data), ctypes.c_double(b), ctypes.c_double(c), ctypes.c_double(g.dt),
                 nat.ptr(phi_hbya.single_buffer), nat.stream())
    nat.call("gf_constrain_bnd_flux", g.mesh_ref, nat.ptr(st.psi_b), nat.ptr(st.nonassignable),
             nat.ptr(phi_hbya.bnd_buffer), nat.stream())
    return phi_hbya


def adjust_phi(phi: FaceField, U: CellField, p: Optional[CellField] = None) -> None:
    """fv/adjust_phi.py:44-127 (decided and applied on the device, no sync)."""
    if p is not None and any(isinstance(bc, DirichletBC) for bc in p.bcs.values()):
        return
    st = U.boundary_state()
    nat.call("gf_adjust_phi", U.grid.mesh_ref, nat.ptr(st.adjust_kind), nat.ptr(phi.bnd_buffer), None, nat.stream())


def solve_pressure_poisson(p: CellField, rAtU: CellField, div_phi_hbya: torch.Tensor, solver: LinearSolver, *,
                           n_non_orthogonal_correctors: int, p_needs_ref: bool, p_ref_cell: Optional[int] = None,
                           p_ref_value: Optional[float] = None, final_solver: Optional[LinearSolver] = None,
                           compute_residual: bool = True) -> PressureSolveResult:
    """pressure_correction.py:68-160."""
    g = p.grid
    mat = None
    stats = None
    first = last = None
    for corr in range(n_non_orthogonal_correctors + 1):
        mat = fvm.assemble(p, lap=-1.0, gamma=rAtU.data)
        nat.call("gf_predictor_source", ctypes.c_int32(g.num_cells), ctypes.c_int32(1), nat.ptr(mat._source),
                 nat.ptr(g.cell_volumes), nat.ptr(div_phi_hbya), nat.ptr(mat._source), nat.stream())
        if p_needs_ref:
            if p_ref_cell is None or p_ref_value is None:
                raise ValueError("p_ref_cell and p_ref_value are required when p_needs_ref is True")
            set_reference_value(mat, p_ref_cell, p_ref_value)
        eq = Equation(p, mat)
        if compute_residual:
            last = field_initial_residual(eq.fv_matrix, p)
            if first is None:
                first = last
        s = final_solver if (final_solver is not None and corr == n_non_orthogonal_correctors) else solver
        res = s.solve(eq)
        p.data = res.solution
        stats = res.stats
    return PressureSolveResult(mat, stats, first, last)


def correct_phi(phi: FaceField, phi_hbya: FaceField, p_eqn: FvMatrix, p: CellField, rAtU: CellField) -> None:
    """pressure_correction.py:270-335: phi = phiHbyA - pEqn.flux(p) with the
    boundary normal-gradient flux, then the mass-balance closure of cells
    fixed by immersed snapping."""
    _face_flux_correction(phi, phi_hbya, p_eqn, p, rAtU)
    close_fixed_cell_flux(phi, p)


def _face_flux_correction(phi: FaceField, phi_hbya: FaceField, p_eqn: FvMatrix, p: CellField,
                          rAtU: CellField) -> None:
    """phi = phiHbyA - pEqn.flux(p) on internal faces; boundary faces
    phiHbyA_b - rAtU_P |Sf| snGrad(p)_b (pressure_correction.py:197-217, :297-304)."""
    g = p.grid
    st = p.boundary_state()
    nat.call("gf_correct_phi", g.mesh_ref, nat.ptr(phi_hbya.single_buffer), nat.ptr(phi_hbya.bnd_buffer),
             nat.ptr(p_eqn._upper), nat.ptr(p_eqn._lower), nat.ptr(p_eqn.face_flux_correction), nat.ptr(p.data),
             nat.ptr(rAtU.data), nat.ptr(st.frac), nat.ptr(st.ref_v), nat.ptr(st.ref_g), nat.ptr(st.covered),
             nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer), nat.stream())


def simplec_rAtU(UEqn: FvMatrix, rAU: CellField, *, bounded: bool = True) -> torch.Tensor:
    """pressure_correction.py:163-193: 1/max(1/rAU - H1, 0.1/rAU)."""
    h1 = UEqn.H1()
    out = torch.empty_like(rAU.data)
    nat.call("gf_simplec_rautu", ctypes.c_int32(rAU.grid.num_cells), nat.ptr(rAU.data), nat.ptr(h1),
             ctypes.c_int32(int(bounded)), nat.ptr(out), nat.stream())
    return out


def apply_simplec(phi_hbya: FaceField, HbyA: CellField, p: CellField, rAU: CellField, rAtU: CellField) -> None:
    """pressure_correction.py:222-267: phiHbyA += (rAtU-rAU)_f |Sf| snGrad(p) on
    every face (boundary faces with the pressure condition's normal
    gradient) and HbyA -= (rAU - rAtU) grad(p).  The face term is the flux
    of laplacian(rAtU - rAU, p), added by the same kernel that corrects phi."""
    g = p.grid
    d_r = CellField(g, "rAtU-rAU")
    d_r.data = rAtU.data - rAU.data
    m = fvm.assemble(p, lap=1.0, gamma=d_r.data, corrected=sn_grad_corrected(g, p.name))
    neg = CellField(g, "rAU-rAtU")
    neg.data = -d_r.data
    _face_flux_correction(phi_hbya, phi_hbya, m, p, neg)
    gp = fvc.grad(p)
    nat.call("gf_correct_velocity", ctypes.c_int32(g.num_cells), nat.ptr(HbyA.data), nat.ptr(neg.data), nat.ptr(gp),
             nat.ptr(HbyA.data), nat.stream())


def correct_velocity(U: CellField, HbyA: CellField, rAtU: CellField, p: CellField) -> None:
    """pressure_correction.py:338-363: U = HbyA - rAtU grad(p), then the
    prescribed values of snapped immersed cells."""
    gp = fvc.grad(p)
    nat.call("gf_correct_velocity", ctypes.c_int32(U.grid.num_cells), nat.ptr(HbyA.data), nat.ptr(rAtU.data),
             nat.ptr(gp), nat.ptr(U.data), nat.stream())
    apply_immersed_dirichlet_values(U)


def compute_continuity_error(phi: FaceField, delta_t: float) -> Tuple[float, float]:
    """post/diagnostics/continuity.py:12-49: (local, global) volume-weighted
    continuity errors of the face flux, computed in FP64 on the GPU."""
    g = phi.grid
    div = torch.empty(g.num_cells, dtype=torch.float64, device=g.device)
    out = torch.empty(2, dtype=torch.float64, device=g.device)
    nat.call("gf_continuity_error", g.mesh_ref, nat.ptr(phi.single_buffer), nat.ptr(phi.bnd_buffer),
             ctypes.c_double(delta_t), nat.ptr(div), nat.ptr(out), nat.stream())
    local, glob = out.tolist()
    return local, glob


def continuity_residual(phi: FaceField) -> float:
    """algorithms/utils/residual.py:39-53: |div(phi)|_2."""
    return float(torch.linalg.vector_norm(fvc.div(phi)))


def residual_satisfied(residual: float, tolerance: float, rel_tolerance: float, initial: float) -> bool:
    """algorithms/utils/residual.py:56-87."""
    return residual < tolerance or (rel_tolerance > 0.0 and residual < rel_tolerance * initial)


def normalize_residual_control(value: dict) -> Dict[str, Tuple[float, float]]:
    """meta/config.py:364-388: a float is an absolute tolerance."""
    out = {}
    for name, entry in (value or {}).items():
        if isinstance(entry, (int, float)):
            out[name] = (float(entry), 0.0)
        else:
            out[name] = (float(entry["tolerance"]), float(entry.get("rel_tolerance", 0.0)))
    return out


def _predictor(UEqn: FvMatrix, U: CellField, grad_p: torch.Tensor) -> FvMatrix:
    """UEqn.with_source(source - V grad p): shares the LDU arrays (fvmatrix.py:213-240)."""
    g = U.grid
    pred = FvMatrix.__new__(FvMatrix)
    pred.field, pred.grid = U, g
    pred._diag, pred._upper, pred._lower = UEqn._diag, UEqn._upper, UEqn._lower
    pred.face_flux_correction = None
    pred._source = torch.empty_like(UEqn._source)
    nat.call("gf_predictor_source", ctypes.c_int32(g.num_cells), ctypes.c_int32(3), nat.ptr(UEqn._source),
             nat.ptr(g.cell_volumes), nat.ptr(grad_p), nat.ptr(pred._source), nat.stream())
    return pred


class AlgorithmBase:
    """algorithms/base.py:46-100 (field set-up shared by SIMPLE and PISO)."""

    def __init__(self, grid: OctreeGrid, transient: bool):
        self._grid = grid
        cfg = grid.sim_config
        cond = cfg.conditions
        self.U = CellField(grid, "U", (3,), transient=transient, condition=cond.get("U"))
        self.p = CellField(grid, "p", (), condition=cond.get("p"))
        self.phi = FaceField(grid, "phi")
        self.phi_hbya = FaceField(grid, "phiHbyA")
        self.rAU = CellField(grid, "rAU")
        self.rAtU = self.rAU  # a field of its own with the SIMPLEC (consistent) formulation
        self.consistent = False
        self.HbyA = CellField(grid, "HbyA", (3,))
        self._solvers: Dict[str, LinearSolver] = {k: create_solver(v) for k, v in cfg.fvSolution.solvers.items()}
        self.nu = float(cfg.properties.transport.nu)
        self.adjust_phi_enabled = cfg.fvSolution.adjustPhi
        alg = cfg.fvSolution.algorithm
        self.n_non_orthogonal_correctors = alg.nNonOrthogonalCorrectors
        self.p_needs_ref = needs_reference_value(self.p)
        if self.p_needs_ref:
            if alg.pRefCell is None:
                raise ValueError("pRefCell is required when p_needs_ref is True")
            if alg.pRefValue is None:
                raise ValueError("pRefValue is required when p_needs_ref is True")
            self.p_ref_cell, self.p_ref_value = alg.pRefCell, alg.pRefValue
        else:
            self.p_ref_cell = self.p_ref_value = None
        self.solve_stats: Dict[str, Tuple[SolveStats, ...]] = {}
        self._diagnostics = None
        self._diagnostics_step = 0
        self.last_p_residual: Optional[torch.Tensor] = None
        self.last_p_matrix: Optional[FvMatrix] = None
        self.p_solve_history: list = []

    @property
    def grid(self) -> OctreeGrid:
        return self._grid

    @property
    def solvers(self) -> Dict[str, LinearSolver]:
        return self._solvers

    def has_simulation_converged(self) -> bool:
        """algorithms/base.py:62-64: only a steady SIMPLE run ends early."""
        return False

    def attach_diagnostics(self, diagnostics) -> None:
        """algorithms/base.py:66-75: a DiagnosticsCollector recording every step."""
        self._diagnostics = diagnostics

    def _finalize_diagnostics(self) -> None:
        """algorithms/base.py:77-100: one record per completed step."""
        if self._diagnostics is None:
            return
        self._diagnostics_step += 1
        self._diagnostics.record_step(self._diagnostics_step, self.phi, self.solve_stats)

    def _grad_p_into(self) -> torch.Tensor:
        return fvc.grad(self.p)

    def _pressure_corrector(self, UEqn: FvMatrix, *, transient: bool, abc, final: bool) -> PressureSolveResult:
        g = self.grid
        nat.call("gf_rau_hbya", g.mesh_ref, nat.ptr(UEqn._diag), nat.ptr(UEqn._upper), nat.ptr(UEqn._lower),
                 nat.ptr(UEqn._source), nat.ptr(self.U.data), nat.ptr(self.rAU.data), nat.ptr(self.HbyA.data),
                 nat.stream())
        if transient:
            compute_phi_hbya(self.phi_hbya, self.HbyA, self.U, rAU=self.rAU, phi=self.phi, abc=abc)
        else:
            compute_phi_hbya(self.phi_hbya, self.HbyA, self.U)
        if self.adjust_phi_enabled:
            adjust_phi(self.phi_hbya, self.U, self.p)
        if self.consistent:  # SIMPLEC: rAtU = 1/max(1/rAU - H1, 0.1/rAU), consistent phiHbyA / HbyA
            self.rAtU.data = simplec_rAtU(UEqn, self.rAU)
            apply_simplec(self.phi_hbya, self.HbyA, self.p, self.rAU, self.rAtU)
        div_phi_hbya = fvc.div(self.phi_hbya)
        p_solver = resolve_solver(self.solvers, self.p.name, is_final=final) if transient else self.solvers["p"]
        res = solve_pressure_poisson(self.p, self.rAtU, div_phi_hbya, self.solvers["p"],
                                     n_non_orthogonal_correctors=self.n_non_orthogonal_correctors,
                                     p_needs_ref=self.p_needs_ref, p_ref_cell=self.p_ref_cell,
                                     p_ref_value=self.p_ref_value, final_solver=p_solver if transient else None)
        self.solve_stats[self.p.name] = res.stats
        self.last_p_residual = res.initial_residual
        self.last_p_matrix = res.matrix
        self.p_solve_history.extend(res.stats)
        return res


class SIMPLE(AlgorithmBase):
    """algorithms/simple.py:141-408 (steady; SIMPLEC with consistent: true)."""

    def __init__(self, grid: OctreeGrid):
        super().__init__(grid, transient=False)
        alg = grid.sim_config.fvSolution.algorithm
        assert isinstance(alg, SIMPLEAlgorithm)
        if alg.consistent:
            self.consistent = True
            self.rAtU = CellField(grid, "rAtU")
        self.alpha_U = alg.relaxationFactors.equations["U"]
        self.alpha_p = alg.relaxationFactors.equations["p"]
        self._residual_control = normalize_residual_control(alg.residualControl)
        self._initial_residuals: Dict[str, float] = {}
        self._current_residuals: Dict[str, float] = {}
        correct_flux(self.phi, self.U, update_internal=True)

    def _record_residual(self, name: str, residual) -> None:
        """simple.py:277-285: the first recorded value stays the initial one."""
        r = float(residual)
        self._initial_residuals.setdefault(name, r)
        self._current_residuals[name] = r

    def has_converged(self) -> bool:
        """simple.py:241-271: every residualControl entry below its absolute
        tolerance (rel_tolerance is ignored); False when none is set."""
        if not self._residual_control:
            return False
        for name, (tol, _rtol) in self._residual_control.items():
            res, init = self._current_residuals.get(name), self._initial_residuals.get(name)
            if res is None or init is None:
                return False
            if not residual_satisfied(res, tol, 0.0, init):
                return False
        return True

    def has_simulation_converged(self) -> bool:
        """simple.py:273-275: a converged steady iteration ends the run."""
        return self.has_converged()

    def step(self) -> None:
        g, U, p = self.grid, self.U, self.p
        UEqn = fvm.assemble(U, div=1.0, lap=-1.0, phi=self.phi, gamma=self.nu)
        nat.call("gf_relax_momentum", ctypes.c_int32(g.num_cells), ctypes.c_double(self.alpha_U),
                 nat.ptr(UEqn._diag), nat.ptr(UEqn._source), nat.ptr(U.data), nat.stream())
        gp = fvc.grad(p)
        pred = FvMatrix.__new__(FvMatrix)
        pred.field, pred.grid = U, g
        pred._diag, pred._upper, pred._lower = UEqn._diag, UEqn._upper, UEqn._lower
        pred.face_flux_correction = None
        pred._source = torch.empty_like(UEqn._source)
        nat.call("gf_predictor_source", ctypes.c_int32(g.num_cells), ctypes.c_int32(3), nat.ptr(UEqn._source),
                 nat.ptr(g.cell_volumes), nat.ptr(gp), nat.ptr(pred._source), nat.stream())
        u_eq = Equation(U, pred)
        if U.name in self._residual_control:
            self._record_residual(U.name, field_initial_residual(u_eq.fv_matrix, U))
        res = self.solvers["U"].solve(u_eq)
        U.data = res.solution
        self.solve_stats[U.name] = res.stats
        UEqn = Equation(U, UEqn).fv_matrix  # simple.py:335: constrained pressure-free equation for H/A
        p_old = p.data.clone()
        pres = self._pressure_corrector(UEqn, transient=False, abc=None, final=False)
        if p.name in self._residual_control:
            self._record_residual(p.name, pres.initial_residual)
        correct_phi(self.phi, self.phi_hbya, pres.matrix, p, self.rAtU)
        nat.call("gf_relax_field", ctypes.c_int64(g.num_cells), ctypes.c_double(self.alpha_p), nat.ptr(p_old),
                 nat.ptr(p.data), nat.stream())
        apply_immersed_dirichlet_values(p)
        correct_velocity(U, self.HbyA, self.rAtU, p)
        if "phi" in self._residual_control:
            self._record_residual("phi", continuity_residual(self.phi))
        self._finalize_diagnostics()


class PISO(AlgorithmBase):
    """algorithms/piso.py:86-326."""

    def __init__(self, grid: OctreeGrid):
        super().__init__(grid, transient=True)
        alg = grid.sim_config.fvSolution.algorithm
        assert isinstance(alg, PISOAlgorithm)
        self.n_correctors = alg.nCorrectors
        correct_flux(self.phi, self.U, update_internal=True)
        self.U.update_history(reset=True)
        self.phi.update_history(reset=True)

    def step(self) -> None:
        g, U, p = self.grid, self.U, self.p
        abc = ddt_coefficients(U)
        UEqn = fvm.assemble(U, ddt=1.0, div=1.0, lap=-1.0, phi=self.phi, gamma=self.nu, abc=abc)
        gp = fvc.grad(p)
        pred = FvMatrix.__new__(FvMatrix)
        pred.field, pred.grid = U, g
        pred._diag, pred._upper, pred._lower = UEqn._diag, UEqn._upper, UEqn._lower
        pred.face_flux_correction = None
        pred._source = torch.empty_like(UEqn._source)
        nat.call("gf_predictor_source", ctypes.c_int32(g.num_cells), ctypes.c_int32(3), nat.ptr(UEqn._source),
                 nat.ptr(g.cell_volumes), nat.ptr(gp), nat.ptr(pred._source), nat.stream())
        res = self.solvers["U"].solve(Equation(U, pred))
        U.data = res.solution
        self.solve_stats[U.name] = res.stats
        UEqn = Equation(U, UEqn).fv_matrix  # piso.py:257
        for i in range(self.n_correctors):
            pres = self._pressure_corrector(UEqn, transient=True, abc=abc, final=i == self.n_correctors - 1)
            correct_phi(self.phi, self.phi_hbya, pres.matrix, p, self.rAtU)
            correct_velocity(U, self.HbyA, self.rAtU, p)
        U.update_history()
        self.phi.update_history()
        self._finalize_diagnostics()


class PIMPLE(AlgorithmBase):
    """algorithms/pimple.py:136-454 (laminar): outer correctors around the
    PISO loop; the outer loop stops one (final) pass after residualControl
    is met; SIMPLEC with consistent: true."""

    def __init__(self, grid: OctreeGrid):
        super().__init__(grid, transient=True)
        alg = grid.sim_config.fvSolution.algorithm
        assert isinstance(alg, PIMPLEAlgorithm)
        self.n_outer_correctors = alg.nOuterCorrectors
        self.n_correctors = alg.nCorrectors
        if alg.consistent:
            self.consistent = True
            self.rAtU = CellField(grid, "rAtU")
        self._residual_control = normalize_residual_control(alg.residualControl)
        self._initial_residuals: Dict[str, float] = {}
        self._current_residuals: Dict[str, float] = {}
        self._outer_converged = False
        self.outer_iterations = 0
        correct_flux(self.phi, self.U, update_internal=True)
        self.U.update_history(reset=True)
        self.phi.update_history(reset=True)

    def has_converged(self) -> bool:
        """pimple.py:246-257."""
        return self._outer_converged

    def _record_residual(self, name: str, residual) -> None:
        r = float(residual)
        self._initial_residuals.setdefault(name, r)
        self._current_residuals[name] = r

    def _check_outer_convergence(self, *, allow_relative: bool = True) -> bool:
        """pimple.py:269-287."""
        if not self._residual_control:
            return False
        for name, (tol, rtol) in self._residual_control.items():
            res, init = self._current_residuals.get(name), self._initial_residuals.get(name)
            if res is None or init is None:
                return False
            if not residual_satisfied(res, tol, rtol if allow_relative else 0.0, init):
                return False
        return True

    def step(self) -> None:
        g, U, p = self.grid, self.U, self.p
        abc = ddt_coefficients(U)
        self._outer_converged = False
        self._initial_residuals.clear()
        self._current_residuals.clear()
        for outer in range(self.n_outer_correctors):
            final_after = 0 < outer < self.n_outer_correctors - 1 and \
                self._check_outer_convergence(allow_relative=outer > 1)
            if final_after:
                self._outer_converged = True
            UEqn = fvm.assemble(U, ddt=1.0, div=1.0, lap=-1.0, phi=self.phi, gamma=self.nu, abc=abc)
            u_eq = Equation(U, _predictor(UEqn, U, fvc.grad(p)))
            if U.name in self._residual_control:
                self._record_residual(U.name, field_initial_residual(u_eq.fv_matrix, U))
            res = self.solvers["U"].solve(u_eq)
            U.data = res.solution
            self.solve_stats[U.name] = res.stats
            UEqn = Equation(U, UEqn).fv_matrix  # pimple.py:346
            for i in range(self.n_correctors):
                pres = self._pressure_corrector(UEqn, transient=True, abc=abc, final=i == self.n_correctors - 1)
                if p.name in self._residual_control:
                    if i == 0:
                        self._record_residual(p.name, pres.initial_residual)
                    self._current_residuals[p.name] = float(pres.last_initial_residual)
                correct_phi(self.phi, self.phi_hbya, pres.matrix, p, self.rAtU)
                correct_velocity(U, self.HbyA, self.rAtU, p)
            if "phi" in self._residual_control:
                self._record_residual("phi", continuity_residual(self.phi))
            self.outer_iterations = outer + 1
            if final_after:
                break
        U.update_history()
        self.phi.update_history()
        self._finalize_diagnostics()


def create_algorithm(grid: OctreeGrid):
    """algorithms/factory.py:9-38."""
    if grid.sim_config.control.precision != PrecisionType.FLOAT64:
        # the fields, assembly and true residuals are FP64 here; an FP32 inner
        # pressure solve is the solver's mixed_precision option
        raise ValueError("control.precision float32 is not supported: fields and assembly run in FP64 "
                         "(use mixed_precision on the pressure solver for FP32 inner solves)")
    alg = grid.sim_config.fvSolution.algorithm
    if isinstance(alg, SIMPLEAlgorithm):
        return SIMPLE(grid)
    if isinstance(alg, PISOAlgorithm):
        return PISO(grid)
    if isinstance(alg, PIMPLEAlgorithm):
        return PIMPLE(grid)
    if isinstance(alg, ManualAlgorithm):
        raise ValueError("Manual algorithm has no step sequence")  # algorithms/factory.py:37-38
    raise ValueError(f"unsupported algorithm {alg}")


def create_case(config, device: str = "cuda"):
    """Grid + algorithm from a validated GridfoamConfig."""
    grid = create_grid(config, device=device)
    grid.sim_config = config.simulator
    return grid, create_algorithm(grid)
