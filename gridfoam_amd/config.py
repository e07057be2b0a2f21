"""YAML case configuration (subset of src/gridfoam/meta/config.py, Python 3.10).

Field names, defaults and validation follow the reference schema for the
parts the pressure-velocity path reads: fluxel (domain, root resolution,
refinement regions, leaf refinement), control, fvSchemes, fvSolution
(algorithm, solvers, adjustPhi), conditions and properties.  Two solver keys
are additions of this build: ``mixed_precision`` (FP32 inner Krylov solve in
an FP64 refinement loop) and ``inner_rel_tolerance``.
"""

from __future__ import annotations

import pathlib
from enum import Enum
from typing import Any, Dict, List, Literal, Optional, Union

import yaml
from pydantic import BaseModel, Field, field_validator, model_validator


class _StrEnum(str, Enum):
    def __str__(self) -> str:  # behave like enum.StrEnum (3.11+)
        return str(self.value)


class SolverType(_StrEnum):  # meta/enums.py:186-196
    CG = "cg"
    BiCGSTAB = "bicgstab"
    PyAMG = "pyamg"


class PreconditionerType(_StrEnum):  # meta/enums.py:222-228
    NONE = "none"
    JACOBI = "jacobi"


class NormType(_StrEnum):  # meta/enums.py:199-219
    L_inf = "l_inf"
    L_2 = "l_2"

    def to_norm_order(self) -> Union[int, float]:
        return float("inf") if self is NormType.L_inf else 2


class DivScheme(_StrEnum):  # meta/enums.py:126-146 (upwind/linear implemented)
    UPWIND = "upwind"
    LINEAR = "linear"
    VANLEER = "vanleer"
    MINMOD = "minmod"
    SUPERBEE = "superbee"
    MONOTONIZED_CENTRAL = "monotonized_central"
    LIMITEDLINEAR = "limitedlinear"
    KOREN = "koren"


class DdtScheme(_StrEnum):
    EULER = "euler"
    BACKWARD = "backward"


class GradScheme(_StrEnum):
    LINEAR = "linear"
    LEASTSQUARE = "leastsquare"


class LaplacianScheme(_StrEnum):  # meta/enums.py:162-183
    LINEAR = "linear"
    CORRECTED = "corrected"
    UNCORRECTED = "uncorrected"
    GAUSS_LINEAR_CORRECTED = "Gauss linear corrected"
    GAUSS_LINEAR_UNCORRECTED = "Gauss linear uncorrected"
    GAUSS_HARMONIC_CORRECTED = "Gauss harmonic corrected"
    GAUSS_HARMONIC_UNCORRECTED = "Gauss harmonic uncorrected"


class SnGradScheme(_StrEnum):
    UNCORRECTED = "uncorrected"
    CORRECTED = "corrected"


class BoundaryConditionType(_StrEnum):  # meta/enums.py:265-275
    DIRICHLET = "dirichlet"
    NEUMANN = "neumann"
    EMPTY = "empty"
    SLIP = "slip"
    INLET_OUTLET = "inlet_outlet"
    FIXED_FLUX_PRESSURE = "fixed_flux_pressure"


class DomainBoundaryPatch(_StrEnum):  # meta/enums.py:53-79
    X_MINUS = "x_minus"
    X_PLUS = "x_plus"
    Y_MINUS = "y_minus"
    Y_PLUS = "y_plus"
    Z_MINUS = "z_minus"
    Z_PLUS = "z_plus"

    def to_direction(self) -> int:
        """dir id with axis = dir // 2 and sign = 2 (dir % 2) - 1."""
        return list(DomainBoundaryPatch).index(self)


class AlgorithmType(_StrEnum):
    SIMPLE = "simple"
    PISO = "piso"
    PIMPLE = "pimple"
    MANUAL = "manual"


class PrecisionType(_StrEnum):
    FLOAT32 = "float32"
    FLOAT64 = "float64"


class DomainConfig(BaseModel, frozen=True):
    lower: List[float]
    upper: List[float]


class RefinementRegionConfig(BaseModel, frozen=True):  # meta/config.py:47-87
    name: Optional[str] = None
    min: List[float]
    max: List[float]
    level: int

    @field_validator("min", "max")
    @classmethod
    def _corner(cls, v: List[float]) -> List[float]:
        if len(v) != 3:
            raise ValueError("refinement region corners must have 3 values")
        return v

    @field_validator("level")
    @classmethod
    def _level(cls, v: int) -> int:
        if v < 0:
            raise ValueError("refinement region level must be non-negative")
        return v

    @model_validator(mode="after")
    def _bounds(self) -> "RefinementRegionConfig":
        if any(lo >= hi for lo, hi in zip(self.min, self.max)):
            raise ValueError("refinement region min must be less than max")
        return self


class ImmersedBodyConfig(BaseModel, frozen=True):
    """An analytic immersed body; it stands in for the surface mesh
    (`mesh_path`) the reference hands to fluxel's AP-IBM builder, and its
    `name` is the patch both sides of every cut face carry.
    kind "plane": normal to `axis` at `offset` (an open surface);
    "box": axis-aligned, `min` - `max`;
    "fourier": prism along z over r(theta) = radius (1 + sum_k a_k cos k theta
    + b_k sin k theta), k = 1..len(a), around `center` (x, y);
    "bluff": the box `min` - `max` with its nose (low-x) edges rounded by
    `radius`, the top rear edge cut by a plane at `slant_angle` degrees over
    the last `slant_length`, standing on up to four vertical `struts` (x, y)
    of radius `strut_radius` that reach `strut_height` below it."""

    name: str
    kind: str
    axis: int = 0
    offset: float = 0.0
    min: Optional[List[float]] = None
    max: Optional[List[float]] = None
    center: Optional[List[float]] = None
    radius: float = 0.0
    a: List[float] = Field(default_factory=list)
    b: List[float] = Field(default_factory=list)
    slant_angle: float = 0.0
    slant_length: float = 0.0
    struts: List[List[float]] = Field(default_factory=list)
    strut_radius: float = 0.0
    strut_height: float = 0.0

    @model_validator(mode="after")
    def _check(self) -> "ImmersedBodyConfig":
        if self.kind not in ("plane", "box", "fourier", "bluff"):
            raise ValueError(f"unknown immersed body kind {self.kind!r}")
        if self.kind == "bluff":
            if self.min is None or self.max is None or len(self.min) != 3 or len(self.max) != 3:
                raise ValueError("a bluff body needs min and max (3 values each)")
            ext = [hi - lo for lo, hi in zip(self.min, self.max)]
            if min(ext) <= 0.0 or not 0.0 <= self.radius <= 0.5 * min(ext[1], ext[2]) or self.radius > ext[0]:
                raise ValueError("a bluff body needs min < max and 0 <= radius <= half its width and height")
            if not 0.0 <= self.slant_angle < 90.0 or not 0.0 <= self.slant_length <= ext[0]:
                raise ValueError("slant_angle must be in [0, 90) degrees and slant_length within the length")
            if len(self.struts) > 4 or any(len(st) != 2 for st in self.struts):
                raise ValueError("at most 4 struts, each (x, y)")
            if self.struts and (self.strut_radius <= 0.0 or self.strut_height <= 0.0):
                raise ValueError("struts need strut_radius > 0 and strut_height > 0")
        if self.name in {p.value for p in DomainBoundaryPatch}:
            raise ValueError(f"immersed body name {self.name!r} is a domain patch name")
        if self.kind == "box" and (self.min is None or self.max is None or len(self.min) != 3 or len(self.max) != 3):
            raise ValueError("a box needs min and max (3 values each)")
        if self.kind == "fourier" and (self.center is None or len(self.center) < 2 or len(self.a) > 8
                                       or len(self.b) > 8 or self.radius <= 0.0):
            raise ValueError("a fourier body needs center (x, y), radius > 0 and at most 8 a_k and b_k")
        return self


class FluxelConfig(BaseModel, frozen=True):  # meta/config.py:90-145
    domain: DomainConfig
    root_resolution: List[int]
    target_level: int = 0
    n_leaf_refinement: int = 0
    refinement_regions: List[RefinementRegionConfig] = Field(default_factory=list)
    mesh_path: Optional[pathlib.Path] = None
    ibm_type: str = "axis_projected"
    motion: str = "static"
    immersed_bodies: List[ImmersedBodyConfig] = Field(default_factory=list)

    @model_validator(mode="after")
    def _no_surface(self) -> "FluxelConfig":
        if self.mesh_path is not None:
            raise ValueError("surface (STL/OBJ) immersed boundaries are not supported by this build")
        return self


class ControlConfig(BaseModel, frozen=True):
    deltaT: float
    endTime: float
    writeInterval: int = 1
    precision: PrecisionType = PrecisionType.FLOAT64
    output: Optional[Dict[str, Any]] = None


class fvSchemesConfig(BaseModel, frozen=True):
    ddtSchemes: Optional[Dict[str, DdtScheme]] = None
    gradSchemes: Optional[Dict[str, GradScheme]] = None
    divSchemes: Optional[Dict[str, DivScheme]] = None
    laplacianSchemes: Optional[Dict[str, LaplacianScheme]] = None
    snGradSchemes: Optional[Dict[str, SnGradScheme]] = None


class SolverConfig(BaseModel, frozen=True):  # meta/config.py:269-316
    method: SolverType
    preconditioner: PreconditionerType = PreconditionerType.NONE
    tolerance: float = 1e-6
    rel_tolerance: float = 1e-6
    max_iter: int = 1000
    norm_type: NormType = NormType.L_2
    max_restart: int = 5
    log_interval: int = 25
    mixed_precision: bool = False
    # mixed: refine once the FP32 residual fell to this fraction of the last
    # true one (or after 128 FP32 iterations, whichever comes first)
    inner_rel_tolerance: float = 0.01


class RelaxationFactorsConfig(BaseModel, frozen=True):
    equations: Dict[str, float] = Field(default_factory=dict)


class SIMPLEAlgorithm(BaseModel, frozen=True):
    type: AlgorithmType = AlgorithmType.SIMPLE
    nNonOrthogonalCorrectors: int = Field(default=0, ge=0)
    residualControl: Dict[str, Any] = Field(default_factory=dict)
    consistent: bool = False
    relaxationFactors: RelaxationFactorsConfig = Field(default_factory=RelaxationFactorsConfig)
    pRefCell: Optional[int] = Field(default=None, ge=0)
    pRefValue: Optional[float] = None


class PISOAlgorithm(BaseModel, frozen=True):
    type: AlgorithmType = AlgorithmType.PISO
    nNonOrthogonalCorrectors: int = Field(default=0, ge=0)
    nCorrectors: int = Field(default=2, ge=1)
    pRefCell: Optional[int] = Field(default=None, ge=0)
    pRefValue: Optional[float] = None


class PIMPLEAlgorithm(BaseModel, frozen=True):  # meta/config.py:467-515
    type: AlgorithmType = AlgorithmType.PIMPLE
    nNonOrthogonalCorrectors: int = Field(default=0, ge=0)
    nCorrectors: int = Field(default=2, ge=1)
    nOuterCorrectors: int = Field(default=1, ge=1)
    residualControl: Dict[str, Any] = Field(default_factory=dict)
    consistent: bool = False
    pRefCell: Optional[int] = Field(default=None, ge=0)
    pRefValue: Optional[float] = None


class ManualAlgorithm(BaseModel, frozen=True):  # meta/config.py:515-516: equations driven by the user
    type: AlgorithmType = AlgorithmType.MANUAL


class fvSolutionConfig(BaseModel, frozen=True):
    algorithm: Union[SIMPLEAlgorithm, PISOAlgorithm, PIMPLEAlgorithm, ManualAlgorithm]
    solvers: Dict[str, SolverConfig] = Field(default_factory=dict)
    adjustPhi: bool = True

    @field_validator("algorithm", mode="before")
    @classmethod
    def _pick(cls, v: Any) -> Any:
        if isinstance(v, dict):
            kind = str(v.get("type", "")).lower()
            if kind == "simple":
                return SIMPLEAlgorithm(**v)
            if kind == "piso":
                return PISOAlgorithm(**v)
            if kind == "pimple":
                return PIMPLEAlgorithm(**v)
            if kind == "manual":
                return ManualAlgorithm(**v)
            raise ValueError(f"unsupported algorithm type {kind!r}")
        return v


class BoundaryConditionConfig(BaseModel, frozen=True):  # meta/config.py:552-585
    type: BoundaryConditionType
    patches: List[str]  # domain patches (DomainBoundaryPatch values) or immersed body names
    value: Optional[Union[float, List[float]]] = None


class ConditionConfig(BaseModel, frozen=True):  # meta/config.py:741-766
    dimension: Optional[Dict[str, Union[float, int]]] = None
    internal: Union[float, List[float]] = 0.0
    boundary: Dict[str, BoundaryConditionConfig] = Field(default_factory=dict)
    export: bool = True


class TransportConfig(BaseModel, frozen=True):
    type: str = "newtonian"
    nu: float


class TurbulenceConfig(BaseModel, frozen=True):
    type: str = "laminar"


class PropertiesConfig(BaseModel, frozen=True):
    transport: TransportConfig
    turbulence: TurbulenceConfig = Field(default_factory=TurbulenceConfig)


class ContinuityErrorConfig(BaseModel, frozen=True):  # meta/config.py:767-778
    phi: str = "phi"
    writeInterval: int = Field(default=1, ge=1)


class SolverInfoConfig(BaseModel, frozen=True):  # meta/config.py:781-792
    fields: Optional[List[str]] = None
    writeInterval: int = Field(default=1, ge=1)


class DragLiftCoord(BaseModel, frozen=True, extra="forbid"):  # meta/config.py:625-638
    mode: Literal["drag_lift"] = "drag_lift"
    drag_dir: List[float]
    lift_dir: List[float]
    center_of_rotation: List[float]

    @model_validator(mode="after")
    def _three(self) -> "DragLiftCoord":
        if any(len(v) != 3 for v in (self.drag_dir, self.lift_dir, self.center_of_rotation)):
            raise ValueError("drag_dir, lift_dir and center_of_rotation must have 3 values")
        return self


class DragPitchCoord(BaseModel, frozen=True, extra="forbid"):  # meta/config.py:641-656
    mode: Literal["drag_pitch"] = "drag_pitch"
    drag_dir: List[float]
    pitch_axis: List[float]
    center_of_rotation: List[float]

    @model_validator(mode="after")
    def _three(self) -> "DragPitchCoord":
        if any(len(v) != 3 for v in (self.drag_dir, self.pitch_axis, self.center_of_rotation)):
            raise ValueError("drag_dir, pitch_axis and center_of_rotation must have 3 values")
        return self


class ForceCoeffConfig(BaseModel, frozen=True):  # meta/config.py:665-740
    local_coord: Union[DragLiftCoord, DragPitchCoord] = Field(discriminator="mode")
    patches: List[str]
    rho: float
    magU_ref: float
    A_ref: float
    L_ref: float

    @model_validator(mode="before")
    @classmethod
    def _infer_mode(cls, data: Any) -> Any:
        """meta/config.py:710-740: the mode follows from lift_dir or pitch_axis."""
        if not isinstance(data, dict):
            return data
        lc = data.get("local_coord")
        if not isinstance(lc, dict):
            raise ValueError("local_coord must be a dictionary")
        if "mode" in lc:
            return data
        has_lift, has_pitch = "lift_dir" in lc, "pitch_axis" in lc
        if has_lift and has_pitch:
            raise ValueError("local_coord is ambiguous: provide either lift_dir or pitch_axis, not both.")
        if not (has_lift or has_pitch):
            raise ValueError("local_coord needs lift_dir or pitch_axis")
        return {**data, "local_coord": {**lc, "mode": "drag_lift" if has_lift else "drag_pitch"}}


class PostProcessingConfig(BaseModel, frozen=True):  # meta/config.py:794-810
    forceCoeff: Optional[ForceCoeffConfig] = None
    continuityError: Optional[ContinuityErrorConfig] = None
    solverInfo: Optional[SolverInfoConfig] = None


class SimulatorConfig(BaseModel, frozen=True):
    control: ControlConfig
    fvSchemes: fvSchemesConfig = Field(default_factory=fvSchemesConfig)
    fvSolution: fvSolutionConfig
    conditions: Dict[str, ConditionConfig] = Field(default_factory=dict)
    properties: PropertiesConfig
    post_processing: Optional[PostProcessingConfig] = None
    device: str = "cuda"

    @model_validator(mode="after")
    def _simple_time(self) -> "SimulatorConfig":  # meta/config.py:852-865
        if isinstance(self.fvSolution.algorithm, SIMPLEAlgorithm):
            if self.control.deltaT != 1:
                raise ValueError("SIMPLE requires deltaT = 1")
        return self


class GridfoamConfig(BaseModel, frozen=True):
    version: Optional[str] = None
    fluxel: FluxelConfig
    simulator: SimulatorConfig


def load_config(path: Union[str, pathlib.Path, dict]) -> GridfoamConfig:
    """Parse and validate a case YAML file (or an already-parsed dict)."""
    data = path if isinstance(path, dict) else yaml.safe_load(pathlib.Path(path).read_text())
    return GridfoamConfig.model_validate(data)


CASES = pathlib.Path(__file__).resolve().parent / "cases"
