"""Constitutive, forcing, phase-field and extension closures, mirroring the
constructors (and their validation) of the reference:

  src/TransportConstitutive.jl  ConstantViscosity :89-97, ConstantBodyForce,
                                KolmogorovBodyForce :340-355
  src/StructuralConstitutive.jl MaxwellRelaxation :22-30,
                                LinearInternalTensorStress :40-48
  src/InterfaceRegularization.jl ConservativeDiffuseInterfaceFlux :23-63
  src/PhaseFieldMethods.jl      CurrentCDIPhaseField :37-51, JainACDIPhaseField
                                :53-91, PirozzoliUWCDIPhaseField :93-119
  src/StructureTensor.jl        IdentityOutsideStructure :71-83,
                                GeometricStructureExtension :106-131
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from .errors import ArgumentError, UnsupportedError


class UnsupportedModel(UnsupportedError):
    pass


def _finite(x, msg):
    x = float(x)
    if not math.isfinite(x):
        raise ArgumentError(msg)
    return x


# ---- viscosity, elasticity, relaxation -------------------------------------
class NoViscosity:
    pass


@dataclass(frozen=True)
class ConstantViscosity:
    value: float

    def __post_init__(self):
        v = _finite(self.value, "Viscosity must be finite.")
        if not v > 0:
            raise ArgumentError("Viscosity must be strictly positive.")


class NoElasticStress:
    pass


@dataclass(frozen=True)
class LinearInternalTensorStress:
    coefficient: float

    def __post_init__(self):
        v = _finite(self.coefficient, "Elastic-stress coefficient must be finite.")
        if v < 0:
            raise ArgumentError("Elastic-stress coefficient must be non-negative.")


class NoRelaxation:
    pass


@dataclass(frozen=True)
class MaxwellRelaxation:
    rate: float

    def __post_init__(self):
        v = _finite(self.rate, "Maxwell relaxation rate must be finite.")
        if v < 0:
            raise ArgumentError("Maxwell relaxation rate must be non-negative.")


# ---- forcing ------------------------------------------------------------------
class NoBodyForce:
    pass


@dataclass(frozen=True)
class ConstantBodyForce:
    fx: float
    fy: float

    def __post_init__(self):
        _finite(self.fx, "Body-force x-component must be finite.")
        _finite(self.fy, "Body-force y-component must be finite.")


@dataclass(frozen=True)
class KolmogorovBodyForce:
    amplitude: float
    wavenumber: float
    phase: float = 0.0

    def __post_init__(self):
        _finite(self.amplitude, "Kolmogorov-force amplitude must be finite.")
        k = _finite(self.wavenumber, "Kolmogorov-force wavenumber must be finite.")
        if not k > 0:
            raise ArgumentError("Kolmogorov-force wavenumber must be positive.")
        _finite(self.phase, "Kolmogorov-force phase must be finite.")


class NoCapillarity:
    pass


@dataclass
class DiffuseAreaCapillarity:
    """DiffuseAreaCapillarity(coefficient; delta_normal, smoothing_passes)
    (Capillarity.jl:21-41): f = coefficient * kappa * grad(phi), kappa = -div(n)."""
    coefficient: float
    delta_normal: float = 1e-6
    smoothing_passes: int = 0

    def __post_init__(self):
        c = _finite(self.coefficient, "coefficient must be finite.")
        if not c >= 0.0:
            raise ArgumentError("coefficient must be nonnegative.")
        d = _finite(self.delta_normal, "delta_normal must be finite.")
        if not d > 0.0:
            raise ArgumentError("delta_normal must be positive.")
        if int(self.smoothing_passes) != self.smoothing_passes or self.smoothing_passes < 0:
            raise ArgumentError("smoothing_passes must be nonnegative.")
        self.coefficient, self.delta_normal, self.smoothing_passes = float(c), float(d), int(self.smoothing_passes)


# ---- growth / resource (GrowthModels.jl, TransportConstitutive.jl:181-335) -----
class NoGrowth:
    """G = 0 (GrowthModels.jl:214-223)."""


@dataclass
class ConstantIntrinsicGrowth:
    """G = rate * phi (GrowthModels.jl:55-75, 225-235)."""
    rate: float

    def __post_init__(self):
        self.rate = _finite(self.rate, "Intrinsic growth rate must be finite.")


@dataclass
class SaturatingResourceGrowth:
    """G = rate * phi * c / (half_saturation + c) (GrowthModels.jl:77-104, 237-252)."""
    rate: float
    half_saturation: float

    def __post_init__(self):
        self.rate = _finite(self.rate, "Resource-limited growth rate must be finite.")
        self.half_saturation = _finite(self.half_saturation, "Growth half-saturation constant must be finite.")
        if not self.half_saturation > 0.0:
            raise ArgumentError("Growth half-saturation constant must be strictly positive.")


class IsotropicGrowthTensor:
    """K_g = G/2 I (GrowthModels.jl:114-130, 268-289)."""


@dataclass
class DirectionalGrowthTensor:
    """K_g = G t (x) t with t normalised by the constructor (GrowthModels.jl:132-166)."""
    direction_x: float
    direction_y: float

    def __post_init__(self):
        tx, ty = float(self.direction_x), float(self.direction_y)
        if not (math.isfinite(tx) and math.isfinite(ty)):
            raise ArgumentError("Directional-growth components must be finite.")
        mag = math.hypot(tx, ty)
        if not mag > 0.0:
            raise ArgumentError("Directional growth requires a non-zero direction vector.")
        self.direction_x, self.direction_y = tx / mag, ty / mag


class NoDiffusion:
    """No resource diffusion; no implicit resource solve (TransportConstitutive.jl:183-184)."""


@dataclass
class ConstantDiffusivity:
    """TransportConstitutive.jl:186-196."""
    value: float

    def __post_init__(self):
        self.value = _finite(self.value, "Diffusivity must be finite.")
        if not self.value > 0.0:
            raise ArgumentError("Diffusivity must be strictly positive.")


@dataclass
class HarmonicPhaseDiffusivity:
    """D = 1 / ((1 - phi)/D_f + phi/D_s), phi unclamped (TransportConstitutive.jl:198-225, 255-263)."""
    fluid_value: float
    structure_value: float

    def __post_init__(self):
        self.fluid_value = _finite(self.fluid_value, "Fluid diffusivity must be finite.")
        self.structure_value = _finite(self.structure_value, "Structure diffusivity must be finite.")
        if not self.fluid_value > 0.0:
            raise ArgumentError("Fluid diffusivity must be strictly positive.")
        if not self.structure_value > 0.0:
            raise ArgumentError("Structure diffusivity must be strictly positive.")


class NoResourceConsumption:
    """C = 0 (TransportConstitutive.jl:267-268)."""


@dataclass
class SaturatingResourceConsumption:
    """C = phi * rate * c / (half_saturation + c) (TransportConstitutive.jl:270-295, 303-335)."""
    rate: float
    half_saturation: float

    def __post_init__(self):
        self.rate = _finite(self.rate, "Consumption rate must be finite.")
        self.half_saturation = _finite(self.half_saturation, "Consumption half-saturation constant must be finite.")
        if not self.rate >= 0.0:
            raise ArgumentError("Consumption rate must be non-negative.")
        if not self.half_saturation > 0.0:
            raise ArgumentError("Consumption half-saturation constant must be strictly positive.")


class NoInterfaceFlux:
    pass


@dataclass(frozen=True)
class ConservativeDiffuseInterfaceFlux:
    inverse_peclet: float
    epsilon: float
    delta_normal: float = 1e-6

    def __post_init__(self):
        ip = _finite(self.inverse_peclet, "inverse_peclet must be finite.")
        ep = _finite(self.epsilon, "epsilon must be finite.")
        dn = _finite(self.delta_normal, "delta_normal must be finite.")
        if ip < 0:
            raise ArgumentError("inverse_peclet must be non-negative.")
        if not ep > 0:
            raise ArgumentError("epsilon must be strictly positive.")
        if not dn > 0:
            raise ArgumentError("delta_normal must be strictly positive.")


class AbstractPhaseFieldMethod:
    pass


@dataclass(frozen=True)
class LegacySplitPhaseFieldMethod(AbstractPhaseFieldMethod):
    interface_flux: object


class NoPhaseFieldRegularization(AbstractPhaseFieldMethod):
    pass


class CurrentCDIPhaseField(AbstractPhaseFieldMethod):
    def __init__(self, inverse_peclet, epsilon, delta_normal=1e-6):
        self.flux = ConservativeDiffuseInterfaceFlux(inverse_peclet, epsilon, delta_normal)


@dataclass(frozen=True)
class JainACDIPhaseField(AbstractPhaseFieldMethod):
    gamma: float
    epsilon: float
    psi_floor: float = 1e-100
    delta_normal: float = 1e-14
    dynamic_gamma: bool = True

    def __post_init__(self):
        g = float(self.gamma)
        if not (math.isfinite(g) and g >= 0):
            raise ArgumentError("gamma must be finite and non-negative.")
        e = float(self.epsilon)
        if not (math.isfinite(e) and e > 0):
            raise ArgumentError("epsilon must be finite and strictly positive.")
        f = float(self.psi_floor)
        if not (math.isfinite(f) and 0 < f < 0.5):
            raise ArgumentError("psi_floor must lie strictly between 0 and 0.5.")
        d = float(self.delta_normal)
        if not (math.isfinite(d) and d > 0):
            raise ArgumentError("delta_normal must be finite and strictly positive.")


@dataclass(frozen=True)
class PirozzoliUWCDIPhaseField(AbstractPhaseFieldMethod):
    alpha: float = 1.0
    delta_normal: float = 1e-14
    normal_reconstruction: str = "standard"

    def __post_init__(self):
        a = float(self.alpha)
        if not (math.isfinite(a) and a >= 1.0):
            raise ArgumentError("alpha must be finite and >= 1 for the published UWCDI stability range.")
        d = float(self.delta_normal)
        if not (math.isfinite(d) and d > 0):
            raise ArgumentError("delta_normal must be finite and strictly positive.")
        if self.normal_reconstruction not in ("standard", "youngs", "smoothed9"):
            raise ArgumentError("normal_reconstruction must be :standard, :youngs, or :smoothed9.")


def jain_equilibrium_initialization_epsilon(model: JainACDIPhaseField) -> float:
    """PhaseFieldMethods.jl:121"""
    return 2.0 * model.epsilon


def phase_field_label(model) -> str:
    """PhaseFieldMethods.jl:123-129"""
    if isinstance(model, NoPhaseFieldRegularization):
        return "upwind_no_regularization"
    if isinstance(model, CurrentCDIPhaseField):
        return "current_cdi"
    if isinstance(model, JainACDIPhaseField):
        return "jain_acdi"
    if isinstance(model, PirozzoliUWCDIPhaseField):
        return "pirozzoli_uwcdi"
    if isinstance(model, LegacySplitPhaseFieldMethod):
        return "legacy_split_" + type(model.interface_flux).__name__
    raise ArgumentError(f"unknown phase-field method {model!r}")


# ---- structure-tensor extension --------------------------------------------------------
class NoStructureExtension:
    pass


@dataclass(frozen=True)
class IdentityOutsideStructure:
    cutoff: float

    def __post_init__(self):
        c = _finite(self.cutoff, "Structure-extension cutoff must be finite.")
        if not 0 <= c <= 1:
            raise ArgumentError("Structure-extension cutoff must lie in [0, 1].")


@dataclass(frozen=True)
class GeometricStructureExtension:
    max_rate: float
    exponent: float = 4.0
    safety_cutoff: float = 1e-6

    def __post_init__(self):
        r = _finite(self.max_rate, "Geometric extension rate must be finite.")
        if r < 0:
            raise ArgumentError("Geometric extension rate must be non-negative.")
        e = _finite(self.exponent, "Geometric extension exponent must be finite.")
        if not e > 0:
            raise ArgumentError("Geometric extension exponent must be positive.")
        c = _finite(self.safety_cutoff, "Geometric extension safety cutoff must be finite.")
        if not 0 <= c < 1:
            raise ArgumentError("Geometric extension safety cutoff must lie in [0, 1).")
