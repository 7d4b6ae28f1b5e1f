"""Driver diagnostics of the reference's lid-driven-cavity benchmark, computed
on the device-resident state.

Mirrors scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl: the
"DIAGNOSTICS" functions (:483-849), the timestep selection
comparison_timestep_bounds (:852-934) and the CSV writers (:938-1135), with
tensor_cell_diagnostics_2d from scripts/benchmarks/BenchmarkUtilities.jl:62-88.
The reductions run in libfsi_hip.so (fsi_velocity_extrema,
fsi_phase_centroid, fsi_diagnostics_compute), so a driver loop reads a few
scalars per step instead of the fields; the CSV rows are formatted exactly as
the reference's @printf / @sprintf calls format them, so files written here
feed the reference's post-processing unchanged.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import NamedTuple

from . import _abi
from .imex import FSIState, IMEXContext

EPS = 2.220446049250313e-16  # eps(Float64)


@dataclass
class BenchmarkConfig:
    """The BenchmarkConfig fields the diagnostics and the timestep selection
    read, with the reference's defaults (:40-128)."""
    U_lid: float = 1.0
    density: float = 1.0
    elastic_modulus: float = 0.10
    Pe_gamma: float = 0.25
    acdi_gamma: float = 1.0
    uwcdi_alpha: float = 1.0
    sigma_safety_cutoff: float = 1e-6
    T: float = 20.0
    cfl_hyperbolic: float = 0.10
    cfl_interface_compression: float = 0.25
    cfl_interface_diffusion: float = 0.50
    dt_max: float = 5e-3
    dt_min_physical: float = 1e-8
    trajectory_interval: float = 0.02
    diagnostics_interval: float = 0.10
    deformation_mode_max: int = 12
    sigma_core_cutoff: float = 0.90
    sigma_interface_cutoff: float = 0.10


# ---- device reductions -------------------------------------------------------
def velocity_extrema(state: FSIState, ctx: IMEXContext) -> tuple[float, float]:
    """(max |U| over every x-face, max |V| over every y-face) of the state."""
    mu, mv = ctypes.c_double(), ctypes.c_double()
    st = ctx.lib.fsi_velocity_extrema(ctx.handle, state._device(ctx), ctypes.byref(mu), ctypes.byref(mv))
    _abi.check(ctx.lib, ctx.handle, st, "velocity_extrema")
    return mu.value, mv.value


def phase_centroid(state: FSIState, ctx: IMEXContext) -> tuple[float, float, float]:
    """phase_centroid (:496-509): (x_centroid, y_centroid, phase_integral)."""
    xc, yc, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    st = ctx.lib.fsi_phase_centroid(ctx.handle, state._device(ctx), ctypes.byref(xc), ctypes.byref(yc),
                                    ctypes.byref(m))
    _abi.check(ctx.lib, ctx.handle, st, "phase_centroid")
    return xc.value, yc.value, m.value


class TensorRegion(NamedTuple):
    """tensor_region_diagnostics (:753-822), TENSOR_REGION_FIELDS order."""
    cells: int
    area: float
    determinant_min: float
    determinant_max: float
    max_abs_determinant_error: float
    mean_abs_determinant_error: float
    lambda_min: float
    lambda_max: float
    non_spd_cells: int
    max_effective_stress: float
    mean_effective_stress: float


class PhaseGeometry(NamedTuple):
    """phase_geometry_diagnostics (:632-707)."""
    x_centroid: float
    y_centroid: float
    phase_integral: float
    perimeter_proxy: float
    mixed_layer_measure: float
    effective_interface_width: float
    effective_interface_cells: float
    covariance_major: float
    covariance_minor: float
    aspect_ratio: float
    orientation_radians: float
    R0: float
    Rc: tuple
    Rs: tuple
    amplitudes: tuple


class PhaseBounds(NamedTuple):
    """phase_bound_diagnostics (:709-722)."""
    minimum: float
    maximum: float
    lower_violation: float
    upper_violation: float
    out_of_bounds_cells: int


class FlowDiagnostics(NamedTuple):
    """flow_diagnostics (:724-751)."""
    maximum_speed: float
    kinetic_energy: float
    divergence_linf: float
    divergence_l2: float


class TensorDiagnostics(NamedTuple):
    """tensor_diagnostics_detailed (:824-849)."""
    global_region: TensorRegion
    core: TensorRegion
    interface: TensorRegion
    tail: TensorRegion


@dataclass
class Diagnostics:
    """Everything the driver writes at a diagnostics event, from one call."""
    geometry: PhaseGeometry
    bounds: PhaseBounds
    flow: FlowDiagnostics
    tensor: TensorDiagnostics
    max_abs_u: float
    max_abs_v: float


def _region(r) -> TensorRegion:
    return TensorRegion(int(r.cells), r.area, r.determinant_min, r.determinant_max, r.max_abs_determinant_error,
                        r.mean_abs_determinant_error, r.lambda_min, r.lambda_max, int(r.non_spd_cells),
                        r.max_effective_stress, r.mean_effective_stress)


def compute_diagnostics(state: FSIState, ctx: IMEXContext, cfg: BenchmarkConfig, h_interface: float) -> Diagnostics:
    """The interval diagnostics of the driver (:1777-1782) in one device call;
    h_interface is the grid spacing that normalises the interface width."""
    prm = _abi.DiagParamsC(cfg.density, cfg.elastic_modulus, cfg.sigma_core_cutoff, cfg.sigma_interface_cutoff,
                           cfg.sigma_safety_cutoff, h_interface, cfg.deformation_mode_max)
    out = _abi.DiagnosticsC()
    st = ctx.lib.fsi_diagnostics_compute(ctx.handle, state._device(ctx), ctypes.byref(prm), ctypes.byref(out))
    _abi.check(ctx.lib, ctx.handle, st, "diagnostics")
    n = out.nmax
    geometry = PhaseGeometry(out.x_centroid, out.y_centroid, out.phase_integral, out.perimeter_proxy,
                             out.mixed_layer_measure, out.effective_interface_width, out.effective_interface_cells,
                             out.covariance_major, out.covariance_minor, out.aspect_ratio, out.orientation_radians,
                             out.R0, tuple(out.Rc[:n]), tuple(out.Rs[:n]), tuple(out.amplitudes[:n]))
    bounds = PhaseBounds(out.phi_min, out.phi_max, out.lower_violation, out.upper_violation,
                         int(out.out_of_bounds_cells))
    flow = FlowDiagnostics(out.maximum_speed, out.kinetic_energy, out.divergence_linf, out.divergence_l2)
    tensor = TensorDiagnostics(_region(out.global_region), _region(out.core), _region(out.interface_region),
                               _region(out.tail))
    return Diagnostics(geometry, bounds, flow, tensor, out.max_abs_u, out.max_abs_v)


# ---- timestep selection --------------------------------------------------------
class TimestepBounds(NamedTuple):
    hyperbolic: float
    cdi_compression: float
    cdi_diffusion: float
    acdi_compression: float
    acdi_diffusion: float
    uwcdi: float
    selected: float
    limiting_constraint: str
    max_advection_speed: float
    elastic_wave_speed: float
    gamma_acdi: float


_LIMIT_LABELS = ("hyperbolic", "current_cdi_compression", "current_cdi_diffusion", "acdi_compression",
                 "acdi_diffusion", "uwcdi", "dt_max")


def timestep_bounds_from_speed(cfg: BenchmarkConfig, hx: float, hy: float, max_abs_u: float, max_abs_v: float,
                               cdi_epsilon: float, acdi_epsilon: float) -> TimestepBounds:
    """The arithmetic of comparison_timestep_bounds (:852-934) given the
    velocity extrema (uniform grid: hx, hy are the cell widths)."""
    hmin = min(hx, hy)
    max_advection_speed = _jmax(cfg.U_lid, max_abs_u, max_abs_v)
    elastic_wave_speed = math.sqrt(cfg.elastic_modulus / cfg.density)
    hyperbolic_speed = _jmax(max_advection_speed, elastic_wave_speed, EPS)
    dt_hyperbolic = cfg.cfl_hyperbolic * hmin / hyperbolic_speed
    inverse_pe_gamma = 1.0 / cfg.Pe_gamma
    dt_cdi_compression = cfg.cfl_interface_compression * hmin / inverse_pe_gamma
    nu_cdi = 0.5 * cdi_epsilon * inverse_pe_gamma
    dt_cdi_diffusion_limit = _fdiv(1.0, 2.0 * nu_cdi * (1.0 / hx ** 2 + 1.0 / hy ** 2))
    dt_cdi_diffusion = cfg.cfl_interface_diffusion * dt_cdi_diffusion_limit
    gamma_acdi = cfg.acdi_gamma
    dt_acdi_compression = cfg.cfl_interface_compression * hmin / _jmax(gamma_acdi, EPS)
    nu_acdi = gamma_acdi * acdi_epsilon
    dt_acdi_diffusion = (cfg.cfl_interface_diffusion / (2.0 * nu_acdi * (1.0 / hx ** 2 + 1.0 / hy ** 2))
                         if nu_acdi > 0.0 else math.inf)
    uwcdi_cfl_limit = 1.0 / (1.0 + 1.0 / cfg.uwcdi_alpha)
    dt_uwcdi = uwcdi_cfl_limit * hmin / _jmax(max_advection_speed, EPS)
    values = (dt_hyperbolic, dt_cdi_compression, dt_cdi_diffusion, dt_acdi_compression, dt_acdi_diffusion,
              dt_uwcdi, cfg.dt_max)
    index = _findmin(values)
    return TimestepBounds(dt_hyperbolic, dt_cdi_compression, dt_cdi_diffusion, dt_acdi_compression,
                          dt_acdi_diffusion, dt_uwcdi, values[index], _LIMIT_LABELS[index], max_advection_speed,
                          elastic_wave_speed, gamma_acdi)


def comparison_timestep_bounds(cfg: BenchmarkConfig, state: FSIState, ctx: IMEXContext, setup,
                               extrema: tuple[float, float] | None = None) -> TimestepBounds:
    """comparison_timestep_bounds(cfg, state, metrics, setup) (:852-934);
    `setup` carries cdi_epsilon and acdi_epsilon as the driver's setup does.
    `extrema` = (max |U|, max |V|) of the state when the caller has them (the
    step statistics of the step that produced the state carry them); else
    they are reduced on the device."""
    mu, mv = velocity_extrema(state, ctx) if extrema is None else extrema
    m = ctx.metrics
    return timestep_bounds_from_speed(cfg, float(min(m.dx_c)), float(min(m.dy_c)), mu, mv, setup.cdi_epsilon,
                                      setup.acdi_epsilon)


def _fdiv(a, b):
    """a / b with IEEE semantics for b == 0 (Julia returns +-Inf or NaN)."""
    if b == 0.0:
        return math.nan if a == 0.0 or a != a else math.copysign(math.inf, a) * math.copysign(1.0, b)
    return a / b


def _jmax(*xs):
    """Julia's max: NaN wins."""
    if any(x != x for x in xs):
        return math.nan
    return max(xs)


def _findmin(values):
    """Julia's findmin over a tuple: the first index of the minimum, NaN
    counting as the smallest."""
    best = 0
    for k in range(1, len(values)):
        a, b = values[k], values[best]
        if b != b:
            break
        if a != a or a < b:
            best = k
    return best


# ---- CSV output (:938-1135) ------------------------------------------------------
def fmt_e(x: float, digits: int) -> str:
    """%.{digits}e as Julia's @printf writes it (NaN, Inf, -Inf spelled out)."""
    if x != x:
        return "NaN"
    if x == math.inf:
        return "Inf"
    if x == -math.inf:
        return "-Inf"
    return f"{x:.{digits}e}"


def _e10(x):
    return fmt_e(x, 10)


def _e16(x):
    return fmt_e(x, 16)


def write_trajectory_header(io):
    io.write("time,x_centroid,y_centroid,phase_integral,phase_integral_ratio\n")


def write_trajectory_row(io, t, xc, yc, mass, mass0):
    io.write(",".join(_e10(v) for v in (t, xc, yc, mass, mass / mass0)) + "\n")
    io.flush()


def write_diagnostics_header(io):
    io.write("step,time,dt,x_centroid,y_centroid,phase_integral,relative_phase_integral_error,"
             "phi_min,phi_max,max_speed,divergence_linf,min_sigma_eigenvalue,"
             "max_abs_sigma_det_error,non_spd_cells,projection_constraint_linf\n")


def write_diagnostics_row(io, step, t, dt, xc, yc, mass, mass0, phi_min, phi_max, max_speed, divergence_linf,
                          min_sigma_eigenvalue, max_abs_sigma_det_error, non_spd_cells, projection_constraint_linf):
    rel = (mass - mass0) / mass0
    io.write(",".join([str(int(step))] + [_e10(v) for v in (
        t, dt, xc, yc, mass, rel, phi_min, phi_max, max_speed, divergence_linf, min_sigma_eigenvalue,
        max_abs_sigma_det_error)] + [str(int(non_spd_cells)), _e10(projection_constraint_linf)]) + "\n")
    io.flush()


def write_phase_diagnostics_header(io):
    io.write("time,phase_integral,relative_phase_integral_error,geometric_area_reference,"
             "phase_area_bias,relative_phase_area_bias,phi_min,phi_max,"
             "lower_bound_violation,upper_bound_violation,out_of_bounds_cells,"
             "perimeter_proxy,mixed_layer_measure,effective_interface_width,"
             "effective_interface_cells,aspect_ratio,orientation_radians,"
             "covariance_major,covariance_minor\n")


def write_phase_diagnostics_row(io, t, geometry: PhaseGeometry, bounds: PhaseBounds, mass0, geometric_area):
    area_bias = geometry.phase_integral - geometric_area
    vals = [_e16(v) for v in (t, geometry.phase_integral, (geometry.phase_integral - mass0) / mass0, geometric_area,
                              area_bias, area_bias / geometric_area, bounds.minimum, bounds.maximum,
                              bounds.lower_violation, bounds.upper_violation)]
    vals.append(str(int(bounds.out_of_bounds_cells)))
    vals += [_e16(v) for v in (geometry.perimeter_proxy, geometry.mixed_layer_measure,
                               geometry.effective_interface_width, geometry.effective_interface_cells,
                               geometry.aspect_ratio, geometry.orientation_radians, geometry.covariance_major,
                               geometry.covariance_minor)]
    io.write(",".join(vals) + "\n")
    io.flush()


def write_deformation_modes_header(io, nmax: int):
    names = ["time", "R0"]
    for n in range(1, nmax + 1):
        names += [f"Rc{n}", f"Rs{n}", f"R{n}"]
    names += ["high_mode_amplitude_n6_to_nmax", "high_mode_ratio_n6_to_nmax"]
    io.write(",".join(names) + "\n")


def write_deformation_modes_row(io, t, geometry: PhaseGeometry, nmax: int):
    values = [_e16(t), _e16(geometry.R0)]
    for n in range(nmax):
        values += [_e16(geometry.Rc[n]), _e16(geometry.Rs[n]), _e16(geometry.amplitudes[n])]
    high = math.sqrt(sum(a * a for a in geometry.amplitudes[5:nmax])) if nmax >= 6 else 0.0
    ratio = high / abs(geometry.R0) if abs(geometry.R0) > EPS else math.nan
    values += [_e16(high), _e16(ratio)]
    io.write(",".join(values) + "\n")
    io.flush()


TENSOR_REGION_FIELDS = TensorRegion._fields


def write_tensor_diagnostics_header(io):
    names = ["time"]
    for region in ("global", "core", "interface", "tail"):
        names += [f"{region}_{f}" for f in TENSOR_REGION_FIELDS]
    io.write(",".join(names) + "\n")


def tensor_region_csv_values(region: TensorRegion):
    return ([str(int(region.cells))] + [_e16(v) for v in region[1:8]] + [str(int(region.non_spd_cells))]
            + [_e16(region.max_effective_stress), _e16(region.mean_effective_stress)])


def write_tensor_diagnostics_row(io, t, tensor: TensorDiagnostics):
    values = [_e16(t)]
    for region in (tensor.global_region, tensor.core, tensor.interface, tensor.tail):
        values += tensor_region_csv_values(region)
    io.write(",".join(values) + "\n")
    io.flush()


def write_flow_diagnostics_header(io):
    io.write("time,max_speed,kinetic_energy,divergence_linf,divergence_l2\n")


def write_flow_diagnostics_row(io, t, flow: FlowDiagnostics):
    io.write(",".join(_e16(v) for v in (t, flow.maximum_speed, flow.kinetic_energy, flow.divergence_linf,
                                        flow.divergence_l2)) + "\n")
    io.flush()


def write_numerics_diagnostics_header(io):
    io.write("step,time,dt_actual,dt_stability,limiting_constraint,"
             "dt_hyperbolic,dt_cdi_compression,dt_cdi_diffusion,"
             "dt_acdi_compression,dt_acdi_diffusion,dt_uwcdi,"
             "max_advection_speed,elastic_wave_speed,gamma_acdi_scale,gamma_acdi_effective,"
             "projection_constraint_linf\n")


def write_numerics_diagnostics_row(io, step, t, dt, timestep: TimestepBounds, projection_constraint,
                                   cfg: BenchmarkConfig):
    vals = [str(int(step)), _e16(t), _e16(dt), _e16(timestep.selected), timestep.limiting_constraint]
    vals += [_e16(v) for v in (timestep.hyperbolic, timestep.cdi_compression, timestep.cdi_diffusion,
                               timestep.acdi_compression, timestep.acdi_diffusion, timestep.uwcdi,
                               timestep.max_advection_speed, timestep.elastic_wave_speed, cfg.acdi_gamma,
                               timestep.gamma_acdi, projection_constraint)]
    io.write(",".join(vals) + "\n")
    io.flush()
