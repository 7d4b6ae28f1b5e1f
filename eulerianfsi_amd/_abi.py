"""ctypes mirror of include/fsi_api.h and the loader of libfsi_hip.so.

The structures here are the C-ABI configuration blocks; `abi_config` turns
the Python closure objects (the mirror of the reference's FSIModels /
FSIBoundaryConditions / IMEXNumerics) into them.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p
from pathlib import Path

from . import boundary as _bc
from . import closures as _cl
from .errors import raise_for_status

# ---- enums (fsi_api.h) ------------------------------------------------------
FSI_OK = 0
BC_PERIODIC, BC_DIRICHLET, BC_NEUMANN = 0, 1, 2
PHASE_UPWIND, PHASE_CDI, PHASE_ACDI, PHASE_UWCDI = 0, 1, 2, 3
VISC_NONE, VISC_CONSTANT = 0, 1
ELASTIC_NONE, ELASTIC_LINEAR = 0, 1
RELAX_NONE, RELAX_MAXWELL = 0, 1
FORCE_NONE, FORCE_CONSTANT, FORCE_KOLMOGOROV = 0, 1, 2
EXT_NONE, EXT_IDENTITY, EXT_GEOMETRIC = 0, 1, 2
PRECISION_FP64, PRECISION_MIXED = 0, 1
UWCDI_NORMALS = {"standard": 0, "youngs": 1, "smoothed9": 2}
CAPILLARITY_NONE, CAPILLARITY_DIFFUSE_AREA = 0, 1
SOLVER_DIRECT, SOLVER_GMRES = 0, 1
GROWTH_NONE, GROWTH_CONSTANT, GROWTH_SATURATING = 0, 1, 2
GROWTH_TENSOR_ISOTROPIC, GROWTH_TENSOR_DIRECTIONAL = 0, 1
DIFFUSION_NONE, DIFFUSION_CONSTANT, DIFFUSION_HARMONIC = 0, 1, 2
CONSUMPTION_NONE, CONSUMPTION_SATURATING = 0, 1


class BcSide(ctypes.Structure):
    _fields_ = [("kind", c_int32), ("value", c_double)]


class BcSetup(ctypes.Structure):
    _fields_ = [("left", BcSide), ("right", BcSide), ("bottom", BcSide), ("top", BcSide)]


class Bcs(ctypes.Structure):
    _fields_ = [(n, BcSetup) for n in ("U", "V", "correction", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy")]


class Grid(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("lx", c_double), ("ly", c_double),
                ("periodic_x", c_int32), ("periodic_y", c_int32)]


class Models(ctypes.Structure):
    _fields_ = [
        ("phase_method", c_int32),
        ("cdi_inverse_peclet", c_double), ("cdi_epsilon", c_double), ("cdi_delta_normal", c_double),
        ("acdi_gamma", c_double), ("acdi_epsilon", c_double), ("acdi_psi_floor", c_double),
        ("acdi_delta_normal", c_double), ("acdi_dynamic_gamma", c_int32),
        ("uwcdi_alpha", c_double), ("uwcdi_delta_normal", c_double),
        ("viscosity_kind", c_int32), ("viscosity", c_double),
        ("elastic_kind", c_int32), ("shear_modulus", c_double),
        ("relaxation_kind", c_int32), ("relaxation_rate", c_double),
        ("body_force_kind", c_int32),
        ("force_fx", c_double), ("force_fy", c_double),
        ("force_amplitude", c_double), ("force_wavenumber", c_double), ("force_phase", c_double),
        ("uwcdi_normal_method", c_int32),
        ("capillarity_kind", c_int32), ("capillary_coefficient", c_double), ("capillary_delta_normal", c_double),
        ("capillary_smoothing_passes", c_int32),
        ("growth_source_kind", c_int32), ("growth_rate", c_double), ("growth_half_saturation", c_double),
        ("growth_tensor_kind", c_int32), ("growth_direction_x", c_double), ("growth_direction_y", c_double),
        ("diffusivity_kind", c_int32), ("diffusivity_fluid", c_double), ("diffusivity_structure", c_double),
        ("consumption_kind", c_int32), ("consumption_rate", c_double), ("consumption_half_saturation", c_double),
    ]


class Numerics(ctypes.Structure):
    _fields_ = [
        ("extension_kind", c_int32),
        ("ext_max_rate", c_double), ("ext_exponent", c_double), ("ext_cutoff", c_double),
        ("velocity_rtol", c_double), ("velocity_maxiter", c_int32),
        ("projection_rtol", c_double), ("projection_maxiter", c_int32),
        ("stage1_atol", c_double), ("stage1_rtol", c_double), ("check_stage1", c_int32),
        ("precision", c_int32),
        ("velocity_method", c_int32), ("direct_fallback", c_int32), ("fallback_rtol", c_double),
        ("resource_method", c_int32),
    ]


class StepReportC(ctypes.Structure):
    _fields_ = [
        ("maximum_projection_constraint_linf", c_double),
        ("maximum_growth_trace_error", c_double),
        ("minimum_phi", c_double),
        ("maximum_phi", c_double),
        ("minimum_resource", c_double),
        ("minimum_sigma_eigenvalue", c_double),
        ("maximum_abs_sigma_determinant_error", c_double),
        ("non_spd_cells", c_int64),
    ]


class ProjectionResultC(ctypes.Structure):
    _fields_ = [
        ("linear_relative_residual", c_double),
        ("constraint_linf", c_double),
        ("constraint_l2", c_double),
        ("compatibility_defect", c_double),
        ("iterations", c_int32),
    ]


class StepStatsC(ctypes.Structure):
    _fields_ = [("velocity_iterations", c_int32 * 3), ("projection_iterations", c_int32 * 3),
                ("planned", c_int32), ("max_abs_u", c_double), ("max_abs_v", c_double),
                ("velocity_residual", c_double * 3), ("velocity_reference_residual", c_double * 3),
                ("velocity_cycles", c_int32 * 3)]


# ---- configuration conversion ------------------------------------------------
def _side(s) -> BcSide:
    if isinstance(s, _bc.Periodic):
        return BcSide(BC_PERIODIC, 0.0)
    if isinstance(s, _bc.Dirichlet):
        return BcSide(BC_DIRICHLET, _bc.scalar_value(s))
    if isinstance(s, _bc.Neumann):
        return BcSide(BC_NEUMANN, _bc.scalar_value(s))
    raise _bc.UnsupportedBoundary(f"{type(s).__name__} boundaries are outside the device path")


def _setup(b) -> BcSetup:
    return BcSetup(_side(b.left), _side(b.right), _side(b.bottom), _side(b.top))


def abi_config(metrics, bcs, models, numerics):
    """(Grid, Bcs, Models, Numerics) for a MeshMetrics / FSIBoundaryConditions /
    FSIModels / IMEXNumerics quadruple."""
    grid = Grid(metrics.Nx, metrics.Ny, metrics.Lx, metrics.Ly, int(metrics.periodic_x), int(metrics.periodic_y))
    cb = Bcs(
        _setup(bcs.U), _setup(bcs.V), _setup(bcs.correction), _setup(bcs.phi), _setup(bcs.resource),
        _setup(bcs.sigma["xx"]), _setup(bcs.sigma["yy"]), _setup(bcs.sigma["xy"]),
    )
    m = Models()
    pf = models.phase_field_method
    if isinstance(pf, _cl.LegacySplitPhaseFieldMethod):
        flux = pf.interface_flux
        if isinstance(flux, _cl.NoInterfaceFlux):
            m.phase_method = PHASE_UPWIND
        else:
            m.phase_method = PHASE_CDI
            m.cdi_inverse_peclet, m.cdi_epsilon, m.cdi_delta_normal = flux.inverse_peclet, flux.epsilon, flux.delta_normal
    elif isinstance(pf, _cl.NoPhaseFieldRegularization):
        m.phase_method = PHASE_UPWIND
    elif isinstance(pf, _cl.CurrentCDIPhaseField):
        f = pf.flux
        m.phase_method = PHASE_CDI
        m.cdi_inverse_peclet, m.cdi_epsilon, m.cdi_delta_normal = f.inverse_peclet, f.epsilon, f.delta_normal
    elif isinstance(pf, _cl.JainACDIPhaseField):
        m.phase_method = PHASE_ACDI
        m.acdi_gamma, m.acdi_epsilon, m.acdi_psi_floor = pf.gamma, pf.epsilon, pf.psi_floor
        m.acdi_delta_normal, m.acdi_dynamic_gamma = pf.delta_normal, int(pf.dynamic_gamma)
    elif isinstance(pf, _cl.PirozzoliUWCDIPhaseField):
        m.phase_method = PHASE_UWCDI
        m.uwcdi_alpha, m.uwcdi_delta_normal = pf.alpha, pf.delta_normal
        m.uwcdi_normal_method = UWCDI_NORMALS[pf.normal_reconstruction]
    else:
        raise _cl.UnsupportedModel(f"{type(pf).__name__} is outside the device path")

    v = models.viscosity
    if isinstance(v, _cl.ConstantViscosity):
        m.viscosity_kind, m.viscosity = VISC_CONSTANT, v.value
    elif isinstance(v, _cl.NoViscosity):
        m.viscosity_kind = VISC_NONE
    else:
        raise _cl.UnsupportedModel(f"{type(v).__name__} is outside the device path")
    e = models.elastic_stress
    if isinstance(e, _cl.LinearInternalTensorStress):
        m.elastic_kind, m.shear_modulus = ELASTIC_LINEAR, e.coefficient
    elif isinstance(e, _cl.NoElasticStress):
        m.elastic_kind = ELASTIC_NONE
    else:
        raise _cl.UnsupportedModel(f"{type(e).__name__} is outside the device path")
    r = models.tensor_relaxation
    if isinstance(r, _cl.MaxwellRelaxation):
        m.relaxation_kind, m.relaxation_rate = RELAX_MAXWELL, r.rate
    elif isinstance(r, _cl.NoRelaxation):
        m.relaxation_kind = RELAX_NONE
    else:
        raise _cl.UnsupportedModel(f"{type(r).__name__} is outside the device path")
    f = models.body_force
    if isinstance(f, _cl.KolmogorovBodyForce):
        m.body_force_kind = FORCE_KOLMOGOROV
        m.force_amplitude, m.force_wavenumber, m.force_phase = f.amplitude, f.wavenumber, f.phase
    elif isinstance(f, _cl.ConstantBodyForce):
        m.body_force_kind, m.force_fx, m.force_fy = FORCE_CONSTANT, f.fx, f.fy
    elif isinstance(f, _cl.NoBodyForce):
        m.body_force_kind = FORCE_NONE
    else:
        raise _cl.UnsupportedModel(f"{type(f).__name__} is outside the device path")
    cap = models.capillarity
    if isinstance(cap, _cl.DiffuseAreaCapillarity):
        m.capillarity_kind = CAPILLARITY_DIFFUSE_AREA
        m.capillary_coefficient, m.capillary_delta_normal = cap.coefficient, cap.delta_normal
        m.capillary_smoothing_passes = cap.smoothing_passes
    elif not isinstance(cap, _cl.NoCapillarity):
        raise _cl.UnsupportedModel(f"capillarity={type(cap).__name__} is outside the device path")
    gs = models.growth_source
    if isinstance(gs, _cl.SaturatingResourceGrowth):
        m.growth_source_kind, m.growth_rate, m.growth_half_saturation = GROWTH_SATURATING, gs.rate, gs.half_saturation
    elif isinstance(gs, _cl.ConstantIntrinsicGrowth):
        m.growth_source_kind, m.growth_rate = GROWTH_CONSTANT, gs.rate
    elif not isinstance(gs, _cl.NoGrowth):
        raise _cl.UnsupportedModel(f"growth_source={type(gs).__name__} is outside the device path")
    gt = models.growth_tensor
    if isinstance(gt, _cl.DirectionalGrowthTensor):
        m.growth_tensor_kind = GROWTH_TENSOR_DIRECTIONAL
        m.growth_direction_x, m.growth_direction_y = gt.direction_x, gt.direction_y
    elif not isinstance(gt, _cl.IsotropicGrowthTensor):
        raise _cl.UnsupportedModel(f"growth_tensor={type(gt).__name__} is outside the device path")
    d = models.diffusivity
    if isinstance(d, _cl.HarmonicPhaseDiffusivity):
        m.diffusivity_kind, m.diffusivity_fluid, m.diffusivity_structure = (
            DIFFUSION_HARMONIC, d.fluid_value, d.structure_value)
    elif isinstance(d, _cl.ConstantDiffusivity):
        m.diffusivity_kind, m.diffusivity_fluid = DIFFUSION_CONSTANT, d.value
    elif not isinstance(d, _cl.NoDiffusion):
        raise _cl.UnsupportedModel(f"diffusivity={type(d).__name__} is outside the device path")
    rc = models.resource_consumption
    if isinstance(rc, _cl.SaturatingResourceConsumption):
        m.consumption_kind, m.consumption_rate, m.consumption_half_saturation = (
            CONSUMPTION_SATURATING, rc.rate, rc.half_saturation)
    elif not isinstance(rc, _cl.NoResourceConsumption):
        raise _cl.UnsupportedModel(f"resource_consumption={type(rc).__name__} is outside the device path")

    n = Numerics()
    ext = numerics.structure_extension
    if isinstance(ext, _cl.GeometricStructureExtension):
        n.extension_kind, n.ext_max_rate, n.ext_exponent, n.ext_cutoff = (
            EXT_GEOMETRIC, ext.max_rate, ext.exponent, ext.safety_cutoff)
    elif isinstance(ext, _cl.IdentityOutsideStructure):
        n.extension_kind, n.ext_cutoff = EXT_IDENTITY, ext.cutoff
    else:
        n.extension_kind = EXT_NONE
    ls = numerics.linear_solvers
    n.velocity_rtol, n.velocity_maxiter = ls.reltol, ls.maxiter
    n.velocity_method = SOLVER_GMRES if ls.velocity_method == "gmres" else SOLVER_DIRECT
    n.direct_fallback, n.fallback_rtol = int(ls.direct_fallback), ls.fallback_reltol
    n.resource_method = SOLVER_GMRES if ls.resource_method == "gmres" else SOLVER_DIRECT
    n.projection_rtol, n.projection_maxiter = numerics.projection_reltol, numerics.projection_maxiter
    n.stage1_atol, n.stage1_rtol = numerics.stage1_constraint_atol, numerics.stage1_constraint_rtol
    n.check_stage1 = int(numerics.check_stage1_constraint)
    n.precision = PRECISION_MIXED if numerics.precision == "mixed" else PRECISION_FP64
    return grid, cb, m, n


# ---- driver diagnostics (fsi_api.h) -------------------------------------------
DIAG_MAX_MODES = 32


class DiagParamsC(ctypes.Structure):
    _fields_ = [("density", c_double), ("elastic_modulus", c_double), ("sigma_core_cutoff", c_double),
                ("sigma_interface_cutoff", c_double), ("sigma_safety_cutoff", c_double),
                ("h_interface", c_double), ("deformation_mode_max", c_int32)]


class TensorRegionC(ctypes.Structure):
    _fields_ = [("cells", c_int64), ("area", c_double), ("determinant_min", c_double),
                ("determinant_max", c_double), ("max_abs_determinant_error", c_double),
                ("mean_abs_determinant_error", c_double), ("lambda_min", c_double), ("lambda_max", c_double),
                ("non_spd_cells", c_int64), ("max_effective_stress", c_double),
                ("mean_effective_stress", c_double)]


class DiagnosticsC(ctypes.Structure):
    _fields_ = ([(n, c_double) for n in (
        "x_centroid", "y_centroid", "phase_integral", "perimeter_proxy", "mixed_layer_measure",
        "effective_interface_width", "effective_interface_cells", "covariance_major", "covariance_minor",
        "aspect_ratio", "orientation_radians", "R0")]
        + [("nmax", c_int32), ("Rc", c_double * DIAG_MAX_MODES), ("Rs", c_double * DIAG_MAX_MODES),
           ("amplitudes", c_double * DIAG_MAX_MODES)]
        + [(n, c_double) for n in ("phi_min", "phi_max", "lower_violation", "upper_violation")]
        + [("out_of_bounds_cells", c_int64)]
        + [(n, c_double) for n in ("maximum_speed", "kinetic_energy", "divergence_linf", "divergence_l2",
                                   "max_abs_u", "max_abs_v")]
        + [(n, TensorRegionC) for n in ("global_region", "core", "interface_region", "tail")])


# ---- discrete adjoint (fsi_api.h) ----------------------------------------------
class AdjointStatsC(ctypes.Structure):
    _fields_ = [("steps", c_int64), ("advances", c_int64), ("recordings", c_int64), ("adjoint_steps", c_int64),
                ("snapshots_used", c_int32), ("repetitions", c_int32), ("device_bytes", c_double),
                ("replanned", c_int64), ("last_shortfall", c_int32), ("solve_iterations", c_int32 * 3)]


# ---- library -----------------------------------------------------------------
_LIB = None
LIB_PATH = Path(__file__).resolve().parent / "libfsi_hip.so"

_dp = POINTER(c_double)


def _declare(lib):
    P = c_void_p
    sig = {
        "fsi_context_create": (c_int, [POINTER(Grid), POINTER(Bcs), POINTER(Models), POINTER(Numerics), POINTER(P)]),
        "fsi_context_destroy": (None, [P]),
        "fsi_last_error": (c_char_p, [P]),
        "fsi_state_create": (c_int, [P, POINTER(P)]),
        "fsi_state_destroy": (None, [P]),
        "fsi_state_upload": (c_int, [P, P] + [_dp] * 7),
        "fsi_state_download": (c_int, [P, P] + [_dp] * 7),
        "fsi_workspace_create": (c_int, [P, POINTER(P)]),
        "fsi_workspace_destroy": (None, [P]),
        "fsi_workspace_download_correction": (c_int, [P, P, _dp]),
        "fsi_workspace_upload_correction": (c_int, [P, P, _dp]),
        "fsi_make_velocity_compatible": (c_int, [P, P, P, POINTER(ProjectionResultC)]),
        "fsi_integrate_step": (c_int, [P, P, P, c_double, c_double, POINTER(StepReportC), POINTER(StepStatsC)]),
        "fsi_ensemble_step": (c_int, [c_int32, POINTER(P), POINTER(P), POINTER(P), _dp, _dp, POINTER(StepReportC),
                                      POINTER(StepStatsC), POINTER(c_int32)]),
        "fsi_project": (c_int, [P, P, P, POINTER(ProjectionResultC)]),
        "fsi_eval_tendencies": (c_int, [P, P, P, c_double] + [_dp] * 8),
        "fsi_uwcdi_strang_step": (c_int, [P, P, P, c_double, c_int32]),
        "fsi_benchmark_kernel": (c_int, [P, P, P, c_char_p, c_int32, _dp, _dp]),
        "fsi_velocity_extrema": (c_int, [P, P, _dp, _dp]),
        "fsi_phase_centroid": (c_int, [P, P, _dp, _dp, _dp]),
        "fsi_phase_centroid_noalloc": (c_int, [P, P, c_void_p, c_void_p]),
        "fsi_diagnostics_compute": (c_int, [P, P, POINTER(DiagParamsC), POINTER(DiagnosticsC)]),
        "fsi_version": (c_char_p, []),
        "fsi_adjoint_create": (c_int, [P, P, c_int32, POINTER(P)]),
        "fsi_adjoint_destroy": (None, [P]),
        "fsi_adjoint_forward": (c_int, [P, P, c_double, c_double, c_int32]),
        "fsi_adjoint_backward": (c_int, [P, P, _dp]),
        "fsi_adjoint_get_stats": (c_int, [P, POINTER(AdjointStatsC)]),
        "fsi_adjoint_record_step": (c_int, [P, P, c_double, c_double]),
        "fsi_adjoint_step": (c_int, [P, P, _dp]),
        "fsi_adjoint_centroid_objective": (c_int, [P, P, c_double, c_double, P, _dp]),
        "fsi_adjoint_tendencies": (c_int, [P, P, POINTER(_dp), POINTER(_dp), _dp]),
        "fsi_adjoint_primal_tendencies": (c_int, [P, P, POINTER(_dp)]),
        "fsi_adjoint_plan": (c_int, [c_int32, c_int32, POINTER(c_int64), POINTER(c_int32), POINTER(c_int32),
                                     POINTER(c_int32), c_int64, POINTER(c_int64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


EXPORTED = (
    "fsi_context_create", "fsi_context_destroy", "fsi_last_error", "fsi_state_create", "fsi_state_destroy",
    "fsi_state_upload", "fsi_state_download", "fsi_workspace_create", "fsi_workspace_destroy",
    "fsi_workspace_download_correction", "fsi_workspace_upload_correction", "fsi_make_velocity_compatible",
    "fsi_integrate_step", "fsi_ensemble_step", "fsi_project", "fsi_eval_tendencies", "fsi_uwcdi_strang_step", "fsi_benchmark_kernel",
    "fsi_velocity_extrema", "fsi_phase_centroid", "fsi_phase_centroid_noalloc", "fsi_diagnostics_compute",
    "fsi_version", "fsi_adjoint_create", "fsi_adjoint_destroy", "fsi_adjoint_forward", "fsi_adjoint_backward",
    "fsi_adjoint_get_stats", "fsi_adjoint_record_step", "fsi_adjoint_step", "fsi_adjoint_centroid_objective",
    "fsi_adjoint_tendencies", "fsi_adjoint_primal_tendencies", "fsi_adjoint_plan",
)


def load_library(path: str | os.PathLike | None = None):
    """Load libfsi_hip.so.  Raises ImportError when the library has not been
    built: there is no CPU fallback behind this API."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise ImportError(f"{p} is missing: build it with `python -m eulerianfsi_amd.build`")
    lib = _declare(ctypes.CDLL(str(p)))
    if path is None:
        _LIB = lib
    return lib


def check(lib, ctx, status: int, what: str) -> None:
    if status != FSI_OK:
        msg = lib.fsi_last_error(ctx)
        raise_for_status(status, f"{what}: {msg.decode() if msg else ''}")


def as_ptr(arr):
    return arr.ctypes.data_as(_dp) if arr is not None else None
