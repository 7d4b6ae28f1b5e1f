"""The IMEX step API of EulerianFSI.jl, backed by libfsi_hip.so on gfx950.

Mirrors src/TimeIntegrationIMEX.jl:
  FSIState :103-118, FSIBoundaryConditions :127-147, FSIModels :156-208,
  LinearSolverOptions :211-252, IMEXNumerics :262-301, IMEXContext :304-311,
  StepReport :318-327, IMEXWorkspace :333-441, build_imex_context :468-493,
  allocate_imex_workspace :389-441, make_velocity_compatible! :1088-1109,
  integrate_step! :1152-1256; and project! (src/Projection.jl:170-307).

Arrays use the reference's index order: U has shape (Nx+1, Ny), V (Nx, Ny+1),
cell fields (Nx, Ny), Fortran-ordered float64 (element [i, j] is Julia's
[i+1, j+1]).  The state lives in device memory once a step has touched it;
reading an attribute of FSIState copies the device state back lazily.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from . import closures as C
from .boundary import is_periodic_x, is_periodic_y
from .errors import ArgumentError, DimensionMismatch
from .mesh import MeshMetrics

A_EXPLICIT = ((0.0, 0.0, 0.0, 0.0), (8 / 15, 0.0, 0.0, 0.0), (1 / 4, 5 / 12, 0.0, 0.0), (1 / 4, 0.0, 3 / 4, 0.0))
A_IMPLICIT = ((0.0, 0.0, 0.0, 0.0), (4 / 15, 4 / 15, 0.0, 0.0), (4 / 15, 1 / 3, 1 / 15, 0.0),
              (4 / 15, 1 / 3, 7 / 30, 1 / 6))
STAGE_ABSCISSA = (0.0, 8 / 15, 2 / 3, 1.0)

_FIELDS = ("U", "V", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy")


class FourStageIMEXFractionalStep:
    """Marker of the inherited four-stage tableau (TimeIntegrationIMEX.jl:72-90)."""


# ---------------------------------------------------------------------------
class FSIState:
    """Dynamic variables (U, V, phi, resource, sigma_xx, sigma_yy, sigma_xy).

    The 6-argument form (no resource) zero-fills the resource, as the
    reference's compatibility constructor does (:113-118).
    """

    def __init__(self, U, V, phi, *rest):
        if len(rest) == 3:
            resource = np.zeros_like(np.asarray(phi, dtype=np.float64))
            sxx, syy, sxy = rest
        elif len(rest) == 4:
            resource, sxx, syy, sxy = rest
        else:
            raise TypeError("FSIState takes 6 or 7 arrays")
        host = {}
        for name, a in zip(_FIELDS, (U, V, phi, resource, sxx, syy, sxy)):
            host[name] = np.asfortranarray(np.asarray(a, dtype=np.float64))
        object.__setattr__(self, "_host", host)
        object.__setattr__(self, "_dev", None)      # (context, handle)
        object.__setattr__(self, "_host_stale", False)
        object.__setattr__(self, "_dev_stale", True)

    def __getattr__(self, name):
        if name in _FIELDS:
            self._sync_host()
            return self._host[name]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _FIELDS:
            self._sync_host()
            self._host[name] = np.asfortranarray(np.asarray(value, dtype=np.float64))
            object.__setattr__(self, "_dev_stale", True)
        else:
            object.__setattr__(self, name, value)

    def mark_host_modified(self):
        """Call after writing into an array returned by an attribute in place."""
        self._sync_host()
        object.__setattr__(self, "_dev_stale", True)

    # -- device residency -----------------------------------------------------
    def _sync_host(self):
        if self._host_stale and self._dev is not None:
            ctx, handle = self._dev
            h = self._host
            st = ctx.lib.fsi_state_download(ctx.handle, handle, *(_abi.as_ptr(h[n]) for n in _FIELDS))
            _abi.check(ctx.lib, ctx.handle, st, "fsi_state_download")
            object.__setattr__(self, "_host_stale", False)

    def _device(self, ctx: "IMEXContext"):
        _validate_shapes(self._host, ctx.metrics)
        if self._dev is None or self._dev[0] is not ctx:
            self._sync_host()
            self._release()
            handle = ctypes.c_void_p()
            _abi.check(ctx.lib, ctx.handle, ctx.lib.fsi_state_create(ctx.handle, ctypes.byref(handle)),
                       "fsi_state_create")
            object.__setattr__(self, "_dev", (ctx, handle))
            object.__setattr__(self, "_dev_stale", True)
        if self._dev_stale:
            h = self._host
            st = ctx.lib.fsi_state_upload(ctx.handle, self._dev[1], *(_abi.as_ptr(h[n]) for n in _FIELDS))
            _abi.check(ctx.lib, ctx.handle, st, "fsi_state_upload")
            object.__setattr__(self, "_dev_stale", False)
        return self._dev[1]

    def _device_modified(self):
        object.__setattr__(self, "_host_stale", True)

    def _release(self):
        if self._dev is not None:
            ctx, handle = self._dev
            if ctx.handle:
                ctx.lib.fsi_state_destroy(handle)
            object.__setattr__(self, "_dev", None)

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


def _validate_shapes(h, metrics: MeshMetrics):
    """_validate_state (:495-523), shapes only; finiteness is checked on the device."""
    nx, ny = metrics.Nx, metrics.Ny
    exp = {"U": (nx + 1, ny), "V": (nx, ny + 1)}
    for n in _FIELDS:
        shape = exp.get(n, (nx, ny))
        if h[n].shape != shape:
            raise DimensionMismatch(f"state.{n} must have size {shape}; received {h[n].shape}.")


# ---------------------------------------------------------------------------
class FSIBoundaryConditions:
    """Per-field boundary setups; `sigma` maps xx/yy/xy (:127-147).  The
    5-argument form reuses the phi setup for the resource (:144-147)."""

    def __init__(self, U, V, correction, phi, *rest):
        if len(rest) == 1:
            resource, sigma = phi, rest[0]
        elif len(rest) == 2:
            resource, sigma = rest
        else:
            raise TypeError("FSIBoundaryConditions takes 5 or 6 arguments")
        sigma = dict(sigma)
        for k in ("xx", "yy", "xy"):
            if k not in sigma:
                raise ArgumentError(f"sigma boundary data require an `{k}` entry.")
        self.U, self.V, self.correction, self.phi, self.resource, self.sigma = U, V, correction, phi, resource, sigma


@dataclass
class FSIModels:
    """Closure bundle (:156-208).  `interface_flux` is the legacy keyword."""
    growth_source: object = field(default_factory=C.NoGrowth)
    growth_tensor: object = field(default_factory=C.IsotropicGrowthTensor)
    viscosity: object = field(default_factory=C.NoViscosity)
    diffusivity: object = field(default_factory=C.NoDiffusion)
    resource_consumption: object = field(default_factory=C.NoResourceConsumption)
    tensor_relaxation: object = field(default_factory=C.NoRelaxation)
    elastic_stress: object = field(default_factory=C.NoElasticStress)
    capillarity: object = field(default_factory=C.NoCapillarity)
    phase_field_method: object = None
    interface_flux: object = None
    body_force: object = field(default_factory=C.NoBodyForce)

    def __post_init__(self):
        flux = self.interface_flux if self.interface_flux is not None else C.NoInterfaceFlux()
        if self.phase_field_method is None:
            self.phase_field_method = C.LegacySplitPhaseFieldMethod(flux)
        elif not isinstance(flux, C.NoInterfaceFlux):
            raise ArgumentError("Specify either phase_field_method or the legacy interface_flux keyword, not both.")
        if not isinstance(self.phase_field_method, C.AbstractPhaseFieldMethod):
            raise ArgumentError("phase_field_method must be an AbstractPhaseFieldMethod.")


@dataclass
class LinearSolverOptions:
    """:211-252.  The device solves the SPD velocity predictor by CG with
    the stopping rules of `_solve_linear_system!` (:602-641): "direct" to a
    relative residual of 1e-14 (the exact solve), "gmres" to reltol times
    the initial residual from uv_n within maxiter iterations, then accepted
    below fallback_reltol, else the direct target (direct_fallback) or an
    ErrorException.  The resource stage solve follows the same rules with
    `resource_method`.  `restart` has no effect on the device path (no
    restarted Krylov method)."""
    velocity_method: str = "direct"
    resource_method: str = "gmres"
    reltol: float = 1e-8
    maxiter: int = 800
    restart: int = 100
    direct_fallback: bool = True
    fallback_reltol: float = 1e-6
    verbose: bool = False

    def __post_init__(self):
        if self.velocity_method not in ("direct", "gmres"):
            raise ArgumentError("velocity_method must be :direct or :gmres.")
        if self.resource_method not in ("direct", "gmres"):
            raise ArgumentError("resource_method must be :direct or :gmres.")
        if not self.reltol > 0:
            raise ArgumentError("Linear-solver reltol must be positive.")
        if not self.maxiter > 0:
            raise ArgumentError("Linear-solver maxiter must be positive.")
        if not self.restart > 0:
            raise ArgumentError("Linear-solver restart must be positive.")
        if not self.fallback_reltol > 0:
            raise ArgumentError("fallback_reltol must be positive.")


@dataclass
class IMEXNumerics:
    """:262-301, plus `precision` ("fp64" | "mixed") of the device path."""
    scheme: object = field(default_factory=FourStageIMEXFractionalStep)
    linear_solvers: LinearSolverOptions = field(default_factory=LinearSolverOptions)
    structure_extension: object = field(default_factory=C.NoStructureExtension)
    projection_reltol: float = 1e-10
    projection_maxiter: int = 3000
    projection_restart: int = 50
    stage1_constraint_atol: float = 5e-8
    stage1_constraint_rtol: float = 1e-8
    check_stage1_constraint: bool = True
    precision: str = "fp64"

    def __post_init__(self):
        if not self.projection_reltol > 0:
            raise ArgumentError("projection_reltol must be positive.")
        if not self.projection_maxiter > 0:
            raise ArgumentError("projection_maxiter must be positive.")
        if not self.projection_restart > 0:
            raise ArgumentError("projection_restart must be positive.")
        if self.stage1_constraint_atol < 0 or self.stage1_constraint_rtol < 0:
            raise ArgumentError("stage-1 constraint tolerances must be non-negative.")
        if self.precision not in ("fp64", "mixed"):
            raise ArgumentError("precision must be 'fp64' or 'mixed'.")


@dataclass(frozen=True)
class StepReport:
    """:318-327"""
    maximum_projection_constraint_linf: float
    maximum_growth_trace_error: float
    minimum_phi: float
    maximum_phi: float
    minimum_resource: float
    minimum_sigma_eigenvalue: float
    maximum_abs_sigma_determinant_error: float
    non_spd_cells: int


@dataclass(frozen=True)
class ProjectionResult:
    """src/Projection.jl:31-37"""
    correction: np.ndarray
    linear_relative_residual: float
    constraint_linf: float
    constraint_l2: float
    compatibility_defect: float
    iterations: int


@dataclass(frozen=True)
class StepStats:
    velocity_iterations: tuple
    projection_iterations: tuple
    planned: bool = False
    # max |U|, max |V| over the faces of the new state (NaN if a face is NaN)
    max_abs_u: float = float("nan")
    max_abs_v: float = float("nan")
    # per predictor solve: final |b - A x| / max(|b|, eps) and the initial
    # residual from uv_n on the same scale (the :gmres rule's reference)
    velocity_residual: tuple = (0.0, 0.0, 0.0)
    velocity_reference_residual: tuple = (0.0, 0.0, 0.0)
    # mixed precision: FP64 refinement cycles per predictor solve (the
    # iterations above are then FP32 inner iterations); 0 in FP64
    velocity_cycles: tuple = (0, 0, 0)


# ---------------------------------------------------------------------------
class IMEXContext:
    """Preassembled device context (:304-311): grid, closures, multigrid
    hierarchy of the projection, HIP stream."""

    def __init__(self, metrics, bcs, models, numerics):
        self.metrics, self.bcs, self.models, self.numerics = metrics, bcs, models, numerics
        self.lib = _abi.load_library()
        self._cfg = _abi.abi_config(metrics, bcs, models, numerics)
        handle = ctypes.c_void_p()
        g, b, m, n = self._cfg
        st = self.lib.fsi_context_create(ctypes.byref(g), ctypes.byref(b), ctypes.byref(m), ctypes.byref(n),
                                         ctypes.byref(handle))
        self.handle = handle
        if st != _abi.FSI_OK:
            msg = self.lib.fsi_last_error(handle) if handle else b""
            self.lib.fsi_context_destroy(handle)
            self.handle = ctypes.c_void_p()
            from .errors import raise_for_status
            raise_for_status(st, f"build_imex_context: {msg.decode() if msg else ''}")

    def close(self):
        if self.handle:
            self.lib.fsi_context_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IMEXWorkspace:
    """Device workspace (:333-441): stage buffers, histories, solver scratch
    and the projection warm start `correction`."""

    def __init__(self, ctx: IMEXContext):
        self.ctx = ctx
        handle = ctypes.c_void_p()
        _abi.check(ctx.lib, ctx.handle, ctx.lib.fsi_workspace_create(ctx.handle, ctypes.byref(handle)),
                   "allocate_imex_workspace")
        self.handle = handle
        self.last_stats = None

    @property
    def correction(self) -> np.ndarray:
        m = self.ctx.metrics
        out = np.zeros((m.Nx, m.Ny), order="F")
        _abi.check(self.ctx.lib, self.ctx.handle,
                   self.ctx.lib.fsi_workspace_download_correction(self.ctx.handle, self.handle, _abi.as_ptr(out)),
                   "download correction")
        return out

    @correction.setter
    def correction(self, value):
        m = self.ctx.metrics
        a = np.asfortranarray(np.asarray(value, dtype=np.float64))
        if a.shape != (m.Nx, m.Ny):
            raise DimensionMismatch(f"correction must have size {(m.Nx, m.Ny)}; received {a.shape}.")
        _abi.check(self.ctx.lib, self.ctx.handle,
                   self.ctx.lib.fsi_workspace_upload_correction(self.ctx.handle, self.handle, _abi.as_ptr(a)),
                   "upload correction")

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                self.ctx.lib.fsi_workspace_destroy(self.handle)
        except Exception:
            pass


def _validate_periodicity(metrics, bcs):
    """:447-466"""
    pairs = [("U", bcs.U), ("V", bcs.V), ("correction", bcs.correction), ("phi", bcs.phi),
             ("resource", bcs.resource), ("sigma_xx", bcs.sigma["xx"]), ("sigma_yy", bcs.sigma["yy"]),
             ("sigma_xy", bcs.sigma["xy"])]
    for name, bc in pairs:
        if metrics.periodic_x != is_periodic_x(bc):
            raise ArgumentError(f"{name} boundary conditions do not match MeshMetrics periodicity in x.")
        if metrics.periodic_y != is_periodic_y(bc):
            raise ArgumentError(f"{name} boundary conditions do not match MeshMetrics periodicity in y.")


def build_imex_context(metrics, bcs, models, numerics=None, use_projection_amg=True) -> IMEXContext:
    """build_imex_context (:468-493).  The projection preconditioner is always
    the device multigrid hierarchy; `use_projection_amg` is accepted for API
    compatibility."""
    _validate_periodicity(metrics, bcs)
    return IMEXContext(metrics, bcs, models, numerics if numerics is not None else IMEXNumerics())


def allocate_imex_workspace(ctx: IMEXContext) -> IMEXWorkspace:
    """allocate_imex_workspace (:389-441); the device workspace is sized by the context."""
    return IMEXWorkspace(ctx)


def make_velocity_compatible(state: FSIState, ctx: IMEXContext, work: IMEXWorkspace) -> ProjectionResult:
    """make_velocity_compatible! (:1088-1109): project the state's velocity onto
    div v = G (G = 0 under NoGrowth), in place."""
    return _project(state, ctx, work, ctx.lib.fsi_make_velocity_compatible, "make_velocity_compatible!")


def project(state: FSIState, ctx: IMEXContext, work: IMEXWorkspace) -> ProjectionResult:
    """project! (src/Projection.jl:170-307) of the state's velocity, warm-started
    from the workspace correction, in place."""
    return _project(state, ctx, work, ctx.lib.fsi_project, "project!")


def _project(state, ctx, work, fn, what):
    h = state._device(ctx)
    res = _abi.ProjectionResultC()
    st = fn(ctx.handle, h, work.handle, ctypes.byref(res))
    state._device_modified()
    _abi.check(ctx.lib, ctx.handle, st, what)
    return ProjectionResult(work.correction, res.linear_relative_residual, res.constraint_linf,
                            res.constraint_l2, res.compatibility_defect, res.iterations)


def _step_stats(s) -> StepStats:
    return StepStats(tuple(s.velocity_iterations), tuple(s.projection_iterations), bool(s.planned),
                     float(s.max_abs_u), float(s.max_abs_v), tuple(s.velocity_residual),
                     tuple(s.velocity_reference_residual), tuple(s.velocity_cycles))


def integrate_step(state: FSIState, t: float, dt: float, ctx: IMEXContext, work: IMEXWorkspace) -> StepReport:
    """integrate_step! (:1152-1256): advance `state` from t to t+dt in place.
    On error the state is unchanged and the mapped exception is raised."""
    if not isinstance(ctx.numerics.scheme, FourStageIMEXFractionalStep):
        raise ArgumentError(f"Unsupported IMEX scheme {type(ctx.numerics.scheme).__name__}.")
    h = state._device(ctx)
    rep = _abi.StepReportC()
    stats = _abi.StepStatsC()
    st = ctx.lib.fsi_integrate_step(ctx.handle, h, work.handle, float(t), float(dt), ctypes.byref(rep),
                                    ctypes.byref(stats))
    _abi.check(ctx.lib, ctx.handle, st, "integrate_step!")
    state._device_modified()
    work.last_stats = _step_stats(stats)
    return StepReport(rep.maximum_projection_constraint_linf, rep.maximum_growth_trace_error, rep.minimum_phi,
                      rep.maximum_phi, rep.minimum_resource, rep.minimum_sigma_eigenvalue,
                      rep.maximum_abs_sigma_determinant_error, int(rep.non_spd_cells))


def ensemble_step(states, ts, dts, ctxs, works) -> list:
    """integrate_step! on n independent instances (states[k], ts[k], dts[k],
    ctxs[k], works[k]; pairwise distinct contexts) with their device work in
    flight together (fsi_ensemble_step).  Every instance advances exactly as
    integrate_step would advance it; the first failing instance's error is
    raised after all instances have been processed (the others have
    advanced)."""
    n = len(states)
    if not (len(ts) == len(dts) == len(ctxs) == len(works) == n):
        raise DimensionMismatch("ensemble_step needs one t, dt, context and workspace per state.")
    if len({id(c) for c in ctxs}) != n:
        raise ArgumentError("ensemble_step needs pairwise distinct contexts.")
    if n == 0:
        return []
    lib = ctxs[0].lib
    for c in ctxs:
        if not isinstance(c.numerics.scheme, FourStageIMEXFractionalStep):
            raise ArgumentError(f"Unsupported IMEX scheme {type(c.numerics.scheme).__name__}.")
    P = ctypes.c_void_p
    hc = (P * n)(*(c.handle for c in ctxs))
    hs = (P * n)(*(s._device(c) for s, c in zip(states, ctxs)))
    hw = (P * n)(*(w.handle for w in works))
    t_arr = (ctypes.c_double * n)(*map(float, ts))
    dt_arr = (ctypes.c_double * n)(*map(float, dts))
    reps = (_abi.StepReportC * n)()
    stats = (_abi.StepStatsC * n)()
    status = (ctypes.c_int32 * n)()
    lib.fsi_ensemble_step(n, hc, hs, hw, t_arr, dt_arr, reps, stats, status)
    out, first = [], None
    for k in range(n):
        if status[k] == 0:
            states[k]._device_modified()
            st = stats[k]
            works[k].last_stats = _step_stats(st)
            r = reps[k]
            out.append(StepReport(r.maximum_projection_constraint_linf, r.maximum_growth_trace_error,
                                  r.minimum_phi, r.maximum_phi, r.minimum_resource, r.minimum_sigma_eigenvalue,
                                  r.maximum_abs_sigma_determinant_error, int(r.non_spd_cells)))
        else:
            out.append(None)
            if first is None:
                first = k
    if first is not None:
        _abi.check(ctxs[first].lib, ctxs[first].handle, status[first], f"ensemble_step (instance {first})")
    return out


def evaluate_stage_tendencies(state: FSIState, ctx: IMEXContext, work: IMEXWorkspace, t: float = 0.0) -> dict:
    """Stage tendencies of the state as _evaluate_stage_tendencies! forms
    them (:944-1071): compact E_u, E_v, I_u, I_v (Nx, Ny) and E_phi,
    E_sigma_xx/yy/xy (Nx, Ny)."""
    h = state._device(ctx)
    m = ctx.metrics
    out = {k: np.zeros((m.Nx, m.Ny), order="F") for k in
           ("E_u", "E_v", "I_u", "I_v", "E_phi", "E_sigma_xx", "E_sigma_yy", "E_sigma_xy")}
    st = ctx.lib.fsi_eval_tendencies(ctx.handle, h, work.handle, float(t), *(_abi.as_ptr(out[k]) for k in out))
    _abi.check(ctx.lib, ctx.handle, st, "evaluate_stage_tendencies")
    return out


def uwcdi_strang_step(state: FSIState, dt: float, ctx: IMEXContext, work: IMEXWorkspace,
                      first_direction: str = "x") -> np.ndarray:
    """uwcdi_strang_step (PhaseFieldMethods.jl:793-850) of the context's
    PirozzoliUWCDIPhaseField: the Strang-split directional update of
    state.phi by the state's velocity.  state.phi is replaced by the result,
    which is also returned."""
    if first_direction not in ("x", "y"):
        raise ArgumentError(f"first_direction must be :x or :y; received {first_direction}.")
    h = state._device(ctx)
    st = ctx.lib.fsi_uwcdi_strang_step(ctx.handle, h, work.handle, float(dt), 0 if first_direction == "x" else 1)
    state._device_modified()
    _abi.check(ctx.lib, ctx.handle, st, "uwcdi_strang_step")
    return state.phi
