"""Benchmark configurations of the step, restated from the reference drivers.

  lid-driven cavity with a deformable disk:
      scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl
      (build_boundary_conditions :309-333, build_models :387-422,
       build_numerics :423-448, initial_state :459-481, phase setups :335-385)
  wall-supported filament in Kolmogorov flow:
      scripts/benchmarks/run_wall_supported_filament_kolmogorov_flow.jl
      (BCs :371-410, models :412-448, forcing :13-46, dt policy :875-934)
  initial phase fields: src/InterfaceInitialization.jl (:46-95, :139-154)

Seeded draws use SplitMix64 so that any host (C, Python, Julia) can
regenerate the same inputs from the recorded seed.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import closures as C
from .boundary import BoundaryConditionSetup, Dirichlet, Neumann, Periodic
from .imex import FSIBoundaryConditions, FSIModels, FSIState, IMEXNumerics, LinearSolverOptions
from .mesh import MACGrid, build_mesh_metrics, generate_grid


class SplitMix64:
    """Counter-based 64-bit generator (Steele, Lea & Flood 2014)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / 9007199254740992.0))


# ---- initial conditions --------------------------------------------------------
def _profile_from_distance(d, eps):
    """InterfaceInitialization.jl:46-49"""
    return np.clip(0.5 * (1.0 - np.tanh(d / eps)), 0.0, 1.0)


def initialize_circle(grid: MACGrid, center, radius, epsilon):
    """initialize_circle (InterfaceInitialization.jl:74-95), non-periodic."""
    X, Y = np.meshgrid(grid.x_p, grid.y_p, indexing="ij")
    d = np.hypot(X - center[0], Y - center[1]) - radius
    return np.asfortranarray(_profile_from_distance(d, epsilon))


def circle_parameter_derivatives(grid: MACGrid, center, radius, epsilon):
    """Derivatives of initialize_circle's phase field (InterfaceInitialization.jl:46-49,
    74-95) with respect to the radius and the two centre coordinates: with
    d = |x - c| - r and phi = (1 - tanh(d/eps))/2, dphi/dd = -(1 - tanh^2)/(2 eps),
    dphi/dr = -dphi/dd and dphi/dc = -dphi/dd (x - c)/|x - c| (zero at the
    centre, and wherever the clamp to [0, 1] is active)."""
    X, Y = np.meshgrid(grid.x_p, grid.y_p, indexing="ij")
    rho = np.hypot(X - center[0], Y - center[1])
    t = np.tanh((rho - radius) / epsilon)
    raw = 0.5 * (1.0 - t)
    dphi_dd = np.where((raw > 0.0) & (raw < 1.0), -0.5 * (1.0 - t * t) / epsilon, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        ex = np.where(rho > 0.0, (X - center[0]) / rho, 0.0)
        ey = np.where(rho > 0.0, (Y - center[1]) / rho, 0.0)
    return tuple(np.asfortranarray(a) for a in (-dphi_dd, -dphi_dd * ex, -dphi_dd * ey))


def sharp_circle_indicator(grid: MACGrid, center, radius):
    """run_sugiyama_lid_driven_cavity.jl:450-457"""
    X, Y = np.meshgrid(grid.x_p, grid.y_p, indexing="ij")
    return np.asfortranarray((np.hypot(X - center[0], Y - center[1]) <= radius).astype(np.float64))


def sharp_wall_filament(grid: MACGrid, center_x, half_width, height):
    """Sharp rectangle [cx-w, cx+w] x [0, height] on the bottom wall, periodic in x
    (indicator form of _signed_distance_to_box, InterfaceInitialization.jl:139-154)."""
    X, Y = np.meshgrid(grid.x_p, grid.y_p, indexing="ij")
    dx = np.mod(X - center_x + 0.5 * grid.Lx, grid.Lx) - 0.5 * grid.Lx
    inside = (np.abs(dx) <= half_width) & (Y <= height)
    return np.asfortranarray(inside.astype(np.float64))


# ---- configurations ----------------------------------------------------------------
@dataclass
class Case:
    name: str
    grid: MACGrid
    metrics: object
    bcs: FSIBoundaryConditions
    models: FSIModels
    numerics: IMEXNumerics
    state: FSIState
    dt: float
    params: dict


def _noflux(px=False):
    if px:
        return BoundaryConditionSetup(Periodic(), Periodic(), Neumann(0.0), Neumann(0.0))
    return BoundaryConditionSetup(Neumann(0.0), Neumann(0.0), Neumann(0.0), Neumann(0.0))


def cavity_numerics(velocity_rtol=1e-8, projection_rtol=1e-10, precision="fp64", velocity_method="gmres"):
    """build_numerics (run_sugiyama_lid_driven_cavity.jl:423-448).  The
    driver's velocity solve is GMRES to reltol 1e-8; velocity_method="direct"
    is the reference's default (the exact solve), used for tight parity."""
    return IMEXNumerics(
        linear_solvers=LinearSolverOptions(velocity_method=velocity_method, reltol=velocity_rtol, maxiter=800,
                                           restart=100),
        structure_extension=C.GeometricStructureExtension(20.0, 4.0, safety_cutoff=1e-6),
        projection_reltol=projection_rtol,
        projection_maxiter=3000,
        projection_restart=50,
        precision=precision,
    )


def phase_method(method: str, h: float, acdi_eps_over_h: float = 1.5):
    """Phase models and initial interface widths of the cavity driver
    (run_sugiyama_lid_driven_cavity.jl:335-385, :397-410):
      CDI   = CurrentCDIPhaseField(1/Pe=4, eps=2h; delta 1e-6), tanh init eps=2h
      ACDI  = JainACDIPhaseField(1, eps_J; delta 1e-14, fixed gamma), tanh init 2 eps_J
      UWCDI = PirozzoliUWCDIPhaseField(1; delta 1e-14), sharp init
    Returns (model, init_epsilon or None for a sharp indicator)."""
    if method == "cdi":
        return C.CurrentCDIPhaseField(1.0 / 0.25, 2.0 * h, delta_normal=1e-6), 2.0 * h
    if method == "acdi":
        eps = acdi_eps_over_h * h
        return C.JainACDIPhaseField(1.0, eps, delta_normal=1e-14, dynamic_gamma=False), 2.0 * eps
    if method == "uwcdi":
        return C.PirozzoliUWCDIPhaseField(1.0, delta_normal=1e-14), None
    if method == "upwind":
        return C.NoPhaseFieldRegularization(), 2.0 * h
    raise ValueError(method)


def lid_driven_cavity(N=128, method="acdi", radius=0.2, center=(0.6, 0.5), shear_modulus=0.1, viscosity=0.01,
                      dt=None, velocity_rtol=1e-8, projection_rtol=1e-10, acdi_eps_over_h=1.5,
                      precision="fp64", velocity_method="gmres") -> Case:
    """Sugiyama lid-driven cavity with a deformable disk on the unit square."""
    grid = generate_grid(1.0, 1.0, N, N)
    metrics = build_mesh_metrics(grid)
    wall, lid = Dirichlet(0.0), Dirichlet(1.0)
    noflux = _noflux()
    bcs = FSIBoundaryConditions(
        BoundaryConditionSetup(wall, wall, wall, lid), BoundaryConditionSetup(wall, wall, wall, wall),
        noflux, noflux, {"xx": noflux, "yy": noflux, "xy": noflux})
    h = 1.0 / N
    pf, eps0 = phase_method(method, h, acdi_eps_over_h)
    models = FSIModels(viscosity=C.ConstantViscosity(viscosity), tensor_relaxation=C.NoRelaxation(),
                       elastic_stress=C.LinearInternalTensorStress(shear_modulus), capillarity=C.NoCapillarity(),
                       phase_field_method=pf, body_force=C.NoBodyForce())
    phi = sharp_circle_indicator(grid, center, radius) if eps0 is None else initialize_circle(grid, center, radius, eps0)
    state = FSIState(np.zeros((N + 1, N), order="F"), np.zeros((N, N + 1), order="F"), phi,
                     np.ones((N, N), order="F"), np.ones((N, N), order="F"), np.zeros((N, N), order="F"))
    if dt is None:
        dt = 1e-3 * 128.0 / N
    return Case(f"cavity_N{N}_{method}", grid, metrics, bcs, models,
                cavity_numerics(velocity_rtol, projection_rtol, precision, velocity_method), state, dt,
                dict(N=N, method=method, radius=radius, center=tuple(center), shear_modulus=shear_modulus,
                     viscosity=viscosity, init_epsilon=eps0))


def seeded_cavity(N, method, seed, **kw) -> Case:
    """Cavity with radius ~ U[0.15, 0.25] and centre ~ U[0.4, 0.6]^2 drawn from `seed`."""
    rng = SplitMix64(seed)
    r = rng.uniform(0.15, 0.25)
    cx = rng.uniform(0.4, 0.6)
    cy = rng.uniform(0.4, 0.6)
    case = lid_driven_cavity(N, method, radius=r, center=(cx, cy), **kw)
    case.params["seed"] = seed
    return case


def kolmogorov_filament(N=1024, method="uwcdi", shear_modulus=0.5, viscosity=0.01, amplitude=0.5,
                        velocity_rtol=1e-8, projection_rtol=1e-11, precision="fp64",
                        velocity_method="gmres") -> Case:
    """Wall-supported filament in Kolmogorov flow on the unit square: x-periodic,
    no-slip walls at y=0,1, f_x = amplitude*sin(4 pi y), 0.05 x 0.4 filament at
    x=0.5 on the bottom wall.  phi uses Neumann(0) walls (no support layer)."""
    grid = generate_grid(1.0, 1.0, N, N)
    metrics = build_mesh_metrics(grid, periodic_x=True)
    px, wall = Periodic(), Dirichlet(0.0)
    bcs = FSIBoundaryConditions(
        BoundaryConditionSetup(px, px, wall, wall), BoundaryConditionSetup(px, px, wall, wall),
        _noflux(True), _noflux(True), _noflux(True),
        {"xx": BoundaryConditionSetup(px, px, Dirichlet(1.0), Dirichlet(1.0)),
         "yy": BoundaryConditionSetup(px, px, Dirichlet(1.0), Dirichlet(1.0)),
         "xy": BoundaryConditionSetup(px, px, Dirichlet(0.0), Dirichlet(0.0))})
    h = 1.0 / N
    pf, eps0 = phase_method(method, h)
    models = FSIModels(viscosity=C.ConstantViscosity(viscosity), tensor_relaxation=C.NoRelaxation(),
                       elastic_stress=C.LinearInternalTensorStress(shear_modulus), capillarity=C.NoCapillarity(),
                       phase_field_method=pf, body_force=C.KolmogorovBodyForce(amplitude, 4.0 * math.pi))
    phi = sharp_wall_filament(grid, 0.5, 0.025, 0.4)
    if eps0 is not None and method != "uwcdi":
        X, Y = np.meshgrid(grid.x_p, grid.y_p, indexing="ij")
        dx = np.abs(np.mod(X - 0.5 + 0.5, 1.0) - 0.5) - 0.025
        dy = np.abs(Y - 0.2) - 0.2
        d = np.hypot(np.maximum(dx, 0), np.maximum(dy, 0)) + np.minimum(np.maximum(dx, dy), 0)
        phi = np.asfortranarray(_profile_from_distance(d, eps0))
    state = FSIState(np.zeros((N + 1, N), order="F"), np.zeros((N, N + 1), order="F"), phi,
                     np.ones((N, N), order="F"), np.ones((N, N), order="F"), np.zeros((N, N), order="F"))
    # fixed dt from the driver's hyperbolic bound at t=0 (:875-890):
    # 0.1 h / max(U_ref = 1, sqrt(G)) with U_ref the driver default
    speed = max(1.0, math.sqrt(shear_modulus))
    dt = 0.1 * h / speed
    num = IMEXNumerics(
        linear_solvers=LinearSolverOptions(velocity_method=velocity_method, reltol=velocity_rtol, maxiter=800,
                                           restart=100),
        structure_extension=C.GeometricStructureExtension(20.0, 4.0, safety_cutoff=1e-6),
        projection_reltol=projection_rtol, precision=precision)
    return Case(f"filament_N{N}_{method}", grid, metrics, bcs, models, num, state, dt,
                dict(N=N, method=method, shear_modulus=shear_modulus, viscosity=viscosity))


def ensemble_cases(n=64, seed=4, sizes=(128, 192, 256, 384), method="acdi", **kw) -> list:
    """Batched-ensemble workload: n cavities with N drawn uniformly from
    `sizes`, shear modulus log-uniform in [0.05, 0.5], disk radius ~ U[0.15,
    0.25], centre ~ U[0.4, 0.6]^2, and dt = 1e-3 * 128 / N each."""
    rng = SplitMix64(seed)
    out = []
    for k in range(n):
        N = sizes[min(int(rng.uniform() * len(sizes)), len(sizes) - 1)]
        G = math.exp(rng.uniform(math.log(0.05), math.log(0.5)))
        r = rng.uniform(0.15, 0.25)
        cx, cy = rng.uniform(0.4, 0.6), rng.uniform(0.4, 0.6)
        case = lid_driven_cavity(N, method, radius=r, center=(cx, cy), shear_modulus=G, **kw)
        case.params["instance"] = k
        out.append(case)
    return out


def growing_disk(N=64, tensor=None, acdi=True, numerics=None):
    """An elastic disk growing on a transported resource (the growth API of
    README.md:48-94 in a closed box): a walled unit box whose top is opened by
    the projection (correction q = 0 there, zero flux elsewhere, so the volume
    the disk produces leaves through the top), an ACDI disk of radius 0.2 at
    (0.5, 0.4), SaturatingResourceGrowth(0.5, 0.5), a directional growth
    tensor t = (1, 2)/sqrt(5) (or `tensor`), HarmonicPhaseDiffusivity(0.05,
    0.01), SaturatingResourceConsumption(0.2, 0.5), viscosity 0.01, shear
    modulus 0.1; the resource starts at 1 + 0.2 cos(2 pi x) sin(pi y) and is
    held at 1 on the top.  Returns (grid, metrics, bcs, models, numerics, state)."""
    grid = generate_grid(1.0, 1.0, N, N)
    metrics = build_mesh_metrics(grid)
    w, n0 = Dirichlet(0.0), Neumann(0.0)
    wall = BoundaryConditionSetup(w, w, w, w)
    nf = BoundaryConditionSetup(n0, n0, n0, n0)
    q = BoundaryConditionSetup(n0, n0, n0, Dirichlet(0.0))
    bc_c = BoundaryConditionSetup(n0, n0, n0, Dirichlet(1.0))
    bcs = FSIBoundaryConditions(wall, wall, q, nf, bc_c, {"xx": nf, "yy": nf, "xy": nf})
    h = 1.0 / N
    pm, eps = phase_method("acdi" if acdi else "upwind", h)
    models = FSIModels(phase_field_method=pm, viscosity=C.ConstantViscosity(0.01),
                       elastic_stress=C.LinearInternalTensorStress(0.1),
                       growth_source=C.SaturatingResourceGrowth(0.5, 0.5),
                       growth_tensor=tensor or C.DirectionalGrowthTensor(1.0, 2.0),
                       diffusivity=C.HarmonicPhaseDiffusivity(0.05, 0.01),
                       resource_consumption=C.SaturatingResourceConsumption(0.2, 0.5))
    phi = initialize_circle(grid, (0.5, 0.4), 0.2, eps)
    x = (np.arange(N) + 0.5) * h
    c0 = 1.0 + 0.2 * np.outer(np.cos(2 * np.pi * x), np.sin(np.pi * x))
    state = FSIState(np.zeros((N + 1, N)), np.zeros((N, N + 1)), phi, c0, np.ones((N, N)), np.ones((N, N)),
                     np.zeros((N, N)))
    return grid, metrics, bcs, models, numerics or IMEXNumerics(), state
