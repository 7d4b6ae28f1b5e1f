"""The reference's lid-driven-cavity driver loop on the device.

Follows scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl:1540-1935: the
timestep of every step is the smallest of the comparison bounds
(comparison_timestep_bounds, from the velocity extrema that each step's
statistics carry) and of the time to the next output event; trajectory rows
are written every trajectory_interval and the full diagnostics every
diagnostics_interval, to the seven CSV files of the reference with its
formats.  Snapshot figures
(the plotting part of the driver) are not produced.  Nothing but a few
scalars per step crosses to the host.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass
from typing import NamedTuple

from . import diagnostics as D
from .imex import allocate_imex_workspace, build_imex_context, integrate_step, make_velocity_compatible

EVENT_TOL = 5e-12  # :1721


class PhaseFieldSetup(NamedTuple):
    """The interface scales comparison_timestep_bounds reads (phase_field_setup
    :334-383): the CDI epsilon (epsilon_over_h * h) and the ACDI epsilon
    (acdi_epsilon_over_h * h), h = conservative_mesh_spacing."""
    cdi_epsilon: float
    acdi_epsilon: float


def phase_field_setup(h_interface: float, epsilon_over_h: float = 2.0, acdi_epsilon_over_h: float = 0.51):
    return PhaseFieldSetup(epsilon_over_h * h_interface, acdi_epsilon_over_h * h_interface)


def conservative_mesh_spacing(metrics) -> float:  # BenchmarkUtilities.jl:19-21
    return float(max(max(metrics.dx_c), max(metrics.dy_c)))


@dataclass
class RunSummary:
    steps: int
    time: float
    wall_seconds: float
    files: dict


FILES = ("centroid_trajectory.csv", "diagnostics.csv", "phase_diagnostics.csv", "deformation_modes.csv",
         "tensor_diagnostics.csv", "flow_diagnostics.csv", "numerics_diagnostics.csv")


def run_cavity(case, cfg: D.BenchmarkConfig, setup: PhaseFieldSetup, output_dir: str, radius: float,
               progress=None) -> RunSummary:
    """Advance `case` (a setups.Case) to cfg.T with the driver's adaptive dt,
    writing the driver's CSV files under output_dir.  `radius` is the disk
    radius of the geometric area reference (pi r^2).  `progress(step, t, dt)`
    is called after every diagnostics event when given."""
    os.makedirs(output_dir, exist_ok=True)
    state = case.state
    ctx = build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = allocate_imex_workspace(ctx)
    init_projection = make_velocity_compatible(state, ctx, work)
    h_interface = conservative_mesh_spacing(case.metrics)
    geometric_area = math.pi * radius ** 2
    paths = {name: os.path.join(output_dir, name) for name in FILES}
    ios = {name: open(p, "w") for name, p in paths.items()}
    traj, diag, phase, modes, tensor, flow, numerics = (ios[n] for n in FILES)
    nmax = cfg.deformation_mode_max
    t0_wall = time.perf_counter()
    try:
        D.write_trajectory_header(traj)
        D.write_diagnostics_header(diag)
        D.write_phase_diagnostics_header(phase)
        D.write_deformation_modes_header(modes, nmax)
        D.write_tensor_diagnostics_header(tensor)
        D.write_flow_diagnostics_header(flow)
        D.write_numerics_diagnostics_header(numerics)

        def diagnostics_event(step, t, dt, timestep, constraint, mass0):
            d = D.compute_diagnostics(state, ctx, cfg, h_interface)
            g = d.tensor.global_region
            D.write_diagnostics_row(diag, step, t, dt, d.geometry.x_centroid, d.geometry.y_centroid,
                                    d.geometry.phase_integral, mass0, d.bounds.minimum, d.bounds.maximum,
                                    d.flow.maximum_speed, d.flow.divergence_linf, g.lambda_min,
                                    g.max_abs_determinant_error, g.non_spd_cells, constraint)
            D.write_phase_diagnostics_row(phase, t, d.geometry, d.bounds, mass0, geometric_area)
            D.write_deformation_modes_row(modes, t, d.geometry, nmax)
            D.write_tensor_diagnostics_row(tensor, t, d.tensor)
            D.write_flow_diagnostics_row(flow, t, d.flow)
            D.write_numerics_diagnostics_row(numerics, step, t, dt, timestep, constraint, cfg)

        # t = 0 (:1621-1657; the sigma columns of the first row come from
        # structure_tensor_diagnostics, whose definitions are the global region's)
        xc0, yc0, mass0 = D.phase_centroid(state, ctx)
        D.write_trajectory_row(traj, 0.0, xc0, yc0, mass0, mass0)
        initial_bounds = D.comparison_timestep_bounds(cfg, state, ctx, setup)
        diagnostics_event(0, 0.0, 0.0, initial_bounds, init_projection.constraint_linf, mass0)

        t, step = 0.0, 0
        next_traj, next_diag = cfg.trajectory_interval, cfg.diagnostics_interval
        last_traj = last_diag = 0.0
        last_report = None
        dt = 0.0
        timestep = initial_bounds
        extrema = None  # velocity extrema of the current state, from the step that made it
        while t < cfg.T - EVENT_TOL:  # :1723-1830
            timestep = D.comparison_timestep_bounds(cfg, state, ctx, setup, extrema)
            dt_physical = timestep.selected
            if not dt_physical >= cfg.dt_min_physical:
                raise RuntimeError(f"Physical timestep bound {dt_physical:.6e} is below dt_min_physical "
                                   f"{cfg.dt_min_physical:.6e} at t={t:.6e}.")
            next_event = min(cfg.T, next_traj, next_diag)
            dt = min(dt_physical, next_event - t)
            if dt <= EVENT_TOL:
                t = next_event
            else:
                last_report = integrate_step(state, t, dt, ctx, work)
                extrema = (work.last_stats.max_abs_u, work.last_stats.max_abs_v)
                t += dt
                step += 1
                if abs(t - next_event) <= EVENT_TOL:
                    t = next_event
            if t >= next_traj - EVENT_TOL:
                xc, yc, mass = D.phase_centroid(state, ctx)
                D.write_trajectory_row(traj, t, xc, yc, mass, mass0)
                last_traj = t
                next_traj += cfg.trajectory_interval
            if t >= next_diag - EVENT_TOL:
                constraint = (init_projection.constraint_linf if last_report is None
                              else last_report.maximum_projection_constraint_linf)
                diagnostics_event(step, t, dt, timestep, constraint, mass0)
                last_diag = t
                next_diag += cfg.diagnostics_interval
                if progress:
                    progress(step, t, dt)
        # final rows at T (:1870-1930)
        if abs(t - cfg.T) <= 1e-10:
            t = cfg.T
        if abs(last_traj - t) > EVENT_TOL:
            xc, yc, mass = D.phase_centroid(state, ctx)
            D.write_trajectory_row(traj, t, xc, yc, mass, mass0)
        if abs(last_diag - t) > EVENT_TOL:
            constraint = (init_projection.constraint_linf if last_report is None
                          else last_report.maximum_projection_constraint_linf)
            diagnostics_event(step, t, 0.0, D.comparison_timestep_bounds(cfg, state, ctx, setup, extrema), constraint,
                              mass0)
    finally:
        for f in ios.values():
            f.close()
    return RunSummary(step, t, time.perf_counter() - t0_wall, paths)
