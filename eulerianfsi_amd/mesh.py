"""Uniform MAC grid and metrics, mirroring src/MeshGeneration.jl.

Only the :uniform stretching is on the device path: the kernels use one
spacing per axis.  The metric vectors are still built exactly as
_axis_metrics (:237-335) does, for host-side use (initial conditions,
diagnostics).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .errors import ArgumentError, UnsupportedError


@dataclass
class MACGrid:
    Lx: float
    Ly: float
    Nx: int
    Ny: int
    x_p: np.ndarray
    y_p: np.ndarray
    x_u: np.ndarray
    y_v: np.ndarray

    @property
    def dx_nominal(self):
        return self.Lx / self.Nx

    @property
    def dy_nominal(self):
        return self.Ly / self.Ny


@dataclass
class MeshMetrics:
    Nx: int
    Ny: int
    Lx: float
    Ly: float
    periodic_x: bool
    periodic_y: bool
    dx_c: np.ndarray
    dy_c: np.ndarray
    dx_f: np.ndarray
    dy_f: np.ndarray


def generate_grid(Lx, Ly, Nx, Ny, stretch_x="uniform", stretch_y="uniform", gamma_x=None, gamma_y=None) -> MACGrid:
    """generate_grid (MeshGeneration.jl:179-233) for the :uniform stretching."""
    if stretch_x != "uniform" or stretch_y != "uniform":
        raise UnsupportedError("stretched grids are outside the device path")
    Lx, Ly, Nx, Ny = float(Lx), float(Ly), int(Nx), int(Ny)
    if not Lx > 0 or not Ly > 0:
        raise ArgumentError("Lx and Ly must be positive.")
    if Nx < 2 or Ny < 2:
        raise ArgumentError("Nx and Ny must be at least 2.")
    x_u = Lx * (np.arange(Nx + 1, dtype=np.float64) / Nx)
    y_v = Ly * (np.arange(Ny + 1, dtype=np.float64) / Ny)
    x_u[0], x_u[-1], y_v[0], y_v[-1] = 0.0, Lx, 0.0, Ly
    x_p = Lx * ((np.arange(Nx, dtype=np.float64) + 0.5) / Nx)
    y_p = Ly * ((np.arange(Ny, dtype=np.float64) + 0.5) / Ny)
    return MACGrid(Lx, Ly, Nx, Ny, x_p, y_p, x_u, y_v)


def build_mesh_metrics(grid: MACGrid, periodic_x=False, periodic_y=False) -> MeshMetrics:
    """build_mesh_metrics (MeshGeneration.jl:340-372), cell widths and centre distances."""

    def axis(c, f, L, periodic):
        h_c = np.diff(f)
        h_f = np.empty(len(c) + 1)
        wrapped = c[0] + L - c[-1]
        h_f[0] = wrapped if periodic else 2.0 * (c[0] - f[0])
        h_f[1:-1] = np.diff(c)
        h_f[-1] = wrapped if periodic else 2.0 * (f[-1] - c[-1])
        return h_c, h_f

    dx_c, dx_f = axis(grid.x_p, grid.x_u, grid.Lx, periodic_x)
    dy_c, dy_f = axis(grid.y_p, grid.y_v, grid.Ly, periodic_y)
    return MeshMetrics(grid.Nx, grid.Ny, grid.Lx, grid.Ly, bool(periodic_x), bool(periodic_y), dx_c, dy_c, dx_f, dy_f)
