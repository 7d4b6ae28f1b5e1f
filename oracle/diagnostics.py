"""NumPy restatement of the reference cavity driver's diagnostics.

TEST INFRASTRUCTURE ONLY: the checker of eulerianfsi_amd.diagnostics (the
device reductions), imported by tests/.  The product never imports it.

Each function follows the Julia function it names, loop for loop, in
scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl and
scripts/benchmarks/BenchmarkUtilities.jl, on a uniform grid (cell widths
hx, hy; centres x_p = lx (i + 1/2) / nx).  Arrays are indexed [i, j] (x
first), as the reference's matrices.  Pinned by closed forms in
tests/test_diagnostics_pins.py.
"""
from __future__ import annotations

import math

import numpy as np

EPS = np.finfo(np.float64).eps


class Grid:
    """The geometry the driver diagnostics read: centres, cell widths and the
    phi boundary conditions of scalar_gradient (kind per side: 'P', 'D' or
    'N' with its value)."""

    def __init__(self, lx, ly, nx, ny, phi_bc=(("N", 0.0),) * 4):
        self.lx, self.ly, self.nx, self.ny = lx, ly, nx, ny
        self.hx, self.hy = lx / nx, ly / ny
        self.x_p = lx * ((np.arange(nx) + 0.5) / nx)       # MeshGeneration.jl:215
        self.y_p = ly * ((np.arange(ny) + 0.5) / ny)
        self.phi_bc = phi_bc  # left, right, bottom, top


def cell_area(g, i, j):  # :485-486
    return g.hx * g.hy


def phase_integral(phi, g):  # :488-494
    return float(np.sum(phi * (g.hx * g.hy)))


def phase_centroid(phi, g):  # :496-509
    mass = phase_integral(phi, g)
    if not mass > EPS:
        raise RuntimeError("Phase integral is too small to define a centroid.")
    w = phi * (g.hx * g.hy)
    return float(np.sum(g.x_p[:, None] * w)) / mass, float(np.sum(g.y_p[None, :] * w)) / mass, mass


def cell_center_velocity(U, V):  # :511-520
    return 0.5 * (U[:-1, :] + U[1:, :]), 0.5 * (V[:, :-1] + V[:, 1:])


def divergence(U, V, g):  # Operators.jl divergence on a uniform grid
    return (U[1:, :] - U[:-1, :]) / g.hx + (V[:, 1:] - V[:, :-1]) / g.hy


def _face_gradient_x(phi, g):
    """x-face gradient of scalar_gradient with the phi BC rows
    (Operators.jl:189-356 on a uniform grid)."""
    nx, ny = phi.shape
    gx = np.zeros((nx + 1, ny))
    gx[1:nx, :] = (phi[1:, :] - phi[:-1, :]) / g.hx
    (kl, vl), (kr, vr) = g.phi_bc[0], g.phi_bc[1]
    if kl == "P":
        gx[0, :] = gx[nx, :] = (phi[0, :] - phi[nx - 1, :]) / g.hx
    else:
        gx[0, :] = (phi[0, :] - vl) / (g.hx / 2) if kl == "D" else -vl
        gx[nx, :] = (vr - phi[nx - 1, :]) / (g.hx / 2) if kr == "D" else vr
    return gx


def _face_gradient_y(phi, g):
    nx, ny = phi.shape
    gy = np.zeros((nx, ny + 1))
    gy[:, 1:ny] = (phi[:, 1:] - phi[:, :-1]) / g.hy
    (kb, vb), (kt, vt) = g.phi_bc[2], g.phi_bc[3]
    if kb == "P":
        gy[:, 0] = gy[:, ny] = (phi[:, 0] - phi[:, ny - 1]) / g.hy
    else:
        gy[:, 0] = (phi[:, 0] - vb) / (g.hy / 2) if kb == "D" else -vb
        gy[:, ny] = (vt - phi[:, ny - 1]) / (g.hy / 2) if kt == "D" else vt
    return gy


def phase_center_gradient(phi, g):  # :612-624 (uniform weights 1/2)
    gxf, gyf = _face_gradient_x(phi, g), _face_gradient_y(phi, g)
    return 0.5 * gxf[:-1, :] + 0.5 * gxf[1:, :], 0.5 * gyf[:, :-1] + 0.5 * gyf[:, 1:]


def phase_geometry_diagnostics(phi, g, h_interface, nmax):  # :632-707
    xc, yc, mass = phase_centroid(phi, g)
    gx, gy = phase_center_gradient(phi, g)
    area = g.hx * g.hy
    pb = np.clip(phi, 0.0, 1.0)
    iw = area * np.hypot(gx, gy)
    perimeter = float(np.sum(iw))
    mixed = float(np.sum(area * 4.0 * pb * (1.0 - pb)))
    dx = g.x_p[:, None] - xc
    dy = g.y_p[None, :] - yc
    mw = area * phi
    cxx = float(np.sum(mw * dx * dx)) / mass
    cyy = float(np.sum(mw * dy * dy)) / mass
    cxy = float(np.sum(mw * dx * dy)) / mass
    theta = np.arctan2(np.broadcast_to(dy, phi.shape), np.broadcast_to(dx, phi.shape))
    rc = np.array([float(np.sum(iw * np.cos(n * theta) / math.pi)) for n in range(1, nmax + 1)])
    rs = np.array([float(np.sum(iw * np.sin(n * theta) / math.pi)) for n in range(1, nmax + 1)])
    gap = math.hypot(cxx - cyy, 2.0 * cxy)
    major, minor = 0.5 * (cxx + cyy + gap), 0.5 * (cxx + cyy - gap)
    width = mixed / perimeter if perimeter > EPS else math.nan
    return dict(x_centroid=xc, y_centroid=yc, phase_integral=mass, perimeter_proxy=perimeter,
                mixed_layer_measure=mixed, effective_interface_width=width,
                effective_interface_cells=width / h_interface, covariance_major=major, covariance_minor=minor,
                aspect_ratio=math.sqrt(major / minor) if minor > 0.0 else math.inf,
                orientation_radians=0.5 * math.atan2(2.0 * cxy, cxx - cyy) if gap > 100.0 * EPS else 0.0,
                R0=perimeter / (2.0 * math.pi), Rc=rc, Rs=rs, amplitudes=np.hypot(rc, rs))


def phase_bound_diagnostics(phi):  # :709-722
    lo, hi = float(np.min(phi)), float(np.max(phi))
    return dict(minimum=lo, maximum=hi, lower_violation=max(0.0, -lo), upper_violation=max(0.0, hi - 1.0),
                out_of_bounds_cells=int(np.count_nonzero((phi < 0.0) | (phi > 1.0))))


def flow_diagnostics(U, V, g, density):  # :724-751
    u, v = cell_center_velocity(U, V)
    s2 = u ** 2 + v ** 2
    area = g.hx * g.hy
    div = divergence(U, V, g)
    volume = float(np.sum(np.full(div.shape, area)))
    return dict(maximum_speed=max(0.0, float(np.max(np.sqrt(s2)))),
                kinetic_energy=float(np.sum(0.5 * density * s2 * area)),
                divergence_linf=float(np.max(np.abs(div))),
                divergence_l2=math.sqrt(float(np.sum(area * div ** 2)) / volume))


def tensor_cell_diagnostics_2d(sxx, syy, sxy, phi, G):  # BenchmarkUtilities.jl:62-88
    det = sxx * syy - sxy * sxy
    tr = sxx + syy
    gap = np.hypot(sxx - syy, 2.0 * sxy)
    stress = np.abs(phi * G) * np.sqrt((sxx - 1.0) ** 2 + (syy - 1.0) ** 2 + 2.0 * sxy ** 2)
    return det, det - 1.0, 0.5 * (tr - gap), 0.5 * (tr + gap), stress


def tensor_region_diagnostics(phi, sxx, syy, sxy, g, G, selector):  # :753-822
    sel = selector(phi)
    if not np.any(sel):
        return dict(cells=0, area=0.0, determinant_min=math.nan, determinant_max=math.nan,
                    max_abs_determinant_error=math.nan, mean_abs_determinant_error=math.nan,
                    lambda_min=math.nan, lambda_max=math.nan, non_spd_cells=0, max_effective_stress=math.nan,
                    mean_effective_stress=math.nan)
    det, derr, lmin, lmax, stress = tensor_cell_diagnostics_2d(sxx[sel], syy[sel], sxy[sel], phi[sel], G)
    area = g.hx * g.hy
    area_total = float(np.sum(np.full(det.shape, area)))
    return dict(cells=int(det.size), area=area_total, determinant_min=float(np.min(det)),
                determinant_max=float(np.max(det)), max_abs_determinant_error=float(np.max(np.abs(derr))),
                mean_abs_determinant_error=float(np.sum(area * np.abs(derr))) / area_total,
                lambda_min=float(np.min(lmin)), lambda_max=float(np.max(lmax)),
                non_spd_cells=int(np.count_nonzero(~(np.isfinite(lmin) & (lmin > 0.0)))),
                max_effective_stress=float(np.max(stress)),
                mean_effective_stress=float(np.sum(area * stress)) / area_total)


def tensor_diagnostics_detailed(phi, sxx, syy, sxy, g, G, core, interface, safety):  # :824-849
    region = lambda f: tensor_region_diagnostics(phi, sxx, syy, sxy, g, G, f)  # noqa: E731
    return dict(global_region=region(lambda p: np.ones(p.shape, dtype=bool)),
                core=region(lambda p: p >= core),
                interface=region(lambda p: (interface < p) & (p < core)),
                tail=region(lambda p: (safety < p) & (p <= interface)))
