"""Pins of the driver-diagnostics restatement (oracle/diagnostics.py) by
closed forms, and of the host-side parts of eulerianfsi_amd.diagnostics (the
timestep-bound arithmetic and the CSV formats of
scripts/benchmarks/run_sugiyama_lid_driven_cavity.jl:938-1135).  CPU only."""
import io
import math

import numpy as np
import pytest

from eulerianfsi_amd import diagnostics as D
from oracle import diagnostics as O


def _tanh_shape(g, center, radius_of_theta, eps):
    """phi = (1 + tanh((R(theta) - rho) / (2 eps))) / 2 about `center`."""
    X, Y = np.meshgrid(g.x_p, g.y_p, indexing="ij")
    dx, dy = X - center[0], Y - center[1]
    rho, theta = np.hypot(dx, dy), np.arctan2(dy, dx)
    return 0.5 * (1.0 + np.tanh((radius_of_theta(theta) - rho) / (2.0 * eps)))


def test_tanh_disk_closed_forms():
    """Perimeter proxy -> 2 pi r, interface width -> 4 eps (int sech^2),
    mass -> pi r^2 + pi^3 eps^2 / 3 (Sommerfeld), second moments -> r^2 / 4,
    aspect ratio 1, no deformation modes."""
    g = O.Grid(1.2, 1.0, 600, 500)  # rectangular, so x_p and y_p differ
    r, eps, c = 0.2, 0.01, (0.7, 0.45)
    phi = _tanh_shape(g, c, lambda t: r, eps)
    geo = O.phase_geometry_diagnostics(phi, g, g.hx, 12)
    assert abs(geo["x_centroid"] - c[0]) < 1e-8 and abs(geo["y_centroid"] - c[1]) < 1e-8
    assert abs(geo["phase_integral"] / (math.pi * r * r + math.pi ** 3 * eps * eps / 3.0) - 1.0) < 1e-4
    assert abs(geo["perimeter_proxy"] / (2.0 * math.pi * r) - 1.0) < 5e-3
    assert abs(geo["R0"] / r - 1.0) < 5e-3
    assert abs(geo["effective_interface_width"] / (4.0 * eps) - 1.0) < 5e-3
    assert abs(geo["effective_interface_cells"] - geo["effective_interface_width"] / g.hx) < 1e-12
    # second moment of the smoothed disk (Sommerfeld, int pi rho^3 f):
    # (r^4/4 + pi^2 eps^2 r^2 / 2) / (r^2 + pi^2 eps^2 / 3), next term
    # (7 pi^4 / 360) 24 (eps / r)^4 = 2.8e-4
    pe2 = math.pi ** 2 * eps * eps
    cxx = (r ** 4 / 4.0 + pe2 * r * r / 2.0) / (r * r + pe2 / 3.0)
    assert abs(geo["covariance_major"] / cxx - 1.0) < 1e-3 and abs(geo["covariance_minor"] / cxx - 1.0) < 1e-3
    assert abs(geo["aspect_ratio"] - 1.0) < 1e-6
    assert np.max(geo["amplitudes"]) < 1e-3 * geo["R0"]


@pytest.mark.parametrize("theta0", [0.0, 0.4])
def test_perturbed_circle_mode(theta0):
    """r(theta) = r0 (1 + delta cos 3 (theta - theta0)): R3 / R0 = delta to
    O(delta^2), with (Rc3, Rs3) in the direction 3 theta0; other modes ~ 0."""
    g = O.Grid(1.0, 1.0, 512, 512)
    r0, delta, eps = 0.2, 0.04, 0.006
    phi = _tanh_shape(g, (0.5, 0.5), lambda t: r0 * (1.0 + delta * np.cos(3.0 * (t - theta0))), eps)
    geo = O.phase_geometry_diagnostics(phi, g, g.hx, 6)
    R0, Rc, Rs = geo["R0"], geo["Rc"], geo["Rs"]
    assert abs(math.hypot(Rc[2], Rs[2]) / R0 / delta - 1.0) < 0.1
    assert abs(math.atan2(Rs[2], Rc[2]) - 3.0 * theta0) < 1e-2
    for n in (0, 1, 3, 4):
        assert math.hypot(Rc[n], Rs[n]) < 0.05 * delta * R0


def test_rotated_ellipse_covariance():
    """Uniform elliptical disk: covariance eigenvalues a^2/4, b^2/4, aspect
    a / b, orientation the rotation angle."""
    g = O.Grid(1.0, 1.0, 1024, 1024)
    a, b, th = 0.25, 0.12, 0.3
    X, Y = np.meshgrid(g.x_p, g.y_p, indexing="ij")
    xr = (X - 0.5) * math.cos(th) + (Y - 0.5) * math.sin(th)
    yr = -(X - 0.5) * math.sin(th) + (Y - 0.5) * math.cos(th)
    phi = ((xr / a) ** 2 + (yr / b) ** 2 <= 1.0).astype(float)
    geo = O.phase_geometry_diagnostics(phi, g, g.hx, 4)
    assert abs(geo["covariance_major"] / (a * a / 4) - 1.0) < 1e-2
    assert abs(geo["covariance_minor"] / (b * b / 4) - 1.0) < 1e-2
    assert abs(geo["aspect_ratio"] - a / b) < 1e-2 * a / b
    assert abs(geo["orientation_radians"] - th) < 5e-3


def test_solid_body_rotation_flow():
    """u = -w (y - 1/2), v = w (x - 1/2) on the faces of the unit square:
    the MAC divergence vanishes exactly, KE = w^2 / 12 (1 - 1/N^2) (midpoint
    sums), max speed = w sqrt(2) (1/2 - h/2)."""
    n, w = 64, 1.7
    g = O.Grid(1.0, 1.0, n, n)
    U = np.repeat((-w * (g.y_p - 0.5))[None, :], n + 1, axis=0)
    V = np.repeat((w * (g.x_p - 0.5))[:, None], n + 1, axis=1)
    f = O.flow_diagnostics(U, V, g, 1.0)
    assert f["divergence_linf"] == 0.0 and f["divergence_l2"] == 0.0
    assert abs(f["kinetic_energy"] - w * w / 12.0 * (1.0 - 1.0 / n ** 2)) < 1e-14
    assert abs(f["maximum_speed"] - w * math.sqrt(2.0) * (0.5 - 0.5 / n)) < 1e-14
    # U = x on the x-faces: divergence 1 everywhere
    x_u = np.arange(n + 1) / n
    f = O.flow_diagnostics(np.repeat(x_u[:, None], n, axis=1), np.zeros((n, n + 1)), g, 2.0)
    assert abs(f["divergence_linf"] - 1.0) < 1e-12 and abs(f["divergence_l2"] - 1.0) < 1e-12


def test_tensor_regions_partition_and_closed_forms():
    """Regions on a ramp that hits the cutoffs exactly: global = every cell and
    core + interface + tail + (phi <= safety) = every cell; a constant tensor
    gives det = ab - c^2, eigenvalues (a + b -+ hypot(a - b, 2c)) / 2."""
    n = 40
    g = O.Grid(1.0, 1.0, n, n)
    levels = np.array([0.0, 1e-6, 0.05, 0.1, 0.5, 0.9, 0.95, 1.0])  # the cutoffs themselves included
    phi = np.repeat(levels[np.arange(n) % levels.size][:, None], n, axis=1)
    a, b, c, G = 1.3, 0.8, 0.2, 0.1
    sxx, syy, sxy = np.full((n, n), a), np.full((n, n), b), np.full((n, n), c)
    t = O.tensor_diagnostics_detailed(phi, sxx, syy, sxy, g, G, 0.9, 0.1, 1e-6)
    assert t["global_region"]["cells"] == n * n
    below = int(np.count_nonzero(phi <= 1e-6))
    assert t["core"]["cells"] + t["interface"]["cells"] + t["tail"]["cells"] + below == n * n
    assert t["core"]["cells"] == int(np.count_nonzero(phi >= 0.9))
    assert t["tail"]["cells"] == int(np.count_nonzero((phi == 0.1) | (phi == 0.05)))
    r = t["global_region"]
    assert abs(r["determinant_min"] - (a * b - c * c)) < 1e-15
    assert abs(r["lambda_min"] - 0.5 * (a + b - math.hypot(a - b, 2 * c))) < 1e-15
    assert abs(r["lambda_max"] - 0.5 * (a + b + math.hypot(a - b, 2 * c))) < 1e-15
    frob = math.sqrt((a - 1) ** 2 + (b - 1) ** 2 + 2 * c * c)
    assert abs(r["max_effective_stress"] - G * frob) < 1e-15
    assert abs(r["mean_effective_stress"] - G * frob * np.mean(phi)) < 1e-14
    assert r["non_spd_cells"] == 0
    sxx[3, 5] = -0.5  # one cell with a negative eigenvalue
    assert O.tensor_diagnostics_detailed(phi, sxx, syy, sxy, g, G, 0.9, 0.1, 1e-6)["global_region"][
        "non_spd_cells"] == 1
    empty = O.tensor_region_diagnostics(phi, sxx, syy, sxy, g, G, lambda p: p > 2.0)
    assert empty["cells"] == 0 and math.isnan(empty["lambda_min"])


def test_phase_bounds():
    phi = np.array([[0.2, -0.1], [1.25, 0.5]])
    b = O.phase_bound_diagnostics(phi)
    assert b == dict(minimum=-0.1, maximum=1.25, lower_violation=0.1, upper_violation=0.25, out_of_bounds_cells=2)


def test_timestep_bounds_hand_values():
    """comparison_timestep_bounds (:852-934) with the driver defaults,
    h = 1/128, CDI eps = 2h, ACDI eps = 0.51h: the CDI diffusion bound h/32
    limits."""
    cfg = D.BenchmarkConfig()
    h = 1.0 / 128
    tb = D.timestep_bounds_from_speed(cfg, h, h, 0.8, 0.3, 2.0 * h, 0.51 * h)
    assert tb.max_advection_speed == 1.0
    assert abs(tb.elastic_wave_speed - math.sqrt(0.1)) < 1e-16
    assert abs(tb.hyperbolic - 0.1 * h) < 1e-18
    assert abs(tb.cdi_compression - h / 16) < 1e-18
    assert abs(tb.cdi_diffusion - h / 32) < 1e-18
    assert abs(tb.acdi_compression - h / 4) < 1e-18
    assert abs(tb.acdi_diffusion - 0.5 * h / (4.0 * 0.51)) < 1e-17
    assert abs(tb.uwcdi - 0.5 * h) < 1e-18
    assert tb.selected == tb.cdi_diffusion and tb.limiting_constraint == "current_cdi_diffusion"
    # a fast lid makes the hyperbolic bound the smallest
    tb = D.timestep_bounds_from_speed(D.BenchmarkConfig(U_lid=40.0), h, h, 0.0, 0.0, 2.0 * h, 0.51 * h)
    assert tb.limiting_constraint == "hyperbolic" and abs(tb.selected - 0.1 * h / 40.0) < 1e-18


def test_csv_formats():
    """Headers and rows as the reference's println / @printf write them."""
    out = io.StringIO()
    D.write_flow_diagnostics_header(out)
    D.write_flow_diagnostics_row(out, 0.5, D.FlowDiagnostics(1.0, math.nan, math.inf, -0.0))
    assert out.getvalue() == ("time,max_speed,kinetic_energy,divergence_linf,divergence_l2\n"
                              "5.0000000000000000e-01,1.0000000000000000e+00,NaN,Inf,-0.0000000000000000e+00\n")
    out = io.StringIO()
    D.write_trajectory_row(out, 0.02, 0.6, 0.5, 0.125, 0.125)
    assert out.getvalue() == "2.0000000000e-02,6.0000000000e-01,5.0000000000e-01,1.2500000000e-01,1.0000000000e+00\n"
    out = io.StringIO()
    D.write_diagnostics_row(out, 7, 0.1, 1e-3, 0.6, 0.5, 0.126, 0.125, -1e-3, 1.0, 0.9, 1e-12, 0.99, 1e-9, 3, 2e-12)
    assert out.getvalue().startswith("7,1.0000000000e-01,1.0000000000e-03,")
    assert out.getvalue().endswith(",3,2.0000000000e-12\n")
    out = io.StringIO()
    D.write_deformation_modes_header(out, 2)
    assert out.getvalue() == "time,R0,Rc1,Rs1,R1,Rc2,Rs2,R2,high_mode_amplitude_n6_to_nmax,high_mode_ratio_n6_to_nmax\n"
    geo = D.PhaseGeometry(*([0.0] * 11), 0.2, (0.1,) * 7, (0.0,) * 7, (0.1,) * 7)
    out = io.StringIO()
    D.write_deformation_modes_row(out, 1.0, geo, 7)
    tail = out.getvalue().strip().split(",")[-2:]
    assert float(tail[0]) == pytest.approx(math.sqrt(2) * 0.1) and float(tail[1]) == pytest.approx(math.sqrt(2) / 2)
    out = io.StringIO()
    D.write_tensor_diagnostics_header(out)
    names = out.getvalue().strip().split(",")
    assert len(names) == 1 + 4 * 11 and names[1] == "global_cells" and names[-1] == "tail_mean_effective_stress"
    region = D.TensorRegion(5, 0.5, *([1.0] * 6), 2, 0.0, math.nan)
    vals = D.tensor_region_csv_values(region)
    assert vals[0] == "5" and vals[8] == "2" and vals[-1] == "NaN"
