"""The cavity driver loop (eulerianfsi_amd.driver.run_cavity) on the device:
adaptive dt from the comparison bounds, output events at the exact interval
times, the reference's CSV files, phase volume conserved along the
trajectory, and the last diagnostics row equal to the diagnostics of the
final state."""
import csv
import math
import os

import pytest

from eulerianfsi_amd import diagnostics as D
from eulerianfsi_amd import driver, setups

pytestmark = pytest.mark.gpu


def _rows(path):
    with open(path) as f:
        return list(csv.DictReader(f))


def test_short_cavity_run(tmp_path):
    n = 64
    case = setups.lid_driven_cavity(n, "acdi", acdi_eps_over_h=0.51)
    cfg = D.BenchmarkConfig(T=0.06, trajectory_interval=0.01, diagnostics_interval=0.025)
    h = driver.conservative_mesh_spacing(case.metrics)
    setup = driver.phase_field_setup(h)
    out = str(tmp_path / "run")
    summary = driver.run_cavity(case, cfg, setup, out, radius=0.2)
    print("steps", summary.steps, "t", summary.time, "wall", summary.wall_seconds, flush=True)
    assert abs(summary.time - cfg.T) < 1e-12
    for name in driver.FILES:
        assert os.path.exists(os.path.join(out, name))
    traj = _rows(os.path.join(out, "centroid_trajectory.csv"))
    times = [float(r["time"]) for r in traj]
    assert len(times) == 7 and all(abs(t - 0.01 * k) < 1e-10 for k, t in enumerate(times))
    for r in traj:  # ACDI conserves the phase volume
        assert abs(float(r["phase_integral_ratio"]) - 1.0) < 1e-12
    diag = _rows(os.path.join(out, "diagnostics.csv"))
    assert [round(float(r["time"]), 10) for r in diag] == [0.0, 0.025, 0.05, 0.06]
    num = _rows(os.path.join(out, "numerics_diagnostics.csv"))
    for r in num[1:-1]:
        assert float(r["dt_actual"]) <= float(r["dt_stability"]) * (1 + 1e-12)
        assert r["limiting_constraint"] == "current_cdi_diffusion"
        assert abs(float(r["dt_stability"]) - h / 32.0) < 1e-15  # 2h CDI width, 1/Pe = 4: h / 32
    # the CDI diffusion bound h/32 sets the step count, plus at most one
    # shortened step per output event (6 trajectory + 3 diagnostics times)
    full = cfg.T / (h / 32.0)
    assert full <= summary.steps <= full + 9
    # the last row holds the diagnostics of the final state
    d = D.compute_diagnostics(case.state, _ctx_of(case), cfg, h)
    last = diag[-1]
    assert float(last["x_centroid"]) == pytest.approx(d.geometry.x_centroid, rel=1e-9)
    assert float(last["max_speed"]) == pytest.approx(d.flow.maximum_speed, rel=1e-9)
    modes = _rows(os.path.join(out, "deformation_modes.csv"))
    assert len(modes) == 4 and "R12" in modes[0]
    tensor = _rows(os.path.join(out, "tensor_diagnostics.csv"))
    assert int(tensor[-1]["global_cells"]) == n * n


def _ctx_of(case):
    """The state's device context (the driver built it)."""
    return case.state._dev[0]
