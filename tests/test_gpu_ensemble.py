"""The batched ensemble entry point (fsi_ensemble_step): instances stepped
together advance exactly as each does alone (same kernels on its own
stream), including the planned steps that follow the first polled ones; an
instance's error is reported without stopping the others."""
import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import setups

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")


def _cases():
    return [setups.seeded_cavity(64, "acdi", seed=1),
            setups.seeded_cavity(128, "cdi", seed=2),
            setups.kolmogorov_filament(64, "uwcdi")]


def _prepare(cases):
    out = []
    for c in cases:
        ctx = fsi.build_imex_context(c.metrics, c.bcs, c.models, c.numerics)
        work = fsi.allocate_imex_workspace(ctx)
        fsi.make_velocity_compatible(c.state, ctx, work)
        out.append((ctx, work))
    return out


def test_ensemble_steps_equal_single_steps():
    nsteps = 6
    alone = _cases()
    for c, (ctx, work) in zip(alone, _prepare(alone)):
        for k in range(nsteps):
            fsi.integrate_step(c.state, k * c.dt, c.dt, ctx, work)
    together = _cases()
    env = _prepare(together)
    planned = []
    for k in range(nsteps):
        reps = fsi.ensemble_step([c.state for c in together], [k * c.dt for c in together],
                                 [c.dt for c in together], [e[0] for e in env], [e[1] for e in env])
        assert all(r is not None for r in reps)
        planned.append([e[1].last_stats.planned for e in env])
    print("planned", planned, flush=True)
    assert sum(map(sum, planned)) >= 6  # the later steps ran planned, in flight together
    for a, b in zip(alone, together):
        for f in FIELDS:
            assert np.array_equal(getattr(a.state, f), getattr(b.state, f)), f


def test_ensemble_reports_an_instance_error_and_advances_the_others():
    cases = _cases()
    env = _prepare(cases)
    before = [c.state.U.copy() for c in cases]
    dts = [c.dt for c in cases]
    dts[1] = -1.0  # invalid for the second instance only
    with pytest.raises(fsi.ArgumentError, match="instance 1"):
        fsi.ensemble_step([c.state for c in cases], [0.0] * 3, dts, [e[0] for e in env], [e[1] for e in env])
    assert np.array_equal(cases[1].state.U, before[1])
    ref = _cases()
    for i in (0, 2):  # the valid instances advanced exactly one step
        ctx = fsi.build_imex_context(ref[i].metrics, ref[i].bcs, ref[i].models, ref[i].numerics)
        work = fsi.allocate_imex_workspace(ctx)
        fsi.make_velocity_compatible(ref[i].state, ctx, work)
        fsi.integrate_step(ref[i].state, 0.0, ref[i].dt, ctx, work)
        assert np.array_equal(cases[i].state.U, ref[i].state.U)
        assert not np.array_equal(cases[i].state.U, before[i])


def test_config4_ensemble_equals_single_instance_steps():
    """The 64-cavity workload of the batched ensemble (N ~ U{128, 192, 256,
    384}, G log-uniform [0.05, 0.5], disk r ~ U[0.15, 0.25], centre ~
    U[0.4, 0.6]^2, ACDI, seed 4): three batched steps, every instance
    bitwise equal to the same steps taken one instance at a time."""
    nsteps = 3
    alone = setups.ensemble_cases()
    assert len(alone) == 64 and {c.params["N"] for c in alone} == {128, 192, 256, 384}
    env = _prepare(alone)
    for c, (ctx, work) in zip(alone, env):
        for k in range(nsteps):
            fsi.integrate_step(c.state, k * c.dt, c.dt, ctx, work)
    ref = [{f: getattr(c.state, f).copy() for f in FIELDS} for c in alone]
    del env
    together = setups.ensemble_cases()
    env = _prepare(together)
    for k in range(nsteps):
        reps = fsi.ensemble_step([c.state for c in together], [k * c.dt for c in together],
                                 [c.dt for c in together], [e[0] for e in env], [e[1] for e in env])
        assert all(r is not None for r in reps)
    for i, c in enumerate(together):
        for f in FIELDS:
            assert np.array_equal(getattr(c.state, f), ref[i][f]), (i, f)
