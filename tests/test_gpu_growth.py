"""The growth and resource path on the device (libfsi_hip.so through the C
ABI) against the CPU oracle and against closed forms.  Both sides' solves are
tight (velocity and resource :direct, projection 1e-13), so the per-step gate
is a relative L2 of 1e-10 on U, V, phi, resource and sigma."""
import os

import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import closures as C
from tests import cases
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy")


def tight():
    return fsi.IMEXNumerics(projection_reltol=1e-13, linear_solvers=fsi.LinearSolverOptions(
        velocity_method="direct", resource_method="direct"))


def _device(cfg):
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("dry run: oracle half only")
    _, metrics, bcs, models, numerics = cfg
    ctx = fsi.build_imex_context(metrics, bcs, models, numerics)
    return ctx, fsi.allocate_imex_workspace(ctx)


def _pair(cfg, state, nsteps, dt, tol=1e-10):
    from oracle.oracle import state_dict
    o = cases.oracle_for(cfg)
    ref = state_dict(state)
    res_o = o.make_velocity_compatible(ref)
    traj = []
    for k in range(nsteps):
        rep_o, _ = o.integrate_step(ref, k * dt, dt)
        traj.append(({f: ref[f].copy() for f in FIELDS}, rep_o))
    ctx, work = _device(cfg)
    res = fsi.make_velocity_compatible(state, ctx, work)
    assert res.constraint_linf < 1e-9 and res_o.constraint_linf < 1e-9
    worst = 0.0
    for k in range(nsteps):
        rep = fsi.integrate_step(state, k * dt, dt, ctx, work)
        ref_k, rep_o = traj[k]
        # velocity components relative to the whole velocity, sigma components
        # to the whole tensor (a component that is exactly zero holds only
        # round-off on either side)
        vel = np.sqrt(np.sum(ref_k["U"] ** 2) + np.sum(ref_k["V"] ** 2))
        sig = np.sqrt(sum(np.sum(ref_k[f] ** 2) for f in ("sigma_xx", "sigma_yy", "sigma_xy")))
        for f in FIELDS:
            d = np.sqrt(np.sum((np.asarray(getattr(state, f)) - ref_k[f]) ** 2))
            e = d / vel if f in ("U", "V") else (d / sig if f.startswith("sigma") else rel_l2(getattr(state, f), ref_k[f]))
            worst = max(worst, e)
            assert e <= tol, (k, f, e)
        assert rep.maximum_growth_trace_error == pytest.approx(rep_o.maximum_growth_trace_error, abs=1e-15)
        assert rep.minimum_resource == pytest.approx(rep_o.minimum_resource, rel=1e-12)
        assert rep.minimum_phi == pytest.approx(rep_o.minimum_phi, rel=1e-10, abs=1e-14)
        assert rep.maximum_phi == pytest.approx(rep_o.maximum_phi, rel=1e-10)
        assert rep.non_spd_cells == rep_o.non_spd_cells
        assert rep.maximum_projection_constraint_linf < 1e-8
    print(f"worst relative L2 over {nsteps} steps: {worst:.3e}", flush=True)
    return state, rep


def test_growth_smoke_matches_oracle_and_reference_invariants():
    """test/growth_tests.jl:96-141 on the device: saturating growth with a
    diffused, consumed resource in an x-periodic strip opened at the bottom."""
    cfg = cases.growth_smoke(8, numerics=tight())
    state, rep = _pair(cfg, cases.growth_smoke_state(8), 5, 1e-4)
    assert np.max(np.abs(state.U[-1] - state.U[0])) < 1e-12
    assert np.max(np.abs(state.V[:, -1])) < 1e-12
    assert rep.maximum_growth_trace_error < 1e-12 and rep.minimum_resource > 0.0


@pytest.mark.parametrize("N,steps", [(64, 4), (256, 2)])
def test_growing_disk_matches_oracle(N, steps):
    """Viscous, elastic ACDI disk growing on a resource with harmonic
    diffusivity and consumption, a directional growth tensor (shear part of
    the growth stress nonzero), the top opened by the projection."""
    cfg, state = cases.growth_box(N, numerics=tight())
    _pair(cfg, state, steps, 2e-3)


def test_isotropic_growth_with_periodic_resource_matches_oracle():
    cfg, state = cases.growth_box(64, tensor=C.IsotropicGrowthTensor(), acdi=False, numerics=tight())
    _pair(cfg, state, 3, 2e-3)


@pytest.mark.parametrize("tensor,rates", [(C.IsotropicGrowthTensor(), (-1.0, 1.0)),
                                          (C.DirectionalGrowthTensor(0.0, 1.0), (0.0, 0.0))])
def test_uniaxial_growth_closed_form_on_device(tensor, rates):
    cfg, state = cases.uniaxial_growth(128, 0.3, tensor)
    ctx, work = _device(cfg)
    V0 = state.V.copy()
    dt = 0.05
    fsi.integrate_step(state, 0.0, dt, ctx, work)
    rk3 = lambda z: 1.0 + z + z * z / 2.0 + z ** 3 / 6.0
    assert np.allclose(state.V, V0, rtol=0, atol=1e-12)
    assert np.allclose(state.sigma_xx, rk3(rates[0] * 0.3 * dt), rtol=0, atol=1e-12)
    assert np.allclose(state.sigma_yy, rk3(rates[1] * 0.3 * dt), rtol=0, atol=1e-12)
    assert np.allclose(state.phi, 1.0, rtol=0, atol=1e-12)


def _unchanged_after_error(cfg, state, exc, dt=1e-3):
    ctx, work = _device(cfg)
    before = {f: np.array(getattr(state, f), copy=True) for f in FIELDS}
    with pytest.raises(exc):
        fsi.integrate_step(state, 0.0, dt, ctx, work)
    for f in FIELDS:
        assert np.array_equal(np.asarray(getattr(state, f)), before[f]), f
    return ctx, work


def test_neumann_resource_inflow_is_an_argument_error():
    """Flow entering through a side whose resource condition is Neumann:
    ArgumentError (ResourceTransport.jl:111-116), the state unchanged."""
    cfg, state = cases.uniaxial_growth(32, 0.3)
    grid, metrics, bcs, models, numerics = cfg
    px, n0 = fsi.Periodic(), fsi.Neumann(0.0)
    bcs.V = fsi.BoundaryConditionSetup(px, px, fsi.Dirichlet(0.2), fsi.Dirichlet(0.5))
    state.V = state.V + 0.2
    _unchanged_after_error((grid, metrics, bcs, models, numerics), state, fsi.ArgumentError)


def test_nonpositive_harmonic_diffusivity_is_an_argument_error():
    """phi = -0.5 with D_f = 0.1, D_s = 0.01 makes the harmonic law negative
    (TransportConstitutive.jl:255-263, :53-59): ArgumentError, state unchanged."""
    N = 16
    models = fsi.FSIModels(diffusivity=C.HarmonicPhaseDiffusivity(0.1, 0.01))
    cfg = cases.resource_box(N, models=models)
    phi = np.full((N, N), 0.5)
    phi[3, 4] = -0.5
    state = fsi.FSIState(np.zeros((N + 1, N)), np.zeros((N, N + 1)), phi, np.ones((N, N)), np.ones((N, N)),
                         np.ones((N, N)), np.zeros((N, N)))
    ctx, work = _unchanged_after_error(cfg, state, fsi.ArgumentError)
    # a valid phase field then steps normally
    state.phi = np.full((N, N), 0.5)
    fsi.integrate_step(state, 0.0, 1e-3, ctx, work)
