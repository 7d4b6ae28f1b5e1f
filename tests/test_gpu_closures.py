"""Parity of the device's UWCDI normal reconstructions (:youngs,
:smoothed9), uwcdi_strang_step and DiffuseAreaCapillarity with the CPU
oracle, through the C ABI (libfsi_hip.so).  Tolerances as in the step
parity tests: relative L2 <= 1e-12 for single right-hand sides and
tendencies, <= 1e-10 per step with both sides' solves tightened; the Strang
update and its increment within 5e-11."""
import dataclasses
import os

import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import closures as C
from eulerianfsi_amd import setups
from tests.golden import make_phase_golden as G
from tests.helpers import compact_u, compact_v, rel_l2
from tests.test_gpu_parity import FIELDS, TIGHT, _perturbed, _run_pair

pytestmark = pytest.mark.gpu


def _oracle(cfg):
    from oracle.oracle import Oracle
    return Oracle(*cfg)


def _dry():
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("dry run: oracle half only")


def _with_models(case, **kw):
    case.models = dataclasses.replace(case.models, **kw)
    return case


def _cfg_with(cfg, **kw):
    metrics, bcs, models, num = cfg
    return metrics, bcs, dataclasses.replace(models, **kw), num


NORMAL_CASES = [(c, m) for c in ("cavity", "xperiodic", "periodic", "channel", "dirichlet")
                for m in ("uwcdi_youngs", "uwcdi_smoothed9") if G.model_applies(c, m)]


@pytest.mark.parametrize("n", [48, 160])
@pytest.mark.parametrize("cname,mname", NORMAL_CASES)
def test_uwcdi_normal_reconstructions_match_oracle(cname, mname, n):
    """E_phi of the two reconstructions on every boundary branch (inflow and
    outflow walls, Dirichlet and Neumann phi, periodic sides, anisotropic
    cells); n = 160 spans several tiles in both directions."""
    cfg = G.configuration(cname, G.models()[mname], n=n, walls=True)
    U, V, phi, _ = G.inputs(cname, 9, n=n, walls=True)
    one, zero = np.ones_like(phi), np.zeros_like(phi)
    F = np.asfortranarray
    ref = _oracle(cfg).tendencies(dict(U=F(U), V=F(V), phi=F(phi), sigma_xx=F(one), sigma_yy=F(one),
                                       sigma_xy=F(zero)))["E_phi"]
    std = _oracle(_cfg_with(cfg, phase_field_method=C.PirozzoliUWCDIPhaseField(
        G.models()[mname].alpha, delta_normal=G.models()[mname].delta_normal))).tendencies(
        dict(U=F(U), V=F(V), phi=F(phi), sigma_xx=F(one), sigma_yy=F(one), sigma_xy=F(zero)))["E_phi"]
    _dry()
    ctx = fsi.build_imex_context(*cfg)
    work = fsi.allocate_imex_workspace(ctx)
    dev = fsi.evaluate_stage_tendencies(fsi.FSIState(U, V, phi, one, one, zero), ctx, work)["E_phi"]
    err = rel_l2(dev, ref)
    print(cname, mname, n, f"{err:.2e}", f"(vs :standard {rel_l2(std, ref):.2e})", flush=True)
    assert rel_l2(std, ref) > 1e-4  # the reconstruction changes the result
    assert err <= 1e-12, err


@pytest.mark.parametrize("d", ["x", "y"])
@pytest.mark.parametrize("cname,mname", [("cavity", "uwcdi"), ("xperiodic", "uwcdi_smoothed9"),
                                         ("channel", "uwcdi_youngs"), ("dirichlet", "uwcdi"),
                                         ("periodic", "uwcdi_youngs")])
def test_strang_step_matches_oracle(cname, mname, d):
    n, dt = 96, 0.004
    cfg = G.configuration(cname, G.models()[mname], n=n, walls=True)
    U, V, phi, _ = G.inputs(cname, 21, n=n, walls=True)
    ref = _oracle(cfg).uwcdi_strang_step(U, V, phi, dt, d)
    _dry()
    ctx = fsi.build_imex_context(*cfg)
    work = fsi.allocate_imex_workspace(ctx)
    one, zero = np.ones_like(phi), np.zeros_like(phi)
    state = fsi.FSIState(U, V, phi, one, one, zero)
    out = fsi.uwcdi_strang_step(state, dt, ctx, work, d)
    # phi and its increment to 5e-11 (relative L2).  The seeded
    # fields have neighbouring cells with nearly opposite normals: their face
    # average cancels to ~1e-5, which turns last-bit differences of the
    # centre normals (fused multiply-adds in |g|) into ~1e-11 relative
    # differences of that face's flux, three sweeps in a row
    err = rel_l2(out - phi, ref - phi)
    print(cname, mname, d, f"{err:.2e}", f"max {np.max(np.abs(out - ref)):.1e}", flush=True)
    assert rel_l2(out, ref) <= 5e-11
    assert err <= 5e-11, err
    assert np.max(np.abs(out - ref)) <= 1e-9
    # dt = 0 leaves phi as it is; the state's other fields are untouched
    before = state.phi.copy()
    fsi.uwcdi_strang_step(state, 0.0, ctx, work, d)
    assert np.array_equal(state.phi, before) and np.array_equal(state.U, U)


@pytest.mark.parametrize("N", [48, 256])
@pytest.mark.parametrize("passes", [0, 2])
def test_capillary_tendencies_match_oracle(passes, N):
    from oracle.oracle import state_dict
    case = _with_models(_perturbed(setups.lid_driven_cavity(N, "acdi", **TIGHT), seed=3),
                        capillarity=C.DiffuseAreaCapillarity(0.05, smoothing_passes=passes))
    ref = _oracle((case.metrics, case.bcs, case.models, case.numerics)).tendencies(state_dict(case.state))
    _dry()
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    dev = fsi.evaluate_stage_tendencies(case.state, ctx, work)
    for name, a, b in (("E_u", dev["E_u"], compact_u(ref["E_u"])), ("E_v", dev["E_v"], compact_v(ref["E_v"])),
                       ("E_phi", dev["E_phi"], ref["E_phi"])):
        err = rel_l2(a, b)
        print(passes, N, name, f"{err:.2e}", flush=True)
        assert err <= 1e-12, (name, err)


@pytest.mark.parametrize("variant", ["youngs", "smoothed9", "capillary_acdi", "capillary_uwcdi_smoothed9"])
def test_cavity_steps_with_closures_match_oracle(variant):
    method = "acdi" if variant == "capillary_acdi" else "uwcdi"
    case = setups.seeded_cavity(64, method, seed=11, **TIGHT)
    if variant in ("youngs", "smoothed9"):
        _with_models(case, phase_field_method=C.PirozzoliUWCDIPhaseField(1.0, delta_normal=1e-14,
                                                                         normal_reconstruction=variant))
    elif variant == "capillary_acdi":
        _with_models(case, capillarity=C.DiffuseAreaCapillarity(0.02, smoothing_passes=1))
    else:
        _with_models(case, capillarity=C.DiffuseAreaCapillarity(0.01),
                     phase_field_method=C.PirozzoliUWCDIPhaseField(1.0, delta_normal=1e-14,
                                                                   normal_reconstruction="smoothed9"))
    vol0 = case.state.phi.sum()
    errs, rep, rep_o, ref, _, _ = _run_pair(case, 4)
    for k, e in enumerate(errs):
        print(variant, k, {f: f"{v:.2e}" for f, v in e.items()}, flush=True)
        for f, v in e.items():
            assert v <= 1e-10, (k, f, v)
    assert abs(case.state.phi.sum() - vol0) <= 1e-13 * vol0
    assert abs(rep.maximum_phi - rep_o.maximum_phi) <= 1e-12


def test_periodic_filament_steps_with_youngs_normals_match_oracle():
    case = setups.kolmogorov_filament(64, "uwcdi", **TIGHT)
    _with_models(case, phase_field_method=C.PirozzoliUWCDIPhaseField(1.0, delta_normal=1e-14,
                                                                     normal_reconstruction="youngs"))
    errs, *_ = _run_pair(case, 3)
    for k, e in enumerate(errs):
        print("filament youngs", k, {f: f"{v:.2e}" for f, v in e.items()}, flush=True)
        for f in FIELDS:
            if f != "V":
                assert e[f] <= 1e-10, (k, f, e[f])
        assert e["UV"] <= 1e-10


def test_closure_argument_errors():
    _dry()
    # smoothed9 needs homogeneous Neumann or periodic phase sides (PhaseFieldMethods.jl:347-362)
    cfg = G.configuration("channel", G.models()["uwcdi_smoothed9"], n=32)
    with pytest.raises(fsi.ArgumentError, match="smoothed9"):
        fsi.build_imex_context(*cfg)
    # uwcdi_strang_step needs the UWCDI phase field, dt >= 0 and a known direction
    case = setups.lid_driven_cavity(32, "acdi")
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    with pytest.raises(fsi.ArgumentError):
        fsi.uwcdi_strang_step(case.state, 0.01, ctx, work)
    case = setups.lid_driven_cavity(32, "uwcdi")
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    with pytest.raises(fsi.ArgumentError):
        fsi.uwcdi_strang_step(case.state, -0.01, ctx, work)
    with pytest.raises(fsi.ArgumentError):
        fsi.uwcdi_strang_step(case.state, float("nan"), ctx, work)
    with pytest.raises(fsi.ArgumentError):
        fsi.uwcdi_strang_step(case.state, 0.01, ctx, work, "z")
    # the device adjoint does not differentiate capillary forcing
    case = _with_models(setups.lid_driven_cavity(32, "acdi", velocity_method="direct"),
                        capillarity=C.DiffuseAreaCapillarity(0.01))
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    with pytest.raises(fsi.UnsupportedError):
        fsi.adjoint.StepAdjoint(ctx, work, snapshots=2)
