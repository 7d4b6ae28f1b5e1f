"""The device discrete adjoint against the oracle's tangent and central
differences (lid-driven cavity, ACDI, FP64, tight solves).

* The point functions the adjoint differentiates reproduce the forward
  tendencies (fsi_eval_tendencies).
* Transposes, checked by the dot-product identity <bar, J dx> = <J^T bar, dx>
  with J dx from the oracle's FP64 tangent (oracle/fsi_tangent.cpp, the
  oracle compiled with dual numbers): the stage tendencies alone, and one
  whole integrate_step!, each with the shear-modulus derivative.
* The gradient of J = |x_c(phi(T)) - target|^2 with respect to (shear
  modulus, disk radius, disk centre) over a run with checkpoints, against the
  oracle's tangent along the four parameters and against central differences
  of the device's own forward runs (h = 1e-6).
* Checkpoint counts change nothing: every recomputation reproduces the run;
  keeping every step's tape instead gives the same gradient.
"""
import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import adjoint as A
from eulerianfsi_amd import setups
from oracle import tangent as T
from oracle.oracle import state_dict

pytestmark = pytest.mark.gpu

P0 = (0.1, 0.2, 0.6, 0.5)  # shear modulus, radius, centre
TARGET = (0.56, 0.53)


def _case(N, p=P0, dt=2e-3):
    G, r, cx, cy = p
    return setups.lid_driven_cavity(N, "acdi", radius=r, center=(cx, cy), shear_modulus=G, dt=dt,
                                    velocity_method="direct", projection_rtol=1e-13)


def _stepped(N, steps):
    case = _case(N)
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    for k in range(steps):
        fsi.integrate_step(case.state, k * case.dt, case.dt, ctx, work)
    return case, ctx, work


def _rand_fields(rng, nx, ny, walls_zero):
    f = {"U": rng.standard_normal((nx + 1, ny)), "V": rng.standard_normal((nx, ny + 1))}
    for k in ("phi", "sigma_xx", "sigma_yy", "sigma_xy"):
        f[k] = rng.standard_normal((nx, ny))
    if walls_zero:  # the wall faces carry boundary data, not unknowns
        f["U"][0, :] = f["U"][nx, :] = 0.0
        f["V"][:, 0] = f["V"][:, ny] = 0.0
    return {k: np.asfortranarray(v) for k, v in f.items()}


def _dot(a, b, keys):
    return sum(float(np.sum(a[k] * b[k])) for k in keys)


GRADS = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")


@pytest.mark.parametrize("N", [40, 80])
def test_point_functions_reproduce_the_forward_tendencies(N):
    case, ctx, work = _stepped(N, 4)
    ref = fsi.evaluate_stage_tendencies(case.state, ctx, work)
    got = A.primal_tendencies(case.state, ctx)
    for k in ref:
        a, b = got[k][:N, :N], ref[k]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
        print(f"N={N} {k}: relative difference {err:.2e}", flush=True)
        assert err <= 1e-12, (k, err)


def test_tendency_transpose_matches_the_oracle_tangent():
    N = 40
    case, ctx, work = _stepped(N, 4)
    s = state_dict(case.state)
    rng = np.random.default_rng(11)
    dx = _rand_fields(rng, N, N, walls_zero=False)
    tg = T.Tangent(case.metrics, case.bcs, case.models, case.numerics)
    tg.set_shear_modulus_derivatives([0.0, 1.0, 0.0, 0.0])
    ds = {k: T.pack(s[k], (dx.get(k),)) for k in T.FIELDS}
    tout = tg.tendencies(ds)
    # cotangents of the tendencies on the faces and cells the device forms
    bar = {"E_u": rng.standard_normal((N + 1, N)), "E_v": rng.standard_normal((N, N + 1)),
           "I_u": rng.standard_normal((N + 1, N)), "I_v": rng.standard_normal((N, N + 1))}
    for k in ("E_u", "I_u"):
        bar[k][N, :] = 0.0
    for k in ("E_v", "I_v"):
        bar[k][:, N] = 0.0
    for k in ("E_phi", "E_sigma_xx", "E_sigma_yy", "E_sigma_xy"):
        bar[k] = rng.standard_normal((N, N))
    bar = {k: np.asfortranarray(v) for k, v in bar.items()}
    lhs = sum(float(np.sum(bar[k] * T.unpack(tout[k])[1][0])) for k in bar)
    lhs_G = sum(float(np.sum(bar[k] * T.unpack(tout[k])[1][1])) for k in bar)
    # the sums cancel (random signs): the bound is 1e-12 of the sum of the
    # terms' magnitudes, a per-term agreement far below any transposition
    # error (observed 2e-13 at most, on sigma_xy: the device's fused
    # multiply-adds against the oracle's separate roundings)
    mag = sum(float(np.sum(np.abs(bar[k] * T.unpack(tout[k])[1][0]))) for k in bar)
    mag_G = sum(float(np.sum(np.abs(bar[k] * T.unpack(tout[k])[1][1]))) for k in bar)
    grads, dG = A.tendencies_transpose(case.state, ctx, bar)
    rhs = _dot(grads, dx, GRADS)
    print(f"<bar, J dx> = {lhs:.15e}, <J^T bar, dx> = {rhs:.15e} (terms {mag:.3e}); dG {lhs_G:.15e} vs "
          f"{dG:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-12 * max(mag, 1.0)
    assert abs(lhs_G - dG) <= 1e-12 * max(mag_G, 1.0)
    # each input field on its own
    for k in GRADS:
        only = {kk: (dx[kk] if kk == k else np.zeros_like(dx[kk])) for kk in dx}
        ds = {kk: T.pack(s[kk], (only.get(kk),)) for kk in T.FIELDS}
        out = tg.tendencies(ds)
        l = sum(float(np.sum(bar[q] * T.unpack(out[q])[1][0])) for q in bar)
        r = float(np.sum(grads[k] * dx[k]))
        m = float(np.sum(np.abs(grads[k] * dx[k])))
        print(f"  {k}: {l:.15e} vs {r:.15e} (terms {m:.3e}, {abs(l - r) / max(m, 1.0):.1e})", flush=True)
        assert abs(l - r) <= 1e-12 * max(m, 1.0), k


def test_step_transpose_matches_the_oracle_tangent():
    N = 40
    case, ctx, work = _stepped(N, 5)
    s = state_dict(case.state)
    rng = np.random.default_rng(5)
    dx = _rand_fields(rng, N, N, walls_zero=True)
    lam = _rand_fields(rng, N, N, walls_zero=True)
    # the oracle's tangent of one step from the same state
    tg = T.Tangent(case.metrics, case.bcs, case.models, case.numerics)
    tg.set_shear_modulus_derivatives([0.0, 1.0, 0.0, 0.0])
    ds = {k: T.pack(s[k], (dx.get(k),)) for k in T.FIELDS}
    t = 5 * case.dt
    tg.integrate_step(ds, t, case.dt)
    lhs = sum(float(np.sum(lam[k] * T.unpack(ds[k])[1][0])) for k in GRADS)
    lhs_G = sum(float(np.sum(lam[k] * T.unpack(ds[k])[1][1])) for k in GRADS)
    # the device: taped step, then its transpose
    adj = A.StepAdjoint(ctx, work, snapshots=1)
    adj.record_step(case.state, t, case.dt)
    new = state_dict(case.state)
    for k in GRADS:
        v, _ = T.unpack(ds[k])
        err = np.linalg.norm(new[k] - v) / max(np.linalg.norm(v), 1e-300)
        assert err <= 1e-10, (k, err)  # the taped step is the oracle's step
    bar = A.zero_cotangent(ctx)
    for k in GRADS:
        setattr(bar, k, lam[k])
    dG = adj.step(bar)
    rhs = sum(float(np.sum(getattr(bar, k) * dx[k])) for k in GRADS)
    print(f"<lam, J dx> = {lhs:.15e}, <J^T lam, dx> = {rhs:.15e}; dG {lhs_G:.15e} vs {dG:.15e}", flush=True)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
    assert abs(lhs_G - dG) <= 1e-9 * abs(lhs_G)
    adj.close()


def _objective_tangent(N, steps, dt):
    """dJ/d(G, r, cx, cy) from the oracle's tangent (four directions at once)."""
    case = _case(N, dt=dt)
    tg = T.Tangent(case.metrics, case.bcs, case.models, case.numerics)
    tg.set_shear_modulus_derivatives([1.0, 0.0, 0.0, 0.0])
    s = state_dict(case.state)
    d = setups.circle_parameter_derivatives(case.grid, P0[2:], P0[1], case.params["init_epsilon"])
    ds = {k: T.pack(s[k]) for k in T.FIELDS}
    ds["phi"] = T.pack(s["phi"], (None, *d))
    for k in range(steps):
        tg.integrate_step(ds, k * dt, dt)
    phi, dphi = T.unpack(ds["phi"])
    X, Y = np.meshgrid(case.grid.x_p, case.grid.y_p, indexing="ij")
    m = phi.sum()
    xc, yc = (X * phi).sum() / m, (Y * phi).sum() / m
    J = (xc - TARGET[0]) ** 2 + (yc - TARGET[1]) ** 2
    g = []
    for dp in dphi:
        dm = dp.sum()
        dxc = ((X * dp).sum() - xc * dm) / m
        dyc = ((Y * dp).sum() - yc * dm) / m
        g.append(2 * (xc - TARGET[0]) * dxc + 2 * (yc - TARGET[1]) * dyc)
    return J, np.array(g)


def _device_J(N, steps, dt, p):
    case = _case(N, p, dt=dt)
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    for k in range(steps):
        fsi.integrate_step(case.state, k * dt, dt, ctx, work)
    J, _ = A.centroid_objective(case.state, ctx, TARGET)
    return J


def test_run_gradient_matches_the_oracle_tangent_and_differences():
    N, steps, dt = 32, 20, 2e-3
    out = A.disk_gradient(_case(N, dt=dt), steps, TARGET, snapshots=4)
    J_t, g_t = _objective_tangent(N, steps, dt)
    print(f"J device {out['J']:.15e} oracle {J_t:.15e}", flush=True)
    print(f"gradient device {out['grad']}\n         oracle  {g_t}", flush=True)
    assert abs(out["J"] - J_t) <= 1e-10 * abs(J_t)
    assert np.linalg.norm(out["grad"] - g_t) <= 1e-8 * np.linalg.norm(g_t)
    h = 1e-6
    fd = []
    for k in range(4):
        pp, pm = list(P0), list(P0)
        pp[k] += h
        pm[k] -= h
        fd.append((_device_J(N, steps, dt, pp) - _device_J(N, steps, dt, pm)) / (2 * h))
    fd = np.array(fd)
    rel = np.abs(fd - out["grad"]) / np.abs(out["grad"])
    print(f"central differences {fd}\nrelative differences {rel}", flush=True)
    # 1e-6 relative, or the differences' rounding (a few ulps of J over h)
    # where a derivative is that small (d/dG after a few steps)
    assert np.all(np.abs(fd - out["grad"]) <= 1e-6 * np.abs(out["grad"]) + 1e-14 * abs(out["J"]) / h)
    st = out["stats"]
    assert st.steps == steps and st.adjoint_steps == steps and st.recordings == steps


def test_checkpoint_counts_give_the_same_gradient():
    N, steps, dt = 32, 12, 2e-3
    ref = None
    for snaps in (1, 3, steps):
        out = A.disk_gradient(_case(N, dt=dt), steps, TARGET, snapshots=snaps)
        st = out["stats"]
        expect = A.plan(steps, snaps)["advances"]
        print(f"snapshots {snaps}: gradient {out['grad']}, recomputed steps {st.advances} (plan {expect})", flush=True)
        assert st.advances == expect
        if ref is None:
            ref = out
        else:
            assert np.array_equal(out["grad"], ref["grad"])
            assert np.array_equal(out["bar"].phi, ref["bar"].phi)
    # every step's tape kept (no recomputation; the forward keeps its warm starts,
    # so the trajectory agrees to the solver tolerance, not bit for bit)
    out = A.disk_gradient(_case(N, dt=dt), steps, TARGET, snapshots=0)
    st = out["stats"]
    print(f"all tapes kept: gradient {out['grad']}, recomputed steps {st.advances}", flush=True)
    assert st.advances == 0 and st.recordings == steps and st.adjoint_steps == steps
    assert np.linalg.norm(out["grad"] - ref["grad"]) <= 1e-10 * np.linalg.norm(ref["grad"])


def test_unsupported_configurations_are_rejected():
    case = setups.kolmogorov_filament(32, velocity_method="direct")
    ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
    work = fsi.allocate_imex_workspace(ctx)
    with pytest.raises(fsi.UnsupportedError):
        A.StepAdjoint(ctx, work, snapshots=2)
