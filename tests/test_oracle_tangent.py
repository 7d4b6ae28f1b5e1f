"""The oracle's FP64 tangent (oracle/fsi_tangent.cpp) against the oracle itself.

The tangent build is fsi_oracle.c with a dual-number type for double, so its
values must be the oracle's bit for bit, and its derivatives must be those
of the oracle's step: checked here against central differences of the plain
oracle along the four parameters of the differentiation benchmark (shear
modulus, disk radius, disk centre) and along a random direction of the
state.  Tight solves (the exact predictor solve, projection 1e-13) keep the
differences' noise below the tolerance.
"""
import numpy as np
import pytest

from eulerianfsi_amd import setups
from oracle import tangent as T
from oracle.oracle import Oracle, state_dict

N, DT, STEPS = 24, 2e-3, 3
P0 = (0.1, 0.2, 0.6, 0.5)  # shear modulus, radius, centre x, centre y
CHECKED = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")


def _case(G, r, cx, cy):
    return setups.lid_driven_cavity(N, "acdi", radius=r, center=(cx, cy), shear_modulus=G, dt=DT,
                                    velocity_method="direct", projection_rtol=1e-13)


def _run(p, perturb=None, eps=0.0):
    case = _case(*p)
    o = Oracle(case.metrics, case.bcs, case.models, case.numerics)
    s = state_dict(case.state)
    if perturb is not None:
        for k, d in perturb.items():
            s[k] = np.asfortranarray(s[k] + eps * d)
    for k in range(STEPS):
        o.integrate_step(s, k * DT, DT)
    return s


def _tangent_run(seeds_phi=(), seeds_state=None, shear=None):
    case = _case(*P0)
    tg = T.Tangent(case.metrics, case.bcs, case.models, case.numerics)
    s0 = state_dict(case.state)
    ds = {k: T.pack(s0[k], (seeds_state or {}).get(k, ())) for k in T.FIELDS}
    if seeds_phi:
        ds["phi"] = T.pack(s0["phi"], seeds_phi)
    tg.set_shear_modulus_derivatives(shear)
    for k in range(STEPS):
        tg.integrate_step(ds, k * DT, DT)
    return ds


def _close(fd, tg, rtol):
    err = np.linalg.norm(fd - tg)
    assert err <= rtol * np.linalg.norm(tg) + 1e-9 * np.sqrt(tg.size), (err, np.linalg.norm(tg))


def test_tangent_values_are_the_oracles():
    case = _case(*P0)
    d = setups.circle_parameter_derivatives(case.grid, P0[2:], P0[1], case.params["init_epsilon"])
    ds = _tangent_run(seeds_phi=(None, *d), shear=[1.0, 0, 0, 0])
    ref = _run(P0)
    for f in CHECKED:
        v, _ = T.unpack(ds[f])
        assert np.array_equal(v, ref[f]), f


def test_tangent_matches_central_differences_of_the_parameters():
    case = _case(*P0)
    d = setups.circle_parameter_derivatives(case.grid, P0[2:], P0[1], case.params["init_epsilon"])
    ds = _tangent_run(seeds_phi=(None, *d), shear=[1.0, 0, 0, 0])
    h = 1e-6
    for k in range(4):
        pp, pm = list(P0), list(P0)
        pp[k] += h
        pm[k] -= h
        sp, sm = _run(pp), _run(pm)
        for f in CHECKED:
            _, der = T.unpack(ds[f])
            _close((sp[f] - sm[f]) / (2 * h), der[k], 1e-5)


def test_tangent_matches_central_differences_of_the_state():
    rng = np.random.default_rng(3)
    case = _case(*P0)
    s0 = state_dict(case.state)
    # a smooth perturbation of phi and sigma (velocity must stay solenoidal)
    X, Y = np.meshgrid(case.grid.x_p, case.grid.y_p, indexing="ij")
    bump = np.asfortranarray(np.exp(-((X - 0.6) ** 2 + (Y - 0.5) ** 2) / 0.02))
    pert = {"phi": bump * rng.uniform(0.5, 1.0), "sigma_xx": bump * 0.3, "sigma_xy": bump * 0.2}
    seeds = {k: (v,) for k, v in pert.items()}
    ds = _tangent_run(seeds_state=seeds)
    h = 1e-6
    sp, sm = _run(P0, pert, h), _run(P0, pert, -h)
    for f in CHECKED:
        _, der = T.unpack(ds[f])
        _close((sp[f] - sm[f]) / (2 * h), der[0], 1e-5)
    assert np.array_equal(T.unpack(ds["phi"])[0], _run(P0)["phi"])
    assert s0["phi"].shape == (N, N)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_circle_parameter_derivatives_match_differences(k):
    case = _case(*P0)
    grid, eps = case.grid, case.params["init_epsilon"]
    d = setups.circle_parameter_derivatives(grid, P0[2:], P0[1], eps)[k]
    h = 1e-7
    p = [P0[1], P0[2], P0[3]]
    pp, pm = list(p), list(p)
    pp[k] += h
    pm[k] -= h
    fd = (setups.initialize_circle(grid, pp[1:], pp[0], eps) - setups.initialize_circle(grid, pm[1:], pm[0], eps)) / (2 * h)
    assert np.abs(fd - d).max() <= 1e-6 * np.abs(d).max()
