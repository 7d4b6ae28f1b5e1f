"""Writes tests/golden/phase_rhs.npz: seeded inputs (U, V, phi, S) and the
phase right-hand side of each (configuration, phase model) pair below,
evaluated by the NumPy restatement in tests/reference_phase.py; the
DiffuseAreaCapillarity force of seeded phase fields (capillary/...) and
uwcdi_strang_step of seeded states (strang/...) from the same restatement.

The configurations cover every boundary branch of the reference's phase
operators (src/InterfaceTransport.jl, src/InterfaceRegularization.jl,
src/PhaseFieldMethods.jl):
  cavity     closed box, homogeneous Neumann phi, no normal wall velocity
  xperiodic  periodic in x; Neumann y walls with OUTFLOW normal velocity
             (the boundary value is the interior state)
  channel    inflow through a Dirichlet phi = 0.8 left wall, Neumann
             outflow on the right, periodic in y
  dirichlet  Dirichlet phi on all four walls, normal wall velocity of both
             signs (inflow and outflow through every wall)
  periodic   doubly periodic
Anisotropic cells (hx != hy) exercise the UWCDI alpha correction.  A few phi
values sit at or outside [0, 1] to reach the ACDI psi clamp.

Usage: python -m tests.golden.make_phase_golden   (from the repository root)
"""
import os

import numpy as np

import eulerianfsi_amd as fsi
from eulerianfsi_amd import closures as C

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "phase_rhs.npz")

CONFIGS = {
    "cavity": dict(Lx=1.0, Ly=0.6, nx=12, ny=10, px=False, py=False, phi=("N", "N", "N", "N")),
    "xperiodic": dict(Lx=1.0, Ly=0.75, nx=12, ny=10, px=True, py=False, phi=("P", "P", "N", "N")),
    "channel": dict(Lx=0.8, Ly=1.0, nx=10, ny=12, px=False, py=True, phi=(0.8, "N", "P", "P")),
    "dirichlet": dict(Lx=0.9, Ly=1.0, nx=12, ny=11, px=False, py=False, phi=(0.8, 0.1, 0.3, 0.95)),
    "periodic": dict(Lx=1.0, Ly=1.0, nx=12, ny=12, px=True, py=True, phi=("P", "P", "P", "P")),
}


def models():
    """Phase closures, keyed by name (fresh objects on every call)."""
    return {
        "upwind": C.LegacySplitPhaseFieldMethod(C.NoInterfaceFlux()),
        "legacy_cdi": C.LegacySplitPhaseFieldMethod(C.ConservativeDiffuseInterfaceFlux(2.0, 0.15, 1e-6)),
        "cdi": C.CurrentCDIPhaseField(4.0, 0.12, 1e-6),
        "acdi_fixed": C.JainACDIPhaseField(0.7, 0.06, dynamic_gamma=False),
        "acdi_dynamic": C.JainACDIPhaseField(1.3, 0.06, psi_floor=1e-12, delta_normal=1e-10),
        "uwcdi": C.PirozzoliUWCDIPhaseField(1.0),
        "uwcdi_alpha": C.PirozzoliUWCDIPhaseField(1.6, delta_normal=1e-8),
        "uwcdi_youngs": C.PirozzoliUWCDIPhaseField(1.0, normal_reconstruction="youngs"),
        "uwcdi_smoothed9": C.PirozzoliUWCDIPhaseField(1.3, delta_normal=1e-8, normal_reconstruction="smoothed9"),
    }


def model_applies(cname, mname):
    """:smoothed9 accepts homogeneous Neumann or periodic phase sides only."""
    return mname != "uwcdi_smoothed9" or all(s in ("N", "P") for s in CONFIGS[cname]["phi"])


def _side(s):
    if s == "P":
        return fsi.Periodic()
    if s == "N":
        return fsi.Neumann(0.0)
    return fsi.Dirichlet(float(s))


# normal wall velocities as Dirichlet data, for runs that take the wall
# velocity from the boundary conditions (the device stores interior faces only)
WALLS = {
    "cavity": dict(left=0.0, right=0.0, bottom=0.0, top=0.0),
    "xperiodic": dict(bottom=-0.5, top=0.6),
    "channel": dict(left=0.6, right=0.5),
    "dirichlet": dict(left=0.5, right=-0.4, bottom=0.3, top=0.2),
    "periodic": dict(),
}


def configuration(name, model, n=None, walls=False):
    """(metrics, bcs, models, numerics) of one configuration; `n` overrides
    the cell counts (n x n), `walls` prescribes the WALLS normal velocities."""
    c = CONFIGS[name]
    nx, ny = (n, n) if n else (c["nx"], c["ny"])
    grid = fsi.generate_grid(c["Lx"], c["Ly"], nx, ny)
    metrics = fsi.build_mesh_metrics(grid, periodic_x=c["px"], periodic_y=c["py"])
    phi_bc = fsi.BoundaryConditionSetup(*(_side(s) for s in c["phi"]))
    w = WALLS[name] if walls else {}
    wall = lambda k: fsi.Dirichlet(w.get(k, 0.0))
    xs = (fsi.Periodic(),) * 2 if c["px"] else (wall("left"), wall("right"))
    ys = (fsi.Periodic(),) * 2 if c["py"] else (wall("bottom"), wall("top"))
    xt = (fsi.Periodic(),) * 2 if c["px"] else (fsi.Dirichlet(0.0),) * 2
    yt = (fsi.Periodic(),) * 2 if c["py"] else (fsi.Dirichlet(0.0),) * 2
    bcU = fsi.BoundaryConditionSetup(*xs, *yt)
    bcV = fsi.BoundaryConditionSetup(*xt, *ys)
    xn = fsi.Periodic() if c["px"] else fsi.Neumann(0.0)
    yn = fsi.Periodic() if c["py"] else fsi.Neumann(0.0)
    sc = fsi.BoundaryConditionSetup(xn, xn, yn, yn)
    sigma = {"xx": sc, "yy": sc, "xy": sc}
    if walls:  # sigma = I enters through the inflow walls
        for k, v in (("xx", 1.0), ("yy", 1.0), ("xy", 0.0)):
            xd = fsi.Periodic() if c["px"] else fsi.Dirichlet(v)
            yd = fsi.Periodic() if c["py"] else fsi.Dirichlet(v)
            sigma[k] = fsi.BoundaryConditionSetup(xd, xd, yd, yd)
    bcs = fsi.FSIBoundaryConditions(bcU, bcV, sc, phi_bc, sigma)
    return metrics, bcs, fsi.FSIModels(phase_field_method=model), fsi.IMEXNumerics()


def inputs(name, seed, n=None, walls=False):
    """Seeded U (nx+1, ny), V (nx, ny+1), phi, S respecting the boundary
    rules of the configuration (with `walls`, the wall faces carry the WALLS
    velocities)."""
    c = CONFIGS[name]
    nx, ny = (n, n) if n else (c["nx"], c["ny"])
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    phi = 0.5 + 0.4 * np.tanh(np.subtract.outer(np.cos(2 * np.pi * x), np.sin(2 * np.pi * y)) * 2.0)
    phi = np.clip(phi + 0.05 * rng.standard_normal((nx, ny)), 0.02, 0.98)
    phi[2, 3], phi[5, 5], phi[7, 2] = -0.01, 1.0, 1.03
    U = rng.uniform(-1.0, 1.0, (nx + 1, ny))
    V = rng.uniform(-1.0, 1.0, (nx, ny + 1))
    if c["px"]:
        U[-1] = U[0]
    elif name == "channel":
        U[0] = rng.uniform(0.2, 1.0, ny)    # inflow through the Dirichlet wall
        U[-1] = rng.uniform(0.2, 1.0, ny)   # outflow through the Neumann wall
    elif name == "dirichlet":
        pass                                # both signs on both x walls
    else:
        U[0] = U[-1] = 0.0
    if c["py"]:
        V[:, -1] = V[:, 0]
    elif name == "xperiodic":
        V[:, 0] = -rng.uniform(0.2, 1.0, nx)   # outflow at the bottom
        V[:, -1] = rng.uniform(0.2, 1.0, nx)   # outflow at the top
    elif name == "dirichlet":
        pass
    else:
        V[:, 0] = V[:, -1] = 0.0
    if walls and not c["px"]:
        U[0], U[-1] = WALLS[name]["left"], WALLS[name]["right"]
    if walls and not c["py"]:
        V[:, 0], V[:, -1] = WALLS[name]["bottom"], WALLS[name]["top"]
    S = 0.3 * rng.standard_normal((nx, ny)) if seed % 2 else np.zeros((nx, ny))
    return U, V, phi, S


def cases():
    """(key, configuration name, model name, seed) of every fixture."""
    out = []
    for ci, cname in enumerate(CONFIGS):
        for mi, mname in enumerate(models()):
            if model_applies(cname, mname):
                out.append((f"{cname}/{mname}", cname, mname, 100 * ci + mi))
    return out


# DiffuseAreaCapillarity(coefficient; delta_normal, smoothing_passes) of the capillary fixtures
CAPILLARITY = {"plain": (0.7, 1e-6, 0), "filtered": (0.45, 1e-5, 2)}
STRANG_DT = 0.01


def capillary_cases():
    """(key, configuration, capillarity name, seed): the force of the seeded phi."""
    return [(f"capillary/{cname}/{k}", cname, k, 1000 + 10 * ci + n)
            for ci, cname in enumerate(CONFIGS) for n, k in enumerate(CAPILLARITY)]


def strang_cases():
    """(key, configuration, UWCDI model name, first direction, seed)."""
    out = []
    for ci, cname in enumerate(CONFIGS):
        for mname in ("uwcdi", "uwcdi_youngs", "uwcdi_smoothed9"):
            if model_applies(cname, mname):
                for d in ("x", "y"):
                    out.append((f"strang/{cname}/{mname}/{d}", cname, mname, d, 2000 + 10 * ci + len(out) % 10))
    return out


def _geometry(cname, mname):
    from tests import reference_assembly as RA
    from tests import reference_phase as RP
    from oracle.oracle import Oracle
    metrics, bcs, mdl, num = configuration(cname, models()[mname])
    met = RA.Met(Oracle(metrics, bcs, mdl, num))
    return RP, RP.PhaseGeometry(met, bcs.phi, metrics.periodic_x, metrics.periodic_y), mdl


def restated_rhs(cname, mname, U, V, phi, S):
    RP, g, mdl = _geometry(cname, mname)
    return RP.phase_rhs(g, mdl.phase_field_method, U, V, phi, S)


def restated_capillary(cname, kname, phi):
    RP, g, _ = _geometry(cname, "uwcdi")
    coef, delta, passes = CAPILLARITY[kname]
    return RP.capillary_force(g, phi, coef, delta, passes)


def restated_strang(cname, mname, d, U, V, phi):
    RP, g, mdl = _geometry(cname, mname)
    return RP.uwcdi_strang_step(g, mdl.phase_field_method, U, V, phi, STRANG_DT, d)


def main():
    data = {}
    for key, cname, mname, seed in cases():
        U, V, phi, S = inputs(cname, seed)
        rhs = restated_rhs(cname, mname, U, V, phi, S)
        for f, a in (("U", U), ("V", V), ("phi", phi), ("S", S), ("rhs", rhs)):
            data[f"{key}/{f}"] = a
        print(key, f"max|rhs| = {np.max(np.abs(rhs)):.4e}")
    for key, cname, kname, seed in capillary_cases():
        phi = inputs(cname, seed)[2]
        fx, fy = restated_capillary(cname, kname, phi)
        for f, a in (("phi", phi), ("fx", fx), ("fy", fy)):
            data[f"{key}/{f}"] = a
        print(key, f"max|f| = {max(np.max(np.abs(fx)), np.max(np.abs(fy))):.4e}")
    for key, cname, mname, d, seed in strang_cases():
        U, V, phi, _ = inputs(cname, seed)
        out = restated_strang(cname, mname, d, U, V, phi)
        for f, a in (("U", U), ("V", V), ("phi", phi), ("out", out)):
            data[f"{key}/{f}"] = a
        print(key, f"max|dphi| = {np.max(np.abs(out - phi)):.4e}")
    np.savez_compressed(PATH, **data)
    print("wrote", PATH)


if __name__ == "__main__":
    main()
