"""The OpenMP build of the oracle (the all-cores CPU baseline of bench.py)
computes the same steps as the serial build: its dot products sum per thread
and its multigrid smoother runs in red-black order, so the two agree to the
solver tolerances (velocity 1e-8 of the initial residual, projection 1e-10),
not bitwise."""
import numpy as np
import pytest

from eulerianfsi_amd import setups
from oracle import oracle as O

FIELDS = ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")


def _run(case, lib_path, nsteps):
    o = O.Oracle(case.metrics, case.bcs, case.models, case.numerics, lib_path=lib_path)
    s = O.state_dict(case.state)
    o.make_velocity_compatible(s)
    for k in range(nsteps):
        o.integrate_step(s, k * case.dt, case.dt)
    return s


@pytest.mark.parametrize("method", ["acdi", "uwcdi"])
def test_openmp_build_matches_serial_build(method):
    O.build_oracle()
    if not O.OMP_LIB_PATH.exists():
        pytest.skip("the OpenMP build of the oracle is not built")
    # 128 x 128 reaches the threaded smoother, restriction and reductions
    # (levels of at least 128 x 128 cells run in parallel)
    case = setups.seeded_cavity(128, method, seed=5)
    a = _run(case, None, 3)
    b = _run(case, O.OMP_LIB_PATH, 3)
    for f in FIELDS:
        scale = np.abs(a[f]).max()
        err = np.abs(a[f] - b[f]).max() / scale
        print(f"{method} {f}: max relative difference {err:.2e}", flush=True)
        assert err <= 1e-9, f
