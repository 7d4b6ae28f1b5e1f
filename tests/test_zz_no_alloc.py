"""fsi_phase_centroid_noalloc (the allocation-free centroid, no counterpart
in the reference): through the C ABI it gives the same bits as
fsi_phase_centroid on two calls in a row, the second on a smaller grid,
writing into caller-owned device memory on a caller-owned stream; grids
beyond its stated maximum are rejected."""
import ctypes
import os

import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import setups

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime libfsi_hip.so runs on (for the caller-side device
    buffer and stream of the test)."""
    lib = ctypes.CDLL("libamdhip64.so.7")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    lib.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    lib.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    lib.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return lib


def _centroid_noalloc(ctx, state, hip, buf, stream):
    h = state._device(ctx)
    st = ctx.lib.fsi_phase_centroid_noalloc(ctx.handle, h, buf, stream)
    if st != 0:
        raise fsi.ArgumentError(ctx.lib.fsi_last_error(ctx.handle).decode())
    assert hip.hipStreamSynchronize(stream) == 0
    out = np.zeros(3)
    assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), buf, 24, 2) == 0  # device to host
    return out


def _centroid(ctx, state):
    x, y, m = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    st = ctx.lib.fsi_phase_centroid(ctx.handle, state._device(ctx), ctypes.byref(x), ctypes.byref(y),
                                    ctypes.byref(m))
    assert st == 0
    return np.array([x.value, y.value, m.value])


def test_noalloc_centroid_equals_centroid():
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("device only")
    hip = _hip()
    buf, stream = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(buf), 24) == 0
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for N, seed in ((512, 3), (64, 4)):  # the second call on a smaller grid
            case = setups.seeded_cavity(N, "acdi", seed=seed)
            ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
            work = fsi.allocate_imex_workspace(ctx)
            fsi.make_velocity_compatible(case.state, ctx, work)
            for k in range(2):
                fsi.integrate_step(case.state, k * case.dt, case.dt, ctx, work)
            assert hip.hipMemset(buf, 0xFF, 24) == 0  # NaN: the result must be written by the call
            got = _centroid_noalloc(ctx, case.state, hip, buf, stream)
            ref = _centroid(ctx, case.state)
            print(N, got, ref, flush=True)
            assert np.array_equal(got, ref), (N, got, ref)
            # the context's own stream (stream = NULL)
            got0 = _centroid_noalloc(ctx, case.state, hip, buf, None)
            assert np.array_equal(got0, ref)
    finally:
        hip.hipStreamDestroy(stream)
        hip.hipFree(buf)


def test_noalloc_centroid_without_mass_is_nan():
    """Where fsi_phase_centroid raises (phase integral not above eps), the
    allocation-free variant writes NaN for the centroid (and the integral
    it found)."""
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("device only")
    hip = _hip()
    buf = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(buf), 24) == 0
    try:
        case = setups.seeded_cavity(64, "acdi", seed=1)
        case.state.phi = np.zeros((64, 64))
        ctx = fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)
        with pytest.raises(fsi.SolverError):
            fsi.diagnostics.phase_centroid(case.state, ctx)
        got = _centroid_noalloc(ctx, case.state, hip, buf, None)
        assert np.isnan(got[0]) and np.isnan(got[1]) and got[2] == 0.0
    finally:
        hip.hipFree(buf)


def test_noalloc_centroid_rejects_grids_beyond_the_scratch():
    if os.environ.get("FSI_TEST_DRYRUN") == "1":
        pytest.skip("device only")
    grid = fsi.generate_grid(1.0, 1.0, 8192, 2064)  # 128 x 129 tiles > 16384
    metrics = fsi.build_mesh_metrics(grid)
    w, n0 = fsi.Dirichlet(0.0), fsi.Neumann(0.0)
    wall = fsi.BoundaryConditionSetup(w, w, w, w)
    nf = fsi.BoundaryConditionSetup(n0, n0, n0, n0)
    bcs = fsi.FSIBoundaryConditions(wall, wall, nf, nf, {"xx": nf, "yy": nf, "xy": nf})
    ctx = fsi.build_imex_context(metrics, bcs, fsi.FSIModels(), fsi.IMEXNumerics())
    one = np.ones((8192, 2064))
    state = fsi.FSIState(np.zeros((8193, 2064)), np.zeros((8192, 2065)), 0.5 * one, one, one, 0 * one)
    hip = _hip()
    buf = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(buf), 24) == 0
    try:
        st = ctx.lib.fsi_phase_centroid_noalloc(ctx.handle, state._device(ctx), buf, None)
        assert st == 1  # FSI_E_ARGUMENT, before anything is launched
        assert "at most 16384 tiles" in ctx.lib.fsi_last_error(ctx.handle).decode()
    finally:
        hip.hipFree(buf)
