"""CPU tests of the C-ABI boundary and the host-side mirror of the API.
No device compute is called here."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import eulerianfsi_amd as fsi
from eulerianfsi_amd import _abi, setups
from eulerianfsi_amd import closures as C

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "fsi_api.h").read_text()
    return sorted(set(re.findall(r"\b(fsi_[a-z_]+)\s*\(", text)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_abi.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 16
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_abi.EXPORTED)


def test_version_string():
    lib = _abi.load_library()
    assert b"gfx950" in lib.fsi_version()


def test_struct_layout_matches_header():
    # fsi_bc_side = {int32, double}: 16 bytes; 8 setups of 4 sides
    assert ctypes.sizeof(_abi.BcSide) == 16
    assert ctypes.sizeof(_abi.Bcs) == 8 * 4 * 16
    assert ctypes.sizeof(_abi.StepReportC) == 8 * 8
    assert _abi.StepReportC.non_spd_cells.offset == 56


def test_context_without_device_is_a_device_error():
    case = setups.lid_driven_cavity(16, "acdi")
    with pytest.raises(fsi.DeviceError):
        fsi.build_imex_context(case.metrics, case.bcs, case.models, case.numerics)


def test_abi_config_of_cavity():
    case = setups.lid_driven_cavity(64, "acdi")
    g, b, m, n = _abi.abi_config(case.metrics, case.bcs, case.models, case.numerics)
    assert (g.nx, g.ny, g.periodic_x, g.periodic_y) == (64, 64, 0, 0)
    assert b.U.top.kind == _abi.BC_DIRICHLET and b.U.top.value == 1.0
    assert b.phi.left.kind == _abi.BC_NEUMANN
    assert m.phase_method == _abi.PHASE_ACDI and m.acdi_epsilon == pytest.approx(1.5 / 64)
    assert m.acdi_dynamic_gamma == 0 and m.viscosity_kind == _abi.VISC_CONSTANT and m.viscosity == 0.01
    assert m.elastic_kind == _abi.ELASTIC_LINEAR and m.shear_modulus == 0.1
    assert n.extension_kind == _abi.EXT_GEOMETRIC and n.ext_max_rate == 20.0 and n.ext_cutoff == 1e-6
    assert n.velocity_rtol == 1e-8 and n.projection_rtol == 1e-10


def test_abi_config_of_filament():
    case = setups.kolmogorov_filament(32, "uwcdi")
    g, b, m, n = _abi.abi_config(case.metrics, case.bcs, case.models, case.numerics)
    assert g.periodic_x == 1 and g.periodic_y == 0
    assert b.U.left.kind == _abi.BC_PERIODIC and b.sigma_xx.bottom.kind == _abi.BC_DIRICHLET
    assert m.body_force_kind == _abi.FORCE_KOLMOGOROV and m.force_wavenumber == pytest.approx(4 * np.pi)
    assert m.phase_method == _abi.PHASE_UWCDI


def test_unsupported_closures_are_rejected():
    case = setups.lid_driven_cavity(16, "uwcdi")

    class SomeGrowth:
        pass

    for kw in (dict(growth_source=SomeGrowth()), dict(capillarity=SomeGrowth())):
        models = fsi.FSIModels(phase_field_method=C.PirozzoliUWCDIPhaseField(1.0), **kw)
        with pytest.raises(fsi.UnsupportedError):
            _abi.abi_config(case.metrics, case.bcs, models, case.numerics)


def test_uwcdi_normals_and_capillarity_reach_the_abi():
    case = setups.lid_driven_cavity(16, "uwcdi")
    for name, code in (("standard", 0), ("youngs", 1), ("smoothed9", 2)):
        models = fsi.FSIModels(phase_field_method=C.PirozzoliUWCDIPhaseField(1.0, normal_reconstruction=name))
        m = _abi.abi_config(case.metrics, case.bcs, models, case.numerics)[2]
        assert m.uwcdi_normal_method == code and m.capillarity_kind == 0
    models = fsi.FSIModels(capillarity=C.DiffuseAreaCapillarity(0.3, delta_normal=1e-5, smoothing_passes=2))
    m = _abi.abi_config(case.metrics, case.bcs, models, case.numerics)[2]
    assert (m.capillarity_kind, m.capillary_coefficient, m.capillary_delta_normal,
            m.capillary_smoothing_passes) == (1, 0.3, 1e-5, 2)
    for bad in (dict(coefficient=-1.0), dict(coefficient=float("nan")), dict(coefficient=1.0, delta_normal=0.0),
                dict(coefficient=1.0, smoothing_passes=-1)):
        with pytest.raises(fsi.ArgumentError):
            C.DiffuseAreaCapillarity(**bad)


def test_constructor_validation_mirrors_reference():
    with pytest.raises(fsi.ArgumentError):
        C.PirozzoliUWCDIPhaseField(0.9)
    with pytest.raises(fsi.ArgumentError):
        C.PirozzoliUWCDIPhaseField(1.0, normal_reconstruction="invalid")
    with pytest.raises(fsi.ArgumentError):
        C.JainACDIPhaseField(-1.0, 0.01)
    with pytest.raises(fsi.ArgumentError):
        C.GeometricStructureExtension(-1.0)
    with pytest.raises(fsi.ArgumentError):
        C.GeometricStructureExtension(1.0, 0.0)
    with pytest.raises(fsi.ArgumentError):
        C.GeometricStructureExtension(1.0, 4.0, safety_cutoff=1.0)
    with pytest.raises(fsi.ArgumentError):
        fsi.BoundaryConditionSetup(fsi.Periodic(), fsi.Neumann(0.0), fsi.Neumann(0.0), fsi.Neumann(0.0))
    with pytest.raises(fsi.ArgumentError):
        fsi.FSIModels(phase_field_method=C.PirozzoliUWCDIPhaseField(), interface_flux=C.ConservativeDiffuseInterfaceFlux(1, 1))
    assert C.jain_equilibrium_initialization_epsilon(C.JainACDIPhaseField(0.2, 0.51 / 16)) == pytest.approx(1.02 / 16)
    assert C.phase_field_label(C.CurrentCDIPhaseField(4.0, 0.1)) == "current_cdi"


def test_periodicity_mismatch_is_an_argument_error():
    case = setups.lid_driven_cavity(16, "acdi")
    metrics = fsi.build_mesh_metrics(case.grid, periodic_x=True)
    with pytest.raises(fsi.ArgumentError):
        fsi.build_imex_context(metrics, case.bcs, case.models, case.numerics)


def test_state_constructors_and_shapes():
    st = fsi.FSIState(np.zeros((5, 4)), np.zeros((4, 5)), np.full((4, 4), 0.5), np.ones((4, 4)), np.ones((4, 4)),
                      np.zeros((4, 4)))
    assert np.all(st.resource == 0.0) and st.U.flags["F_CONTIGUOUS"]
    st.phi = np.full((4, 4), 0.25)
    assert np.all(st.phi == 0.25)


def test_splitmix_is_deterministic():
    a, b = setups.SplitMix64(7), setups.SplitMix64(7)
    xs = [a.uniform() for _ in range(5)]
    assert xs == [b.uniform() for _ in range(5)]
    assert all(0.0 <= x < 1.0 for x in xs)
    # first output of SplitMix64(0) (published reference value)
    assert setups.SplitMix64(0).next_u64() == 0xE220A8397B1DCDAF
