"""eulerianfsi_amd -- MI355X (gfx950) backend of EulerianFSI.jl's IMEX step.

The device work happens in libfsi_hip.so (hand-written HIP kernels behind the
C ABI declared in include/fsi_api.h).  This package is the Python host side:
it mirrors the reference's step API (FSIState, FSIBoundaryConditions,
FSIModels, IMEXNumerics, build_imex_context, allocate_imex_workspace,
make_velocity_compatible!, integrate_step!) and calls the library through
ctypes.  There is no CPU fallback: without the built library the API raises.
"""
from .boundary import (AbstractBoundaryCondition, BoundaryConditionSetup, Dirichlet, Neumann, Periodic,
                       has_dirichlet, is_periodic_x, is_periodic_y)
from .closures import (ConservativeDiffuseInterfaceFlux, ConstantBodyForce, ConstantViscosity, CurrentCDIPhaseField,
                       GeometricStructureExtension, IdentityOutsideStructure, IsotropicGrowthTensor,
                       JainACDIPhaseField, KolmogorovBodyForce, LegacySplitPhaseFieldMethod,
                       LinearInternalTensorStress, MaxwellRelaxation, NoBodyForce, NoCapillarity, DiffuseAreaCapillarity, NoDiffusion,
                       NoElasticStress, NoGrowth, NoInterfaceFlux, NoPhaseFieldRegularization, NoRelaxation,
                       NoResourceConsumption, NoStructureExtension, NoViscosity, PirozzoliUWCDIPhaseField,
                       ConstantIntrinsicGrowth, SaturatingResourceGrowth, DirectionalGrowthTensor,
                       ConstantDiffusivity, HarmonicPhaseDiffusivity, SaturatingResourceConsumption,
                       jain_equilibrium_initialization_epsilon, phase_field_label)
from .errors import ArgumentError, DeviceError, DimensionMismatch, DomainError, SolverError, UnsupportedError
from .imex import (A_EXPLICIT, A_IMPLICIT, STAGE_ABSCISSA, FourStageIMEXFractionalStep, FSIBoundaryConditions,
                   FSIModels, FSIState, IMEXContext, IMEXNumerics, IMEXWorkspace, LinearSolverOptions,
                   ProjectionResult, StepReport, allocate_imex_workspace, build_imex_context,
                   ensemble_step, evaluate_stage_tendencies, integrate_step, make_velocity_compatible, project,
                   uwcdi_strang_step)
from .mesh import MACGrid, MeshMetrics, build_mesh_metrics, generate_grid
from . import adjoint, diagnostics

__all__ = [n for n in dir() if not n.startswith("_")]
