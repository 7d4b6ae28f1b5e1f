"""Boundary conditions, mirroring src/BoundaryConditions.jl.

Only scalar data are carried to the device.  MixedBoundary and
AdvectiveOutflow exist in the reference (:33-46) but no configuration of
this path uses them; they are rejected with UnsupportedBoundary.
"""
from __future__ import annotations

from dataclasses import dataclass
from numbers import Real

from .errors import ArgumentError, UnsupportedError


class UnsupportedBoundary(UnsupportedError):
    pass


class AbstractBoundaryCondition:
    pass


@dataclass(frozen=True)
class Periodic(AbstractBoundaryCondition):
    """Periodic boundary (BoundaryConditions.jl:26)."""


@dataclass(frozen=True)
class Dirichlet(AbstractBoundaryCondition):
    """Prescribed value (BoundaryConditions.jl:29-31)."""
    value: float


@dataclass(frozen=True)
class Neumann(AbstractBoundaryCondition):
    """Prescribed outward normal derivative (BoundaryConditions.jl:33-35)."""
    value: float


def scalar_value(bc) -> float:
    v = bc.value
    if not isinstance(v, Real):
        raise UnsupportedBoundary("vector- or function-valued boundary data are outside the device path")
    return float(v)


class BoundaryConditionSetup:
    """Four-sided boundary specification (BoundaryConditions.jl:51-72)."""

    def __init__(self, left, right, bottom, top):
        _validate_periodic_pair(left, right, "x")
        _validate_periodic_pair(bottom, top, "y")
        self.left, self.right, self.bottom, self.top = left, right, bottom, top

    def __repr__(self):
        return f"BoundaryConditionSetup({self.left!r}, {self.right!r}, {self.bottom!r}, {self.top!r})"


def _validate_periodic_pair(first, second, direction):
    """BoundaryConditions.jl:74-84"""
    if isinstance(first, Periodic) != isinstance(second, Periodic):
        raise ArgumentError(f"Periodic boundaries in the {direction}-direction must be specified as an opposing pair.")


def is_periodic_x(bc: BoundaryConditionSetup) -> bool:
    return isinstance(bc.left, Periodic)


def is_periodic_y(bc: BoundaryConditionSetup) -> bool:
    return isinstance(bc.bottom, Periodic)


def has_dirichlet(bc: BoundaryConditionSetup) -> bool:
    return any(isinstance(s, Dirichlet) for s in (bc.left, bc.right, bc.bottom, bc.top))
