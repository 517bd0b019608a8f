"""MI355X-native bond-based peridynamics time stepping.

A from-scratch HIP (gfx950) implementation of PeriPy's time-stepping path --
per-node bond-force gather over the neighbour list with in-place bond
breaking, n-linear damage, damage, and the Euler / Euler-Cromer /
Velocity-Verlet updates -- behind PeriPy's ``Model`` / ``Integrator`` API.
"""
from .integrators import (Integrator, Euler, EulerHIP, EulerCromerHIP,
                          VelocityVerletHIP, EulerCL, EulerCromerCL,
                          VelocityVerletCL, HIPContext, ContextError,
                          get_context)
from .model import (Model, initial_crack_helper, vectorised,
                    DimensionalityError, FamilyError, DamageModelError,
                    InvalidIntegrator)
from .crack import LineCrack
from .predicates import Boxes, PlaneBondTypes

__all__ = [
    "Model", "initial_crack_helper", "Integrator", "Euler", "EulerHIP",
    "EulerCromerHIP", "VelocityVerletHIP", "EulerCL", "EulerCromerCL",
    "VelocityVerletCL", "HIPContext", "ContextError", "get_context",
    "DimensionalityError", "FamilyError", "DamageModelError",
    "InvalidIntegrator", "LineCrack", "vectorised", "Boxes",
    "PlaneBondTypes",
]
__version__ = "0.1.0"
