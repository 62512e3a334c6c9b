from .interface import BaseStepper
from .implementations import EvolveStateVector
from .torch import DifferentiableStateVectorEvolution

__all__ = ["BaseStepper", "DifferentiableStateVectorEvolution", "EvolveStateVector"]
