from .interface import BaseStepper
from .implementations import EvolveMonteCarlo, EvolveStateVector
from .torch import DifferentiableStateVectorEvolution

__all__ = ["BaseStepper", "DifferentiableStateVectorEvolution", "EvolveMonteCarlo",
           "EvolveStateVector"]
