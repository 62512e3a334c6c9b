from .interface import BaseStepper
from .implementations import EvolveDensityMatrix, EvolveMonteCarlo, EvolveStateVector
from .torch import DifferentiableStateVectorEvolution

__all__ = ["BaseStepper", "DifferentiableStateVectorEvolution", "EvolveDensityMatrix",
           "EvolveMonteCarlo", "EvolveStateVector"]
