"""State-vector backend: drop-in for the reference's emu_sv on MI355X."""

from .dense_state import DenseState
from .density_matrix_state import DensityMatrix
from .matrix_operator import DenseOperator, SparseOperator
from .hamiltonian import RydbergHamiltonian
from .lindblad_operator import RydbergLindbladian
from .observables import (
    BitStrings,
    CorrelationMatrix,
    Energy,
    EnergySecondMoment,
    EnergyVariance,
    Expectation,
    Fidelity,
    Observable,
    Occupation,
    Results,
    StateResult,
)
from .state_vector import StateVector, inner
from .sv_backend import SVBackend
from .sv_config import Solver, SVConfig

__all__ = [
    "BitStrings",
    "CorrelationMatrix",
    "DenseOperator",
    "DenseState",
    "DensityMatrix",
    "Energy",
    "EnergySecondMoment",
    "EnergyVariance",
    "Expectation",
    "Fidelity",
    "Observable",
    "Occupation",
    "Results",
    "RydbergHamiltonian",
    "RydbergLindbladian",
    "SVBackend",
    "SVConfig",
    "SparseOperator",
    "Solver",
    "StateResult",
    "StateVector",
    "inner",
]

__version__ = "0.1.0"
