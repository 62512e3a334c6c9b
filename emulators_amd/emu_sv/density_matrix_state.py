"""Density-matrix states (reference emu_sv/density_matrix_state.py:13-136)."""

from __future__ import annotations

from typing import Mapping, Sequence

import torch

from emulators_amd.emu_sv.dense_state import DenseState
from emulators_amd.emu_sv.state_vector import StateVector

dtype = torch.complex128


class DensityMatrix(DenseState):
    """An n-qubit density matrix (2^n x 2^n complex128) in the (g, r) qubit
    basis; Hermitian, positive semidefinite and of unit trace by contract
    (not checked, as in the reference)."""

    @classmethod
    def make(cls, n_atoms: int, gpu: bool = True) -> "DensityMatrix":
        """|g...g><g...g|."""
        result = torch.zeros(2**n_atoms, 2**n_atoms, dtype=dtype)
        result[0, 0] = 1.0
        return cls(result, gpu=gpu)

    def __add__(self, other) -> "DensityMatrix":
        raise NotImplementedError("Not implemented")

    def __rmul__(self, scalar: complex) -> "DensityMatrix":
        return DensityMatrix(self.data * scalar, gpu=self.data.is_cuda)

    def _normalize(self) -> None:
        tr = torch.trace(self.data).real
        if not torch.allclose(tr.cpu(), torch.tensor(1.0, dtype=torch.float64)):
            self.data = self.data / tr

    def norm(self) -> torch.Tensor:
        """Tr(rho), on the CPU."""
        return torch.trace(self.data).real.cpu()

    @classmethod
    def from_state_vector(cls, state: StateVector) -> "DensityMatrix":
        """|psi><psi|."""
        return cls(torch.outer(state.data, state.data.conj()), gpu=state.data.is_cuda)

    @classmethod
    def _from_state_amplitudes(cls, *, eigenstates: Sequence[str], n_qudits: int,
                               amplitudes: Mapping[str, complex]):
        sv, _ = StateVector._from_state_amplitudes(eigenstates=eigenstates, n_qudits=n_qudits,
                                                   amplitudes=amplitudes)
        return DensityMatrix.from_state_vector(sv), amplitudes
