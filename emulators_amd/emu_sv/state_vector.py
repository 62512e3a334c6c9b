"""StateVector (reference emu_sv/state_vector.py:17-210)."""

from __future__ import annotations

from typing import Sequence

import torch

from emulators_amd.emu_sv.dense_state import DenseState

dtype = torch.complex128


class StateVector(DenseState):
    """A 2^N complex128 state vector (qubit 0 = most significant bit)."""

    def _normalize(self) -> None:
        """Rescale to unit norm unless already within 1e-12 in norm^4
        (state_vector.py:33-47)."""
        norm = torch.linalg.vector_norm(self.data)
        if abs(self.norm() ** 4 - 1.0) > 1e-12:
            self.data = self.data / norm

    @classmethod
    def zero(cls, num_sites: int, gpu: bool = True,
             eigenstates: Sequence[str] = ("r", "g")) -> "StateVector":
        return cls(torch.zeros(2**num_sites, dtype=dtype), gpu=gpu, eigenstates=eigenstates)

    @classmethod
    def make(cls, num_sites: int, gpu: bool = True) -> "StateVector":
        """|g...g> (index 0)."""
        result = cls.zero(num_sites=num_sites, gpu=gpu)
        result.data[0] = 1.0
        return result

    def inner(self, other: "StateVector") -> torch.Tensor:
        """<self|other>, returned on the CPU."""
        assert isinstance(other, StateVector), "Other state must be a StateVector"
        assert self.data.shape == other.data.shape, "States do not have the same shape"
        return torch.vdot(self.data, other.data.to(self.data.device)).cpu()

    def __add__(self, other: "StateVector") -> "StateVector":
        assert isinstance(other, StateVector), "`Other` state can only be a StateVector"
        assert self.eigenstates == other.eigenstates
        return StateVector(self.data + other.data, gpu=self.data.is_cuda,
                           eigenstates=self.eigenstates)

    def __rmul__(self, scalar: complex) -> "StateVector":
        return StateVector(scalar * self.data, gpu=self.data.is_cuda, eigenstates=self.eigenstates)

    def norm(self) -> torch.Tensor:
        """||psi|| on the CPU."""
        return torch.linalg.vector_norm(self.data).cpu()

    def __repr__(self) -> str:
        return repr(self.data)


def inner(left: StateVector, right: StateVector) -> torch.Tensor:
    assert left.data.shape == right.data.shape and left.data.dim() == 1
    return left.inner(right)
