"""Stepper interface (reference emu_sv/time_evolution/interface.py:7-37)."""

from __future__ import annotations

from abc import ABC, abstractmethod

import torch


class BaseStepper(ABC):
    def __init__(self, max_krylov_dim: int) -> None:
        self.max_krylov_dim = max_krylov_dim

    @staticmethod
    @abstractmethod
    def get_hamiltonian(omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
                        pulser_lindblads: list, interaction_matrix: torch.Tensor,
                        device: torch.device):
        ...

    @abstractmethod
    def apply(self, dt: float, omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
              full_interaction_matrix: torch.Tensor, state: torch.Tensor,
              krylov_tolerance: float, pulser_lindblads: list):
        ...
