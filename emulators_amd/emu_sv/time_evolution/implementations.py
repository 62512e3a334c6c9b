"""Noiseless state-vector stepper (reference emu_sv/time_evolution/implementations.py:16-60)."""

from __future__ import annotations

import torch

from emulators_amd.emu_base.math.krylov_exp import DEFAULT_MAX_KRYLOV_DIM
from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
from emulators_amd.emu_sv.time_evolution.interface import BaseStepper
from emulators_amd.emu_sv.time_evolution.torch import DifferentiableStateVectorEvolution


class EvolveStateVector(BaseStepper):
    """psi <- exp(-i dt H) psi with the fused GPU Lanczos engine.

    The state may be complex128 (the reference's arithmetic) or complex64
    (FP64 Lanczos scalars); it is evolved in its own dtype."""

    def __init__(self, max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM) -> None:
        super().__init__(max_krylov_dim)

    @staticmethod
    def get_hamiltonian(omegas, deltas, phis, pulser_lindblads, interaction_matrix,
                        device) -> RydbergHamiltonian:
        return DifferentiableStateVectorEvolution.get_hamiltonian(
            omegas=omegas, deltas=deltas, phis=phis, interaction_matrix=interaction_matrix,
            device=device)

    def apply(self, dt: float, omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
              full_interaction_matrix: torch.Tensor, state: torch.Tensor,
              krylov_tolerance: float, pulser_lindblads: list):
        if pulser_lindblads:
            raise NotImplementedError("Lindblad noise is not part of the GPU state-vector path")
        return DifferentiableStateVectorEvolution.apply(
            dt, omegas, deltas, phis, full_interaction_matrix, state, krylov_tolerance,
            self.max_krylov_dim)
