"""Steppers (reference emu_sv/time_evolution/implementations.py): the
noiseless state vector (:16-60), the Lindblad density matrix (:63-120) and
the Monte-Carlo quantum-jump state vector (:123-250)."""

from __future__ import annotations

import math
import random

import torch

from emulators_amd import _native as native
from emulators_amd.emu_base.jump_lindblad_operators import compute_noise_from_lindbladians
from emulators_amd.emu_base.math.brents_root_finding import BrentsRootFinder
from emulators_amd.emu_base.math.krylov_exp import (DEFAULT_MAX_KRYLOV_DIM, ScaledHamiltonian,
                                                     krylov_exp)
from emulators_amd.emu_sv.algebra import apply, expect_batch
from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
from emulators_amd.emu_sv.lindblad_operator import RydbergLindbladian, ScaledLindbladian
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
            raise NotImplementedError(
                "EvolveStateVector is noiseless: Lindblad noise runs on EvolveMonteCarlo")
        return DifferentiableStateVectorEvolution.apply(
            dt, omegas, deltas, phis, full_interaction_matrix, state, krylov_tolerance,
            self.max_krylov_dim)


class EvolveDensityMatrix(BaseStepper):
    """rho <- exp(-i dt L) rho for the Lindbladian L (implementations.py:63-120):
    the device Arnoldi engine on the 4^N-element density matrix."""

    def __init__(self, max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM) -> None:
        super().__init__(max_krylov_dim)

    @staticmethod
    def get_hamiltonian(omegas, deltas, phis, pulser_lindblads, interaction_matrix,
                        device) -> RydbergLindbladian:
        return RydbergLindbladian(omegas=omegas, deltas=deltas, phis=phis,
                                  pulser_lindblads=pulser_lindblads,
                                  interaction_matrix=interaction_matrix, device=device)

    def apply(self, dt: float, omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
              full_interaction_matrix: torch.Tensor, state: torch.Tensor,
              krylov_tolerance: float, pulser_lindblads: list):
        ham = self.get_hamiltonian(omegas=omegas, deltas=deltas, phis=phis,
                                   pulser_lindblads=pulser_lindblads,
                                   interaction_matrix=full_interaction_matrix,
                                   device=state.device)
        out = krylov_exp(ScaledLindbladian(ham, -1j * dt), state,
                         norm_tolerance=krylov_tolerance, exp_tolerance=krylov_tolerance,
                         is_hermitian=False, max_krylov_dim=self.max_krylov_dim)
        return out, ham


JUMP_TIME_TOLERANCE_FRACTION = 0.1


def _norm2(state: torch.Tensor) -> float:
    return float(native.dot(native.nqubits_of(state), state)[0].item())


class EvolveMonteCarlo(BaseStepper):
    """One step of a quantum-jump trajectory (implementations.py:123-250).

    The state is evolved under the effective Hamiltonian H - i/2 sum L^+L
    (non-Hermitian; the device Arnoldi engine) until its squared norm falls
    to the trajectory's jump threshold; the crossing time is located with
    Brent's method (tolerance dt/10), a jump operator L_l on qubit q is drawn
    with weight <psi|L_l^+ L_l|psi> on that qubit, applied, the state
    renormalised and a new threshold drawn.  Random numbers come from
    Python's ``random`` in the reference's order (threshold draws and
    ``random.choices``), so a seeded run follows the same trajectory."""

    def __init__(self, max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM) -> None:
        super().__init__(max_krylov_dim)
        self.jump_threshold = random.uniform(0.0, 1.0)

    @staticmethod
    def get_hamiltonian(omegas, deltas, phis, pulser_lindblads, interaction_matrix,
                        device) -> RydbergHamiltonian:
        return RydbergHamiltonian(omegas=omegas, deltas=deltas, phis=phis,
                                  interaction_matrix=interaction_matrix, device=device,
                                  noise=compute_noise_from_lindbladians(pulser_lindblads))

    def evolve(self, dt: float, hamiltonian: RydbergHamiltonian, state: torch.Tensor,
               krylov_tolerance: float) -> torch.Tensor:
        return krylov_exp(ScaledHamiltonian(hamiltonian, -1j * dt), state,
                          norm_tolerance=krylov_tolerance, exp_tolerance=krylov_tolerance,
                          is_hermitian=False, max_krylov_dim=self.max_krylov_dim)

    def do_quantum_jump(self, state: torch.Tensor, lindblad_ops: list,
                        aggregated_lindblad_ops: torch.Tensor, n_qubits: int) -> torch.Tensor:
        weights = expect_batch(state, aggregated_lindblad_ops, n_qubits).real
        qubit, op = random.choices([(q, L) for q in range(n_qubits) for L in lindblad_ops],
                                   weights=weights.reshape(-1).tolist())[0]
        state = apply(state, qubit, op)
        native.scale(n_qubits, state, 1.0 / math.sqrt(_norm2(state)))
        self.jump_threshold = random.uniform(0.0, _norm2(state))
        return state

    def apply(self, dt: float, omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
              full_interaction_matrix: torch.Tensor, state: torch.Tensor,
              krylov_tolerance: float, pulser_lindblads: list):
        ham = self.get_hamiltonian(omegas=omegas, deltas=deltas, phis=phis,
                                   pulser_lindblads=pulser_lindblads,
                                   interaction_matrix=full_interaction_matrix,
                                   device=state.device)
        n = len(omegas)
        tol = dt * JUMP_TIME_TOLERANCE_FRACTION
        stacked = torch.stack([torch.as_tensor(L, dtype=torch.complex128).cpu()
                               for L in pulser_lindblads])
        aggregated = stacked.conj().transpose(1, 2) @ stacked  # L^+ L per operator
        current_time = 0.0
        new_gap = _norm2(state) - self.jump_threshold
        while abs(dt - current_time) > 1e-5:
            old_gap = new_gap
            state = self.evolve(dt - current_time, ham, state, krylov_tolerance)
            new_gap = _norm2(state) - self.jump_threshold
            if new_gap > 0.0:
                break
            # the norm crossed the threshold inside [current_time, dt]
            finder = BrentsRootFinder(start=current_time, end=dt, f_start=old_gap,
                                      f_end=new_gap, epsilon=tol)
            current_time = dt
            while not finder.is_converged(tolerance=tol):
                target = finder.get_next_abscissa()
                state = self.evolve(target - current_time, ham, state, krylov_tolerance)
                current_time = target
                new_gap = _norm2(state) - self.jump_threshold
                finder.provide_ordinate(current_time, new_gap)
            state = self.do_quantum_jump(state, pulser_lindblads, aggregated, n)
            new_gap = _norm2(state) - self.jump_threshold
        return state, ham
