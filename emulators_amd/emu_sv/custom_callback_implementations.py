"""GPU implementations of the observables.

Replace the reference's torch loops (emu_sv/custom_callback_implementations.py):
occupation and correlation come from one fused |psi|^2 pass over bit masks
(``emu_sv_correlation``), energy moments from one fused H.psi launch with
its reductions.  For a density matrix the same correlation kernel runs on
sqrt(diag(rho)) (the diagonal of a density matrix is a probability vector;
rounding-level negative entries count as zero) and the energy moments use
the Lindbladian's h_eff / expect, as the reference's ``*_den_mat_impl``.
Results are float64 CPU tensors, as in the reference.
"""

from __future__ import annotations

import torch

from emulators_amd import _native as native

dtype = torch.float64


def device_data(state) -> torch.Tensor:
    """The state's amplitudes on the GPU (copied there if the state is on the CPU)."""
    data = state.data if hasattr(state, "data") else state
    if not data.is_cuda:
        if not torch.cuda.is_available():
            raise RuntimeError("emu_sv observables run on the GPU; no GPU is available")
        data = data.cuda()
    return data.contiguous()


def qubit_occupation_sv_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """<psi|n_i|psi> for every qubit (reference :30-47)."""
    data = device_data(state)
    return native.correlation(native.nqubits_of(data), data).diagonal().to("cpu", dtype)


def correlation_matrix_sv_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """<psi|n_i n_j|psi> (reference :77-102)."""
    data = device_data(state)
    return native.correlation(native.nqubits_of(data), data).to("cpu", dtype)


def _strip_noise(hamiltonian):
    """The Hamiltonian without the Monte-Carlo noise term (reference :19-27)."""
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    return RydbergHamiltonian(omegas=hamiltonian.omegas * 2.0, deltas=hamiltonian.deltas,
                              phis=hamiltonian.phis,
                              interaction_matrix=hamiltonian.interaction_matrix,
                              device=hamiltonian.device, noise=None)


def energy_variance_sv_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """<psi|H^2|psi> - <psi|H|psi>^2 (reference :131-147), noise term stripped."""
    if getattr(hamiltonian, "noise", None) is not None:
        hamiltonian = _strip_noise(hamiltonian)
    _, e, h2 = hamiltonian.moments(device_data(state))
    return torch.tensor(h2 - e * e, dtype=dtype)


def energy_second_moment_sv_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """<psi|H^2|psi> = ||H psi||^2 (reference :170-185), noise term stripped."""
    if getattr(hamiltonian, "noise", None) is not None:
        hamiltonian = _strip_noise(hamiltonian)
    _, _, h2 = hamiltonian.moments(device_data(state))
    return torch.tensor(h2, dtype=dtype)


def _diag_amplitudes(state) -> torch.Tensor:
    data = device_data(state)
    p = data.diagonal().real.clamp(min=0.0)
    return torch.sqrt(p).to(torch.complex128).contiguous()


def qubit_occupation_sv_den_mat_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """Tr(rho n_i) for every qubit (reference :50-74)."""
    a = _diag_amplitudes(state)
    return native.correlation(native.nqubits_of(a), a).diagonal().to("cpu", dtype)


def correlation_matrix_sv_den_mat_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """Tr(rho n_i n_j) (reference :105-128)."""
    a = _diag_amplitudes(state)
    return native.correlation(native.nqubits_of(a), a).to("cpu", dtype)


def energy_variance_sv_den_mat_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """Tr(rho H^2) - Tr(rho H)^2 (reference :150-167)."""
    from emulators_amd.emu_sv.density_matrix_state import DensityMatrix
    h_rho = hamiltonian.h_eff(device_data(state))
    h2 = hamiltonian.expect(DensityMatrix(h_rho, gpu=True))
    return (h2 - hamiltonian.expect(state) ** 2).to("cpu", dtype)


def energy_second_moment_den_mat_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """Tr(rho H^2) (reference :188-204)."""
    from emulators_amd.emu_sv.density_matrix_state import DensityMatrix
    h_rho = hamiltonian.h_eff(device_data(state))
    return hamiltonian.expect(DensityMatrix(h_rho, gpu=True)).to("cpu", dtype)
