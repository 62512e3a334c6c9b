"""GPU implementations of the state-vector observables.

Replace the reference's torch loops (emu_sv/custom_callback_implementations.py):
occupation and correlation come from one fused |psi|^2 pass over bit masks
(``emu_sv_correlation``), energy moments from one fused H.psi launch with
its reductions.  Results are float64 CPU tensors, as in the reference.
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


def energy_variance_sv_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """<psi|H^2|psi> - <psi|H|psi>^2 (reference :131-147)."""
    _, e, h2 = hamiltonian.moments(device_data(state))
    return torch.tensor(h2 - e * e, dtype=dtype)


def energy_second_moment_sv_impl(self, *, config, state, hamiltonian) -> torch.Tensor:
    """<psi|H^2|psi> = ||H psi||^2 (reference :170-185)."""
    _, _, h2 = hamiltonian.moments(device_data(state))
    return torch.tensor(h2, dtype=dtype)
