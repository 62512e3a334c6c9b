"""Single-qubit operator helpers of the quantum-jump step
(reference emu_sv/algebra.py:4-37), on the device kernels."""

from __future__ import annotations

import torch

from emulators_amd import _native as native


def reduced_density_matrices(state_vector: torch.Tensor) -> torch.Tensor:
    """rho_q[j, k] = sum over the other qubits of psi(.. j ..) conj(psi(.. k ..))
    for every qubit q (complex128, (n, 2, 2), CPU), from one trace launch."""
    n = native.nqubits_of(state_vector)
    tr = native.parameter_traces(n, state_vector, state_vector)
    norm2 = float(native.dot(n, state_vector)[0].item())
    rho = torch.zeros(n, 2, 2, dtype=torch.complex128)
    rho[:, 0, 1] = torch.from_numpy(tr.up)  # sum_{x_q=1} conj(psi(x)) psi(x with q=0)
    rho[:, 1, 0] = torch.from_numpy(tr.dn)
    rho[:, 1, 1] = torch.from_numpy(tr.occ)
    rho[:, 0, 0] = norm2 - rho[:, 1, 1]
    return rho


def expect_batch(state_vector: torch.Tensor, single_qubit_operators: torch.Tensor,
                 n_qubits: int) -> torch.Tensor:
    """T[q, i] = Tr(A_i rho_q) for each qubit q and operator A_i
    (algebra.py:4-25).  Returns a complex128 (n_qubits, count) CPU tensor."""
    rho = reduced_density_matrices(state_vector)
    if rho.shape[0] != n_qubits:
        raise ValueError("expect_batch: n_qubits does not match the state")
    ops = torch.as_tensor(single_qubit_operators, dtype=torch.complex128).cpu()
    # Tr(A rho) = sum_{j,k} A[j,k] rho[k,j]
    return torch.einsum("ijk,qkj->qi", ops, rho)


def apply(state_vector: torch.Tensor, qubit_index: int,
          single_qubit_operator: torch.Tensor) -> torch.Tensor:
    """The operator on one qubit, as a new device vector; the input is left
    as it is (algebra.py:28-37)."""
    n = native.nqubits_of(state_vector)
    return native.apply_1q(n, state_vector.clone(), qubit_index, single_qubit_operator)
