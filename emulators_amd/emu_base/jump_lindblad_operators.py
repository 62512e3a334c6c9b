"""The effective single-qubit noise term of the Monte-Carlo Hamiltonian
(reference emu_base/jump_lindblad_operators.py:80-93)."""

from __future__ import annotations

import torch


def compute_noise_from_lindbladians(lindbladians: list, dim: int = 2) -> torch.Tensor:
    """K = -i/2 sum_l L_l^+ L_l (complex128, dim x dim) over single-qubit
    Lindblad operators."""
    assert all(tuple(L.shape) == (dim, dim) for L in lindbladians), \
        "Only single-qubit lindblad operators are supported"
    acc = torch.zeros(dim, dim, dtype=torch.complex128)
    for L in lindbladians:
        L = torch.as_tensor(L, dtype=torch.complex128).cpu()
        acc = acc + L.mH @ L
    return -0.5j * acc
