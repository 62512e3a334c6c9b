"""Lindblad jump operators from a noise model, and the effective
single-qubit noise term of the Monte-Carlo Hamiltonian (reference
emu_base/jump_lindblad_operators.py).

The noise model is read by attribute, as pulser's ``NoiseModel`` exposes
it: ``noise_types``, ``relaxation_rate``, ``dephasing_rate``,
``hyperfine_dephasing_rate``, ``depolarizing_rate``, ``eff_noise_rates``,
``eff_noise_opers`` (pulser itself is not a dependency).  Operators are in
the (g, r) basis of the state vector; pulser's ising basis lists r first,
so effective-noise operators have their 2x2 block flipped.
"""

from __future__ import annotations

import math

import torch

dtype = torch.complex128


def get_lindblad_operators(*, noise_type: str, noise_model, interact_type: str = "ising",
                           dim: int = 2) -> list[torch.Tensor]:
    """Jump operators of one noise channel (jump_lindblad_operators.py:8-77):
    relaxation sqrt(g)|g><r|, dephasing sqrt(g/2) Z, depolarizing
    sqrt(g/4) (X, Y, Z), effective noise sqrt(rate) A (pulser basis flipped
    for ising), leakage none."""
    assert noise_type in noise_model.noise_types
    if noise_type == "relaxation":
        op = torch.zeros(dim, dim, dtype=dtype)
        op[0, 1] = math.sqrt(noise_model.relaxation_rate)
        return [op]
    if noise_type == "dephasing":
        if noise_model.hyperfine_dephasing_rate != 0.0:
            raise NotImplementedError(
                "hyperfine_dephasing_rate is supported only in the digital basis")
        c = math.sqrt(noise_model.dephasing_rate / 2)
        op = torch.zeros(dim, dim, dtype=dtype)
        op[0, 0], op[1, 1] = c, -c
        return [op]
    if noise_type == "depolarizing":
        c = math.sqrt(noise_model.depolarizing_rate / 4)
        x = torch.zeros(dim, dim, dtype=dtype)
        x[0, 1] = x[1, 0] = c
        y = torch.zeros(dim, dim, dtype=dtype)
        y[0, 1], y[1, 0] = -1j * c, 1j * c
        z = torch.zeros(dim, dim, dtype=dtype)
        z[0, 0], z[1, 1] = c, -c
        return [x, y, z]
    if noise_type == "eff_noise":
        ops = [torch.as_tensor(op).detach().clone().to(dtype) for op in noise_model.eff_noise_opers]
        if not all(tuple(op.shape) == (dim, dim) for op in ops):
            raise ValueError(f"Only {dim} by {dim} effective noise operator matrices are supported")
        out = [math.sqrt(rate) * op for rate, op in zip(noise_model.eff_noise_rates, ops)]
        if interact_type == "ising":
            for t in out:
                t[:2, :2] = torch.flip(t[:2, :2], (0, 1))
        return out
    if noise_type == "leakage":
        return []
    raise ValueError(f"Unknown noise type: {noise_type}")


NON_LINDBLADIAN_NOISE = {"SPAM", "doppler", "amplitude", "detuning", "register", "dmm_sigma",
                         "dmm_crosstalk"}


def get_all_lindblad_noise_operators(noise_model, dim: int = 2,
                                     interact_type: str = "ising") -> list[torch.Tensor]:
    """Every jump operator of a noise model (pulser_adapter.py:31-58)."""
    if noise_model is None:
        return []
    return [op for kind in noise_model.noise_types if kind not in NON_LINDBLADIAN_NOISE
            for op in get_lindblad_operators(noise_type=kind, noise_model=noise_model, dim=dim,
                                             interact_type=interact_type)]


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
