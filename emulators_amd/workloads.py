"""Synthetic registers and pulses of the benchmark and parity workloads.

Pulses are built as 1 ns samples (pulser's sampling convention: a ramp from
a to b over d ns is linspace(a, b, d) on t = 0..d-1, a constant is repeated)
and turned into per-step midpoint tables by
``emulators_amd.emu_base.sequence_data_from_samples`` (PCHIP), exactly as the
reference adapter does with pulser samples.

Workloads (C6 = 5420158.53 rad um^6/us):
  chain10       N=10 chain, 5 um; Omega = 4 pi; delta -10 -> 10 over 1000 ns
  square20      N=20 4x5 square lattice, 6 um; same pulse shape, 2000 ns
  triangular26  N=26 triangular lattice, 5.5 um, rows of 7/6/7/6 atoms with
                alternate rows offset by half a spacing; Omega = 2 pi with
                sin^2 rise/fall over 200 ns; delta -20 -> 15 over 3000 ns
  random16      N=16 random positions in a 30x30 um box, min spacing 5 um
                (seeded rejection sampling); 100 piecewise-constant 10 ns steps
                with Omega_k ~ U(0, 4 pi), delta_k ~ U(-10, 10) (global)
  batch         seeded registers of N ~ U{8..20} with random pulses
"""

from __future__ import annotations

import math

import numpy as np
import torch

from emulators_amd.emu_base import (
    HamiltonianType,
    InteractionMatrixCallable,
    SequenceData,
    interaction_matrix_from_positions,
    sequence_data_from_samples,
)

C6 = 5420158.53


def chain_positions(n: int, spacing: float) -> np.ndarray:
    return np.stack([np.arange(n) * spacing, np.zeros(n)], axis=1)


def square_positions(rows: int, cols: int, spacing: float) -> np.ndarray:
    return np.array([[c * spacing, r * spacing] for r in range(rows) for c in range(cols)],
                    dtype=np.float64)


def triangular_positions(row_sizes, spacing: float) -> np.ndarray:
    pitch = spacing * math.sqrt(3.0) / 2.0
    pts = []
    for r, size in enumerate(row_sizes):
        off = 0.5 * spacing if r % 2 else 0.0
        pts += [[off + c * spacing, r * pitch] for c in range(size)]
    return np.asarray(pts, dtype=np.float64)


def random_positions(n: int, box: float, min_spacing: float, seed: int) -> np.ndarray:
    """Rejection sampling of n points in [0, box]^2 with pairwise distance >= min_spacing."""
    rng = np.random.default_rng(seed)
    pts: list = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 1_000_000:
            raise RuntimeError("could not place the atoms")
        p = rng.uniform(0.0, box, size=2)
        if all(np.hypot(*(p - q)) >= min_spacing for q in pts):
            pts.append(p)
    return np.asarray(pts)


def _config(observables=(), dt: float = 10.0, **kw):
    from emulators_amd.emu_sv import SVConfig

    return SVConfig(observables=list(observables), dt=dt, log_level=40, **kw)


def global_samples(n: int, amp: np.ndarray, det: np.ndarray, phase: np.ndarray | None = None):
    phase = np.zeros_like(amp) if phase is None else phase
    return {f"q{i}": {"amp": amp, "det": det, "phase": phase} for i in range(n)}


def ramp(a: float, b: float, d: int) -> np.ndarray:
    return np.linspace(a, b, d)


def sin2_envelope(omega_max: float, duration: int, rise: int) -> np.ndarray:
    t = np.arange(duration, dtype=np.float64)
    env = np.full(duration, omega_max)
    up = t < rise
    env[up] = omega_max * np.sin(0.5 * np.pi * t[up] / rise) ** 2
    down = t >= duration - rise
    env[down] = omega_max * np.sin(0.5 * np.pi * (duration - t[down]) / rise) ** 2
    return env


def adiabatic_sequence(positions: np.ndarray, amp: np.ndarray, det: np.ndarray, config,
                       dt: float) -> SequenceData:
    n = positions.shape[0]
    u = interaction_matrix_from_positions(positions, C6)
    return sequence_data_from_samples(global_samples(n, amp, det), [f"q{i}" for i in range(n)], u,
                                      config, dt)


def chain10(config=None) -> SequenceData:
    config = config or _config()
    d = 1000
    return adiabatic_sequence(chain_positions(10, 5.0), np.full(d, 4 * np.pi), ramp(-10, 10, d),
                              config, config.dt)


def square20(config=None) -> SequenceData:
    config = config or _config()
    d = 2000
    return adiabatic_sequence(square_positions(4, 5, 6.0), np.full(d, 4 * np.pi),
                              ramp(-10, 10, d), config, config.dt)


def triangular26(config=None, duration: int = 3000) -> SequenceData:
    config = config or _config()
    return adiabatic_sequence(triangular_positions((7, 6, 7, 6), 5.5),
                              sin2_envelope(2 * np.pi, duration, 200),
                              ramp(-20, 15, duration), config, config.dt)


def piecewise_sequence(positions: np.ndarray, omegas: np.ndarray, deltas: np.ndarray,
                       step_ns: float = 10.0, phis: np.ndarray | None = None) -> SequenceData:
    """SequenceData with given per-step (global) Omega_k, delta_k tables."""
    n = positions.shape[0]
    s = len(omegas)
    om = torch.as_tensor(np.repeat(np.asarray(omegas)[:, None], n, axis=1), dtype=torch.complex128)
    de = torch.as_tensor(np.repeat(np.asarray(deltas)[:, None], n, axis=1), dtype=torch.complex128)
    ph = (torch.zeros_like(om) if phis is None else
          torch.as_tensor(np.repeat(np.asarray(phis)[:, None], n, axis=1), dtype=torch.complex128))
    u = interaction_matrix_from_positions(positions, C6).reshape(1, n, n)
    return SequenceData(
        omega=om, delta=de, phi=ph, interaction_matrix=InteractionMatrixCallable(u, u, 0.0),
        qubit_ids=tuple(f"q{i}" for i in range(n)), bad_atoms=(False,) * n, lindblad_ops=[],
        state_prep_error=0.0, target_times=[step_ns * k for k in range(s + 1)],
        eigenstates=["r", "g"], hamiltonian_type=HamiltonianType.Rydberg,
        has_stochastic_noise=False)


RANDOM16_SEED = 2024


def random16_parameters(seed: int = RANDOM16_SEED):
    """(positions, Omega_k, delta_k) of the 16-atom gradient workload."""
    pos = random_positions(16, 30.0, 5.0, seed)
    rng = np.random.default_rng(seed + 1)
    om = rng.uniform(0.0, 4 * np.pi, size=100)
    de = rng.uniform(-10.0, 10.0, size=100)
    return pos, om, de


def random16() -> SequenceData:
    pos, om, de = random16_parameters()
    return piecewise_sequence(pos, om, de)


def batch_instances(count: int = 256, seed: int = 7):
    """Seeded registers of N ~ U{8..20}: (positions, Omega, delta_start, delta_end,
    duration_ns).  Omega ~ U(1, 4 pi) rad/us (the literal reading of "1-4 pi"),
    delta ramps from -U(5, 20) to +U(5, 20), durations U{500..2000} in 10 ns."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(8, 21))
        pos = random_positions(n, 6.0 * math.sqrt(n) + 10.0, 5.0, seed * 1000 + i)
        out.append(dict(positions=pos, omega=float(rng.uniform(1.0, 4 * np.pi)),
                        delta_start=-float(rng.uniform(5, 20)), delta_end=float(rng.uniform(5, 20)),
                        duration=int(rng.integers(50, 201)) * 10))
    return out


def batch_sequence(inst: dict, config=None) -> SequenceData:
    config = config or _config()
    d = inst["duration"]
    return adiabatic_sequence(inst["positions"], np.full(d, inst["omega"]),
                              ramp(inst["delta_start"], inst["delta_end"], d), config, config.dt)


def relaxation_dephasing(gamma_r: float, gamma_d: float) -> list:
    """Jump operators sqrt(g_r)|g><r| and sqrt(g_d/2) Z (relaxation and
    dephasing channels, (g, r) basis)."""
    import torch
    relax = torch.zeros(2, 2, dtype=torch.complex128)
    relax[0, 1] = math.sqrt(gamma_r)
    deph = torch.zeros(2, 2, dtype=torch.complex128)
    deph[0, 0], deph[1, 1] = math.sqrt(gamma_d / 2), -math.sqrt(gamma_d / 2)
    return [relax, deph]


def noisy_chain(n: int, lindblads: list, stochastic: bool, duration: int = 1000,
                config=None) -> SequenceData:
    """The chain10 pulse (5 um spacing, Omega 4 pi, delta -10 -> 10 over
    ``duration`` ns) on n atoms with Lindblad noise: stochastic=True selects
    Monte-Carlo trajectories, False the density matrix."""
    import dataclasses
    config = config or _config()
    seq = adiabatic_sequence(chain_positions(n, 5.0), np.full(duration, 4 * np.pi),
                             ramp(-10, 10, duration), config, config.dt)
    return dataclasses.replace(seq, lindblad_ops=list(lindblads), has_stochastic_noise=stochastic)
