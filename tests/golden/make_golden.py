"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py

Every fixture is cross-checked here against an independent computation
before it is written (dense Hamiltonian matrices and dense matrix
exponentials), so a fixture that exists has passed that check.

  hpsi_random.npz   seeded random Hamiltonians (N=7 with phases, N=9 real)
                    and vectors; H.psi from the dense matrix.
  krylov_step.npz   one Krylov step exp(-i dt H) psi at N=8 (dense expm) and
                    the oracle's Lanczos iteration count.
  chain10.npz       the N=10 chain workload (5 um, Omega = 4 pi, delta
                    -10 -> 10 rad/us over 1000 ns, dt = 10 ns): per-step
                    midpoint tables, interaction matrix, final state, final
                    occupations and energy, Lanczos iterations per step;
                    the final state also from 100 dense eigendecomposition steps.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import emu_sv_oracle as orc  # noqa: E402

C6 = 5420158.53


def random_hamiltonian(n: int, seed: int, with_phase: bool):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, generator=g, dtype=torch.float64) * 12.0
    de = (torch.rand(n, generator=g, dtype=torch.float64) - 0.5) * 20.0
    ph = (torch.rand(n, generator=g, dtype=torch.float64) * 6.0 if with_phase
          else torch.zeros(n, dtype=torch.float64))
    u = torch.rand(n, n, generator=g, dtype=torch.float64) * 30.0
    u = u + u.T
    u.fill_diagonal_(0.0)
    psi = torch.randn(2**n, generator=g, dtype=torch.complex128)
    return om, de, ph, u, psi


def make_hpsi() -> dict:
    out = {}
    for tag, n, seed, phase in (("a", 7, 11, True), ("b", 9, 12, False)):
        om, de, ph, u, psi = random_hamiltonian(n, seed, phase)
        dense = orc.dense_rydberg_hamiltonian(om.to(torch.complex128), de.to(torch.complex128),
                                              ph.to(torch.complex128), u)
        ref = dense @ psi
        got = orc.OracleHamiltonian(om.to(torch.complex128), de.to(torch.complex128),
                                    ph.to(torch.complex128), u) * psi
        assert torch.allclose(got, ref, atol=1e-11, rtol=0), "oracle H.psi disagrees with dense H"
        out.update({f"{tag}_omega": om.numpy(), f"{tag}_delta": de.numpy(),
                    f"{tag}_phi": ph.numpy(), f"{tag}_u": u.numpy(), f"{tag}_psi": psi.numpy(),
                    f"{tag}_hpsi": ref.numpy()})
    return out


def make_krylov() -> dict:
    n, dt, tol = 8, 0.01, 1e-10
    om, de, ph, u, psi = random_hamiltonian(n, 21, True)
    psi = psi / torch.linalg.vector_norm(psi)
    omc, dec, phc = (x.to(torch.complex128) for x in (om, de, ph))
    dense = orc.dense_rydberg_hamiltonian(omc, dec, phc, u)
    ref = torch.linalg.matrix_exp(-1j * dt * dense) @ psi
    ham = orc.OracleHamiltonian(omc, dec, phc, u)
    r = orc.krylov_exp_impl(lambda x: -1j * dt * (ham * x), psi.clone(), is_hermitian=True,
                            exp_tolerance=tol, norm_tolerance=tol)
    assert r.converged and torch.allclose(r.result, ref, atol=tol, rtol=0)
    return {"omega": om.numpy(), "delta": de.numpy(), "phi": ph.numpy(), "u": u.numpy(),
            "psi": psi.numpy(), "dt": np.float64(dt), "tol": np.float64(tol),
            "result": ref.numpy(), "iterations": np.int64(r.iteration_count)}


def make_chain10() -> dict:
    n, duration, dt = 10, 1000, 10
    pos = np.stack([np.arange(n) * 5.0, np.zeros(n)], axis=1)
    dist = np.hypot(pos[:, None, 0] - pos[None, :, 0], pos[:, None, 1] - pos[None, :, 1])
    with np.errstate(divide="ignore"):
        u = np.where(dist > 0, C6 / dist**6, 0.0)
    times = [float(k * dt) for k in range(duration // dt + 1)]
    amp = torch.full((duration,), 4 * np.pi, dtype=torch.float64)
    det = torch.linspace(-10.0, 10.0, duration, dtype=torch.float64)
    om_k = orc.midpoint_samples(amp, times, is_amplitude=True)
    de_k = orc.midpoint_samples(det, times, is_amplitude=False)
    steps = om_k.numel()
    omega = om_k[:, None].repeat(1, n).to(torch.complex128)
    delta = de_k[:, None].repeat(1, n).to(torch.complex128)
    phi = torch.zeros_like(omega)
    ut = torch.as_tensor(u).reshape(1, n, n)
    run = orc.run_sequence(omega, delta, phi, lambda t: ut, times, krylov_tolerance=1e-10)
    final = run.states[steps]
    # independent check: the same steps by dense eigendecomposition,
    # exp(-i dt H) psi = V exp(-i dt L) V^+ psi
    psi = torch.zeros(2**n, dtype=torch.complex128)
    psi[0] = 1.0
    for k in range(steps):
        h = orc.dense_rydberg_hamiltonian(omega[k], delta[k], phi[k], ut[0])
        lam, vec = torch.linalg.eigh(h)
        dtk = (times[k + 1] - times[k]) * 1e-3
        psi = vec @ (torch.exp(-1j * dtk * lam) * (vec.conj().T @ psi))
    fid = float(orc.fidelity(final, psi))
    assert 1.0 - fid < 1e-10, f"oracle vs dense stepping: 1 - F = {1 - fid}"
    return {"omega_k": om_k.numpy(), "delta_k": de_k.numpy(), "u": u, "target_times":
            np.asarray(times), "final_state": final.numpy(),
            "occupation": run.occupation[steps].numpy(), "energy": np.float64(run.energy[steps]),
            "iterations": np.asarray(run.iterations, dtype=np.int64)}


def main() -> None:
    np.savez(os.path.join(HERE, "hpsi_random.npz"), **make_hpsi())
    np.savez(os.path.join(HERE, "krylov_step.npz"), **make_krylov())
    np.savez(os.path.join(HERE, "chain10.npz"), **make_chain10())
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
