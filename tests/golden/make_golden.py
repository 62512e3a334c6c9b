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
  square20.npz      the N=20 lattice workload (4x5, 6 um, the chain10 pulse
                    over 2000 ns): midpoint tables, final occupations and
                    energy, Lanczos iterations per step, and a 128-row sign
                    sketch of the final state (``state_sketch``); ten of the
                    oracle's steps (the first five and the last five) are
                    checked against scipy's expm_multiply on the sparse H.
  random16_grad.npz the 16-atom gradient workload (100 piecewise-constant
                    steps, global Omega_k / delta_k, loss E(T) + sum <n_i>(T)):
                    its inputs, the loss and all 200 gradient components
                    through the oracle's restatement of the reference's
                    per-step autograd (ReferenceStep chained over the pulse);
                    four components are checked against central differences.

Run a single fixture with ``python tests/golden/make_golden.py square20``.
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


SKETCH_ROWS = 128


def sign_sketch(state: torch.Tensor, rows: int = SKETCH_ROWS, seed: int = 20) -> torch.Tensor:
    """S psi with S a (rows x D) matrix of seeded pseudo-random signs scaled by
    1/sqrt(rows): ||S d|| estimates ||d|| (Johnson-Lindenstrauss), so a few
    hundred numbers pin a state of 2^20 amplitudes.  The signs come from an
    integer hash of (row, index), computed the same way on any device."""
    d = state.numel()
    x = torch.arange(d, dtype=torch.int64, device=state.device)
    mask = 0xFFFFFFFF
    out = torch.empty(rows, dtype=torch.complex128, device=state.device)
    for i in range(rows):
        h = (x * 0x9E3779B1 + (i + 1) * 0x85EBCA77 + seed * 0xC2B2AE3D) & mask
        h = h ^ (h >> 16)
        h = (h * 0x7FEB352D) & mask
        h = h ^ (h >> 15)
        h = (h * 0x846CA68B) & mask
        h = h ^ (h >> 16)
        sign = (1 - 2 * (h & 1)).to(torch.float64)
        out[i] = torch.sum(sign * state)
    return out / np.sqrt(rows)


def sparse_hamiltonian(omega_k, delta_k, phi_k, u):
    """scipy CSR matrix of H (qubit q <-> bit n-1-q), from the Hamiltonian's
    definition and independent of the oracle's index_add_ form."""
    import scipy.sparse as sp

    n = u.shape[0]
    d = 2**n
    idx = np.arange(d, dtype=np.int64)
    bits = [(idx >> (n - 1 - q)) & 1 for q in range(n)]
    diag = np.zeros(d)
    for q in range(n):
        diag -= delta_k[q] * bits[q]
        for q2 in range(q + 1, n):
            diag += u[q, q2] * (bits[q] & bits[q2])
    rows, cols, vals = [idx], [idx], [diag.astype(np.complex128)]
    for q in range(n):
        c = 0.5 * omega_k[q] * np.exp(1j * phi_k[q])
        src = idx ^ (1 << (n - 1 - q))
        # <x|H|x^e_q>: x has bit 1 -> c (raising from src), bit 0 -> conj(c)
        vals.append(np.where(bits[q] == 1, c, np.conj(c)))
        rows.append(idx)
        cols.append(src)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                         shape=(d, d))


def make_square20() -> dict:
    from scipy.sparse.linalg import expm_multiply

    rows, cols, sp_um, duration, dt = 4, 5, 6.0, 2000, 10
    n = rows * cols
    pos = np.array([[c * sp_um, r * sp_um] for r in range(rows) for c in range(cols)], dtype=np.float64)
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
    check = list(range(5)) + list(range(steps - 5, steps))
    run = orc.run_sequence(omega, delta, phi, lambda t: ut, times, krylov_tolerance=1e-10,
                           eval_steps=sorted(set([k for k in check] + [k + 1 for k in check])))
    # independent check of ten steps: the oracle's exp(-i dt H) psi_k against
    # a truncated-Taylor action of the exponential on the sparse H
    for k in check:
        h = sparse_hamiltonian(om_k[k].item() * np.ones(n), de_k[k].item() * np.ones(n),
                               np.zeros(n), u)
        dtk = (times[k + 1] - times[k]) * 1e-3
        psi_k = run.states[k].numpy()
        ref = expm_multiply(-1j * dtk * h, psi_k)
        err = float(np.linalg.norm(run.states[k + 1].numpy() - ref))
        assert err < 1e-9, f"square20 step {k}: oracle vs expm_multiply {err}"
    final = run.states[steps]
    return {"omega_k": om_k.numpy(), "delta_k": de_k.numpy(), "u": u, "target_times": np.asarray(times),
            "occupation": run.occupation[steps].numpy(), "energy": np.float64(run.energy[steps]),
            "iterations": np.asarray(run.iterations, dtype=np.int64),
            "state_sketch": sign_sketch(final).numpy(), "sketch_rows": np.int64(SKETCH_ROWS),
            "sketch_seed": np.int64(20), "final_norm": np.float64(final.norm())}


def random16_loss(om_k, de_k, u, tol, grads):
    """Loss E(T) + sum_i <n_i>(T) of the 16-atom workload on the oracle, and
    its gradient through ReferenceStep (the reference's per-step backward)."""
    n, s = u.shape[0], om_k.shape[0]
    om = om_k.clone().requires_grad_(grads)
    de = de_k.clone().requires_grad_(grads)
    psi = torch.zeros(2**n, dtype=torch.complex128)
    psi[0] = 1.0
    zeros = torch.zeros(n, dtype=torch.complex128)
    for k in range(s):
        if grads:
            psi = orc.ReferenceStep.apply(0.01, om[k].expand(n).to(torch.complex128),
                                          de[k].expand(n).to(torch.complex128), zeros, u, psi, tol, 100)
        else:
            psi, _ = orc.step_forward(0.01, om_k[k].expand(n).to(torch.complex128),
                                      de_k[k].expand(n).to(torch.complex128), zeros, u, psi, tol, 100)
    ham = orc.OracleHamiltonian(om_k[-1].expand(n), de_k[-1].expand(n), zeros, u)
    p = psi.detach()
    hpsi = ham * p
    pc = torch.tensor([bin(x).count("1") for x in range(2**n)], dtype=torch.float64)
    loss = float(torch.vdot(p, hpsi).real + ((p.abs() ** 2) * pc).sum())
    if not grads:
        return loss, None, None
    psi.backward(2 * (hpsi + pc * p))
    # explicit dependence of the final energy on the last step's parameters
    e_om, e_de, _, _ = orc.parameter_traces(p[None], p[None], om_k[-1].expand(n), zeros.real, 1.0)
    g_om, g_de = om.grad.clone(), de.grad.clone()
    g_om[-1] += e_om.sum()
    g_de[-1] += e_de.sum()
    return loss, g_om, g_de


def make_random16_grad() -> dict:
    from emulators_amd import workloads

    pos, om_np, de_np = workloads.random16_parameters()
    u = torch.as_tensor(workloads.interaction_matrix_from_positions(pos, workloads.C6))
    om_k, de_k = torch.as_tensor(om_np), torch.as_tensor(de_np)
    tol = 1e-10
    loss, g_om, g_de = random16_loss(om_k, de_k, u, tol, True)
    scale = max(float(g_om.abs().max()), float(g_de.abs().max()))
    h = 1e-6
    for which, k in (("om", 0), ("om", 77), ("de", 31), ("de", 99)):
        vals = []
        for sgn in (1.0, -1.0):
            o, d = om_k.clone(), de_k.clone()
            (o if which == "om" else d)[k] += sgn * h
            vals.append(random16_loss(o, d, u, 1e-12, False)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        g = float((g_om if which == "om" else g_de)[k])
        assert abs(fd - g) <= 1e-5 * scale, (which, k, fd, g, scale)
        print(f"random16 gradient {which}[{k}]: reference chain {g:.9f}, central difference {fd:.9f}")
    return {"positions": pos, "omega_k": om_np, "delta_k": de_np, "u": u.numpy(),
            "tolerance": np.float64(tol), "loss": np.float64(loss), "grad_omega": g_om.numpy(),
            "grad_delta": g_de.numpy()}


FIXTURES = {
    "hpsi_random": make_hpsi,
    "krylov_step": make_krylov,
    "chain10": make_chain10,
    "square20": make_square20,
    "random16_grad": make_random16_grad,
}


def main() -> None:
    names = sys.argv[1:] or list(FIXTURES)
    for name in names:
        np.savez(os.path.join(HERE, name + ".npz"), **FIXTURES[name]())
        print("wrote", name + ".npz", flush=True)


if __name__ == "__main__":
    main()
