"""The device-driven batch engine against the single-register engine and the
oracle (registers of different N, pulses, step counts and phases)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu


def _random_register(n, steps, seed, with_phase):
    from emulators_amd.emu_sv.batch import BatchRegister

    rng = np.random.default_rng(seed)
    om = rng.uniform(0.0, 12.0, size=(steps, n))
    de = rng.uniform(-8.0, 8.0, size=(steps, n))
    ph = rng.uniform(0.0, 6.0, size=(steps, n)) if with_phase else np.zeros((steps, n))
    u = rng.uniform(0.0, 25.0, size=(n, n))
    u = np.triu(u, 1)
    u = u + u.T
    dts = rng.uniform(0.005, 0.02, size=steps)
    return BatchRegister(om, de, ph, u, dts)


def _sequential(reg, dtype, tol):
    """The single-register fused engine, step by step, with one read-back per
    iteration (no speculative iterations, so its matvec count is the
    algorithm's: m + 1 per step, as the batch engine's)."""
    import importlib

    from emulators_amd.emu_base.math.krylov_exp import KrylovStats, ScaledHamiltonian, krylov_exp
    from emulators_amd.emu_sv import RydbergHamiltonian

    kx = importlib.import_module("emulators_amd.emu_base.math.krylov_exp")
    n = reg.nqubits
    psi = torch.zeros(2**n, dtype=dtype, device="cuda")
    psi[0] = 1.0
    KrylovStats.reset()
    saved, kx.PIPELINE = kx.PIPELINE, False
    try:
        for k in range(reg.steps):
            ham = RydbergHamiltonian(torch.as_tensor(reg.omegas[k]), torch.as_tensor(reg.deltas[k]),
                                     torch.as_tensor(reg.phis[k]), torch.as_tensor(reg.interaction_matrix),
                                     device="cuda")
            psi = krylov_exp(ScaledHamiltonian(ham, -1j * float(reg.dts[k])), psi, tol, tol)
    finally:
        kx.PIPELINE = saved
    return psi, KrylovStats.matvecs


def _oracle(reg, tol):
    n = reg.nqubits
    psi = torch.zeros(2**n, dtype=torch.complex128)
    psi[0] = 1.0
    for k in range(reg.steps):
        ham = orc.OracleHamiltonian(torch.as_tensor(reg.omegas[k]), torch.as_tensor(reg.deltas[k]),
                                    torch.as_tensor(reg.phis[k]), torch.as_tensor(reg.interaction_matrix))
        dt = float(reg.dts[k])
        psi = orc.krylov_exp(lambda x: -1j * dt * (ham * x), psi.clone(), tol, tol)
    return psi


SHAPES = [(3, 4, False), (6, 3, True), (9, 5, False), (11, 2, True), (12, 4, False), (14, 3, False)]


@pytest.mark.parametrize("tile_log,min_contig", [(0, 0), (8, 2)])
def test_batch_c128_matches_single_register_engine(gpu_device, tile_log, min_contig):
    from emulators_amd.emu_sv.batch import evolve_batch

    regs = [_random_register(n, s, 100 + i, ph) for i, (n, s, ph) in enumerate(SHAPES)]
    tol = 1e-10
    res = evolve_batch(regs, dtype=torch.complex128, tolerance=tol, max_krylov_dim=60,
                       tile_log=tile_log, min_contig=min_contig, check_every=4)
    assert res.failed == [], res.phases
    for reg, got, mv in zip(regs, res.states, res.matvecs):
        ref, ref_mv = _sequential(reg, torch.complex128, tol)
        assert mv == ref_mv, (reg.nqubits, mv, ref_mv)
        err = float((got - ref).norm())
        assert err < 1e-12 * max(1.0, float(ref.norm())), (reg.nqubits, err)
        ora = _oracle(reg, tol)
        fid = float(orc.fidelity(got.cpu(), ora))
        assert 1 - fid < 1e-10, (reg.nqubits, 1 - fid)


def test_batch_c64_fidelity(gpu_device):
    from emulators_amd.emu_sv.batch import evolve_batch

    regs = [_random_register(n, s, 200 + i, ph) for i, (n, s, ph) in enumerate(SHAPES)]
    res = evolve_batch(regs, dtype=torch.complex64, tolerance=1e-10)
    assert res.failed == []
    for reg, got in zip(regs, res.states):
        ora = _oracle(reg, 1e-10)
        fid = float(orc.fidelity(got.cpu().to(torch.complex128), ora))
        assert 1 - fid < 1e-6, (reg.nqubits, 1 - fid)


def test_batch_reports_nonconvergence(gpu_device):
    """A Krylov dimension too small for the tolerance stops the register
    (phase 'failed'); the others finish."""
    from emulators_amd.emu_sv.batch import BatchRegister, evolve_batch

    hard = _random_register(10, 2, 300, False)
    hard = BatchRegister(hard.omegas * 50, hard.deltas, hard.phis, hard.interaction_matrix * 50,
                         hard.dts)
    easy = _random_register(2, 2, 301, False)  # dimension 4: breaks down within 4 iterations
    res = evolve_batch([hard, easy], dtype=torch.complex128, tolerance=1e-12, max_krylov_dim=4)
    assert res.phases[0] == "failed"
    assert res.failed == [0]


def test_batch_workload_instances_agree(gpu_device):
    """Four registers of the 256-instance batching workload (N 8..20):
    complex64 batch against the complex128 single-register engine."""
    from emulators_amd import workloads
    from emulators_amd.emu_sv.batch import evolve_batch, registers_from_sequences

    inst = workloads.batch_instances()
    pick = [i for i, x in enumerate(inst) if x["duration"] <= 800][:4]
    seqs = [workloads.batch_sequence(inst[i]) for i in pick]
    regs = registers_from_sequences(seqs)
    res = evolve_batch(regs, dtype=torch.complex64)
    assert res.failed == []
    for reg, got in zip(regs, res.states):
        ref, _ = _sequential(reg, torch.complex128, 1e-10)
        fid = float(torch.abs(torch.vdot(ref, got.to(torch.complex128))) ** 2)
        assert 1 - fid < 1e-6, (reg.nqubits, 1 - fid)


def _oracle_sample(inst, per_n):
    """Indices of the workload's registers sampled for the oracle check:
    the first ``per_n[N]`` registers of each size N (in workload order)."""
    pick, seen = [], {}
    for i, x in enumerate(inst):
        n = len(x["positions"])
        if seen.get(n, 0) < per_n.get(n, 0):
            pick.append(i)
            seen[n] = seen.get(n, 0) + 1
    return pick


def test_batch_256_registers_full_workload(gpu_device):
    """The whole batching workload in one evolve_batch call: 256 seeded
    registers, N ~ U{8..20}, 500-2000 ns, complex64 with FP64 Lanczos
    scalars.  No register fails, every final state keeps unit norm to 1e-6,
    and a seeded sample of 16 registers (N 8..14) matches the oracle's
    step-by-step Krylov evolution to 1 - F <= 1e-6."""
    import os

    from emulators_amd import workloads
    from emulators_amd.emu_sv.batch import evolve_batch, registers_from_sequences

    inst = workloads.batch_instances()
    assert len(inst) == 256
    regs = registers_from_sequences([workloads.batch_sequence(x) for x in inst])
    res = evolve_batch(regs, dtype=torch.complex64)
    assert res.failed == [], res.failed
    assert len(res.states) == 256
    for i, st in enumerate(res.states):
        nrm = float(st.to(torch.complex128).norm())
        assert abs(nrm - 1.0) < 1e-6, (i, nrm)
    torch.set_num_threads(min(16, os.cpu_count() or 1))
    sample = _oracle_sample(inst, {8: 2, 9: 2, 10: 2, 11: 2, 12: 2, 13: 3, 14: 3})
    assert len(sample) == 16
    worst = 0.0
    for i in sample:
        ora = _oracle(regs[i], 1e-10)
        fid = float(orc.fidelity(res.states[i].cpu().to(torch.complex128), ora))
        worst = max(worst, 1.0 - fid)
        assert 1.0 - fid <= 1e-6, (i, regs[i].nqubits, 1.0 - fid)
    print(f"256 registers: 16 sampled against the oracle, worst 1 - F = {worst:.2e}")


def test_batch_per_register_schedules(gpu_device):
    """Registers carrying their own H.psi schedules (mixed in one batch, and
    the autotuned per-size picks) evolve exactly as with one schedule for
    all: the same matvec counts, states within 1e-12 (complex128)."""
    from emulators_amd.emu_sv.batch import BatchRegister, evolve_batch

    shapes = [(12, 3, False), (13, 3, True), (14, 2, False), (12, 2, False)]
    regs = [_random_register(n, s, 400 + i, ph) for i, (n, s, ph) in enumerate(shapes)]
    ref = evolve_batch(regs, dtype=torch.complex128, tile_log=11, min_contig=2)
    mixed = [BatchRegister(r.omegas, r.deltas, r.phis, r.interaction_matrix, r.dts, None,
                           *((10, 2) if i % 2 else (12, 3))) for i, r in enumerate(regs)]
    for res in (evolve_batch(mixed, dtype=torch.complex128, tile_log=11, min_contig=2),
                evolve_batch(regs, dtype=torch.complex128, tiles="auto")):
        assert res.failed == [] and res.matvecs == ref.matvecs
        for a, b in zip(res.states, ref.states):
            assert float((a - b).norm()) < 1e-12
