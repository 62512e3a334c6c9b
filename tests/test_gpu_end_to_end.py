"""Whole backend runs with closed-form answers, after the reference's
test/emu_sv/test_end_to_end.py cases that need no pulser objects: a pi/2
pulse with a phase on one atom (the state cos(pi/4)|g> - i e^{i phi}
sin(pi/4)|r>), and an energy observable at t = 0 under Lindblad noise,
which is real."""

import math

import numpy as np
import pytest
import torch

from emulators_amd import workloads
from emulators_amd.emu_sv import Energy, StateResult, SVBackend, SVConfig

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("phi", [0.0, 0.7, -2.1])
@pytest.mark.parametrize("precision", ["complex128", "complex64"])
def test_pi_half_pulse_with_phase(gpu_device, phi, precision):
    omega, steps, step_ns = 2 * math.pi, 25, 10.0  # Omega * 250 ns = pi/2
    seq = workloads.piecewise_sequence(np.zeros((1, 2)), np.full(steps, omega), np.zeros(steps),
                                       step_ns=step_ns, phis=np.full(steps, phi))
    cfg = SVConfig(observables=[StateResult(evaluation_times=[1.0])], log_level=40,
                   precision=precision)
    res = SVBackend._run_from_sequence_data(seq, cfg)
    got = res.state[-1].data.cpu()
    theta = omega * steps * step_ns * 1e-3 / 2
    expected = torch.tensor([math.cos(theta), -1j * complex(math.cos(phi), math.sin(phi)) * math.sin(theta)],
                            dtype=torch.complex128)
    tol = 1e-8 if precision == "complex128" else 1e-6
    assert float((got - expected).abs().max()) < tol


def test_energy_at_time_zero_is_real_under_lindblad_noise(gpu_device):
    import dataclasses
    rate = 0.2  # depolarizing: sqrt(rate / 4) X, Y, Z
    paulis = [torch.tensor(m, dtype=torch.complex128) for m in
              ([[0, 1], [1, 0]], [[0, -1j], [1j, 0]], [[1, 0], [0, -1]])]
    lind = [math.sqrt(rate / 4) * p for p in paulis]
    seq = workloads.piecewise_sequence(workloads.chain_positions(2, 6.0), np.full(5, 3.0),
                                       np.full(5, -1.0))
    seq = dataclasses.replace(seq, lindblad_ops=lind, has_stochastic_noise=False)
    res = SVBackend._run_from_sequence_data(seq, SVConfig(observables=[Energy(evaluation_times=[0.0, 1.0])],
                                                          log_level=40))
    e0 = torch.as_tensor(res.energy[0])
    assert not torch.is_complex(e0) or float(e0.imag) == 0.0
    assert math.isfinite(float(torch.as_tensor(res.energy[-1]).real))


@pytest.mark.usefixtures("restore_rng_state")
def test_spontaneous_emission_rate(gpu_device):
    """The reference's test_end_to_end_spontaneous_emission_rate
    (test/emu_sv/test_end_to_end.py:459-545) on the same sequence data: two
    atoms 1e10 um apart, no drive, relaxation 0.1 / us (sqrt(0.1) |g><r|)
    for 10 us from |rr>, dt = 100 ns, Krylov tolerance 1e-5.
    * Lindblad solver: each occupation exp(-1) (its 0.3678 at atol 1e-4) and
      the density matrix diag((1-e^-1)^2, e^-1 (1-e^-1) twice, e^-2) (its
      table at atol 1e-4);
    * Monte-Carlo solver, 10 trajectories after random.seed(0xDEADBEEF): its
      occupations [0.5, 0.2] exactly (each trajectory ends with each atom in
      |g> or |r>; the jump thresholds and the jumping qubit are Python
      `random` draws in the reference's order)."""
    import dataclasses
    import random

    from emulators_amd.emu_sv import DensityMatrix, Occupation, StateVector
    from emulators_amd.emu_sv.sv_config import Solver
    torch.manual_seed(31415)
    random.seed(0xDEADBEEF)
    steps, step_ns, rate = 100, 100.0, 0.1
    seq = workloads.piecewise_sequence(np.array([[0.0, 0.0], [1e10, 0.0]]), np.zeros(steps),
                                       np.zeros(steps), step_ns=step_ns)
    relax = [math.sqrt(rate) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128)]
    seq = dataclasses.replace(seq, lindblad_ops=relax, has_stochastic_noise=False)
    sv0 = StateVector.from_state_amplitudes(eigenstates=("r", "g"), amplitudes={"rr": 1.0})

    cfg = SVConfig(initial_state=DensityMatrix.from_state_vector(sv0), dt=step_ns, krylov_tolerance=1e-5,
                   observables=[StateResult(evaluation_times=[1.0]), Occupation(evaluation_times=[1.0])],
                   solver=Solver.LINDBLAD, log_level=40)
    res = SVBackend._run_from_sequence_data(seq, cfg)
    e = math.exp(-rate * steps * step_ns * 1e-3)
    occ = torch.as_tensor(res.occupation[-1]).cpu()
    assert torch.allclose(occ, torch.full((2,), e, dtype=torch.float64), atol=1e-8)
    assert torch.allclose(occ, torch.full((2,), 0.3678, dtype=torch.float64), atol=1e-4)
    rho = res.state[-1].data.cpu()
    want = torch.diag(torch.tensor([(1 - e) ** 2, e * (1 - e), e * (1 - e), e * e], dtype=torch.complex128))
    # the run's Krylov tolerance is 1e-5
    assert float((rho - want).abs().max()) < 1e-5, (rho, want)
    ref_table = torch.diag(torch.tensor([0.3996, 0.2325, 0.2325, 0.1353], dtype=torch.complex128))
    assert torch.allclose(rho, ref_table, atol=1e-4)

    cfg_mc = SVConfig(initial_state=sv0, dt=step_ns, krylov_tolerance=1e-5,
                      observables=[Occupation(evaluation_times=[1.0])], n_trajectories=10,
                      solver=Solver.MONTECARLO, log_level=40)
    runs = [SVBackend._run_from_sequence_data(seq, cfg_mc) for _ in range(10)]
    occ_mc = torch.stack([torch.as_tensor(r.occupation[-1]).cpu() for r in runs]).mean(0)
    assert torch.allclose(occ_mc, torch.tensor([0.5, 0.2], dtype=torch.float64), atol=1e-4)


@pytest.mark.usefixtures("restore_rng_state")
def test_random_effective_noise_trajectories(gpu_device):
    """The reference's test_end_to_end_random_effective_noise
    (test/emu_sv/test_end_to_end.py:548-600) through SVBackend on the same
    sequence data: torch.manual_seed(31415) draws the effective-noise
    operator (torch.rand(2, 2) on the CPU, rate 30, pulser basis flipped),
    two atoms 8 um apart, no drive, 1000 ns at dt = 10 ns from |rr>, 10
    Monte-Carlo trajectories after random.seed(0xDEADBEEF): its mean final
    occupations [0.8508, 0.8536] at its atol 1e-4."""
    import dataclasses
    import random
    from types import SimpleNamespace

    from emulators_amd.emu_base.jump_lindblad_operators import get_all_lindblad_noise_operators
    from emulators_amd.emu_sv import Occupation, StateVector
    from emulators_amd.emu_sv.sv_config import Solver
    torch.manual_seed(31415)
    random.seed(0xDEADBEEF)
    op = torch.rand(2, 2, dtype=torch.float64, device="cpu")
    nm = SimpleNamespace(noise_types=("eff_noise",), eff_noise_rates=(30.0,), eff_noise_opers=(op,),
                         relaxation_rate=0.0, dephasing_rate=0.0, hyperfine_dephasing_rate=0.0,
                         depolarizing_rate=0.0)
    steps = 100
    seq = workloads.piecewise_sequence(np.array([[0.0, 0.0], [8.0, 0.0]]), np.zeros(steps),
                                       np.zeros(steps), step_ns=10.0)
    seq = dataclasses.replace(seq, lindblad_ops=get_all_lindblad_noise_operators(nm),
                              has_stochastic_noise=False)
    sv0 = StateVector.from_state_amplitudes(eigenstates=("r", "g"), amplitudes={"rr": 1.0})
    cfg = SVConfig(initial_state=sv0, dt=10.0, observables=[Occupation(evaluation_times=[1.0])],
                   n_trajectories=10, solver=Solver.MONTECARLO, log_level=40)
    runs = [SVBackend._run_from_sequence_data(seq, cfg) for _ in range(10)]
    occ = torch.stack([torch.as_tensor(r.occupation[-1]).cpu() for r in runs]).mean(0)
    assert torch.allclose(occ, torch.tensor([0.8508, 0.8536], dtype=torch.float64), atol=1e-4)
