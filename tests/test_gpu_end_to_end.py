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


def test_trajectories_with_badly_prepared_atoms(gpu_device):
    """The reference's test_trajectories (test/emu_sv/test_end_to_end.py:
    973-1003): three atoms 1e5 um apart, a pi pulse, state-preparation error
    1/3 with one badly prepared atom per trajectory (the reference mocks the
    draws to make atoms 0, 1, 2 bad in turn).  A bad atom's drive and
    interactions are switched off (sv_backend_impl.py:162-183), so it stays in
    |g> while the others end in |r>: the aggregated occupations are 2/3 each
    (its 0.6667 at atol 1e-4)."""
    import dataclasses

    from emulators_amd.emu_sv import Occupation, Results
    steps, step_ns = 10, 10.0
    omega = math.pi / (steps * step_ns * 1e-3)  # area pi
    seq = workloads.piecewise_sequence(np.array([[-1e5, 0.0], [1e5, 0.0], [0.0, 1e5]]),
                                       np.full(steps, omega), np.zeros(steps), step_ns=step_ns)
    cfg = SVConfig(observables=[Occupation(evaluation_times=[1.0])], log_level=40)
    omega_in = seq.omega.clone()
    runs = []
    for bad in range(3):  # the three runs share seq's tables
        flags = tuple(q == bad for q in range(3))
        sd = dataclasses.replace(seq, bad_atoms=flags, state_prep_error=1.0 / 3.0)
        r = SVBackend._run_from_sequence_data(sd, cfg)
        occ = torch.as_tensor(r.occupation[-1]).cpu()
        want = torch.tensor([0.0 if f else 1.0 for f in flags], dtype=torch.float64)
        assert torch.allclose(occ, want, atol=1e-9), (bad, occ)
        runs.append(r)
    assert torch.equal(seq.omega, omega_in)  # the caller's tables are left alone
    agg = torch.as_tensor(Results.aggregate(runs).occupation[-1]).cpu()
    assert torch.allclose(agg, torch.full((3,), 2.0 / 3.0, dtype=torch.float64), atol=1e-9)
    assert torch.allclose(agg, torch.full((3,), 0.6667, dtype=torch.float64), atol=1e-4)


def test_observable_times_as_in_pulser(gpu_device):
    """The reference's test_end_to_end_observable_time_as_in_pulser
    (test/emu_sv/test_end_to_end.py:949-970): a 20 ns constant pulse
    (Omega = 5) on two atoms 6 um apart, dt = 2.5 ns, occupation requested
    at [0, 1/3, 1]: the run reports exactly those times, 1/3 (6.67 ns) being
    off the dt grid (a shorter step lands on it, pulser_adapter.py:130-152)."""
    from emulators_amd.emu_base.pulser_adapter import (interaction_matrix_from_positions,
                                                        sequence_data_from_samples)
    from emulators_amd.emu_sv import Occupation
    duration, dt = 20, 2.5
    occ = Occupation(evaluation_times=[0.0, 1 / 3, 1.0])
    cfg = SVConfig(dt=dt, observables=[occ], log_level=40)
    samples = {q: {"amp": [5.0] * duration, "det": [0.0] * duration, "phase": [0.0] * duration}
               for q in ("q0", "q1")}
    pos = np.array([[-3.0, 0.0], [3.0, 0.0]])
    seq = sequence_data_from_samples(samples, ("q0", "q1"),
                                     interaction_matrix_from_positions(pos, workloads.C6), cfg, dt)
    assert any(abs(t - duration / 3) < 1e-9 for t in seq.target_times)
    res = SVBackend._run_from_sequence_data(seq, cfg)
    assert np.allclose(res.get_result_times(occ), [0.0, 1 / 3, 1.0])
    # the two atoms are equivalent: equal occupations at every reported time
    for o in res.occupation:
        o = torch.as_tensor(o).cpu()
        assert abs(float(o[0] - o[1])) < 1e-10


def test_initial_state_is_used_and_copied(gpu_device):
    """The reference's test_initial_state (test/emu_sv/test_end_to_end.py:
    336-358): a pulse that does nothing on two atoms 1e10 um apart, initial
    state |rr>: the final StateResult is that state (overlap 1) and a copy."""
    from emulators_amd.emu_sv import StateVector
    steps = 1
    seq = workloads.piecewise_sequence(np.array([[0.0, 0.0], [1e10, 0.0]]), np.zeros(steps),
                                       np.zeros(steps), step_ns=10.0)
    state = StateVector.from_state_amplitudes(eigenstates=("r", "g"), amplitudes={"rr": 1.0})
    assert abs(float(state.norm()) - 1.0) < 1e-12
    sr = StateResult(evaluation_times=[1.0])
    res = SVBackend._run_from_sequence_data(seq, SVConfig(observables=[sr], initial_state=state,
                                                          log_level=40))
    out = res.get_result(sr, 1.0)
    assert abs(complex(out.inner(state)).real - 1.0) < 1e-12
    assert out is not state


def test_initial_density_matrix_under_dephasing(gpu_device):
    """The reference's test_initial_state_with_den_matrix (test/emu_sv/
    test_end_to_end.py:385-419): five atoms 1e10 um apart, no drive,
    dephasing 0.5 / us, initial density matrix |rrrrr><rrrrr| on the CPU: the
    final StateResult equals it (atol 1e-8; dephasing leaves a basis state
    alone) and is a copy."""
    import dataclasses

    from emulators_amd.emu_sv import DensityMatrix
    from emulators_amd.emu_sv.sv_config import Solver
    n, steps = 5, 1
    seq = workloads.piecewise_sequence(np.array([[1e10 * q, 0.0] for q in range(n)]), np.zeros(steps),
                                       np.zeros(steps), step_ns=10.0)
    c = math.sqrt(0.5 / 2)  # dephasing sqrt(rate / 2) Z
    seq = dataclasses.replace(seq, lindblad_ops=[torch.tensor([[c, 0], [0, -c]], dtype=torch.complex128)],
                              has_stochastic_noise=False)
    state = DensityMatrix.from_state_amplitudes(eigenstates=("r", "g"), amplitudes={"r" * n: 1.0})
    state.data = state.data.to("cpu")
    assert abs(complex(state.data.trace()) - 1.0) < 1e-12
    sr = StateResult(evaluation_times=[1.0])
    res = SVBackend._run_from_sequence_data(seq, SVConfig(observables=[sr], initial_state=state,
                                                          solver=Solver.LINDBLAD, log_level=40))
    out = res.get_result(sr, 1.0)
    assert torch.allclose(out.data.cpu(), state.data.cpu(), atol=1e-8)
    assert out is not state
