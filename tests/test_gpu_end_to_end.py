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
