"""Known answers of the reference's observable and jump-helper tests on the
device kernels: test/emu_sv/test_algebra.py:5-29 (expect_batch, apply) and
test/emu_sv/test_custom_callbacks.py:40-117 (occupation of
(|r..r> + |g..g>)/sqrt 2 is 1/2 on every qubit; the correlation of
(|rgrg..> + |grgr..>)/sqrt 2 is 1/2 where i + j is even, 0 elsewhere), for
state vectors and density matrices."""

import pytest
import torch

from emulators_amd.emu_sv import DensityMatrix, StateVector, SVConfig
from emulators_amd.emu_sv.algebra import apply, expect_batch
from emulators_amd.emu_sv.custom_callback_implementations import (
    correlation_matrix_sv_den_mat_impl, correlation_matrix_sv_impl,
    qubit_occupation_sv_den_mat_impl, qubit_occupation_sv_impl)

from test_oracle import (ALGEBRA_EXPECT, ALGEBRA_OPS, ALGEBRA_VEC, APPLY_Q0, APPLY_Q1,
                         APPLY_VEC, HADAMARD_LIKE)

pytestmark = pytest.mark.gpu
c128 = torch.complex128


def test_jump_helpers_known_answers(gpu_device):
    vec = torch.tensor(ALGEBRA_VEC, dtype=c128, device=gpu_device)
    got = expect_batch(vec, torch.tensor(ALGEBRA_OPS, dtype=c128), 2)
    assert torch.allclose(got, torch.tensor(ALGEBRA_EXPECT, dtype=c128), atol=1e-15)
    vec = torch.tensor(APPLY_VEC, dtype=c128, device=gpu_device)
    had = torch.tensor(HADAMARD_LIKE, dtype=c128)
    # apply returns a new vector and leaves its input as it was
    assert torch.allclose(apply(vec, 0, had).cpu(), torch.tensor(APPLY_Q0, dtype=c128))
    assert torch.allclose(apply(vec, 1, had).cpu(), torch.tensor(APPLY_Q1, dtype=c128))
    assert torch.equal(vec.cpu(), torch.tensor(APPLY_VEC, dtype=c128))


@pytest.mark.parametrize("density", [False, True])
def test_occupation_known_answer(gpu_device, density):
    n = 10
    state = StateVector.from_state_amplitudes(eigenstates=("r", "g"),
                                              amplitudes={"r" * n: 1.0, "g" * n: 1.0})
    cfg = SVConfig()
    if density:
        state = DensityMatrix.from_state_vector(state)
        occ = qubit_occupation_sv_den_mat_impl(None, config=cfg, state=state, hamiltonian=None)
    else:
        occ = qubit_occupation_sv_impl(None, config=cfg, state=state, hamiltonian=None)
    assert occ.device.type == "cpu"
    assert torch.allclose(occ, torch.full((n,), 0.5, dtype=torch.float64), atol=1e-14)


@pytest.mark.parametrize("density", [False, True])
def test_correlation_known_answer(gpu_device, density):
    n = 6
    state = StateVector.from_state_amplitudes(eigenstates=("r", "g"),
                                              amplitudes={"rg" * (n // 2): 1.0,
                                                          "gr" * (n // 2): 1.0})
    cfg = SVConfig()
    if density:
        state = DensityMatrix.from_state_vector(state)
        corr = correlation_matrix_sv_den_mat_impl(None, config=cfg, state=state, hamiltonian=None)
    else:
        corr = correlation_matrix_sv_impl(None, config=cfg, state=state, hamiltonian=None)
    i = torch.arange(n)
    want = torch.where((i[:, None] + i[None, :]) % 2 == 0, 0.5, 0.0).to(torch.float64)
    assert torch.allclose(corr, want, atol=1e-14)
