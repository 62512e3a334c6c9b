"""DenseState's overlap and sampling entry, after the reference's
test/emu_sv/test_dense_state.py (CPU states; no kernel runs)."""

from collections import Counter
from unittest.mock import patch

import pytest
import torch

from emulators_amd.emu_sv import DensityMatrix, StateVector
from emulators_amd.emu_sv import dense_state


def test_inner_and_overlap():
    """test_dense_state.py:9-24: i(|000> + |111>) against the density matrix
    of i(|001> + |111>) (neither normalised): |<a|b>|^2 = 1 both ways."""
    dtype = torch.complex128
    t1 = torch.tensor([1j, 0, 0, 0, 0, 0, 0, 1j], dtype=dtype)
    t2 = torch.tensor([0, 1j, 0, 0, 0, 0, 0, 1j], dtype=dtype)
    s1 = StateVector(t1, gpu=False)
    s2 = DensityMatrix.from_state_vector(StateVector(t2, gpu=False))
    expected = torch.tensor([1.0], dtype=dtype)
    assert torch.allclose(s1.overlap(s2).cpu(), expected)
    assert torch.allclose(s2.overlap(s1).cpu(), expected)


@pytest.mark.parametrize("density_matrix", [False, True])
def test_sample_goes_through_sample_indices(density_matrix):
    """test_dense_state.py:27-47: the probabilities |psi|^2 (or |diag rho|)
    reach sample_indices with the shot count, and its indices become the
    bitstring counts."""
    state = StateVector(torch.tensor([0.6, 0, 0, 0, 0, 0, 0, 0.8j], dtype=torch.complex128), gpu=False)
    if density_matrix:
        state = DensityMatrix.from_state_vector(state)
    with patch.object(dense_state, "sample_indices", autospec=True,
                      return_value=torch.tensor([7, 0, 7])) as mocked:
        counts = state.sample(num_shots=3)
    assert counts == Counter({"111": 2, "000": 1})
    mocked.assert_called_once()
    probabilities, num_shots = mocked.call_args.args
    assert num_shots == 3
    expected = torch.zeros(8, dtype=torch.float64)
    expected[0], expected[7] = 0.36, 0.64
    torch.testing.assert_close(probabilities, expected, rtol=0, atol=1e-15)
