"""Sampling (BitStrings) on the host against the reference's own golden
vectors (test/emu_sv/test_state.py:105-205): the seeded multinomial counts
of a state, the seeded counts of the inverse-CDF branch, and its explicit
(weights, draws, indices) tables, for the package's sample_indices and for
the oracle's restatement that the GPU sampler is checked against."""

import math
from collections import Counter
from unittest.mock import patch

import pytest
import torch

from oracle import emu_sv_oracle as orc

SEED = 1337
TABLES = [
    # normalised CDF, exact in float64: [0, 0.1, 0.1, 0.4, 1, 1]; draws on CDF
    # values go to the next bin; zero-width bins are never hit
    ([0.0, 1.0, 0.0, 3.0, 6.0, 0.0], [0.0, 0.05, 0.1, 0.25, 0.4, 0.7, math.nextafter(1.0, 0.0)],
     [1, 1, 3, 3, 4, 4, 4]),
    ([1.0, 3.0], [1.0], [1]),  # a draw of 1.0 is clamped to the last index
    ([1.0, 1e-10, 1.0], [0.5], [1]),  # a bin below float32 resolution
]


@pytest.mark.usefixtures("restore_rng_state")
def test_reference_seeded_state_sample():
    """test_state.py:105-116: seed 1337, (|000> + |111>)/sqrt2 on the CPU,
    1000 shots -> 515 x '000', 485 x '111' (torch.multinomial's stream)."""
    from emulators_amd.emu_sv import StateVector

    torch.manual_seed(SEED)
    f = 1.0 / torch.sqrt(torch.tensor(2.0))
    state = StateVector(torch.tensor([f, 0, 0, 0, 0, 0, 0, f], dtype=torch.complex128), gpu=False)
    counts = state.sample(num_shots=1000)
    assert counts["111"] == 485 and counts["000"] == 515 and counts["001"] == 0


@pytest.mark.usefixtures("restore_rng_state")
def test_reference_seeded_inverse_cdf_counts():
    """test_state.py:186-198: seed 1337, weights 2.5 * [1, 0, 3, 6, 0],
    100000 draws through the inverse-CDF branch -> [10007, 0, 30052, 59941, 0]."""
    from emulators_amd.emu_sv.utils import sample_indices

    w = 2.5 * torch.tensor([1.0, 0.0, 3.0, 6.0, 0.0], dtype=torch.float64)
    torch.manual_seed(SEED)
    got = torch.bincount(sample_indices(w, 100_000, max_categories=1), minlength=5)
    assert got.tolist() == [10_007, 0, 30_052, 59_941, 0]
    torch.manual_seed(SEED)
    u = torch.rand(100_000, dtype=torch.float64)
    assert torch.bincount(orc.inverse_cdf_indices(w, u), minlength=5).tolist() == got.tolist()


@pytest.mark.parametrize("weights,draws,expected", TABLES)
def test_reference_inverse_cdf_tables(weights, draws, expected):
    """test_state.py:140-178: given draws map to the given indices."""
    from emulators_amd.emu_sv.utils import sample_indices

    p = 2.5 * torch.tensor(weights, dtype=torch.float64)
    u = torch.tensor(draws, dtype=torch.float64)
    assert orc.inverse_cdf_indices(p, u).tolist() == expected
    with patch.object(torch, "rand", autospec=True, return_value=u) as rand:
        out = sample_indices(p, len(draws), max_categories=1)
    rand.assert_called_once_with(len(draws), dtype=torch.float64, device=p.device)
    assert out.tolist() == expected


def test_oracle_inverse_cdf_brute_force():
    """The oracle against a plain scan of the normalised cumulative weights."""
    g = torch.Generator().manual_seed(3)
    p = torch.rand(300, dtype=torch.float64, generator=g)
    p[torch.rand(300, generator=g) < 0.3] = 0.0
    u = torch.rand(500, dtype=torch.float64, generator=g)
    cdf = torch.cumsum(p, 0)
    cdf = cdf / cdf[-1]
    for k in range(u.numel()):
        ref = next(i for i in range(300) if cdf[i] > u[k])
        assert int(orc.inverse_cdf_indices(p, u[k:k + 1])[0]) == ref


@pytest.mark.parametrize("size", [2**24, 2**24 + 1])
def test_multinomial_up_to_max_categories(size):
    """utils.py:38-39 (test_state.py:119-134): torch.multinomial up to 2**24
    categories, the inverse-CDF branch beyond."""
    from emulators_amd.emu_sv.utils import sample_indices

    p = torch.ones(1, dtype=torch.float64).expand(size)
    with patch.object(torch, "multinomial", autospec=True) as multinomial:
        sample_indices(p, 10)
    if size <= 2**24:
        multinomial.assert_called_once_with(p, 10, replacement=True)
    else:
        multinomial.assert_not_called()


@pytest.mark.parametrize("density_matrix", [False, True])
def test_dense_state_sample_goes_through_multinomial_on_small_states(density_matrix):
    """DenseState.sample below 2**24 categories draws with torch.multinomial
    on the probabilities |psi|^2 (|diag rho|), as dense_state.py:75-95."""
    from emulators_amd.emu_sv import StateVector
    from emulators_amd.emu_sv.density_matrix_state import DensityMatrix

    amplitudes = torch.tensor([0.6, 0, 0, 0, 0, 0, 0, 0.8j], dtype=torch.complex128)
    state = StateVector(amplitudes, gpu=False)
    if density_matrix:
        state = DensityMatrix.from_state_vector(state)
    with patch.object(torch, "multinomial", autospec=True,
                      return_value=torch.tensor([7, 0, 7])) as multinomial:
        counts = state.sample(num_shots=3)
    assert counts == Counter({"111": 2, "000": 1})
    probabilities, shots = multinomial.call_args.args
    assert shots == 3 and multinomial.call_args.kwargs == {"replacement": True}
    expected = torch.zeros(8, dtype=torch.float64)
    expected[0], expected[7] = 0.36, 0.64
    assert torch.allclose(probabilities, expected, atol=1e-15, rtol=0)
