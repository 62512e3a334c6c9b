"""The native inverse-CDF sampler (emu_sv_sample) against the reference's
inverse-CDF branch of sample_indices (emu_sv/utils.py:41-49): normalised
float64 cumulative weights, searchsorted(cdf / total, u, right=True).

Bars:
* the reference's own golden vectors (test/emu_sv/test_state.py:140-198):
  its (weights, draws, indices) tables and its seed-1337 counts, fed to the
  device sampler with the same draws: identical indices and counts;
* dyadic probabilities, whose cumulative sums are exact in any summation
  order: identical indices for every draw, including draws that sit exactly
  on a cumulative boundary;
* random probabilities: the device sums chunks in a different order than
  the oracle's sequential cumsum, so a draw within 1e-12 of a boundary may
  land on either side; every other draw gives the oracle's index;
* zero-probability indices are never drawn; 200k draws within 5 sigma.
"""

import math

import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu
SEED = 1337


def _check(idx, p, u):
    ref = orc.inverse_cdf_indices(p, u)
    cdf = torch.cumsum(p, 0) / p.sum()
    near = torch.min((cdf.unsqueeze(0) - u.unsqueeze(1)).abs(), dim=1).values < 1e-12
    ok = (idx == ref) | near
    assert bool(ok.all()), f"{int((~ok).sum())} draws differ"
    assert int(near.sum()) <= 2, f"{int(near.sum())} draws within 1e-12 of a boundary"
    assert bool((p[idx] > 0).all())


def _diag_sample(weights, u, device):
    """Native sampler in density-matrix mode on diag(weights) (padded with
    zero weights to a power of two): |diag rho| is exact for these weights."""
    from emulators_amd import _native as native
    w = torch.as_tensor(weights, dtype=torch.float64)
    n = max(1, (w.numel() - 1).bit_length())
    d = torch.zeros(2**n, dtype=torch.float64)
    d[: w.numel()] = w
    rho = torch.diag(d).to(torch.complex128)
    return native.sample(n, rho.to(device), u.to(device), density_matrix=True).cpu()


@pytest.mark.parametrize("weights,draws,expected", [
    ([0.0, 1.0, 0.0, 3.0, 6.0, 0.0], [0.0, 0.05, 0.1, 0.25, 0.4, 0.7, math.nextafter(1.0, 0.0)],
     [1, 1, 3, 3, 4, 4, 4]),
    ([1.0, 3.0], [1.0], [1]),
    ([1.0, 1e-10, 1.0], [0.5], [1]),
])
def test_reference_inverse_cdf_tables(gpu_device, weights, draws, expected):
    """test_state.py:140-178 (weights scaled by 2.5 as there)."""
    got = _diag_sample([2.5 * x for x in weights], torch.tensor(draws, dtype=torch.float64), gpu_device)
    assert got.tolist() == expected


@pytest.mark.usefixtures("restore_rng_state")
def test_reference_seeded_inverse_cdf_counts(gpu_device):
    """test_state.py:186-198: the reference's seed-1337 draws (its CPU
    stream) through the device sampler give its counts exactly."""
    torch.manual_seed(SEED)
    u = torch.rand(100_000, dtype=torch.float64)
    got = _diag_sample([2.5, 0.0, 7.5, 15.0, 0.0], u, gpu_device)
    assert torch.bincount(got, minlength=5).tolist() == [10_007, 0, 30_052, 59_941, 0]


@pytest.mark.parametrize("n", [3, 10, 13, 16])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_exact_on_dyadic_probabilities(gpu_device, n, dtype):
    """Amplitudes (a + ib) / 64 with small integers a, b: every |psi|^2 and
    every cumulative sum is exact in FP64, so the device's decision for each
    draw is the reference's, with no tolerance at boundaries."""
    from emulators_amd import _native as native
    g = torch.Generator().manual_seed(50 + n)
    re = torch.randint(-7, 8, (2**n,), generator=g).to(torch.float64)
    im = torch.randint(-7, 8, (2**n,), generator=g).to(torch.float64)
    re[torch.rand(2**n, generator=g) < 0.3] = 0.0
    im[re == 0.0] = 0.0
    re[0] = 1.0
    psi = torch.complex(re, im) / 64.0
    u = torch.rand(8192, dtype=torch.float64, generator=g)
    idx = native.sample(n, psi.to(gpu_device, dtype), u.to(gpu_device)).cpu()
    p = psi.abs() ** 2
    assert torch.equal(idx, orc.inverse_cdf_indices(p, u))


def test_exact_on_draws_at_cumulative_boundaries(gpu_device):
    """Integer weights summing to 2^20: every normalised cumulative weight is
    exact, so draws placed on them (and one ulp either side) test the
    right-continuous rule (a draw on a boundary goes to the next bin)."""
    g = torch.Generator().manual_seed(9)
    w = torch.randint(0, 40, (256,), generator=g).to(torch.float64)
    w[::7] = 0.0
    w[-1] += 2**20 - w.sum()
    cdf = torch.cumsum(w, 0) / 2**20
    on = cdf[:-1][w[:-1] > 0]
    u = torch.cat([on, torch.nextafter(on, torch.zeros_like(on)), torch.nextafter(on, torch.ones_like(on))])
    u = u[u < 1.0]
    got = _diag_sample(w.tolist(), u, gpu_device)
    assert torch.equal(got, orc.inverse_cdf_indices(w, u))


@pytest.mark.usefixtures("restore_rng_state")
def test_dense_state_sample_draws_like_torch_multinomial(gpu_device):
    """Below 2^24 categories DenseState.sample on the GPU is torch.multinomial
    on |psi|^2 (emu_sv/utils.py:38-39): the same seed gives the same counts."""
    from collections import Counter

    from emulators_amd.emu_sv import StateVector
    g = torch.Generator().manual_seed(4)
    psi = torch.randn(2**6, dtype=torch.complex128, generator=g)
    psi /= psi.norm()
    torch.manual_seed(21)
    counts = StateVector(psi, gpu=True).sample(num_shots=500)
    torch.manual_seed(21)
    ref = torch.multinomial(psi.to(gpu_device).abs() ** 2, 500, replacement=True).cpu()
    assert counts == Counter(format(int(i), "06b") for i in ref.tolist())


@pytest.mark.parametrize("n", [1, 3, 6, 12, 13, 16])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_state_vector_sampling(gpu_device, n, dtype):
    from emulators_amd import _native as native
    print(f"sample n={n} {dtype}", flush=True)
    g = torch.Generator().manual_seed(n)
    psi = torch.randn(2**n, dtype=torch.complex128, generator=g)
    psi[torch.rand(2**n, generator=g) < 0.3] = 0  # zero-probability entries
    psi[0] = 1.0
    psi = psi.to(dtype)
    u = torch.rand(4096, dtype=torch.float64, generator=g)
    u[:3] = torch.tensor([0.0, 0.5, 1 - 2**-53], dtype=torch.float64)
    idx = native.sample(n, psi.to(gpu_device), u.to(gpu_device)).cpu()
    p = psi.to(torch.complex128).abs() ** 2
    _check(idx, p, u)


@pytest.mark.parametrize("n", [2, 5, 7])
def test_density_matrix_sampling(gpu_device, n):
    from emulators_amd import _native as native
    g = torch.Generator().manual_seed(100 + n)
    a = torch.randn(2**n, 2**n, dtype=torch.complex128, generator=g)
    rho = a @ a.conj().T
    rho = rho / rho.trace()
    u = torch.rand(2048, dtype=torch.float64, generator=g)
    idx = native.sample(n, rho.to(gpu_device), u.to(gpu_device), density_matrix=True).cpu()
    _check(idx, rho.diagonal().abs(), u)


def test_sampling_statistics(gpu_device):
    from emulators_amd import _native as native
    n = 4
    g = torch.Generator().manual_seed(7)
    psi = torch.randn(2**n, dtype=torch.complex128, generator=g)
    psi /= psi.norm()
    u = torch.rand(200_000, dtype=torch.float64, generator=g)
    idx = native.sample(n, psi.to(gpu_device), u.to(gpu_device)).cpu()
    counts = torch.bincount(idx, minlength=2**n).double()
    p = psi.abs() ** 2
    exp_ = 200_000 * p
    sigma = torch.sqrt(200_000 * p * (1 - p))
    assert bool(((counts - exp_).abs() <= 5 * sigma + 1).all())


def test_bitstrings_observable(gpu_device):
    from emulators_amd.emu_sv import StateVector
    psi = torch.zeros(8, dtype=torch.complex128)
    psi[0b101] = 1.0
    counts = StateVector(psi, gpu=True).sample(num_shots=100)
    assert counts == {"101": 100}
