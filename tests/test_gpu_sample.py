"""The native inverse-CDF sampler (emu_sv_sample) against the oracle's
searchsorted form of the reference's sample_indices, with the same draws.

Bar: identical indices for every draw (draws that land within 1e-12 of a
cumulative boundary are skipped: the two FP64 summation orders may place
them on either side); zero-probability indices are never drawn; the counts
of 200k draws within 5 sigma of the distribution.
"""

import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu


def _check(idx, p, u):
    ref = orc.inverse_cdf_indices(p, u)
    cdf = torch.cumsum(p, 0) / p.sum()
    near = torch.min((cdf.unsqueeze(0) - u.unsqueeze(1)).abs(), dim=1).values < 1e-12
    ok = (idx == ref) | near
    assert bool(ok.all()), f"{int((~ok).sum())} draws differ"
    assert bool((p[idx] > 0).all())


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
