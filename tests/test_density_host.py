"""DensityMatrix behaviour on the host, after the reference's
test/emu_sv/test_density_mat.py: overlaps (a closed form for one qubit,
Bell states), construction, states from vectors and amplitudes, the
reference's seeded sampling counts and trace normalisation."""

import math

import pytest
import torch

from emulators_amd.emu_sv import StateVector
from emulators_amd.emu_sv.density_matrix_state import DensityMatrix

C128 = torch.complex128


def _bell(phase=1.0):
    v = torch.zeros(4, dtype=C128)
    v[0], v[3] = 1 / math.sqrt(2), phase / math.sqrt(2)
    return torch.outer(v, v.conj())


def test_overlap_closed_forms():
    # one qubit, rho = [[a, c], [conj(c), 1 - a]]: Tr(rho^2) = 1 - 2a + 2a^2 + 2|c|^2
    a, c = 0.3, 0.2 - 0.1j
    rho = torch.tensor([[a, c], [c.conjugate(), 1 - a]], dtype=C128)
    d = DensityMatrix(rho, gpu=False)
    expected = 1 - 2 * a + 2 * a * a + 2 * abs(c) ** 2
    assert abs(complex(d.overlap(d)) - expected) < 1e-15
    bell = DensityMatrix(_bell(), gpu=False)
    assert complex(bell.overlap(bell)) == pytest.approx(1.0, abs=1e-15)
    ground = DensityMatrix.make(2, gpu=False)
    assert complex(bell.overlap(ground)) == pytest.approx(0.5, abs=1e-15)


@pytest.mark.parametrize("n", [2, 7])
def test_make(n):
    expected = torch.zeros(2**n, 2**n, dtype=C128)
    expected[0, 0] = 1.0
    assert torch.equal(DensityMatrix.make(n, gpu=False).data, expected)


def test_from_state_vector_and_amplitudes():
    v = torch.tensor([1, 0, 0, 1], dtype=C128) / math.sqrt(2)
    assert torch.allclose(DensityMatrix.from_state_vector(StateVector(v, gpu=False)).data, _bell())
    d, _ = DensityMatrix._from_state_amplitudes(eigenstates=("r", "g"), n_qudits=2,
                                               amplitudes={"gg": 1.0, "rr": 1.0j})
    assert torch.allclose(d.data.cpu(), _bell(1.0j))
    d3, _ = DensityMatrix._from_state_amplitudes(eigenstates=("r", "g"), n_qudits=3,
                                                amplitudes={"rrr": 1.0})
    expected = torch.zeros(8, 8, dtype=C128)
    expected[7, 7] = 1.0
    assert torch.allclose(d3.data.cpu(), expected)


@pytest.mark.usefixtures("restore_rng_state")
def test_seeded_sampling_counts():
    """The reference's counts at seed 1337 (torch.multinomial on |diag rho|)."""
    torch.manual_seed(1337)
    bell = DensityMatrix(_bell(), gpu=False).sample(1000)
    assert bell["11"] == 485 and bell["00"] == 515
    counts = DensityMatrix(torch.eye(8, dtype=C128), gpu=False).sample(1000)
    assert [counts[format(i, "03b")] for i in range(8)] == [144, 105, 130, 122, 103, 128, 150, 118]


def test_normalise_to_unit_trace():
    g = torch.Generator().manual_seed(5)
    a = torch.rand((32, 32), dtype=torch.float64, generator=g).to(C128)
    d = DensityMatrix(a.conj().T @ a, gpu=False)
    d._normalize()
    tr = complex(torch.trace(d.data))
    assert abs(tr.real - 1.0) < 1e-14 and abs(tr.imag) < 1e-16
