"""The oracle's restatement of the Lindblad density-matrix path.

Pins: the textbook Lindblad generator -i[H, rho] + sum_k (A rho A^H -
1/2 {A^H A, rho}) as a dense 4^N superoperator; the reference's own checks
of RydbergLindbladian (test/emu_sv/test_linbladian_operator.py): H @ rho
without jumps equals H rho - (H rho)^H for any rho, and Tr(H rho) for a
diagonal rho under U n n couplings; trace preservation of the generator.
"""

import math

import torch

from oracle import emu_sv_oracle as orc

DT = torch.complex128


def _params(n, seed):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, generator=g, dtype=torch.float64)
    de = torch.rand(n, generator=g, dtype=torch.float64)
    ph = torch.rand(n, generator=g, dtype=torch.float64)
    u = torch.rand(n, n, generator=g, dtype=torch.float64)
    return om, de, ph, (u + u.T) / 2


def _lindblads(seed):
    g = torch.Generator().manual_seed(seed)
    return [math.sqrt(1 / 3) * torch.rand(2, 2, generator=g, dtype=DT),
            math.sqrt(1 / 2) * torch.rand(2, 2, generator=g, dtype=DT)]


def test_oracle_lindbladian_vs_dense_generator():
    n = 3
    om, de, ph, u = _params(n, 1)
    L = _lindblads(2)
    rho = torch.rand(8, 8, dtype=DT)
    rho = 0.5 * (rho + rho.conj().T)
    got = -1j * (orc.OracleLindbladian(om, de, ph, L, u) @ rho)
    sup = orc.dense_lindblad_superoperator(om, de, ph, u, L)
    assert torch.allclose(got, (sup @ rho.reshape(-1)).reshape(8, 8), atol=1e-13, rtol=0)
    # trace preserving: Tr(L(rho)) = 0
    assert abs(complex(got.trace())) < 1e-13


def test_oracle_lindbladian_without_jumps_any_rho():
    """test_linbladian_operator.py:18-48: zero jump operators give
    H rho - (H rho)^H, also for a non-Hermitian rho."""
    n = 4
    om, de, ph, u = _params(n, 3)
    rho = torch.rand(16, 16, dtype=DT)
    got = orc.OracleLindbladian(om, de, ph, [torch.zeros(2, 2, dtype=DT)], u) @ rho
    h = orc.dense_rydberg_hamiltonian(om, de, ph, u)
    assert torch.allclose(got, h @ rho - (h @ rho).conj().T, atol=1e-12, rtol=0)


def test_oracle_lindbladian_expect_closed_form():
    """test_linbladian_operator.py:216-245: U01 = U12 = 1 on three atoms gives
    Tr(H rho) = p3 + p6 + 2 p7 for a diagonal rho."""
    n = 3
    u = torch.zeros(n, n, dtype=torch.float64)
    u[0, 1] = u[1, 0] = u[1, 2] = u[2, 1] = 1.0
    torch.manual_seed(1234)
    p = torch.rand(8, dtype=torch.float64)
    p = p / p.sum()
    z = torch.zeros(n, dtype=torch.float64)
    en = orc.OracleLindbladian(z, z, z, [], u).expect(torch.diag(p).to(DT))
    assert torch.allclose(en, p[3] + p[6] + 2 * p[7])


def test_oracle_density_step_vs_dense_expm():
    n = 3
    om, de, ph, u = _params(n, 5)
    L = _lindblads(6)
    psi = torch.randn(8, dtype=DT)
    psi /= psi.norm()
    rho = torch.outer(psi, psi.conj())
    dt = 0.3
    lo = orc.OracleLindbladian(om, de, ph, L, u)
    got = orc.krylov_exp(lambda x: -1j * dt * (lo @ x), rho.clone(), 1e-12, 1e-12,
                         is_hermitian=False)
    sup = orc.dense_lindblad_superoperator(om, de, ph, u, L)
    exp_ = (torch.linalg.matrix_exp(dt * sup) @ rho.reshape(-1)).reshape(8, 8)
    assert torch.allclose(got, exp_, atol=1e-11, rtol=0)
    assert abs(complex(got.trace()) - 1) < 1e-12
