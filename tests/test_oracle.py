"""Pinning the oracle before it checks anything.

Each oracle function is checked against what the reference's own tests fix
(golden constants, exact iteration counts) or against mathematics that a
plausible restatement slip would violate (dense Kronecker H, dense expm,
closed forms, Hermiticity).
"""

import math

import pytest
import torch

from oracle import emu_sv_oracle as orc


def _randn_params(n, with_phase):
    om = torch.randn(n, dtype=torch.float64)
    de = torch.randn(n, dtype=torch.float64)
    ph = torch.randn(n, dtype=torch.float64) if with_phase else torch.zeros(n, dtype=torch.float64)
    u = torch.zeros(n, n, dtype=torch.float64)
    for i in range(n):
        for j in range(i + 1, n):
            u[i, j] = u[j, i] = torch.rand(1, dtype=torch.float64)
    return om, de, ph, u


# --- H.psi ------------------------------------------------------------------
@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("n,with_phase", [(3, False), (5, False), (8, False), (2, True),
                                          (4, True), (7, True), (9, True)])
def test_hamiltonian_vs_dense(n, with_phase):
    """test_hamiltonian.py:20-85: diagonal and H.psi vs the Kronecker H at 1e-12."""
    torch.manual_seed(1337)
    om, de, ph, u = _randn_params(n, with_phase)
    dense = orc.dense_rydberg_hamiltonian(om, de, ph, u)
    ham = orc.OracleHamiltonian(om, de, ph, u)
    assert torch.allclose(ham.diag, torch.diagonal(dense), atol=1e-12)
    psi = torch.randn(2**n, dtype=torch.complex128)
    assert torch.allclose(ham * psi, dense @ psi, atol=1e-12)


def test_dense_hamiltonian_is_hermitian_and_qubit0_is_msb():
    """A slip in the bit order or a sign of the sigma_y term breaks one of these."""
    n = 3
    om = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64)
    z = torch.zeros(n, dtype=torch.float64)
    ph = torch.tensor([0.4, 0.0, 0.0], dtype=torch.float64)
    h = orc.dense_rydberg_hamiltonian(om, z, ph, torch.zeros(n, n))
    assert torch.allclose(h, h.conj().T)
    # <1|H|0> on qubit 0 couples index 0b000 -> 0b100 (qubit 0 = MSB) with (O/2) e^{i phi}
    assert torch.isclose(h[4, 0], torch.tensor(0.5 * complex(math.cos(0.4), math.sin(0.4)),
                                               dtype=torch.complex128))
    # detuning enters with a minus sign on |r>
    h2 = orc.dense_rydberg_hamiltonian(z, torch.tensor([2.0, 0, 0], dtype=torch.float64), z,
                                       torch.zeros(n, n))
    assert h2[4, 4].real == -2.0 and h2[3, 3].real == 0.0


@pytest.mark.usefixtures("restore_rng_state")
def test_expect_closed_form():
    """test_hamiltonian.py:132-172: <H> of (|rrr>+|rgg>)/sqrt2."""
    torch.manual_seed(1337)
    n = 3
    om = torch.randn(n, dtype=torch.float64)
    de = torch.randn(n, dtype=torch.float64)
    u = torch.randn(n, n, dtype=torch.float64)
    u = u + u.T
    psi = torch.zeros(8, dtype=torch.complex128)
    psi[0b111] = psi[0b100] = 1 / math.sqrt(2)
    ham = orc.OracleHamiltonian(om, de, torch.zeros(n), u)
    expected = 0.5 * (-2 * de[0] - de[1] - de[2] + u[0, 1] + u[0, 2] + u[1, 2])
    assert torch.allclose(ham.expect(psi), expected)


@pytest.mark.usefixtures("restore_rng_state")
def test_energy_moment_constants():
    """test_custom_callbacks.py:120-158 golden constants (seed 1337, N=4)."""
    torch.manual_seed(1337)
    n = 4
    psi = torch.zeros(16, dtype=torch.complex128)
    psi[0b1010] = psi[0b0101] = 1.0
    psi /= psi.norm()
    om = torch.randn(n, dtype=torch.float64).to(torch.complex128)
    de = torch.randn(n, dtype=torch.float64).to(torch.complex128)
    u = torch.randn((n, n))
    ham = orc.OracleHamiltonian(om, de, torch.zeros_like(om), u)
    assert float(orc.energy_variance(ham, psi)) == pytest.approx(3.67378968943955, abs=4e-7)
    assert float(orc.energy_second_moment(ham, psi)) == pytest.approx(4.2188228611101, abs=2e-7)


def test_occupation_and_correlation_patterns():
    """test_custom_callbacks.py:41-109: GHZ occupations 0.5, AFM correlations."""
    n = 10
    psi = torch.zeros(2**n, dtype=torch.complex128)
    psi[0] = psi[-1] = 1 / math.sqrt(2)
    assert torch.allclose(orc.occupation(psi), torch.full((n,), 0.5, dtype=torch.float64))
    n = 6
    psi = torch.zeros(2**n, dtype=torch.complex128)
    psi[int("101010", 2)] = psi[int("010101", 2)] = 1 / math.sqrt(2)
    exp = torch.tensor([[0.5 if (i + j) % 2 == 0 else 0.0 for j in range(n)] for i in range(n)],
                       dtype=torch.float64)
    assert torch.allclose(orc.correlation_matrix(psi), exp)


# --- Krylov -------------------------------------------------------------------
def _uniform(*dims):
    return (torch.rand(*dims, dtype=torch.complex128) + 1j * torch.rand(*dims, dtype=torch.complex128)
            - (0.5 + 0.5j) * torch.ones(*dims))


def _check_krylov(m, v, herm, conv, happy, count, ntol, etol, normalize=True):
    if normalize:
        v /= v.norm()
    r = orc.krylov_exp_impl(lambda x: torch.tensordot(m, x, dims=1), v.clone(), herm, etol, ntol, 80)
    assert (r.converged, r.happy_breakdown, r.iteration_count) == (conv, happy, count)
    expected = torch.tensordot(torch.linalg.matrix_exp(m), v, dims=1)
    assert torch.allclose(r.result, expected, atol=ntol if happy else etol, rtol=0)


def test_krylov_identity_counts():
    """test_krylov_exp.py:73-103: identity -> happy breakdown at iteration 1."""
    for dim in (3, 300):
        _check_krylov(torch.eye(dim, dtype=torch.complex128), torch.arange(dim).to(torch.complex128),
                      True, True, True, 1, 1e-12, 1e-6)


def test_krylov_happy_breakdown_count():
    """test_krylov_exp.py:106-131: minimal polynomial of degree 3 -> breakdown at 3."""
    m = torch.diag(torch.tensor([1.0, 2.0, 3.0, 3.0, 3.0], dtype=torch.complex128))
    _check_krylov(m, torch.ones(5, 1).to(torch.complex128), True, True, True, 3, 1e-12, 1e-6)


@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("herm,count,etol,normalize", [(True, 7, 1e-13, True),
                                                       (False, 7, 1e-15, True),
                                                       (False, 6, 1e-13, False)])
def test_krylov_seeded_counts(herm, count, etol, normalize):
    """test_krylov_exp.py:134-191: 1000x1000 scaled random matrices, seed 1234."""
    torch.random.manual_seed(1234)
    m = 0.001 * _uniform(1000, 1000)
    v = _uniform(1000, 1)
    _check_krylov(m, v, herm, True, False, count, 1e-20, etol, normalize)


@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("n,tol,with_phase", [(n, t, p) for n in (3, 5, 8)
                                              for t in (1e-8, 1e-10, 1e-12) for p in (False, True)])
def test_step_vs_dense(n, tol, with_phase):
    """test_time_evolution.py:58-76: a 1 us step vs expm(-i dt H) at atol = tol."""
    torch.manual_seed(1337)
    om, de, ph, u = _randn_params(n, with_phase)
    psi = torch.randn(2**n, dtype=torch.complex128)
    psi = psi / psi.norm()
    out, _ = orc.step_forward(1.0, om, de, ph, u, psi, tol, 100)
    assert torch.allclose(out, orc.dense_step(1.0, om, de, ph, u, psi), atol=tol)


def test_krylov_nonconvergence_raises():
    torch.manual_seed(3)
    m = 0.5 * _uniform(50, 50)
    with pytest.raises(RecursionError):
        orc.krylov_exp(lambda x: m @ x, _uniform(50), 1e-12, 1e-12, max_krylov_dim=2)


# --- gradients -----------------------------------------------------------------
@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("n,tol", [(3, 1e-8), (3, 1e-12), (4, 1e-10), (5, 1e-12)])
def test_double_krylov_vs_dense_frechet(n, tol):
    """test_double_krylov.py:18-73: Vs^T dS Vg^* vs the dense block-expm Frechet derivative."""
    torch.manual_seed(1337)
    om, de, ph, u = _randn_params(n, True)
    psi = torch.randn(2**n, dtype=torch.complex128)
    psi /= psi.norm()
    g = torch.randn(2**n, dtype=torch.complex128)
    ham = orc.OracleHamiltonian(om, de, ph, u)
    vs, ds, vg = orc.double_krylov(lambda x: -1j * (ham * x), psi, g, tol, 100)
    du = torch.stack(vs).mT @ ds @ torch.stack(vg).conj()
    h = orc.dense_rydberg_hamiltonian(om, de, ph, u)
    exp = orc.frechet_exp(-1j * h, psi.unsqueeze(-1) @ g.conj().unsqueeze(0))
    assert torch.allclose(du, exp, atol=tol, rtol=0)


def _dense_h(om, de, ph, u, n):
    sx = torch.tensor([[0.0, 1.0], [1.0, 0.0]], dtype=torch.complex128)
    sy = torch.tensor([[0.0, -1j], [1j, 0.0]], dtype=torch.complex128)
    nop = torch.tensor([[0.0, 0.0], [0.0, 1.0]], dtype=torch.complex128)
    eye = torch.eye(2, dtype=torch.complex128)

    def gate(i, g):
        out = torch.ones(1, 1, dtype=torch.complex128)
        for k in range(n):
            out = torch.kron(out, g if k == i else eye)
        return out

    h = torch.zeros(2**n, 2**n, dtype=torch.complex128)
    for i in range(n):
        h = h + om[i] * torch.cos(ph[i]) * gate(i, sx) / 2 + om[i] * torch.sin(ph[i]) * gate(i, sy) / 2
        h = h - de[i] * gate(i, nop)
        for j in range(i + 1, n):
            h = h + u[i, j] * (gate(i, nop) @ gate(j, nop))
    return h


@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("n,tol", [(3, 1e-8), (3, 1e-10), (5, 1e-12)])
def test_reference_step_backward_vs_dense_autograd(n, tol):
    """test_time_evolution.py:79-103: the restated custom backward vs autograd
    through the dense matrix_exp, rtol = tol."""
    torch.manual_seed(1337)
    om, de, ph, u0 = _randn_params(n, True)
    params = [x.clone().requires_grad_(True) for x in (om, de, ph, u0)]
    psi = torch.randn(2**n, dtype=torch.complex128)
    psi = (psi / psi.norm()).requires_grad_(True)
    r = torch.randn(2**n, dtype=torch.complex128)
    r *= 0.71 / r.norm()
    out = orc.ReferenceStep.apply(1.0, *params, psi, tol, 100)
    grads = torch.autograd.grad(torch.vdot(r, out).real, (*params, psi))
    params2 = [x.clone().requires_grad_(True) for x in (om, de, ph, u0)]
    psi2 = psi.detach().clone().requires_grad_(True)
    h = _dense_h(*params2, n)
    exp = torch.autograd.grad(torch.vdot(r, torch.linalg.matrix_exp(-1j * h) @ psi2).real,
                              (*params2, psi2))
    for g, e in zip(grads, exp):
        assert torch.allclose(g, e, rtol=tol, atol=tol)


# --- time loop ---------------------------------------------------------------
def test_single_atom_rabi_closed_form():
    """Constant Omega on one atom: psi(t) = cos(Omega t/2)|g> - i sin(Omega t/2)|r>."""
    omega, steps, dt_ns = 2 * math.pi * 1.3, 37, 10.0
    om = torch.full((steps, 1), omega, dtype=torch.complex128)
    z = torch.zeros_like(om)
    run = orc.run_sequence(om, z, z, lambda t: torch.zeros(1, 1, 1, dtype=torch.float64),
                           [dt_ns * k for k in range(steps + 1)])
    t = steps * dt_ns * 1e-3
    expected = torch.tensor([math.cos(omega * t / 2), -1j * math.sin(omega * t / 2)],
                            dtype=torch.complex128)
    assert torch.allclose(run.states[steps], expected, atol=1e-10)


def test_sequence_data_pi_pulse_table():
    """test_sv_backend.py:102-152 Omega table: occupation 0 -> 1."""
    col = [0.7092, 7.8431, 26.2913, 53.0011, 73.0656, 71.8630, 50.3238, 23.9187, 6.6745, 0.4483]
    om = torch.tensor([[c] * 3 for c in col], dtype=torch.complex128)
    z = torch.zeros_like(om)
    run = orc.run_sequence(om, z, z, lambda t: torch.zeros(1, 3, 3, dtype=torch.float64),
                           [10.0 * k for k in range(11)], eval_steps=[0, 10])
    assert torch.allclose(run.occupation[0], torch.zeros(3, dtype=torch.float64), atol=1e-12)
    assert torch.allclose(run.occupation[10], torch.ones(3, dtype=torch.float64), atol=1e-3)


def test_norm_and_energy_conservation():
    """Constant H: norm and <H> are conserved by the Krylov steps."""
    torch.manual_seed(5)
    n, steps = 6, 20
    om, de, ph, u = _randn_params(n, True)
    om, u = om * 5, u * 10
    ham = orc.OracleHamiltonian(om, de, ph, u)
    run = orc.run_sequence(om.repeat(steps, 1).to(torch.complex128),
                           de.repeat(steps, 1).to(torch.complex128),
                           ph.repeat(steps, 1).to(torch.complex128),
                           lambda t: u.reshape(1, n, n), [10.0 * k for k in range(steps + 1)],
                           eval_steps=[1, steps])
    assert abs(float(run.states[steps].norm()) - 1.0) < 1e-12
    assert abs(float(ham.expect(run.states[1])) - float(ham.expect(run.states[steps]))) < 1e-9


# --- PCHIP ----------------------------------------------------------------------
def test_pchip_linear_data_is_exact():
    """PCHIP reproduces linear data (all secants equal -> slopes equal)."""
    x = torch.arange(10, dtype=torch.float64)
    y = 3.0 * x - 2.0
    xq = torch.linspace(-1.0, 10.0, 57, dtype=torch.float64)
    assert torch.allclose(orc.pchip_eval(x, y, xq), 3.0 * xq - 2.0, atol=1e-12)


def test_pchip_knots_and_monotone():
    """Interpolates the knots; stays within each interval's data range for monotone data."""
    x = torch.tensor([0.0, 1.0, 2.5, 3.0, 5.0, 7.0], dtype=torch.float64)
    y = torch.tensor([0.0, 0.5, 0.7, 3.0, 3.1, 9.0], dtype=torch.float64)
    assert torch.allclose(orc.pchip_eval(x, y, x), y, atol=1e-14)
    for i in range(len(x) - 1):
        xq = torch.linspace(float(x[i]), float(x[i + 1]), 25, dtype=torch.float64)
        v = orc.pchip_eval(x, y, xq)
        assert torch.all(v >= y[i] - 1e-12) and torch.all(v <= y[i + 1] + 1e-12)


def test_midpoint_ramp_convention():
    """A ramp from a to b over d ns is linspace(a, b, d) on t=0..d-1; midpoints
    of 10 ns steps are a + (b-a) (10k+5)/(d-1) (test_adapter.py ramp tables)."""
    d = 1000
    samples = torch.linspace(-10, 10, d, dtype=torch.float64)
    tt = [10.0 * k for k in range(101)]
    mid = orc.midpoint_samples(samples, tt, is_amplitude=False)
    k = torch.arange(100, dtype=torch.float64)
    assert torch.allclose(mid, -10 + 20 * (10 * k + 5) / (d - 1), atol=1e-12)


def test_inverse_cdf_indices_brute_force():
    """The oracle's inverse-CDF sampler against a brute-force scan (first index
    whose cumulative weight exceeds u * total), a point mass, and zero-weight
    entries that must never be drawn."""
    g = torch.Generator().manual_seed(11)
    for size in (1, 2, 7, 64):
        p = torch.rand(size, dtype=torch.float64, generator=g)
        if size > 2:
            p[::3] = 0.0
        if float(p.sum()) == 0.0:
            p[-1] = 1.0
        u = torch.rand(200, dtype=torch.float64, generator=g)
        got = orc.inverse_cdf_indices(p, u)
        total = float(p.sum())
        for k in range(u.numel()):
            acc, want = 0.0, size - 1
            for i in range(size):
                acc += float(p[i])
                if acc > float(u[k]) * total:
                    want = i
                    break
            assert int(got[k]) == want
        assert bool((p[got] > 0).all())
    point = torch.zeros(9, dtype=torch.float64)
    point[4] = 2.5
    assert torch.equal(orc.inverse_cdf_indices(point, torch.linspace(0, 0.999, 50, dtype=torch.float64)),
                       torch.full((50,), 4, dtype=torch.int64))


# the reference's known answers for the jump helpers (test/emu_sv/test_algebra.py:5-29)
ALGEBRA_VEC = [0.0, 0.0, 0.5j, 1.0]
ALGEBRA_OPS = [[[0.0, 0.0], [0.0, 1.0j]], [[1.0, 0.0], [0.0, 0.0]], [[0.0, 1.0j], [0.0, 0.0]]]
ALGEBRA_EXPECT = [[1.25j, 0.0, 0.0], [1.0j, 0.25, 0.5]]
HADAMARD_LIKE = [[1.0, -1.0], [1.0, 1.0]]
APPLY_VEC = [0.0, 0.0, 0.5, 1.0]
APPLY_Q0 = [-0.5, -1.0, 0.5, 1.0]
APPLY_Q1 = [0.0, 0.0, -0.5, 1.5]


def test_oracle_jump_helpers_known_answers():
    """test_algebra.py:5-29: Tr(A_i rho_q) on an unnormalised 2-qubit vector,
    and [[1, -1], [1, 1]] on qubit 0 and on qubit 1."""
    c = torch.complex128
    got = orc.expect_batch(torch.tensor(ALGEBRA_VEC, dtype=c), torch.tensor(ALGEBRA_OPS, dtype=c), 2)
    assert torch.allclose(got, torch.tensor(ALGEBRA_EXPECT, dtype=c))
    vec = torch.tensor(APPLY_VEC, dtype=c)
    had = torch.tensor(HADAMARD_LIKE, dtype=c)
    assert torch.allclose(orc.apply_1q(vec, 0, had), torch.tensor(APPLY_Q0, dtype=c))
    assert torch.allclose(orc.apply_1q(vec, 1, had), torch.tensor(APPLY_Q1, dtype=c))
