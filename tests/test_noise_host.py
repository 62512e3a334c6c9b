"""The Monte-Carlo (quantum-jump) pieces that run on the host, and the
oracle's restatement of the noisy path.

Pins: the reference's Brent known answers (roots and exact evaluation counts,
test/emu_base/math/test_brents_root_finding.py), dense Kronecker-product
brute force for the noisy Hamiltonian and the single-qubit helpers, the
reference's check that K = diag(0, -i/2) leaves Re<psi|H|psi> unchanged
(test/emu_sv/test_hamiltonian.py:134-172), and the closed-form jump time of
a decaying atom.
"""

import math
import random

import pytest
import torch

from oracle import emu_sv_oracle as orc

TEST_TOL = 1e-10


def _counted(f):
    calls = [0]

    def g(x):
        calls[0] += 1
        return f(x)

    return g, calls


def _stairs():
    """test_brents_root_finding.py:67-115: a random monotone piecewise-linear
    function on [0, 1] and its exact root."""
    random.seed(0xDEADBEEF)
    n = 100
    diff = [random.random() for _ in range(n)]
    w = 1 / n
    m = sum(diff) * w
    init = -m / 2

    def f(x):
        assert 0 <= x <= 1
        if x >= 1:
            return m
        i = int(x * n)
        return init + sum(diff[:i]) * w + x % w * diff[i]

    acc, i = init, 0
    while acc + w * diff[i] <= 0:
        acc += w * diff[i]
        i += 1
    return f, i * w - acc / diff[i]


CASES = [
    (lambda x: (x + 3) * (x - 1) ** 2, -4.0, 4 / 3.0, -3.0, 12),
    (lambda x: -1.0 if x < 0 else -1 + x * 10, -10.0, 10.0, 0.1, 13),
]


@pytest.mark.parametrize("impl", ["product", "oracle"])
@pytest.mark.parametrize("case", range(3))
@pytest.mark.usefixtures("restore_rng_state")
def test_brent_known_answers(impl, case):
    from emulators_amd.emu_base.math.brents_root_finding import find_root_brents
    if case < 2:
        f, a, b, root, count = CASES[case]
    else:
        f, root = _stairs()
        a, b, count = 0.0, 1.0, 18
    if impl == "product":
        g, calls = _counted(f)
        got = find_root_brents(g, start=a, end=b)
        n = calls[0]
    else:
        got, n = orc.brent_root(f, a, b)
    assert abs(got - root) < TEST_TOL
    assert n == count


@pytest.mark.usefixtures("restore_rng_state")
def test_brent_last_evaluation_is_the_root():
    """test_brents_root_finding.py:118-160: the finder is driven step by step
    on the norm of a non-Hermitian evolution; its last ordinate belongs to the
    returned root, so the caller's state is the state at the root."""
    from emulators_amd.emu_base.math.brents_root_finding import BrentsRootFinder
    n = 10
    torch.manual_seed(0xDEADBEEF)
    lind = torch.rand(n, n, dtype=torch.complex128)
    hami = torch.rand(n, n, dtype=torch.complex128)
    hami = hami + hami.T.conj()
    noisy = hami - 0.5j * 0.1 * lind.T.conj() @ lind
    psi = torch.rand(n, dtype=torch.complex128)
    psi /= psi.norm()

    def gap(t):
        return float((torch.linalg.matrix_exp(-1j * t * noisy) @ psi).norm() ** 2) - 0.5

    f = BrentsRootFinder(start=0.0, end=10.0, f_start=gap(0.0), f_end=gap(10.0), epsilon=1e-6)
    last = None
    while not f.is_converged(1e-6):
        x = f.get_next_abscissa()
        last = x
        f.provide_ordinate(x, gap(x))
    assert f.current_guess == last or abs(gap(f.current_guess)) < 1e-6
    assert abs(gap(f.current_guess)) < 1e-5


def _params(n, seed, phase):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, generator=g, dtype=torch.float64) * 4
    de = torch.randn(n, generator=g, dtype=torch.float64)
    ph = torch.randn(n, generator=g, dtype=torch.float64) if phase else torch.zeros(n, dtype=torch.float64)
    u = torch.rand(n, n, generator=g, dtype=torch.float64)
    return om, de, ph, (u + u.T) / 2


LINDBLADS = [math.sqrt(0.7) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128),
             math.sqrt(0.3) * torch.tensor([[1, 0], [0, -1]], dtype=torch.complex128),
             math.sqrt(0.2) * torch.tensor([[0, 1], [0, 0.5]], dtype=torch.complex128)]


@pytest.mark.parametrize("phase", [False, True])
def test_oracle_noisy_hamiltonian_vs_dense(phase):
    n = 5
    om, de, ph, u = _params(n, 3, phase)
    k = orc.noise_from_lindbladians(LINDBLADS)
    h = orc.OracleHamiltonian(om, de, ph, u, noise=k)
    dense = orc.dense_noisy_hamiltonian(om, de, ph, u, k)
    psi = torch.randn(2**n, dtype=torch.complex128)
    assert torch.allclose(h * psi, dense @ psi, atol=1e-12, rtol=0)
    assert not torch.allclose(dense, dense.conj().T)  # K makes H non-Hermitian


def test_noise_term_closed_form():
    """K = -i/2 L^+L: Hermitian L^+L, so K is anti-Hermitian, and the
    product's helper agrees with the oracle's."""
    from emulators_amd.emu_base.jump_lindblad_operators import compute_noise_from_lindbladians
    k = compute_noise_from_lindbladians(LINDBLADS)
    assert torch.allclose(k, orc.noise_from_lindbladians(LINDBLADS), atol=0, rtol=0)
    assert torch.allclose(k, -k.conj().T, atol=1e-15)
    assert torch.allclose(1j * (k - k.conj().T),
                          sum(L.conj().T @ L for L in LINDBLADS), atol=1e-14)


def test_oracle_noise_leaves_real_energy():
    """test_hamiltonian.py:134-172: a diagonal K = diag(0, -i/2) changes only
    the imaginary part of <psi|H|psi>."""
    n = 3
    torch.manual_seed(0)
    om, de = torch.randn(n, dtype=torch.float64), torch.randn(n, dtype=torch.float64)
    u = torch.randn(n, n, dtype=torch.float64)
    u = u + u.T
    psi = torch.zeros(8, dtype=torch.complex128)
    psi[0b111] = psi[0b100] = 1 / math.sqrt(2)  # (|rrr> + |rgg>)/sqrt 2
    h0 = orc.OracleHamiltonian(om, de, torch.zeros(n), u)
    h = orc.OracleHamiltonian(om, de, torch.zeros(n), u,
                              noise=torch.tensor([[0.0, 0.0], [0.0, -0.5j]], dtype=torch.complex128))
    e0 = h0.expect(psi)
    expected = 0.5 * (-2 * de[0] - de[1] - de[2] + u[0, 1] + u[0, 2] + u[1, 2])
    assert torch.allclose(e0, expected)
    assert h.expect(psi) == e0


def test_oracle_single_qubit_helpers_vs_dense():
    n = 4
    torch.manual_seed(5)
    psi = torch.randn(2**n, dtype=torch.complex128)
    ops = torch.randn(3, 2, 2, dtype=torch.complex128)
    t = orc.expect_batch(psi, ops, n)
    for q in range(n):
        for i in range(3):
            full = orc._single_gate(q, n, ops[i])
            assert abs(t[q, i] - torch.vdot(psi, full @ psi)) < 1e-12
        assert torch.allclose(orc.apply_1q(psi, q, ops[0]), orc._single_gate(q, n, ops[0]) @ psi,
                              atol=1e-13)


@pytest.mark.usefixtures("restore_rng_state")
def test_oracle_monte_carlo_decay_closed_form():
    """One atom, no drive, relaxation L = sqrt(g)|g><r| from |r>: the norm is
    exp(-g t), so the jump happens where exp(-g t) equals the threshold, lands
    in |g> and no further jump follows."""
    g = 2.0
    L = [math.sqrt(g) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128)]
    random.seed(11)
    mc = orc.OracleMonteCarlo()
    thr = mc.jump_threshold
    t_jump = -math.log(thr) / g
    psi = torch.tensor([0, 1], dtype=torch.complex128)
    dt, t = 0.05, 0.0
    zero = torch.zeros(1, dtype=torch.float64)
    while t < t_jump + 3 * dt:
        before = len(mc.jumps)
        psi = mc.step(dt, zero, zero, zero, torch.zeros(1, 1), psi, 1e-12, L)
        t += dt
        if len(mc.jumps) > before:
            assert t - dt <= t_jump <= t + 1e-12
    assert mc.jumps == [0]
    assert abs(abs(psi[0]) - 1) < 1e-12 and abs(psi[1]) < 1e-12


@pytest.mark.usefixtures("restore_rng_state")
def test_oracle_monte_carlo_emission_known_answer():
    """The reference's seeded Monte-Carlo answer of
    test_end_to_end_spontaneous_emission_rate (test/emu_sv/test_end_to_end.py:
    458-545): two non-interacting atoms, no drive, relaxation 0.1 / us for
    10 us (100 steps of 100 ns) from |rr>, Krylov tolerance 1e-5, 10
    trajectories after random.seed(0xDEADBEEF): mean final occupations
    [0.5, 0.2].  Each trajectory ends with each atom in |g> or |r>, so the
    answer is a count of which atoms jumped; it fixes the threshold draws,
    the random.choices of the jumping qubit and the jump-time search to the
    reference's draw order."""
    random.seed(0xDEADBEEF)
    relax = [math.sqrt(0.1) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128)]
    zero = torch.zeros(2, dtype=torch.float64)
    occs = []
    for _ in range(10):
        mc = orc.OracleMonteCarlo()
        psi = torch.zeros(4, dtype=torch.complex128)
        psi[3] = 1.0  # |rr>
        for _ in range(100):
            psi = mc.step(0.1, zero, zero, zero, torch.zeros(2, 2, dtype=torch.float64), psi, 1e-5, relax)
        occs.append(orc.occupation(psi / psi.norm()))
    mean = torch.stack(occs).mean(0)
    assert torch.allclose(mean, torch.tensor([0.5, 0.2], dtype=torch.float64), atol=1e-12)


@pytest.mark.usefixtures("restore_rng_state")
def test_oracle_monte_carlo_random_effective_noise_known_answer():
    """The reference's test_end_to_end_random_effective_noise
    (test/emu_sv/test_end_to_end.py:548-600): torch.manual_seed(31415) draws
    the effective-noise operator (torch.rand(2, 2), rate 30, pulser basis
    flipped to the state-vector one), two atoms 8 um apart (U = C6 / 8^6), no
    drive, 1000 ns in 100 steps from |rr>, Krylov tolerance 1e-10, 10
    trajectories after random.seed(0xDEADBEEF): mean final occupations
    [0.8508, 0.8536] (its atol 1e-4)."""
    from emulators_amd.emu_base.jump_lindblad_operators import get_all_lindblad_noise_operators
    torch.manual_seed(31415)
    random.seed(0xDEADBEEF)
    op = torch.rand(2, 2, dtype=torch.float64, device="cpu")
    ops = get_all_lindblad_noise_operators(_noise_model(noise_types=("eff_noise",),
                                                        eff_noise_rates=(30.0,), eff_noise_opers=(op,)))
    u = 5420158.53 / 8.0**6
    umat = torch.tensor([[0.0, u], [u, 0.0]], dtype=torch.float64)
    zero = torch.zeros(2, dtype=torch.float64)
    occs = []
    for _ in range(10):
        mc = orc.OracleMonteCarlo()
        psi = torch.zeros(4, dtype=torch.complex128)
        psi[3] = 1.0  # |rr>
        for _ in range(100):
            psi = mc.step(0.01, zero, zero, zero, umat, psi, 1e-10, ops)
        occs.append(orc.occupation(psi / psi.norm()))
    mean = torch.stack(occs).mean(0)
    assert torch.allclose(mean, torch.tensor([0.8508, 0.8536], dtype=torch.float64), atol=1e-4)


@pytest.mark.parametrize("solver,ops,stochastic,expected", [
    ("default", [], False, "EvolveStateVector"),
    ("default", [1], True, "EvolveMonteCarlo"),
    ("default", [1], False, "EvolveDensityMatrix"),
    ("montecarlo", [1], False, "EvolveMonteCarlo"),
    ("lindblad", [1], True, "EvolveDensityMatrix"),
])
def test_stepper_selection(solver, ops, stochastic, expected):
    """The solver table of sv_backend_impl.py:84-95 (the reference's
    test_sv_backend.py:158-240)."""
    from types import SimpleNamespace

    from emulators_amd.emu_sv.sv_backend_impl import select_stepper
    from emulators_amd.emu_sv.sv_config import Solver, SVConfig
    data = SimpleNamespace(lindblad_ops=[torch.eye(2, dtype=torch.complex128)] * len(ops),
                           has_stochastic_noise=stochastic)
    stepper, dm = select_stepper(SVConfig(observables=[], log_level=40, solver=Solver(solver)), data)
    assert type(stepper).__name__ == expected
    assert dm == (expected == "EvolveDensityMatrix")


def _noise_model(**kw):
    from types import SimpleNamespace
    base = dict(noise_types=(), relaxation_rate=0.0, dephasing_rate=0.0,
                hyperfine_dephasing_rate=0.0, depolarizing_rate=0.0, eff_noise_rates=(),
                eff_noise_opers=())
    base.update(kw)
    return SimpleNamespace(**base)


def test_lindblad_operators_eff_noise_basis():
    """test_lindblad_operators.py:12-40: pulser's |0><1| in the ising basis
    becomes sqrt(rate) |g><r| in the state-vector basis; in XY it is kept."""
    from emulators_amd.emu_base.jump_lindblad_operators import get_lindblad_operators
    for interaction, b0, b1 in (("ising", [0.0, 1.0], [1.0, 0.0]), ("XY", [1.0, 0.0], [0.0, 1.0])):
        basis0 = torch.tensor(b0, dtype=torch.complex128).reshape(2, 1)
        basis1 = torch.tensor(b1, dtype=torch.complex128).reshape(2, 1)
        nm = _noise_model(noise_types=("eff_noise",), eff_noise_rates=[0.5],
                          eff_noise_opers=[basis0 @ basis1.T])
        ops = get_lindblad_operators(noise_type="eff_noise", noise_model=nm,
                                     interact_type=interaction)
        assert torch.allclose(ops[0], math.sqrt(0.5) * torch.tensor(
            [[0.0, 1.0], [0.0, 0.0]], dtype=torch.complex128))


def test_lindblad_operators_eff_noise_flip_3x3():
    """test_lindblad_operators.py:43-77: for 3x3 operators in the ising basis
    only the leading 2x2 block is flipped (both axes), the rest is scaled by
    sqrt(rate) and kept; a 3x3 numpy operator at dim 2 raises
    (test_lindblad_operators.py:80-91)."""
    import numpy as np
    from emulators_amd.emu_base.jump_lindblad_operators import get_lindblad_operators
    g = torch.Generator().manual_seed(42)
    a0 = torch.rand(3, 3, dtype=torch.complex128, generator=g)
    a1 = torch.rand(3, 3, dtype=torch.complex128, generator=g)
    nm = _noise_model(noise_types=("eff_noise",), eff_noise_rates=[0.5, 0.5],
                      eff_noise_opers=[a0, a1])
    ops = get_lindblad_operators(noise_type="eff_noise", noise_model=nm, interact_type="ising",
                                 dim=3)
    assert len(ops) == 2
    for op, a in zip(ops, (a0, a1)):
        s = math.sqrt(0.5)
        # element by element: (i, j) of the 2x2 block comes from (1-i, 1-j)
        for i in range(3):
            for j in range(3):
                src = (1 - i, 1 - j) if i < 2 and j < 2 else (i, j)
                assert torch.allclose(op[i, j], s * a[src])
    b0 = np.array([0.0, 1.0, 0.0]).reshape(3, 1)
    b1 = np.array([1.0, 0.0, 0.0]).reshape(3, 1)
    with pytest.raises(ValueError, match="Only 2 by 2 effective noise operator matrices"):
        get_lindblad_operators(noise_type="eff_noise", noise_model=_noise_model(
            noise_types=("eff_noise",), eff_noise_rates=[0.5], eff_noise_opers=[b0 @ b1.T]))


def test_lindblad_operators_channels():
    """Relaxation / dephasing / depolarizing closed forms
    (jump_lindblad_operators.py:18-54); non-Lindbladian noise types are
    skipped, an unknown one raises."""
    from emulators_amd.emu_base.jump_lindblad_operators import (get_all_lindblad_noise_operators,
                                                                get_lindblad_operators)
    nm = _noise_model(noise_types=("relaxation", "dephasing", "depolarizing", "SPAM", "doppler"),
                      relaxation_rate=0.4, dephasing_rate=0.6, depolarizing_rate=0.8)
    ops = get_all_lindblad_noise_operators(nm)
    assert len(ops) == 5
    assert torch.allclose(ops[0], torch.tensor([[0, math.sqrt(0.4)], [0, 0]], dtype=torch.complex128))
    assert torch.allclose(ops[1], math.sqrt(0.3) * torch.diag(torch.tensor([1.0, -1.0])).to(torch.complex128))
    c = math.sqrt(0.2)
    assert torch.allclose(ops[2], c * torch.tensor([[0, 1], [1, 0]], dtype=torch.complex128))
    assert torch.allclose(ops[3], c * torch.tensor([[0, -1j], [1j, 0]], dtype=torch.complex128))
    # sum of L^+L of the depolarizing channel = 3/4 rate * I
    s = sum(L.conj().T @ L for L in ops[2:])
    assert torch.allclose(s, 0.6 * torch.eye(2, dtype=torch.complex128))
    with pytest.raises(ValueError, match="Only 2 by 2"):
        get_lindblad_operators(noise_type="eff_noise", noise_model=_noise_model(
            noise_types=("eff_noise",), eff_noise_rates=[0.5], eff_noise_opers=[torch.eye(3)]))
    with pytest.raises(ValueError):
        get_lindblad_operators(noise_type="nope", noise_model=_noise_model(noise_types=("nope",)))
    with pytest.raises(NotImplementedError):
        get_lindblad_operators(noise_type="dephasing", noise_model=_noise_model(
            noise_types=("dephasing",), hyperfine_dephasing_rate=1.0))
