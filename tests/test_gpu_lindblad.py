"""The Lindblad density-matrix path through the C ABI against the oracle.

Bars: L @ rho within 1e-12 (relative to max|rho| scale) of the oracle's
RydbergLindbladian restatement (itself pinned to the dense Lindblad
generator); the pair and anti-Hermitian kernels exactly against dense
torch; density-matrix Krylov steps with the oracle's iteration count and
within the tolerance of the dense 4^N exponential; the single decaying atom
following exp(-gamma t).
"""

import math

import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu
DT = torch.complex128


def _params(n, seed):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, generator=g, dtype=torch.float64) * 3
    de = torch.randn(n, generator=g, dtype=torch.float64)
    ph = torch.rand(n, generator=g, dtype=torch.float64)
    u = torch.rand(n, n, generator=g, dtype=torch.float64)
    return om, de, ph, (u + u.T) / 2


def _lindblads(seed):
    g = torch.Generator().manual_seed(seed)
    return [math.sqrt(1 / 3) * torch.rand(2, 2, generator=g, dtype=DT),
            math.sqrt(1 / 2) * torch.rand(2, 2, generator=g, dtype=DT)]


@pytest.mark.parametrize("n", [1, 2, 4, 6])
@pytest.mark.parametrize("hermitian", [True, False])
def test_lindbladian_matmul_vs_oracle(gpu_device, n, hermitian):
    from emulators_amd.emu_sv.lindblad_operator import RydbergLindbladian
    print(f"lindbladian n={n} hermitian={hermitian}", flush=True)
    om, de, ph, u = _params(n, 10 + n)
    L = _lindblads(20 + n)
    torch.manual_seed(n)
    rho = torch.rand(2**n, 2**n, dtype=DT)
    if hermitian:
        rho = 0.5 * (rho + rho.conj().T)
    ref = orc.OracleLindbladian(om, de, ph, L, u) @ rho
    lind = RydbergLindbladian(om, de, ph, L, u, device=gpu_device)
    got = (lind @ rho.to(gpu_device)).cpu()
    assert torch.allclose(got, ref, atol=1e-12 * max(1.0, float(ref.abs().max())), rtol=0)
    # h_eff and the one-sided helpers
    k = orc.noise_from_lindbladians(L)
    olind = orc.OracleLindbladian(om, de, ph, L, u)
    assert torch.allclose(lind.h_eff(rho.to(gpu_device), k).cpu(), olind.h_eff(rho, k),
                          atol=1e-12, rtol=0)
    q = n - 1
    assert torch.allclose(lind.apply_local_op_to_density_matrix(rho.to(gpu_device), L[0], q).cpu(),
                          olind.left(rho, L[0], q), atol=1e-13, rtol=0)
    assert torch.allclose(lind.apply_density_matrix_to_local_op_T(rho.to(gpu_device), L[1], 0).cpu(),
                          olind.right_dag(rho, L[1], 0), atol=1e-13, rtol=0)


def test_lindbladian_expect_closed_form(gpu_device):
    from emulators_amd.emu_sv import DensityMatrix
    from emulators_amd.emu_sv.lindblad_operator import RydbergLindbladian
    n = 3
    u = torch.zeros(n, n, dtype=torch.float64)
    u[0, 1] = u[1, 0] = u[1, 2] = u[2, 1] = 1.0
    torch.manual_seed(1234)
    p = torch.rand(8, dtype=torch.float64)
    p = p / p.sum()
    z = torch.zeros(n, dtype=torch.float64)
    lind = RydbergLindbladian(z, z, z, torch.rand(2, 2, dtype=DT), u, device=gpu_device)
    en = lind.expect(DensityMatrix(torch.diag(p).to(DT), gpu=True))
    assert abs(float(en) - float(p[3] + p[6] + 2 * p[7])) < 1e-14


@pytest.mark.parametrize("side_log", [0, 1, 4, 5, 7])
@pytest.mark.parametrize("dtype", [DT, torch.complex64])
@pytest.mark.parametrize("inplace", [False, True])
def test_antihermitian_kernel(gpu_device, side_log, dtype, inplace):
    from emulators_amd import _native as native
    torch.manual_seed(side_log)
    m = torch.randn(2**side_log, 2**side_log, dtype=DT).to(gpu_device, dtype)
    s = complex(0.3, -1.7)
    exp_ = s * (m.to(DT) - m.to(DT).conj().T)
    out = m if inplace else torch.empty_like(m)
    native.antihermitian(m, out, s)
    tol = 1e-14 if dtype == DT else 2e-6
    assert float((out.to(DT) - exp_).abs().max()) <= tol * float(exp_.abs().max() + 1)


@pytest.mark.parametrize("n,a,b", [(2, 1, 0), (4, 0, 3), (6, 5, 2), (8, 3, 7)])
def test_pair_kernel_vs_dense(gpu_device, n, a, b):
    from emulators_amd import _native as native
    torch.manual_seed(n)
    s = torch.randn(4, 4, dtype=DT)
    v = torch.randn(2**n, dtype=DT)
    # dense: the 4x4 acts on bits (a, b), local index 2*x_a + x_b
    full = torch.zeros(2**n, 2**n, dtype=DT)
    for x in range(2**n):
        ra, rb = (x >> a) & 1, (x >> b) & 1
        base = x & ~((1 << a) | (1 << b))
        for ca in range(2):
            for cb in range(2):
                y = base | (ca << a) | (cb << b)
                full[x, y] = s[2 * ra + rb, 2 * ca + cb]
    out = torch.randn(2**n, dtype=DT).to(gpu_device)
    init = out.cpu()
    native.apply_pair(n, v.to(gpu_device), out, a, b, s.numpy(), accumulate=True)
    assert torch.allclose(out.cpu(), init + full @ v, atol=1e-12, rtol=0)
    native.apply_pair(n, v.to(gpu_device), out, a, b, s.numpy(), accumulate=False)
    assert torch.allclose(out.cpu(), full @ v, atol=1e-12, rtol=0)


@pytest.mark.parametrize("n,dt,tol", [(2, 0.5, 1e-10), (3, 0.3, 1e-12), (5, 0.2, 1e-10), (7, 0.1, 1e-10)])
def test_density_step_vs_oracle(gpu_device, n, dt, tol):
    from emulators_amd.emu_base.math.krylov_exp import krylov_exp_impl
    from emulators_amd.emu_sv.lindblad_operator import RydbergLindbladian, ScaledLindbladian
    from emulators_amd.emu_sv.time_evolution import EvolveDensityMatrix
    print(f"density step n={n}", flush=True)
    om, de, ph, u = _params(n, 30 + n)
    L = _lindblads(40 + n)
    torch.manual_seed(n)
    psi = torch.randn(2**n, dtype=DT)
    psi /= psi.norm()
    rho = torch.outer(psi, psi.conj())
    olind = orc.OracleLindbladian(om, de, ph, L, u)
    ref = orc.krylov_exp_impl(lambda x: -1j * dt * (olind @ x), rho.clone(), is_hermitian=False,
                              exp_tolerance=tol, norm_tolerance=tol)
    lind = RydbergLindbladian(om, de, ph, L, u, device=gpu_device)
    got = krylov_exp_impl(ScaledLindbladian(lind, -1j * dt), rho.to(gpu_device),
                          is_hermitian=False, exp_tolerance=tol, norm_tolerance=tol)
    assert got.converged and got.iteration_count == ref.iteration_count
    assert torch.allclose(got.result.cpu(), ref.result, atol=10 * tol, rtol=0)
    if n <= 3:
        sup = orc.dense_lindblad_superoperator(om, de, ph, u, L)
        exp_ = (torch.linalg.matrix_exp(dt * sup) @ rho.reshape(-1)).reshape(rho.shape)
        assert torch.allclose(got.result.cpu(), exp_, atol=10 * tol, rtol=0)
    out, _ = EvolveDensityMatrix().apply(dt, om, de, ph, u, rho.to(gpu_device), tol, L)
    assert torch.allclose(out.cpu(), ref.result, atol=10 * tol, rtol=0)
    assert abs(complex(out.trace()) - 1) < 1e-9


def test_decaying_atom_population(gpu_device):
    """One atom, no drive, relaxation sqrt(g)|g><r|: Tr(rho n) = exp(-g t)."""
    from emulators_amd.emu_sv.time_evolution import EvolveDensityMatrix
    g = 1.5
    L = [math.sqrt(g) * torch.tensor([[0, 1], [0, 0]], dtype=DT)]
    rho = torch.tensor([[0, 0], [0, 1]], dtype=DT).to(gpu_device)
    z = torch.zeros(1, dtype=torch.float64)
    st = EvolveDensityMatrix()
    t, dt = 0.0, 0.1
    for _ in range(10):
        rho, _ = st.apply(dt, z, z, z, torch.zeros(1, 1), rho, 1e-12, L)
        t += dt
        assert abs(float(rho[1, 1].real) - math.exp(-g * t)) < 1e-10


def test_lindblad_backend_run(gpu_device):
    """SVBackend on Lindblad noise without stochastic noise runs the density
    matrix (sv_backend_impl.py:84-95); occupations match the oracle's dense
    evolution of the same per-step tables."""
    import dataclasses

    from emulators_amd.emu_base.pulser_adapter import sequence_data_from_samples
    from emulators_amd.emu_sv import Occupation, SVConfig
    from emulators_amd.emu_sv.density_matrix_state import DensityMatrix
    from emulators_amd.emu_sv.sv_backend_impl import SVBackendImpl
    from emulators_amd import workloads
    import numpy as np
    cfg = SVConfig(observables=[Occupation(evaluation_times=[1.0])], log_level=40)
    n, d = 3, 200
    pos = workloads.chain_positions(n, 7.0)
    seq = workloads.adiabatic_sequence(pos, np.full(d, 2 * np.pi), workloads.ramp(-5, 5, d), cfg,
                                       cfg.dt)
    L = [math.sqrt(0.8) * torch.tensor([[0, 1], [0, 0]], dtype=DT)]
    seq = dataclasses.replace(seq, lindblad_ops=L, has_stochastic_noise=False)
    impl = SVBackendImpl(cfg, seq)
    assert impl.density_matrix and isinstance(impl.state, DensityMatrix)
    res = impl._run()
    occ = torch.as_tensor(res.get_result(cfg.observables[0].tag, 1.0))
    # oracle: dense Liouvillian exponentials of the same per-step tables
    rho = torch.zeros(2**n, 2**n, dtype=DT)
    rho[0, 0] = 1.0
    tt = np.asarray(seq.target_times, dtype=np.float64)
    for k in range(seq.omega.shape[0]):
        dt = (tt[k + 1] - tt[k]) * 1e-3
        sup = orc.dense_lindblad_superoperator(seq.omega[k].real, seq.delta[k].real,
                                               seq.phi[k].real, seq.interaction_matrix(tt[k])[0], L)
        rho = (torch.linalg.matrix_exp(dt * sup) @ rho.reshape(-1)).reshape(rho.shape)
    rho = rho / rho.trace()
    p = rho.diagonal().real
    ref = torch.tensor([float(sum(p[x] for x in range(2**n) if (x >> (n - 1 - q)) & 1))
                        for q in range(n)], dtype=torch.float64)
    assert torch.allclose(occ, ref, atol=1e-8, rtol=0)


@pytest.mark.parametrize("n_side", [1, 3, 8])
@pytest.mark.parametrize("dtype", [DT, torch.complex64])
def test_pairs_all_matches_single_pairs(gpu_device, n_side, dtype):
    """The gathered all-qubit launch equals n_side single-pair launches."""
    import numpy as np

    from emulators_amd import _native as native
    torch.manual_seed(n_side)
    s = torch.randn(4, 4, dtype=DT).numpy()
    v = torch.randn(4**n_side, dtype=DT).to(gpu_device, dtype)
    base = torch.randn(4**n_side, dtype=DT).to(gpu_device, dtype)
    a = base.clone()
    native.apply_pairs_all(n_side, v, a, np.ascontiguousarray(s))
    b = base.clone()
    for j in range(n_side):
        native.apply_pair(2 * n_side, v, b, j + n_side, j, s, accumulate=True)
    tol = 1e-12 if dtype == DT else 1e-4
    assert float((a - b).abs().max()) <= tol * float(b.abs().max())
