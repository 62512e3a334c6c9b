"""The Monte-Carlo noisy path through the C ABI against the oracle.

Bars: the noisy H.psi (LDS kernel with the -i/2 sum L^+L term) against the
dense Kronecker-product H at atol 1e-12 (complex128) / 2e-5 relative
(complex64); the device Arnoldi krylov_exp with the oracle's iteration
count and its result within the tolerance of the oracle's; the modified
Gram-Schmidt records against a CPU restatement; the jump helpers against
dense operators; seeded quantum-jump trajectories with the same jumps as the
oracle's and final states within 1e-8.
"""

import math
import random

import numpy as np
import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu

LINDBLADS = [math.sqrt(0.7) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128),
             math.sqrt(0.3) * torch.tensor([[1, 0], [0, -1]], dtype=torch.complex128),
             math.sqrt(0.2) * torch.tensor([[0, 1], [0, 0.5]], dtype=torch.complex128)]


def _params(n, seed, phase):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, generator=g, dtype=torch.float64) * 4
    de = torch.randn(n, generator=g, dtype=torch.float64)
    ph = torch.randn(n, generator=g, dtype=torch.float64) if phase else torch.zeros(n, dtype=torch.float64)
    u = torch.rand(n, n, generator=g, dtype=torch.float64)
    return om, de, ph, (u + u.T) / 2


@pytest.mark.parametrize("n", [2, 5, 9, 12])
@pytest.mark.parametrize("phase", [False, True])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_noisy_hpsi_vs_dense(gpu_device, n, phase, dtype):
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    print(f"noisy hpsi n={n} phase={phase} {dtype}", flush=True)
    om, de, ph, u = _params(n, 17 + n, phase)
    k = orc.noise_from_lindbladians(LINDBLADS)
    h = RydbergHamiltonian(om, de, ph, u, device=gpu_device, noise=k)
    torch.manual_seed(n)
    psi = torch.randn(2**n, dtype=torch.complex128)
    expected = (orc.OracleHamiltonian(om, de, ph, u, noise=k) * psi) if n > 10 else \
        orc.dense_noisy_hamiltonian(om, de, ph, u, k) @ psi
    got = h.apply(psi.to(gpu_device, dtype), sigma=1.0).cpu().to(torch.complex128)
    if dtype == torch.complex128:
        assert torch.allclose(got, expected, atol=1e-12 * max(1.0, float(expected.abs().max())), rtol=0)
    else:
        assert float((got - expected).abs().max()) <= 2e-5 * float(expected.abs().max())
    # the diagonal, with the diagonal entries of K only
    if n <= 9:
        dense = orc.dense_noisy_hamiltonian(torch.zeros(n), de, torch.zeros(n), u,
                                            torch.diag(torch.diag(k)))
        assert torch.allclose(h.diag.cpu(), torch.diagonal(dense), atol=1e-12, rtol=0)


def test_noisy_expectation_real_part(gpu_device):
    """K = diag(0, -i/2) leaves Re<psi|H|psi> unchanged
    (test_hamiltonian.py:134-172)."""
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    n = 3
    torch.manual_seed(0)
    om, de = torch.randn(n, dtype=torch.float64), torch.randn(n, dtype=torch.float64)
    u = torch.randn(n, n, dtype=torch.float64)
    u = u + u.T
    psi = torch.zeros(8, dtype=torch.complex128)
    psi[0b111] = psi[0b100] = 1 / math.sqrt(2)
    psi = psi.to(gpu_device)
    h0 = RydbergHamiltonian(om, de, torch.zeros(n), u, device=gpu_device)
    h = RydbergHamiltonian(om, de, torch.zeros(n), u, device=gpu_device,
                           noise=torch.tensor([[0.0, 0.0], [0.0, -0.5j]], dtype=torch.complex128))
    expected = 0.5 * (-2 * de[0] - de[1] - de[2] + u[0, 1] + u[0, 2] + u[1, 2])
    assert abs(float(h0.expect(psi)) - float(expected)) < 1e-12
    assert abs(float(h.expect(psi)) - float(expected)) < 1e-12


@pytest.mark.parametrize("n,count", [(4, 1), (8, 5), (10, 12), (11, 3), (12, 20), (13, 9)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_arnoldi_records_vs_mgs(gpu_device, n, count, dtype):
    from emulators_amd import _native as native
    print(f"arnoldi records n={n} count={count} {dtype}", flush=True)
    torch.manual_seed(100 + n)
    basis = [torch.randn(2**n, dtype=torch.complex128) for _ in range(count)]
    scales = [1.0 / float(b.norm()) for b in basis]
    w = torch.randn(2**n, dtype=torch.complex128)
    dev = [b.to(gpu_device, dtype) for b in basis]
    wd = w.to(gpu_device, dtype)
    rec = native.arnoldi_orthogonalize(n, dev, scales, wd).cpu().numpy()
    # CPU restatement of the MGS loop on the same (rounded) inputs
    bb = [d.cpu().to(torch.complex128) for d in dev]
    ww = w.to(dtype).to(torch.complex128)
    tol = 1e-12 if dtype == torch.complex128 else 1e-5
    for k in range(count):
        vk = bb[k] * scales[k]
        assert abs(rec[k, 0] - float(ww.norm() ** 2)) <= tol * float(ww.norm() ** 2)
        h = complex(torch.vdot(vk, ww))
        got = scales[k] * complex(rec[k, 1], rec[k, 2])
        assert abs(got - h) <= tol * max(1.0, float(ww.norm()))
        ww = ww - h * vk
    assert abs(rec[count, 0] - float(ww.norm() ** 2)) <= tol * max(1.0, float(w.norm() ** 2))
    out = wd.cpu().to(torch.complex128)
    assert float((out - ww).abs().max()) <= tol * max(1.0, float(w.abs().max()))


@pytest.mark.parametrize("n,count", [(4, 1), (8, 5), (10, 12), (13, 9), (15, 4)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_arnoldi_orthonormalize(gpu_device, n, count, dtype):
    """emu_sv_arnoldi_orthonormalize: the records of emu_sv_arnoldi_orthogonalize
    (norm after the last subtraction unnormalised) and w divided by that norm
    on the device (one-workgroup path for 10 <= n <= 13, launch chain
    otherwise); a zero w stays zero."""
    from emulators_amd import _native as native
    print(f"arnoldi orthonormalize n={n} count={count} {dtype}", flush=True)
    torch.manual_seed(300 + n)
    basis = [torch.randn(2**n, dtype=torch.complex128).to(gpu_device, dtype) for _ in range(count)]
    scales = [1.0 / float(b.norm()) for b in basis]
    w = torch.randn(2**n, dtype=torch.complex128).to(gpu_device, dtype)
    w1, w2 = w.clone(), w.clone()
    rec1 = native.arnoldi_orthogonalize(n, basis, scales, w1).cpu().numpy()
    rec2 = native.arnoldi_orthogonalize(n, basis, scales, w2, normalize=True).cpu().numpy()
    assert np.array_equal(rec1[:, :3], rec2[:, :3])
    unit = w1.to(torch.complex128) / math.sqrt(rec1[count, 0])
    tol = 1e-14 if dtype == torch.complex128 else 2e-7
    assert float((w2.to(torch.complex128) - unit).abs().max()) <= tol
    z = torch.zeros_like(w)
    native.arnoldi_orthogonalize(n, basis[:1], scales[:1], z, normalize=True)
    assert float(z.abs().max()) == 0.0


@pytest.mark.parametrize("n,steps", [(3, 6), (9, 12), (14, 10), (16, 24)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_arnoldi_gram_vs_mgs(gpu_device, n, steps, dtype):
    """emu_sv_arnoldi_gram against the modified Gram-Schmidt launch chain
    over a run of ``steps`` calls on the same inputs: each call's Hessenberg
    column T[:, j] (records) and norms, and the new basis vector (unnormalised
    u_{j+1} times its device scale against the chain's unit vector), within
    1e-12 (complex128) / 2e-5 (complex64) relative; the scaled basis of the
    Gram form stays orthonormal to 1e-13 (complex128)."""
    from emulators_amd import _native as native
    print(f"arnoldi gram n={n} steps={steps} {dtype}", flush=True)
    g = torch.Generator().manual_seed(500 + n)
    size = 2**n
    v0 = torch.randn(size, dtype=torch.complex128, generator=g).to(gpu_device, dtype)
    s0 = 1.0 / math.sqrt(float(native.dot(n, v0)[0]))
    ld = steps + 2
    state = native.arnoldi_gram_state(ld, gpu_device)
    gbasis, gscale = [v0], [s0]
    mbasis, mscale = [v0], [s0]
    tol = 1e-12 if dtype == torch.complex128 else 2e-5
    for j in range(min(steps, size - 1)):
        r = torch.randn(size, dtype=torch.complex128, generator=g).to(gpu_device, dtype)
        wg = r.clone()
        recg = native.arnoldi_gram(n, gbasis, wg, s0, state, ld).cpu().numpy()
        # the chain sees the same vector, scaled as the Gram form scales it
        wm = (r.to(torch.complex128) * gscale[j]).to(dtype)
        recm = native.arnoldi_orthogonalize(n, mbasis, mscale, wm, normalize=True).cpu().numpy()
        colm = np.array([mscale[k] * complex(recm[k, 1], recm[k, 2]) for k in range(j + 1)])
        colg = recg[: j + 1, 1] + 1j * recg[: j + 1, 2]
        ref = max(1.0, float(np.abs(colm).max()))
        assert float(np.abs(colg - colm).max()) <= tol * ref, (j, np.abs(colg - colm).max())
        assert abs(recg[0, 0] - recm[0, 0]) <= tol * recm[0, 0]
        assert abs(recg[j + 1, 0] - recm[j + 1, 0]) <= tol * max(recm[0, 0], 1.0)
        n2 = math.sqrt(recg[j + 1, 0])
        unit_g = wg.to(torch.complex128) / n2
        diff = float((unit_g - wm.to(torch.complex128)).abs().max())
        assert diff <= (1e-11 if dtype == torch.complex128 else 1e-4), (j, diff)
        gbasis.append(wg)
        gscale.append(1.0 / n2)
        mbasis.append(wm)
        mscale.append(1.0)
    if dtype == torch.complex128:
        a = torch.stack([b.to(torch.complex128) * s for b, s in zip(gbasis, gscale)])
        gram = a.conj() @ a.T
        eye = torch.eye(len(gbasis), dtype=torch.complex128, device=gpu_device)
        assert float((gram - eye).abs().max()) < 1e-13


@pytest.mark.parametrize("n,dt,tol", [(8, 0.5, 1e-10), (10, 1.0, 1e-8), (14, 0.3, 1e-10)])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_arnoldi_engines_agree(gpu_device, monkeypatch, n, dt, tol, dtype):
    """The three Arnoldi engines on one noisy step: the one-read-back engine
    (unnormalised basis, host scales), the pipelined one (device-normalised
    basis, host one iteration behind) and the library's driver
    (emu_sv_arnoldi_exp, the same loop in native code).  The same iteration
    count and the same result to rounding: 1e-12 in complex128; in
    complex64 the engines round different basis vectors (unit against
    unnormalised) at 6e-8 each over about 20 iterations, so 1e-5 relative.
    The pipelined and the native engine store the same vectors: 1e-12 /
    1e-6 between them."""
    import importlib

    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    kx = importlib.import_module("emulators_amd.emu_base.math.krylov_exp")
    print(f"arnoldi engines n={n} dt={dt} {dtype}", flush=True)
    om, de, ph, u = _params(n, 60 + n, True)
    h = RydbergHamiltonian(om, de, ph, u, device=gpu_device, noise=orc.noise_from_lindbladians(LINDBLADS))
    torch.manual_seed(n)
    psi = torch.randn(2**n, dtype=torch.complex128).to(gpu_device, dtype)
    op = kx.ScaledHamiltonian(h, -1j * dt)
    monkeypatch.setattr(kx, "PIPELINE", False)
    plain = kx.krylov_exp_impl(op, psi.clone(), is_hermitian=False, exp_tolerance=tol, norm_tolerance=tol)
    piped = kx._pipelined_arnoldi(op, psi.clone(), tol, tol, 100)
    native = kx._native_arnoldi(op, psi.clone(), tol, tol, 100)
    scale = float(plain.result.abs().max())
    for got, bound in ((piped, 1e-12 if dtype == torch.complex128 else 1e-5),
                       (native, 1e-12 if dtype == torch.complex128 else 1e-5)):
        assert got.converged and not got.happy_breakdown
        assert got.iteration_count == plain.iteration_count
        diff = float((got.result - plain.result).abs().max())
        assert diff <= bound * scale, diff
    diff = float((native.result - piped.result).abs().max())
    assert diff <= (1e-12 if dtype == torch.complex128 else 1e-6) * scale, diff
    # the dispatcher takes the native driver (n <= 20)
    monkeypatch.setattr(kx, "PIPELINE", True)
    calls = []
    real = kx._native_arnoldi
    monkeypatch.setattr(kx, "_native_arnoldi", lambda *a, **k: calls.append(1) or real(*a, **k))
    kx.krylov_exp_impl(op, psi.clone(), is_hermitian=False, exp_tolerance=tol, norm_tolerance=tol)
    assert calls


def test_native_arnoldi_happy_breakdown(gpu_device):
    """A basis state under a diagonal noisy H (no drive, dephasing only) is an
    eigenvector: the first Gram-Schmidt leaves nothing, the driver reports a
    happy breakdown after one iteration and returns exp(dt E) psi."""
    import importlib

    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    kx = importlib.import_module("emulators_amd.emu_base.math.krylov_exp")
    n = 6
    _, de, _, u = _params(n, 3, False)
    deph = [math.sqrt(0.4) * torch.tensor([[1, 0], [0, -1]], dtype=torch.complex128)]
    k = orc.noise_from_lindbladians(deph)
    h = RydbergHamiltonian(torch.zeros(n, dtype=torch.float64), de, torch.zeros(n, dtype=torch.float64), u,
                           device=gpu_device, noise=k)
    psi = torch.zeros(2**n, dtype=torch.complex128, device=gpu_device)
    psi[5] = 0.5
    dt = 0.3
    got = kx._native_arnoldi(kx.ScaledHamiltonian(h, -1j * dt), psi.clone(), 1e-10, 1e-10, 100)
    assert got.converged and got.happy_breakdown and got.iteration_count == 1
    dense = orc.dense_noisy_hamiltonian(torch.zeros(n), de, torch.zeros(n), u, k)
    expected = torch.linalg.matrix_exp(-1j * dt * dense) @ psi.cpu()
    assert float((got.result.cpu() - expected).abs().max()) < 1e-12


@pytest.mark.parametrize("n,dt,tol", [(4, 0.3, 1e-10), (8, 0.5, 1e-10), (10, 1.0, 1e-8)])
@pytest.mark.parametrize("phase", [False, True])
def test_arnoldi_krylov_vs_oracle(gpu_device, n, dt, tol, phase):
    from emulators_amd.emu_base.math.krylov_exp import (KrylovStats, ScaledHamiltonian,
                                                         krylov_exp_impl)
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    print(f"arnoldi krylov n={n} dt={dt} phase={phase}", flush=True)
    om, de, ph, u = _params(n, 40 + n, phase)
    k = orc.noise_from_lindbladians(LINDBLADS)
    torch.manual_seed(n)
    psi = torch.randn(2**n, dtype=torch.complex128)
    href = orc.OracleHamiltonian(om, de, ph, u, noise=k)
    ref = orc.krylov_exp_impl(lambda x: -1j * dt * (href * x), psi.clone(), is_hermitian=False,
                              exp_tolerance=tol, norm_tolerance=tol)
    h = RydbergHamiltonian(om, de, ph, u, device=gpu_device, noise=k)
    got = krylov_exp_impl(ScaledHamiltonian(h, -1j * dt), psi.to(gpu_device), is_hermitian=False,
                          exp_tolerance=tol, norm_tolerance=tol)
    assert got.converged and ref.converged
    assert got.iteration_count == ref.iteration_count
    assert float((got.result.cpu() - ref.result).abs().max()) < 10 * tol
    # against the dense exponential too
    dense = torch.linalg.matrix_exp(-1j * dt * orc.dense_noisy_hamiltonian(om, de, ph, u, k)) @ psi \
        if n <= 8 else None
    if dense is not None:
        assert float((got.result.cpu() - dense).abs().max()) < 10 * tol


def test_arnoldi_not_converged_raises(gpu_device):
    from emulators_amd.emu_base.math.krylov_exp import ScaledHamiltonian, krylov_exp
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    n = 8
    om, de, ph, u = _params(n, 5, False)
    h = RydbergHamiltonian(om * 10, de, ph, u, device=gpu_device,
                           noise=orc.noise_from_lindbladians(LINDBLADS))
    psi = torch.randn(2**n, dtype=torch.complex128, device=gpu_device)
    with pytest.raises(RecursionError):
        krylov_exp(ScaledHamiltonian(h, -1j * 2.0), psi, 1e-12, 1e-12, is_hermitian=False,
                   max_krylov_dim=4)
    with pytest.raises(ValueError):  # the Lanczos engine refuses a non-Hermitian H
        krylov_exp(ScaledHamiltonian(h, -1j * 0.1), psi, 1e-10, 1e-10, is_hermitian=True)


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_jump_helpers_vs_dense(gpu_device, dtype):
    from emulators_amd.emu_sv.algebra import apply, expect_batch
    n = 6
    torch.manual_seed(9)
    psi = torch.randn(2**n, dtype=torch.complex128)
    ops = torch.stack([L.conj().T @ L for L in LINDBLADS])
    got = expect_batch(psi.to(gpu_device, dtype), ops, n)
    ref = orc.expect_batch(psi, ops, n)
    tol = 1e-11 if dtype == torch.complex128 else 2e-5 * float(psi.norm() ** 2)
    assert float((got - ref).abs().max()) < tol
    for q in (0, 3, 5):
        src = psi.to(gpu_device, dtype)
        d = apply(src, q, LINDBLADS[2])
        assert torch.equal(src.cpu(), psi.to(dtype))  # the input is left as it was
        exp_ = orc.apply_1q(psi, q, LINDBLADS[2])
        assert float((d.cpu().to(torch.complex128) - exp_).abs().max()) < (
            1e-13 if dtype == torch.complex128 else 1e-5)


@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("seed", [3, 7])
def test_monte_carlo_trajectory_vs_oracle(gpu_device, seed):
    """Seeded trajectories: the same jump sequence (qubits and times) and the
    same states as the oracle's quantum-jump stepper."""
    from emulators_amd.emu_sv.time_evolution import EvolveMonteCarlo
    n, steps, dt, tol = 4, 12, 0.1, 1e-10
    print(f"monte carlo seed={seed}", flush=True)
    om, de, ph, u = _params(n, seed, False)
    lind = [math.sqrt(3.0) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128),
            math.sqrt(1.0) * torch.tensor([[0, 1], [0, 0.5]], dtype=torch.complex128)]
    psi0 = torch.zeros(2**n, dtype=torch.complex128)
    psi0[0] = 1.0
    random.seed(seed)
    ref = orc.OracleMonteCarlo()
    states_ref = []
    psi = psi0.clone()
    for _ in range(steps):
        psi = ref.step(dt, om, de, ph, u, psi, tol, lind)
        states_ref.append(psi.clone())
    random.seed(seed)
    first = random.uniform(0.0, 1.0)
    random.seed(seed)
    mc = EvolveMonteCarlo()
    assert mc.jump_threshold == first  # the construction draw
    psi = psi0.to(gpu_device)
    thresholds = []
    for s in range(steps):
        psi, _ = mc.apply(dt, om, de, ph, u, psi, tol, lind)
        thresholds.append(mc.jump_threshold)
        assert float((psi.cpu() - states_ref[s]).abs().max()) < 1e-8, f"step {s}"
    assert mc.jump_threshold == pytest.approx(ref.jump_threshold, rel=1e-9)
    assert len(ref.jumps) >= 1  # the case exercises at least one jump


@pytest.mark.usefixtures("restore_rng_state")
def test_noisy_backend_run(gpu_device):
    """SVBackend on a SequenceData with stochastic Lindblad noise runs the
    Monte-Carlo stepper (sv_backend_impl.py:84-90) and reports occupations of
    the normalised state."""
    from emulators_amd import workloads
    from emulators_amd.emu_sv import Occupation, SVConfig
    from emulators_amd.emu_sv.time_evolution import EvolveMonteCarlo
    from emulators_amd.emu_sv.sv_backend_impl import SVBackendImpl
    import dataclasses
    random.seed(1)
    cfg = SVConfig(observables=[Occupation(evaluation_times=[1.0])], log_level=40)
    seq = workloads.chain10(cfg)
    seq = dataclasses.replace(seq, lindblad_ops=[math.sqrt(0.5) * torch.tensor(
        [[0, 1], [0, 0]], dtype=torch.complex128)], has_stochastic_noise=True)
    impl = SVBackendImpl(cfg, seq)
    assert isinstance(impl.stepper, EvolveMonteCarlo)
    res = impl._run()
    occ = torch.as_tensor(res.get_result(cfg.observables[0].tag, 1.0))
    assert occ.shape == (10,)
    assert torch.all(occ >= -1e-12) and torch.all(occ <= 1 + 1e-12)


@pytest.mark.parametrize("n,dtype,cmin", [(15, torch.complex128, 2), (16, torch.complex128, 4),
                                          (18, torch.complex128, 2), (16, torch.complex64, 4),
                                          (17, torch.complex64, 5), (19, torch.complex64, 4)])
@pytest.mark.parametrize("phase", [False, True])
def test_noisy_hpsi_btile_vs_oracle(gpu_device, n, dtype, cmin, phase):
    """The big-tile kernel with the noise term (bare H.psi above one tile of
    2^14 complex128 / 2^15 complex64 amplitudes) against the oracle's
    index_add restatement of the noisy H (1e-12 relative in complex128,
    2e-5 in complex64), and the fused dots of its last pass against the
    output."""
    from emulators_amd import _native as native
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    print(f"noisy btile n={n} {dtype} cmin={cmin} phase={phase}", flush=True)
    om, de, ph, u = _params(n, 70 + n, phase)
    k = orc.noise_from_lindbladians(LINDBLADS)
    h = RydbergHamiltonian(om, de, ph, u, device=gpu_device, noise=k)
    torch.manual_seed(n + 1)
    psi = torch.randn(2**n, dtype=torch.complex128)
    psi /= psi.norm()
    sigma = 0.3 - 0.7j
    expected = sigma * (orc.OracleHamiltonian(om, de, ph, u, noise=k) * psi)
    v = psi.to(gpu_device, dtype)
    w = torch.empty_like(v)
    dots = native.hpsi(n, h.device_params(gpu_device), sigma, v, 1.0, w, dots=True, complex_coef=True,
                       noise=True, kernel="btile", min_contig=cmin).cpu()
    got = w.cpu().to(torch.complex128)
    rel = float((got - expected).abs().max()) / float(expected.abs().max())
    assert rel <= (1e-12 if dtype == torch.complex128 else 2e-5), rel
    wc = got
    vc = v.cpu().to(torch.complex128)
    tol = 1e-12 if dtype == torch.complex128 else 1e-5
    assert abs(float(dots[0]) - float(wc.norm() ** 2)) <= tol * float(wc.norm() ** 2)
    ov = complex(torch.vdot(vc, wc))
    assert abs(complex(float(dots[1]), float(dots[2])) - ov) <= tol * float(wc.norm())
    # the LDS kernel with the noise term gives the same vector
    w2 = torch.empty_like(v)
    native.hpsi(n, h.device_params(gpu_device), sigma, v, 1.0, w2, complex_coef=True, noise=True,
                kernel="lds")
    d = float((w2.cpu().to(torch.complex128) - got).abs().max()) / float(expected.abs().max())
    assert d <= (1e-13 if dtype == torch.complex128 else 2e-6), d


_MC16 = {"n": 16, "steps": 2, "dt": 0.04, "tol": 1e-10, "seed": 11}


def _mc16_case():
    c = _MC16
    om, de, ph, u = _params(c["n"], c["seed"], False)
    lind = [math.sqrt(2.0) * torch.tensor([[0, 1], [0, 0]], dtype=torch.complex128),
            math.sqrt(0.5) * torch.tensor([[1, 0], [0, -1]], dtype=torch.complex128)]
    g = torch.Generator().manual_seed(c["seed"])
    psi0 = torch.randn(2 ** c["n"], dtype=torch.complex128, generator=g)
    return om, de, ph, u, lind, psi0 / psi0.norm()


_mc16_ref: dict = {}


def _mc16_oracle():
    """The oracle's states after each step, its jumps and final threshold
    (computed once per session: about 15 s of CPU)."""
    if not _mc16_ref:
        c = _MC16
        om, de, ph, u, lind, psi0 = _mc16_case()
        state = random.getstate()
        random.seed(c["seed"])
        ref = orc.OracleMonteCarlo()
        states, psi = [], psi0.clone()
        for _ in range(c["steps"]):
            psi = ref.step(c["dt"], om, de, ph, u, psi, c["tol"], lind)
            states.append(psi.clone())
        random.setstate(state)
        _mc16_ref.update(states=states, jumps=list(ref.jumps), threshold=ref.jump_threshold)
    return _mc16_ref


@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("schedule", ["c128=btile:0:2", "c128=lds:0:0", ""])
def test_monte_carlo_step_n16_vs_oracle(gpu_device, monkeypatch, schedule):
    """Monte-Carlo steps of a 16-atom register on the tuned noisy kernels
    (big tiles, the LDS kernel, and the autotuner's own pick) against the
    oracle's quantum-jump stepper: a strong decay from a random state, so
    that the seeded threshold is crossed in the first step (Brent search,
    jump draws, renormalisation); the same states within 1e-8 and the same
    final threshold."""
    from emulators_amd.emu_sv.time_evolution import EvolveMonteCarlo
    monkeypatch.setenv("EMU_SV_SCHEDULE", schedule)
    c = _MC16
    print(f"monte carlo n=16 schedule={schedule!r}", flush=True)
    ref = _mc16_oracle()
    assert len(ref["jumps"]) >= 1  # the case exercises a jump
    om, de, ph, u, lind, psi0 = _mc16_case()
    random.seed(c["seed"])
    mc = EvolveMonteCarlo()
    psi = psi0.to(gpu_device)
    for s in range(c["steps"]):
        psi, _ = mc.apply(c["dt"], om, de, ph, u, psi, c["tol"], lind)
        diff = float((psi.cpu() - ref["states"][s]).abs().max())
        assert diff < 1e-8, f"step {s}: {diff}"
    assert mc.jump_threshold == pytest.approx(ref["threshold"], rel=1e-9)


def test_concurrent_streams_arnoldi_drivers(gpu_device):
    """Two host threads, each on its own torch stream, run noisy Krylov steps
    through the library's Arnoldi driver at the same time: one register in
    the Gram-matrix form (n = 14), one in the one-workgroup Gram-Schmidt
    (n = 12).  Every buffer of a run is per stream or per call, so both
    results equal the oracle's."""
    import threading

    from emulators_amd.emu_base.math.krylov_exp import ScaledHamiltonian, krylov_exp
    from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian
    k = orc.noise_from_lindbladians(LINDBLADS)
    dt, tol, steps = 0.05, 1e-10, 4
    cases = []
    for n in (14, 12):
        om, de, ph, u = _params(n, 90 + n, True)
        psi = torch.randn(2**n, dtype=torch.complex128, generator=torch.Generator().manual_seed(n))
        cases.append((n, om, de, ph, u, psi / psi.norm()))
    results: dict = {}
    errors: list = []
    barrier = threading.Barrier(2)

    def worker(i):
        try:
            n, om, de, ph, u, psi = cases[i]
            stream = torch.cuda.Stream(device=gpu_device)
            with torch.cuda.stream(stream):
                ham = RydbergHamiltonian(om, de, ph, u, device=gpu_device, noise=k)
                v = psi.to(gpu_device)
                barrier.wait()
                for _ in range(steps):
                    v = krylov_exp(ScaledHamiltonian(ham, -1j * dt), v, tol, tol, is_hermitian=False)
                stream.synchronize()
                results[i] = v.cpu()
        except Exception as exc:  # reported by the main thread
            errors.append((i, repr(exc)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not errors, errors
    for i, (n, om, de, ph, u, psi) in enumerate(cases):
        ho = orc.OracleHamiltonian(om, de, ph, u, noise=k)
        ref = psi.clone()
        for _ in range(steps):
            ref = orc.krylov_exp(lambda x: -1j * dt * (ho * x), ref.clone(), tol, tol, is_hermitian=False)
        assert float((results[i] - ref).abs().max()) < 1e-9, (n, float((results[i] - ref).abs().max()))
