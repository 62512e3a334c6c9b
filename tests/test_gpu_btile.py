"""The big-tile H.psi kernel (EMU_SV_KERNEL_BTILE: 1024-thread workgroups
holding 2^14 complex128 / 2^15 complex64 amplitudes) against the oracle.

* H.psi of random Hamiltonians (with and without phases) on random states
  for every size from one tile plus one qubit to 24: complex128 within
  1e-12 of the oracle's index_add_ product (relative), complex64 within 2e-6;
* the fused Lanczos iteration (orthogonalise + passes + reductions) against
  the LDS kernel's on the same inputs;
* Krylov steps through the pipelined engine with the btile schedule forced,
  against the oracle's krylov_exp (same iteration count, 1e-10).
"""

import math

import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu


def _params(n, with_phase, seed):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, dtype=torch.float64, generator=g) * 10 + 1
    de = torch.randn(n, dtype=torch.float64, generator=g) * 5
    ph = torch.randn(n, dtype=torch.float64, generator=g) if with_phase else torch.zeros(n, dtype=torch.float64)
    u = torch.triu(torch.rand(n, n, dtype=torch.float64, generator=g) * 30, 1)
    return om, de, ph, u + u.T


@pytest.mark.parametrize("n", [15, 16, 17, 19, 21, 24])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
@pytest.mark.parametrize("phase", [False, True])
def test_btile_hpsi_matches_oracle(gpu_device, n, dtype, phase):
    from emulators_amd import _native as native
    print(f"btile hpsi n={n} {dtype} phase={phase}", flush=True)
    om, de, ph, u = _params(n, phase, 100 * n + phase)
    block, cplx = native.pack_hamiltonian(om.numpy(), de.numpy(), ph.numpy(), u.numpy())
    params = torch.from_numpy(block).to(gpu_device)
    psi = torch.randn(2**n, dtype=torch.complex128, generator=torch.Generator().manual_seed(n))
    v = psi.to(gpu_device, dtype)
    ref = orc.OracleHamiltonian(om, de, ph, u) * v.cpu().to(torch.complex128)
    w = torch.empty_like(v)
    sigma = 0.3 - 0.7j
    d = native.hpsi(n, params, sigma, v, 0.5, w, dots=True, complex_coef=cplx, kernel="btile",
                    min_contig=4 if dtype == torch.complex64 else 2).cpu()
    ref = sigma * 0.5 * ref
    got = w.cpu().to(torch.complex128)
    tol = 1e-12 if dtype == torch.complex128 else 2e-6
    assert float((got - ref).norm() / ref.norm()) < tol
    vs = 0.5 * v.cpu().to(torch.complex128)
    assert math.isclose(float(d[0]), float(ref.norm() ** 2), rel_tol=tol * 10)
    dv = complex(torch.vdot(vs, ref))
    assert abs(complex(d[1], d[2]) - dv) < tol * 10 * float(vs.norm() * ref.norm())


@pytest.mark.parametrize("n", [16, 20])
@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_btile_lanczos_iteration_matches_lds(gpu_device, n, dtype):
    """One fused Lanczos iteration, btile against the LDS kernel: the
    orthogonalised vector z, r = sigma H z and the 12 reductions."""
    from emulators_amd import _native as native
    om, de, ph, u = _params(n, False, 7 + n)
    block, cplx = native.pack_hamiltonian(om.numpy(), de.numpy(), ph.numpy(), u.numpy())
    params = torch.from_numpy(block).to(gpu_device)
    g = torch.Generator(device=gpu_device).manual_seed(n)
    w0, pa, pb = (torch.randn(2**n, dtype=dtype, device=gpu_device, generator=g) for _ in range(3))
    outs = []
    for kw in (dict(kernel="lds", tile_log=11, min_contig=2),
               dict(kernel="btile", min_contig=4 if dtype == torch.complex64 else 2)):
        w = w0.clone()
        r = torch.empty_like(w)
        o = native.lanczos_matvec(n, params, -0.01j, w, 0.9, pa, 1.1, 0.2 - 0.1j, pb, 0.8, 0.3 + 0.05j,
                                  r, complex_coef=cplx, **kw).cpu().clone()
        outs.append((w.cpu().to(torch.complex128), r.cpu().to(torch.complex128), o))
    (z1, r1, o1), (z2, r2, o2) = outs
    tol = 1e-13 if dtype == torch.complex128 else 1e-6
    assert float((z1 - z2).norm() / z1.norm()) < tol
    assert float((r1 - r2).norm() / r1.norm()) < 10 * tol
    for k in (0, 1, 2, 3, 4, 5, 6, 7):  # ||z||^2, <z|pb>, <z|w>, ||r||^2, <z|r>
        assert abs(float(o1[k]) - float(o2[k])) <= 10 * tol * max(1.0, abs(float(o1[k])), float(o1[0]),
                                                                  float(o1[5])), k


@pytest.mark.parametrize("n,dtype", [(16, torch.complex128), (18, torch.complex64),
                                     (20, torch.complex128)])
def test_btile_krylov_steps_match_oracle(gpu_device, monkeypatch, n, dtype):
    """Pipelined Krylov steps with the btile schedule forced: the oracle's
    iteration count and result (complex128 within 1e-10; complex64 to
    1 - F <= 1e-6)."""
    from emulators_amd.emu_base.math.krylov_exp import ScaledHamiltonian, krylov_exp_impl
    from emulators_amd.emu_sv import RydbergHamiltonian
    monkeypatch.setenv("EMU_SV_SCHEDULE", "c128=btile:0:2,c64=btile:0:4")
    om, de, ph, u = _params(n, n == 20, 3 * n)
    psi = torch.randn(2**n, dtype=torch.complex128, generator=torch.Generator().manual_seed(5 * n))
    psi /= psi.norm()
    dt, tol = 0.002, 1e-10
    ham = RydbergHamiltonian(om, de, ph, u, device=gpu_device)
    got = krylov_exp_impl(ScaledHamiltonian(ham, -1j * dt), psi.to(gpu_device, dtype), True, tol, tol)
    ho = orc.OracleHamiltonian(om, de, ph, u)
    ref = orc.krylov_exp_impl(lambda x: -1j * dt * (ho * x), psi.clone(), is_hermitian=True,
                              exp_tolerance=tol, norm_tolerance=tol)
    out = got.result.cpu().to(torch.complex128)
    if dtype == torch.complex128:
        assert got.iteration_count == ref.iteration_count
        assert float((out - ref.result).norm()) < tol
    else:
        assert 1.0 - float(orc.fidelity(out, ref.result)) <= 1e-6
