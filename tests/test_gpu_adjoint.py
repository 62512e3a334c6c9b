"""GPU tests of the hand-written whole-pulse adjoint and its pieces.

Checks: the parameter-trace kernel against the oracle's restatement of the
reference traces; the energy Function's gradients; the pulse gradient
against torch autograd through dense matrix exponentials (small N), against
the oracle's restatement of the reference's per-step autograd chain, and
against central finite differences on the 16-atom workload.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import emu_sv_oracle as orc

pytestmark = pytest.mark.gpu


def _rand_params(n, seed, phases=True):
    g = torch.Generator().manual_seed(seed)
    om = torch.rand(n, generator=g, dtype=torch.float64) * 10.0
    de = (torch.rand(n, generator=g, dtype=torch.float64) - 0.5) * 16.0
    ph = (torch.rand(n, generator=g, dtype=torch.float64) * 6.0 if phases
          else torch.zeros(n, dtype=torch.float64))
    u = torch.rand(n, n, generator=g, dtype=torch.float64) * 20.0
    u = u + u.T
    u.fill_diagonal_(0.0)
    return om, de, ph, u


@pytest.mark.parametrize("dtype", [torch.complex128, torch.complex64])
def test_parameter_traces_vs_oracle(gpu_device, dtype):
    from emulators_amd import _native as native
    from emulators_amd.emu_sv.adjoint import _traces_to_grads

    n, count = 7, 3
    g = torch.Generator().manual_seed(5)
    w = torch.randn(count, 2**n, dtype=torch.complex128, generator=g)
    z = torch.randn(count, 2**n, dtype=torch.complex128, generator=g)
    om, _, ph, _ = _rand_params(n, 6)
    fac = 0.3 - 0.7j
    tr = native.parameter_traces(n, w.to(gpu_device, dtype), z.to(gpu_device, dtype), pairs=True)
    got = _traces_to_grads(tr, fac, om.numpy(), ph.numpy())
    ref = orc.parameter_traces(w.to(dtype).to(torch.complex128), z.to(dtype).to(torch.complex128),
                               om, ph, fac)
    tol = 1e-11 if dtype == torch.complex128 else 1e-4
    for a, b in zip(got, ref):
        b = b.numpy()
        assert np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b))), (a, b)


def test_energy_expectation_gradients(gpu_device):
    from emulators_amd.emu_sv.adjoint import EnergyExpectation

    n = 6
    om, de, ph, u = _rand_params(n, 7)
    g = torch.Generator().manual_seed(8)
    psi = torch.randn(2**n, dtype=torch.complex128, generator=g)
    psi /= psi.norm()
    params = [t.clone().requires_grad_(True) for t in (om, de, ph, u)]
    pd = psi.to(gpu_device).requires_grad_(True)
    e = EnergyExpectation.apply(*params, pd)
    e.backward()
    # oracle: E and d/dpsi from dense H, parameter grads by dense autograd
    qp = [t.clone().requires_grad_(True) for t in (om, de, ph, u)]
    h = orc.dense_rydberg_hamiltonian_autograd(*qp)
    e_ref = torch.vdot(psi, h @ psi).real
    e_ref.backward()
    assert abs(float(e.detach()) - float(e_ref)) < 1e-10 * max(1.0, abs(float(e_ref)))
    for a, b in zip(params, qp):
        ga, gb = a.grad, b.grad
        if b is qp[3]:
            gb = torch.triu(gb, 1)  # H reads U[i, j] for i < j only
        assert torch.allclose(ga, gb, atol=1e-10, rtol=1e-10), (ga, gb)
    hpsi = h @ psi
    assert torch.allclose(pd.grad.cpu(), 2 * hpsi, atol=1e-10)


def _dense_pulse_loss(om, de, ph, u, dts, psi0):
    psi = psi0
    s = om.shape[0]
    h = None
    for k in range(s):
        h = orc.dense_rydberg_hamiltonian_autograd(om[k], de[k], ph[k], u)
        psi = torch.linalg.matrix_exp(-1j * dts[k] * h) @ psi
    n = om.shape[1]
    pc = torch.tensor([bin(x).count("1") for x in range(2**n)], dtype=torch.float64)
    return torch.vdot(psi, h @ psi).real + ((psi.abs() ** 2) * pc).sum()


@pytest.mark.parametrize("mode,budget", [("recompute", 1 << 34), ("recompute", 0),
                                         ("recompute", 3 * 2**6 * 16 * 12), ("reverse", 0),
                                         ("batched", 0)])
def test_pulse_gradient_vs_dense_autograd(gpu_device, mode, budget):
    """Per-atom Omega/delta/phi for every step, U and psi_0 at N=6, 9 steps of
    uneven dt, checkpoints every 4 steps: against autograd through dense
    matrix exponentials.  Recompute mode with every forward basis kept, with
    none kept (all segments recomputed) and with only the newest kept."""
    from emulators_amd.emu_sv.adjoint import EnergyExpectation, evolve_pulse, total_occupation

    n, s = 6, 9
    g = torch.Generator().manual_seed(11)
    om = torch.rand(s, n, generator=g, dtype=torch.float64) * 10.0
    de = (torch.rand(s, n, generator=g, dtype=torch.float64) - 0.5) * 16.0
    ph = torch.rand(s, n, generator=g, dtype=torch.float64) * 6.0
    _, _, _, u = _rand_params(n, 12)
    dts = torch.tensor([0.01 + 0.002 * (k % 3) for k in range(s)], dtype=torch.float64)
    psi0 = torch.randn(2**n, dtype=torch.complex128, generator=g)
    psi0 /= psi0.norm()

    params = [t.clone().requires_grad_(True) for t in (om, de, ph, u)]
    p0 = psi0.to(gpu_device).requires_grad_(True)
    psi = evolve_pulse(*params, p0, dts, 1e-12, 4, 100, mode, budget)
    loss = EnergyExpectation.apply(params[0][-1], params[1][-1], params[2][-1], params[3], psi) \
        + total_occupation(psi)
    loss.backward()

    qp = [t.clone().requires_grad_(True) for t in (om, de, ph, u)]
    q0 = psi0.clone().requires_grad_(True)
    ref = _dense_pulse_loss(*qp, dts, q0)
    ref.backward()
    assert abs(float(loss) - float(ref)) < 1e-9 * abs(float(ref))
    names = ["omegas", "deltas", "phis", "U"]
    for name, a, b in zip(names, params, qp):
        gb = torch.triu(b.grad, 1) if name == "U" else b.grad
        err = float((a.grad - gb).abs().max())
        assert err <= 1e-7 * float(gb.abs().max()), (name, err, float(gb.abs().max()))
    assert torch.allclose(p0.grad.cpu(), q0.grad, atol=1e-7 * float(q0.grad.abs().max()))


def _random16_tensors(steps=None):
    from emulators_amd import workloads

    pos, om, de = workloads.random16_parameters()
    if steps is not None:
        om, de = om[:steps], de[:steps]
    u = torch.as_tensor(workloads.interaction_matrix_from_positions(pos, workloads.C6))
    return torch.as_tensor(om), torch.as_tensor(de), u


def _gpu_loss_and_grads(om_k, de_k, u, tol, mode="recompute", every=10, grads=True,
                        budget=1 << 34):
    from emulators_amd.emu_sv.adjoint import EnergyExpectation, evolve_pulse, total_occupation

    n, s = u.shape[0], om_k.shape[0]
    om = om_k.clone().requires_grad_(grads)
    de = de_k.clone().requires_grad_(grads)
    omegas = om[:, None].expand(s, n)
    deltas = de[:, None].expand(s, n)
    phis = torch.zeros(s, n, dtype=torch.float64)
    psi0 = torch.zeros(2**n, dtype=torch.complex128, device="cuda")
    psi0[0] = 1.0
    dts = torch.full((s,), 0.01, dtype=torch.float64)
    psi = evolve_pulse(omegas, deltas, phis, u, psi0, dts, tol, every, 100, mode, budget)
    loss = EnergyExpectation.apply(omegas[-1], deltas[-1], phis[-1], u, psi) + total_occupation(psi)
    if grads:
        loss.backward()
        return float(loss), om.grad.clone(), de.grad.clone()
    return float(loss), None, None


def test_random16_gradient_vs_reference_autograd(gpu_device):
    """The 16-atom workload (global Omega_k, delta_k), first 12 steps: the
    adjoint against the oracle's restatement of the reference's per-step
    autograd (ReferenceStep chained over the pulse)."""
    om_k, de_k, u = _random16_tensors(12)
    n, s = 16, 12
    loss, g_om, g_de = _gpu_loss_and_grads(om_k, de_k, u, 1e-10, every=5)

    om = om_k.clone().requires_grad_(True)
    de = de_k.clone().requires_grad_(True)
    psi = torch.zeros(2**n, dtype=torch.complex128)
    psi[0] = 1.0
    zeros = torch.zeros(n, dtype=torch.complex128)
    for k in range(s):
        psi = orc.ReferenceStep.apply(0.01, om[k].expand(n).to(torch.complex128),
                                      de[k].expand(n).to(torch.complex128), zeros, u, psi, 1e-10, 100)
    ham = orc.OracleHamiltonian(om_k[-1].expand(n), de_k[-1].expand(n), zeros, u)
    hpsi = ham * psi.detach()
    pc = torch.tensor([bin(x).count("1") for x in range(2**n)], dtype=torch.float64)
    lam = 2 * (hpsi + pc * psi.detach())
    ref_loss = float(torch.vdot(psi.detach(), hpsi).real + ((psi.detach().abs() ** 2) * pc).sum())
    psi.backward(lam)
    # explicit dependence of the final energy on the last step's parameters
    e_om, e_de, _, _ = orc.parameter_traces(psi.detach()[None], psi.detach()[None],
                                            om_k[-1].expand(n), zeros.real, 1.0)
    ref_om = om.grad.clone()
    ref_de = de.grad.clone()
    ref_om[-1] += e_om.sum()
    ref_de[-1] += e_de.sum()
    assert abs(loss - ref_loss) < 1e-9 * abs(ref_loss)
    for a, b in ((g_om, ref_om), (g_de, ref_de)):
        err = float((a - b).abs().max())
        assert err <= 1e-6 * float(b.abs().max()), (err, float(b.abs().max()))


def test_random16_gradient_vs_central_differences(gpu_device):
    """The full 16-atom, 100-step, 200-parameter workload: adjoint (checkpoints
    every 10 steps, all three backward modes) against central differences with
    h = 1e-6 (complex128) on Omega and delta at the first, middle and last step."""
    om_k, de_k, u = _random16_tensors()
    tol = 1e-12
    _, g_om, g_de = _gpu_loss_and_grads(om_k, de_k, u, tol, "recompute")
    scale = max(float(g_om.abs().max()), float(g_de.abs().max()))
    for mode in ("reverse", "batched", "recompute"):
        _, g_om2, g_de2 = _gpu_loss_and_grads(om_k, de_k, u, tol, mode, budget=0)
        assert float((g_om - g_om2).abs().max()) < 1e-7 * scale, mode
        assert float((g_de - g_de2).abs().max()) < 1e-7 * scale, mode
    h = 1e-6
    for which, k in (("om", 0), ("om", 50), ("om", 99), ("de", 0), ("de", 50), ("de", 99)):
        vals = []
        for sgn in (1.0, -1.0):
            o, d = om_k.clone(), de_k.clone()
            (o if which == "om" else d)[k] += sgn * h
            vals.append(_gpu_loss_and_grads(o, d, u, tol, grads=False)[0])
        fd = (vals[0] - vals[1]) / (2 * h)
        adj = float((g_om if which == "om" else g_de)[k])
        # central differences of a loss computed to Krylov tolerance carry
        # noise ~ tol * |L| * steps / h: bound the error against the largest
        # gradient component (1e-5 of it), the usual finite-difference test
        assert abs(fd - adj) <= 1e-5 * scale, (which, k, fd, adj, scale)
