"""One differentiable time step exp(-i dt H) psi.

Forward: the fused GPU Lanczos engine (emulators_amd.emu_base.math.krylov_exp).
Backward (per step, as the reference's custom Function,
emu_sv/time_evolution/torch.py:221-378): the Frechet derivative of exp along
|psi><g| from two Lanczos bases (``double_krylov``), the parameter traces
Tr(-i dt dH/dtheta Vs dS Vg^dag), and exp(+i dt H) g for the input state.
Vectors stay on the device; the matvecs are the fused H.psi kernels.
"""

from __future__ import annotations

import math
from typing import Any

import numpy as np
import torch

from emulators_amd import _native as native
from emulators_amd import autotune
from emulators_amd.emu_base.math.krylov_exp import (
    KrylovStats,
    ScaledHamiltonian,
    _error_estimate,
    krylov_exp,
)
from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian


def lanczos_basis(op: ScaledHamiltonian, v: torch.Tensor, tolerance: float,
                  max_krylov_dim: int) -> tuple[list[torch.Tensor], np.ndarray]:
    """Normalised Lanczos basis and T of ``v`` under ``op`` with the
    reference's stopping rule (emu_base/math/double_krylov.py:61-113)."""
    ham = op.hamiltonian
    n = native.nqubits_of(v)
    params = ham.device_params(v.device)
    sched = autotune.best_schedule(n, v.dtype, v.device, params, ham.complex)
    v0 = v / v.norm()
    vecs = [v0]
    T = np.zeros((max_krylov_dim + 2, max_krylov_dim + 2), dtype=np.complex128)
    converged = False
    for j in range(max_krylov_dim):
        w = torch.empty_like(v0)
        prev = vecs[j - 1] if j > 0 else None
        d = native.hpsi(n, params, op.sigma, vecs[j], 1.0, w, prev=prev, prev_scale=1.0,
                        dots=True, complex_coef=ham.complex, **sched.kwargs()).cpu().numpy()
        KrylovStats.matvecs += 1
        nrm = math.sqrt(max(d[0], 0.0))
        a = complex(d[3], d[4]) if j > 0 else 0j
        if j > 0:
            T[j - 1, j] = a
        b = complex(d[1], d[2]) - a * complex(d[5], d[6])
        T[j, j] = b
        n2 = math.sqrt(max(native.lanczos_update(n, w, prev, 1.0, a, vecs[j], 1.0, b).item(), 0.0))
        T[j + 1, j] = n2
        if n2 < tolerance:
            converged = True
            break
        native.scale(n, w, 1.0 / n2)
        vecs.append(w)
        T[j + 2, j + 1] = 1.0
        expd = native.expm(T[: j + 3, : j + 3])
        if _error_estimate(expd, j, nrm) < tolerance:
            converged = True
            break
    if not converged:
        raise RecursionError(
            "Lanczos iteration did not converge to precision in allotted number of steps."
        )
    size = len(vecs)
    return vecs, T[:size, :size].copy()


def double_krylov(op: ScaledHamiltonian, state: torch.Tensor, grad: torch.Tensor,
                  tolerance: float, max_krylov_dim: int):
    """(Vs, dS, Vg) with dU(op, |state><grad|) = Vs^T dS Vg^* (double_krylov.py:5-58)."""
    vs, ts = lanczos_basis(op, state, tolerance, max_krylov_dim)
    vg, tg = lanczos_basis(op, grad, tolerance, max_krylov_dim)
    ms, mg = len(vs), len(vg)
    big = np.zeros((ms + mg, ms + mg), dtype=np.complex128)
    big[:ms, :ms] = ts
    big[ms:, ms:] = tg
    big[0, ms] = float(state.norm()) * float(grad.norm())
    ds = native.expm(big)[:ms, ms:]
    return vs, ds, vg


def _flip_bilinear(a: torch.Tensor, b: torch.Tensor, q: int, n: int) -> tuple[complex, complex]:
    """(sum_{x_q=1} conj(a) b(x^e_q), sum_{x_q=0} conj(a) b(x^e_q)): <a|s+_q|b>, <a|s-_q|b>."""
    av = a.view(2**q, 2, -1)
    bv = b.view(2**q, 2, -1)
    up = torch.vdot(av[:, 1].reshape(-1), bv[:, 0].reshape(-1))
    dn = torch.vdot(av[:, 0].reshape(-1), bv[:, 1].reshape(-1))
    return complex(up), complex(dn)


class DifferentiableStateVectorEvolution(torch.autograd.Function):
    """exp(-i dt H(omegas, deltas, phis, U)) psi with the reference's backward."""

    @staticmethod
    def get_hamiltonian(omegas, deltas, phis, interaction_matrix, device) -> RydbergHamiltonian:
        return RydbergHamiltonian(omegas=omegas, deltas=deltas, phis=phis,
                                  interaction_matrix=interaction_matrix, device=device, noise=None)

    @staticmethod
    def evolve(dt, omegas, deltas, phis, interaction_matrix, state, krylov_tolerance,
               max_krylov_dim):
        ham = DifferentiableStateVectorEvolution.get_hamiltonian(
            omegas, deltas, phis, interaction_matrix, state.device)
        res = krylov_exp(ScaledHamiltonian(ham, -1j * dt), state.detach(), krylov_tolerance,
                         krylov_tolerance, is_hermitian=True, max_krylov_dim=max_krylov_dim)
        return res, ham

    @staticmethod
    def forward(ctx: Any, dt, omegas, deltas, phis, interaction_matrix, state, krylov_tolerance,
                max_krylov_dim):
        res, ham = DifferentiableStateVectorEvolution.evolve(
            dt, omegas, deltas, phis, interaction_matrix, state, krylov_tolerance, max_krylov_dim)
        ctx.save_for_backward(omegas, deltas, phis, interaction_matrix, state)
        ctx.dt = dt
        ctx.tolerance = krylov_tolerance
        ctx.max_krylov_dim = max_krylov_dim
        return res, ham

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx: Any, grad_state_out: torch.Tensor, gham: None):
        omegas, deltas, phis, interaction_matrix, state = ctx.saved_tensors
        dt, tol = ctx.dt, ctx.tolerance
        n = len(omegas)
        ham = DifferentiableStateVectorEvolution.get_hamiltonian(
            omegas, deltas, phis, interaction_matrix, state.device)
        op = ScaledHamiltonian(ham, -1j * dt)
        g_om = g_de = g_ph = g_u = g_state = None
        need_params = any(ctx.needs_input_grad[1:5])
        if need_params:
            vs, ds, vg = double_krylov(op, state.detach(), grad_state_out.detach(), tol,
                                       ctx.max_krylov_dim)
            ds_t = torch.as_tensor(ds, device=state.device)
            Vs = torch.stack(vs)
            e_l = ds_t.mT.to(Vs.dtype) @ Vs  # (mg, D)
            Vg = torch.stack(vg)
            ph = torch.as_tensor(phis).detach().cpu().to(torch.float64)
            om = torch.as_tensor(omegas).detach().cpu()
            om = (om.real if om.is_complex() else om).to(torch.float64)
            ph = ph.real if ph.is_complex() else ph
            flips_up = np.zeros(n, dtype=np.complex128)
            flips_dn = np.zeros(n, dtype=np.complex128)
            occ = np.zeros(n, dtype=np.complex128)
            pair = np.zeros((n, n), dtype=np.complex128)
            for l in range(Vg.shape[0]):
                a, b = Vg[l], e_l[l]
                prod = a.conj() * b
                for q in range(n):
                    up, dn = _flip_bilinear(a, b, q, n)
                    flips_up[q] += up
                    flips_dn[q] += dn
                    pv = prod.view(2**q, 2, -1)[:, 1]
                    occ[q] += complex(pv.sum())
                    for q2 in range(q + 1, n):
                        pair[q, q2] += complex(pv.reshape(2**q, 2 ** (q2 - q - 1), 2, -1)[:, :, 1].sum())
            fac = -1j * dt
            eiph = np.exp(1j * ph.numpy())
            g_om_np = (fac * 0.5 * (eiph * flips_up + np.conj(eiph) * flips_dn)).real
            g_de_np = (fac * (-occ)).real
            g_ph_np = (fac * 0.5 * om.numpy() * (1j * eiph * flips_up - 1j * np.conj(eiph) * flips_dn)).real
            g_u_np = np.triu((fac * pair).real, 1)
            # each gradient lands on its parameter's device and dtype
            if ctx.needs_input_grad[1]:
                g_om = torch.as_tensor(g_om_np, device=omegas.device).to(omegas.dtype)
            if ctx.needs_input_grad[2]:
                g_de = torch.as_tensor(g_de_np, device=deltas.device).to(deltas.dtype)
            if ctx.needs_input_grad[3]:
                g_ph = torch.as_tensor(g_ph_np, device=phis.device).to(phis.dtype)
            if ctx.needs_input_grad[4]:
                g_u = torch.as_tensor(g_u_np, device=interaction_matrix.device).to(
                    interaction_matrix.dtype)
        if ctx.needs_input_grad[5]:
            g_state = krylov_exp(ScaledHamiltonian(ham, 1j * dt), grad_state_out.detach().clone(),
                                 tol, tol, max_krylov_dim=ctx.max_krylov_dim)
        return None, g_om, g_de, g_ph, g_u, g_state, None, None
