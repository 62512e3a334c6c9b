"""Hand-written adjoint of a whole pulse: gradients of final observables with
respect to every step's (Omega, delta, phi) and the interaction matrix.

The reference differentiates one Krylov step at a time through its custom
torch.autograd.Function (emu_sv/time_evolution/torch.py:221-378), saving
every step's input state and running, per step, two Lanczos bases, a
Frechet-derivative block exponential and a third Krylov exponential for the
input gradient (emu_base/math/double_krylov.py:5-113).  Here the whole
pulse is one Function:

forward   psi_{k+1} = exp(-i dt_k H_k) psi_k with the fused Lanczos engine;
          only every ``checkpoint_every``-th state is kept.

backward  lambda_S = dL/dpsi_S (torch's convention), then for k = S-1 .. 0,
          with B = +i dt_k H_k and L(s) = exp(u B) for u in [0, 1]:
            * one Lanczos decomposition of lambda_{k+1} under B gives
              lambda_k = exp(B) lambda_{k+1} and the adjoint trajectory
              mu(u) = exp(u B) lambda_{k+1};
            * the state trajectory psi(s) = exp(-i s dt H) psi_k equals
              exp(u B) psi_{k+1} (u = 1 - s): one Lanczos decomposition of
              psi_{k+1} under B gives it and, reverse-propagated,
              psi_k = exp(B) psi_{k+1}; at a checkpoint psi_k is the stored
              forward state instead (mode "reverse", the default).  Mode
              "recompute" re-runs each segment forward from its checkpoint,
              keeping the forward Lanczos bases, and needs no psi Lanczos
              in the backward sweep.
            * dL/dtheta = Re <lambda_{k+1}| dU_k/dtheta |psi_k>
                        = Re[-i dt_k int_0^1 <mu(u)| dH/dtheta |psi(u)> du]
              = Re[-i dt_k |lambda||psi| sum_lk X_lk <w_l|dH/dtheta|v_k>]
              with X_lk = int conj(a_l(u)) b_k(u) du from one Van Loan block
              exponential of the two small Lanczos matrices, Z = X V (one
              GEMM on the device) and the per-qubit traces <W|.|Z>
              (emu_sv_parameter_traces).
All vectors stay on the device; the small matrices are FP64 on the host.
"""

from __future__ import annotations

import math
from typing import Any

import numpy as np
import torch

from emulators_amd import _native as native
from emulators_amd.emu_base.math.krylov_exp import (
    DEFAULT_MAX_KRYLOV_DIM,
    LanczosRun,
    ScaledHamiltonian,
    _fused_lanczos,
)
from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian


def frechet_weights(t_bra: np.ndarray, t_ket: np.ndarray, ket_forward: bool = False) -> np.ndarray:
    """X[l, k] = int_0^1 conj([exp(u T_bra) e_0]_l) [exp(g(u) T_ket) e_0]_k du,
    g(u) = u, or 1 - u when ``ket_forward`` (the ket runs forward while the
    bra runs backward).  Van Loan: the upper-right block of
    exp([[P, Q], [0, C]]) is int_0^1 exp((1-u) P) Q exp(u C) du."""
    mb, mk = t_bra.shape[0], t_ket.shape[0]
    big = np.zeros((mb + mk, mb + mk), dtype=np.complex128)
    big[:mb, :mb] = np.conj(t_bra)
    big[0, mb] = 1.0
    if ket_forward:
        # int exp((1-s) conj(Tb)) Q exp(s Tk^T) ds with bra time 1-s, ket time s
        big[mb:, mb:] = t_ket.T
        return native.expm(big)[:mb, mb:]
    big[mb:, mb:] = -t_ket.T
    return native.expm(big)[:mb, mb:] @ native.expm(np.ascontiguousarray(t_ket.T))


def _as_real(x) -> torch.Tensor:
    t = torch.as_tensor(x).detach()
    if t.is_complex():
        t = t.real
    return t.to(device="cpu", dtype=torch.float64)


class _Pulse:
    """Per-step Hamiltonians of a pulse (parameters on the host)."""

    def __init__(self, omegas, deltas, phis, interaction_matrix, dts, device):
        self.om = _as_real(omegas)
        self.de = _as_real(deltas)
        self.ph = _as_real(phis)
        self.u = _as_real(interaction_matrix)
        self.dts = [float(x) for x in _as_real(dts)]
        self.device = device
        self.steps, self.n = self.om.shape
        if self.de.shape != self.om.shape or self.ph.shape != self.om.shape:
            raise ValueError("omegas, deltas and phis must all be (steps, N)")
        if self.u.shape != (self.n, self.n) or len(self.dts) != self.steps:
            raise ValueError("interaction matrix must be (N, N) and dts of length steps")

    def ham(self, k: int) -> RydbergHamiltonian:
        return RydbergHamiltonian(self.om[k], self.de[k], self.ph[k], self.u, device=self.device)


def _traces_to_grads(tr: native.Traces, fac: complex, om: np.ndarray, ph: np.ndarray):
    """(dOmega, ddelta, dphi, dU) per qubit from the traces of <W|.|Z>, times fac."""
    eiph = np.exp(1j * ph)
    g_om = (fac * 0.5 * (eiph * tr.up + np.conj(eiph) * tr.dn)).real
    g_de = (fac * (-tr.occ)).real
    g_ph = (fac * 0.5 * om * (1j * eiph * tr.up - 1j * np.conj(eiph) * tr.dn)).real
    g_u = np.triu((fac * tr.pair).real, 1) if tr.pair is not None else None
    return g_om, g_de, g_ph, g_u


class PulseEvolution(torch.autograd.Function):
    """psi_S = prod_k exp(-i dt_k H(omegas[k], deltas[k], phis[k], U)) psi_0.

    apply(omegas (S,N), deltas (S,N), phis (S,N), interaction_matrix (N,N),
          psi0 (2**N,) device, dts (S,) in us, tolerance, checkpoint_every,
          max_krylov_dim, mode)
    Gradients flow to all parameter tensors and to psi0."""

    @staticmethod
    def forward(ctx: Any, omegas, deltas, phis, interaction_matrix, psi0, dts, tolerance=1e-10,
                checkpoint_every=10, max_krylov_dim=DEFAULT_MAX_KRYLOV_DIM, mode="reverse"):
        if mode not in ("reverse", "recompute"):
            raise ValueError("mode must be 'reverse' or 'recompute'")
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        pulse = _Pulse(omegas, deltas, phis, interaction_matrix, dts, psi0.device)
        psi = psi0.detach().contiguous()
        checkpoints = {}
        for k in range(pulse.steps):
            if k % checkpoint_every == 0:
                checkpoints[k] = psi.clone()
            run = _fused_lanczos(ScaledHamiltonian(pulse.ham(k), -1j * pulse.dts[k]), psi,
                                 tolerance, tolerance, max_krylov_dim)
            if not run.converged:
                raise RecursionError("exponentiation algorithm did not converge to precision "
                                     "in allotted number of steps.")
            psi = run.combine()
        ctx.pulse = pulse
        ctx.checkpoints = checkpoints
        ctx.tolerance = tolerance
        ctx.every = checkpoint_every
        ctx.max_krylov_dim = max_krylov_dim
        ctx.mode = mode
        ctx.param_dtypes = (omegas.dtype, deltas.dtype, phis.dtype, interaction_matrix.dtype)
        ctx.save_for_backward(psi)
        return psi

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx: Any, grad_out: torch.Tensor):
        (psi_final,) = ctx.saved_tensors
        pulse: _Pulse = ctx.pulse
        want_u = ctx.needs_input_grad[3]
        grads = adjoint_sweep(pulse, ctx.checkpoints, psi_final, grad_out.detach().contiguous(),
                              ctx.tolerance, ctx.every, ctx.max_krylov_dim, ctx.mode, want_u)
        g_om, g_de, g_ph, g_u, g_psi = grads
        od, dd, pd, ud = ctx.param_dtypes
        out_om = torch.as_tensor(g_om).to(od) if ctx.needs_input_grad[0] else None
        out_de = torch.as_tensor(g_de).to(dd) if ctx.needs_input_grad[1] else None
        out_ph = torch.as_tensor(g_ph).to(pd) if ctx.needs_input_grad[2] else None
        out_u = torch.as_tensor(g_u).to(ud) if want_u else None
        out_psi = g_psi if ctx.needs_input_grad[4] else None
        return out_om, out_de, out_ph, out_u, out_psi, None, None, None, None, None


class AdjointStats:
    """Lanczos runs of the last adjoint sweep (for reporting)."""

    forward_runs = 0
    adjoint_runs = 0


def _rows(n_rows: int, like: torch.Tensor) -> torch.Tensor:
    return torch.empty((n_rows, like.numel()), dtype=like.dtype, device=like.device)


def _step_grads(pulse: _Pulse, k: int, run_bra: LanczosRun, rows_bra: torch.Tensor,
                run_ket: LanczosRun, rows_ket: torch.Tensor, ket_forward: bool, want_u: bool):
    x = frechet_weights(run_bra.extended(), run_ket.extended(), ket_forward=ket_forward)
    mb, mk = x.shape
    sb = np.asarray(run_bra.scales[:mb])
    sk = np.asarray(run_ket.scales[:mk])
    coef = (sb[:, None] * x * sk[None, :]) * (run_bra.initial_norm * run_ket.initial_norm)
    cdev = torch.as_tensor(coef, device=rows_ket.device).to(rows_ket.dtype)
    z = cdev @ rows_ket[:mk]
    tr = native.parameter_traces(pulse.n, rows_bra[:mb], z, pairs=want_u)
    return _traces_to_grads(tr, -1j * pulse.dts[k], pulse.om[k].numpy(), pulse.ph[k].numpy())


def adjoint_sweep(pulse: _Pulse, checkpoints: dict, psi_final: torch.Tensor, lam: torch.Tensor,
                  tol: float, every: int, max_dim: int, mode: str, want_u: bool):
    """Reverse sweep; returns per-step parameter gradients and dL/dpsi_0."""
    s, n = pulse.steps, pulse.n
    g_om = np.zeros((s, n))
    g_de = np.zeros((s, n))
    g_ph = np.zeros((s, n))
    g_u = np.zeros((n, n)) if want_u else None
    rows_l = _rows(max_dim + 2, lam)
    AdjointStats.forward_runs = 0
    AdjointStats.adjoint_runs = 0

    def adjoint_run(k: int, vec: torch.Tensor) -> LanczosRun:
        run = _fused_lanczos(ScaledHamiltonian(pulse.ham(k), 1j * pulse.dts[k]), vec, tol, tol,
                             max_dim, rows=rows_l)
        AdjointStats.adjoint_runs += 1
        if not run.converged:
            raise RecursionError("adjoint Lanczos did not converge to precision "
                                 "in allotted number of steps.")
        return run

    if mode == "reverse":
        rows_p = _rows(max_dim + 2, lam)
        psi = psi_final
        for k in reversed(range(s)):
            b = ScaledHamiltonian(pulse.ham(k), 1j * pulse.dts[k])
            run_l = adjoint_run(k, lam)
            run_p = _fused_lanczos(b, psi, tol, tol, max_dim, rows=rows_p)
            AdjointStats.adjoint_runs += 1
            if not run_p.converged:
                raise RecursionError("reverse state propagation did not converge")
            gk = _step_grads(pulse, k, run_l, rows_l, run_p, rows_p, False, want_u)
            g_om[k], g_de[k], g_ph[k] = gk[0], gk[1], gk[2]
            if want_u:
                g_u += gk[3]
            lam = run_l.combine()
            if k > 0:
                psi = checkpoints[k].clone() if k in checkpoints else run_p.combine()
        return g_om, g_de, g_ph, g_u, lam

    # mode == "recompute": per segment, re-run forward from its checkpoint
    # keeping each step's Lanczos basis (rows per step of the segment).
    starts = sorted(checkpoints)
    seg_rows = [_rows(max_dim + 2, lam) for _ in range(every)]
    for c in reversed(starts):
        end = min(c + every, s)
        runs = []
        psi = checkpoints[c]
        for k in range(c, end):
            run = _fused_lanczos(ScaledHamiltonian(pulse.ham(k), -1j * pulse.dts[k]), psi, tol,
                                 tol, max_dim, rows=seg_rows[k - c])
            AdjointStats.forward_runs += 1
            if not run.converged:
                raise RecursionError("segment recomputation did not converge")
            runs.append(run)
            if k + 1 < end:
                psi = run.combine()
        for k in reversed(range(c, end)):
            run_l = adjoint_run(k, lam)
            gk = _step_grads(pulse, k, run_l, rows_l, runs[k - c], seg_rows[k - c], True, want_u)
            g_om[k], g_de[k], g_ph[k] = gk[0], gk[1], gk[2]
            if want_u:
                g_u += gk[3]
            lam = run_l.combine()
    return g_om, g_de, g_ph, g_u, lam


class EnergyExpectation(torch.autograd.Function):
    """<psi|H(omegas, deltas, phis, U)|psi> (real), differentiable in psi and in
    the parameters (the observable Hamiltonian of the last step, as
    emu_sv/sv_backend_impl.py:199-209 evaluates it)."""

    @staticmethod
    def forward(ctx: Any, omegas, deltas, phis, interaction_matrix, psi):
        ham = RydbergHamiltonian(omegas, deltas, phis, interaction_matrix, device=psi.device)
        p = psi.detach().contiguous()
        hp = ham.apply(p)
        e = torch.vdot(p, hp).real
        ctx.save_for_backward(p, hp)
        ctx.params = (_as_real(omegas).numpy(), _as_real(phis).numpy())
        ctx.param_dtypes = (omegas.dtype, deltas.dtype, phis.dtype, interaction_matrix.dtype)
        return e

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx: Any, grad_e: torch.Tensor):
        p, hp = ctx.saved_tensors
        g = float(grad_e)
        n = native.nqubits_of(p)
        om, ph = ctx.params
        want_u = ctx.needs_input_grad[3]
        tr = native.parameter_traces(n, p, p, pairs=want_u)
        g_om, g_de, g_ph, g_u = _traces_to_grads(tr, 1.0, om, ph)
        od, dd, pd, ud = ctx.param_dtypes
        outs = [torch.as_tensor(g * g_om).to(od) if ctx.needs_input_grad[0] else None,
                torch.as_tensor(g * g_de).to(dd) if ctx.needs_input_grad[1] else None,
                torch.as_tensor(g * g_ph).to(pd) if ctx.needs_input_grad[2] else None,
                torch.as_tensor(g * g_u).to(ud) if want_u else None,
                (2.0 * g) * hp if ctx.needs_input_grad[4] else None]
        return tuple(outs)


def total_occupation(psi: torch.Tensor) -> torch.Tensor:
    """sum_i <n_i> = sum_x popcount(x) |psi(x)|^2 (torch ops, differentiable)."""
    n = native.nqubits_of(psi)
    idx = torch.arange(psi.numel(), device=psi.device)
    pc = torch.zeros(psi.numel(), dtype=torch.float64, device=psi.device)
    for b in range(n):
        pc += ((idx >> b) & 1).to(torch.float64)
    return ((psi.real.to(torch.float64) ** 2 + psi.imag.to(torch.float64) ** 2) * pc).sum()


def evolve_pulse(omegas, deltas, phis, interaction_matrix, psi0, dts, tolerance=1e-10,
                 checkpoint_every=10, max_krylov_dim=DEFAULT_MAX_KRYLOV_DIM, mode="reverse"):
    """Differentiable whole-pulse evolution (see PulseEvolution)."""
    return PulseEvolution.apply(omegas, deltas, phis, interaction_matrix, psi0, dts, tolerance,
                                checkpoint_every, max_krylov_dim, mode)
