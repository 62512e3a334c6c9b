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
              forward state instead (mode "reverse"; mode "batched" runs the
              two backward recurrences as two registers of the batch engine,
              sharing launches).  Mode "recompute" (the default) needs no
              state recurrence in the backward sweep: it uses the forward
              Lanczos decompositions, kept by the forward pass for the
              newest segments within ``basis_budget_bytes`` and re-run
              from the segment's checkpoint for the others.
            * dL/dtheta = Re <lambda_{k+1}| dU_k/dtheta |psi_k>
                        = Re[-i dt_k int_0^1 <mu(u)| dH/dtheta |psi(u)> du]
              = Re[-i dt_k |lambda||psi| sum_lk X_lk <w_l|dH/dtheta|v_k>]
              with X_lk = int conj(a_l(u)) b_k(u) du in closed form from the
              eigen-decompositions of the two small Lanczos matrices
              (frechet_weights_eig; frechet_weights is the block-exponential
              statement of the same integral), Z = X V (one GEMM on the
              device) and the per-qubit traces <W|.|Z>
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
)
from emulators_amd.emu_base.math.krylov_exp import fused_lanczos as _fused_lanczos
from emulators_amd.emu_sv.hamiltonian import RydbergHamiltonian


BASIS_BUDGET_BYTES = 16 << 30  # forward Lanczos bases kept for the backward sweep


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


def _phi1(z: np.ndarray) -> np.ndarray:
    """(e^z - 1)/z elementwise (series below |z| = 1e-3)."""
    z = np.asarray(z, dtype=np.complex128)
    out = np.empty_like(z)
    small = np.abs(z) < 1e-3
    zs = z[small]
    out[small] = 1.0 + zs / 2.0 * (1.0 + zs / 3.0 * (1.0 + zs / 4.0 * (1.0 + zs / 5.0 * (1.0 + zs / 6.0))))
    zb = z[~small]
    out[~small] = (np.exp(zb) - 1.0) / zb
    return out


def lanczos_trajectory(t: np.ndarray, sigma: complex) -> tuple[np.ndarray, np.ndarray]:
    """(A, r) with [exp(u T) e_0]_k = sum_p A[k, p] exp(u r_p) for the Lanczos
    matrix T = W^-1 (sigma J) W of op = sigma H (J real symmetric
    tridiagonal, W = diag(w^k), w = sigma/|sigma|): J = Z diag(lambda) Z^T
    gives A[k, p] = w^-k Z[k, p] Z[0, p] and r = sigma lambda."""
    m = t.shape[0]
    asig = abs(sigma)
    w = sigma / asig
    d = (np.diagonal(t) / sigma).real
    e = np.real(np.diagonal(t, -1)) / asig
    lam, z = native.tridiag_eig(d, e)
    wk = np.conj(w) ** np.arange(m)
    return wk[:, None] * z * z[0][None, :], sigma * lam


def frechet_weights_eig(t_bra: np.ndarray, sig_bra: complex, t_ket: np.ndarray, sig_ket: complex,
                        ket_forward: bool = False) -> np.ndarray:
    """The weights of frechet_weights from the two eigen-decompositions:
    with a_p, b_q the bra and ket rates,
      X = conj(A_bra) F A_ket^T,  F[p, q] = int_0^1 exp(conj(a_p) g(u) + b_q u') du
    = phi_1(conj(a_p) + b_q) (both at time u) or exp(conj(a_p)) phi_1(b_q -
    conj(a_p)) (bra at 1 - s, ket at s), phi_1(z) = (e^z - 1)/z.  O(m^3)
    small products instead of a 2m x 2m matrix exponential."""
    ab, rb = lanczos_trajectory(t_bra, sig_bra)
    ak, rk = lanczos_trajectory(t_ket, sig_ket)
    a = np.conj(rb)[:, None]
    b = rk[None, :]
    f = np.exp(a) * _phi1(b - a) if ket_forward else _phi1(a + b)
    return np.conj(ab) @ f @ ak.T


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
                checkpoint_every=10, max_krylov_dim=DEFAULT_MAX_KRYLOV_DIM, mode="recompute",
                basis_budget_bytes=BASIS_BUDGET_BYTES):
        if mode not in ("batched", "reverse", "recompute"):
            raise ValueError("mode must be 'recompute', 'reverse' or 'batched'")
        if checkpoint_every < 1:
            raise ValueError("checkpoint_every must be >= 1")
        pulse = _Pulse(omegas, deltas, phis, interaction_matrix, dts, psi0.device)
        psi = psi0.detach().contiguous()
        checkpoints = {}
        stored: dict = {}  # segment start -> forward LanczosRuns (mode recompute)
        stored_bytes = {}
        dropped = set()
        vec_bytes = psi.numel() * psi.element_size()
        for k in range(pulse.steps):
            c = k - k % checkpoint_every
            if k == c:
                checkpoints[k] = psi.clone()
            run = _fused_lanczos(ScaledHamiltonian(pulse.ham(k), -1j * pulse.dts[k]), psi,
                                 tolerance, tolerance, max_krylov_dim)
            if not run.converged:
                raise RecursionError("exponentiation algorithm did not converge to precision "
                                     "in allotted number of steps.")
            psi = run.combine()
            if mode == "recompute" and c not in dropped:
                stored.setdefault(c, []).append(run)
                stored_bytes[c] = stored_bytes.get(c, 0) + len(run.basis) * vec_bytes
                # keep the newest whole segments within the budget; the others
                # are recomputed from their checkpoints in the backward sweep
                while sum(stored_bytes.values()) > basis_budget_bytes:
                    oldest = min(stored)
                    del stored[oldest]
                    del stored_bytes[oldest]
                    dropped.add(oldest)
                    if oldest == c:
                        break
        ctx.pulse = pulse
        ctx.checkpoints = checkpoints
        ctx.stored = stored
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
                              ctx.tolerance, ctx.every, ctx.max_krylov_dim, ctx.mode, want_u,
                              stored=ctx.stored)
        ctx.stored = None
        g_om, g_de, g_ph, g_u, g_psi = grads
        od, dd, pd, ud = ctx.param_dtypes
        out_om = torch.as_tensor(g_om).to(od) if ctx.needs_input_grad[0] else None
        out_de = torch.as_tensor(g_de).to(dd) if ctx.needs_input_grad[1] else None
        out_ph = torch.as_tensor(g_ph).to(pd) if ctx.needs_input_grad[2] else None
        out_u = torch.as_tensor(g_u).to(ud) if want_u else None
        out_psi = g_psi if ctx.needs_input_grad[4] else None
        return out_om, out_de, out_ph, out_u, out_psi, None, None, None, None, None, None


class AdjointStats:
    """Lanczos runs of the last adjoint sweep (for reporting)."""

    forward_runs = 0
    adjoint_runs = 0


def _rows(n_rows: int, like: torch.Tensor) -> torch.Tensor:
    return torch.empty((n_rows, like.numel()), dtype=like.dtype, device=like.device)


def _basis_matrix(run: LanczosRun, rows: torch.Tensor | None, m: int) -> torch.Tensor:
    """The first m basis vectors of a run as an (m, 2**n) tensor."""
    if rows is not None:
        return rows[:m]
    return torch.stack(run.basis[:m])


def _step_grads(pulse: _Pulse, k: int, run_bra: LanczosRun, rows_bra: torch.Tensor,
                run_ket: LanczosRun, rows_ket: torch.Tensor | None, ket_forward: bool, want_u: bool):
    """Parameter gradients of step k from the adjoint decomposition (bra, op =
    +i dt H from lambda_{k+1}) and the state decomposition (ket: +i dt H from
    psi_{k+1}, or with ket_forward the forward -i dt H from psi_k).  The
    Krylov trajectories use the m x m Lanczos matrices and m vectors."""
    dt = pulse.dts[k]
    tb, tk = run_bra.T, run_ket.T
    mb, mk = tb.shape[0], tk.shape[0]
    x = frechet_weights_eig(tb, 1j * dt, tk, -1j * dt if ket_forward else 1j * dt, ket_forward)
    sb = np.asarray(run_bra.scales[:mb])
    sk = np.asarray(run_ket.scales[:mk])
    coef = (sb[:, None] * x * sk[None, :]) * (run_bra.initial_norm * run_ket.initial_norm)
    vk = _basis_matrix(run_ket, rows_ket, mk)
    cdev = torch.as_tensor(coef, device=vk.device).to(vk.dtype)
    z = cdev @ vk
    tr = native.parameter_traces(pulse.n, _basis_matrix(run_bra, rows_bra, mb), z, pairs=want_u)
    return _traces_to_grads(tr, -1j * dt, pulse.om[k].numpy(), pulse.ph[k].numpy())


def adjoint_sweep(pulse: _Pulse, checkpoints: dict, psi_final: torch.Tensor, lam: torch.Tensor,
                  tol: float, every: int, max_dim: int, mode: str, want_u: bool,
                  stored: dict | None = None):
    """Reverse sweep; returns per-step parameter gradients and dL/dpsi_0."""
    if mode == "batched":
        return _adjoint_sweep_batched(pulse, checkpoints, psi_final, lam, tol, min(max_dim, 62),
                                      want_u)
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

    # mode == "recompute": per segment, the forward Lanczos runs kept by the
    # forward pass, or re-run forward from the segment's checkpoint.
    stored = stored or {}
    starts = sorted(checkpoints)
    for c in reversed(starts):
        end = min(c + every, s)
        runs = stored.pop(c, None)
        if runs is None:
            runs = []
            psi = checkpoints[c]
            for k in range(c, end):
                run = _fused_lanczos(ScaledHamiltonian(pulse.ham(k), -1j * pulse.dts[k]), psi, tol,
                                     tol, max_dim)
                AdjointStats.forward_runs += 1
                if not run.converged:
                    raise RecursionError("segment recomputation did not converge")
                runs.append(run)
                if k + 1 < end:
                    psi = run.combine()
        for k in reversed(range(c, end)):
            run_l = adjoint_run(k, lam)
            gk = _step_grads(pulse, k, run_l, rows_l, runs[k - c], None, True, want_u)
            g_om[k], g_de[k], g_ph[k] = gk[0], gk[1], gk[2]
            if want_u:
                g_u += gk[3]
            lam = run_l.combine()
        del runs
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
                 checkpoint_every=10, max_krylov_dim=DEFAULT_MAX_KRYLOV_DIM, mode="recompute",
                 basis_budget_bytes=BASIS_BUDGET_BYTES):
    """Differentiable whole-pulse evolution (see PulseEvolution)."""
    return PulseEvolution.apply(omegas, deltas, phis, interaction_matrix, psi0, dts, tolerance,
                                checkpoint_every, max_krylov_dim, mode, basis_budget_bytes)


def _step_matrix(arr: np.ndarray, lay, m: int, happy: bool) -> np.ndarray:
    """The Lanczos matrix of a finished batch-engine step with m combination
    coefficients: (m-1) x (m-1), or m x m after a happy breakdown."""
    mt = m if happy else m - 1
    diag = arr[lay[0]: lay[0] + 2 * mt].view(np.complex128)
    sub = arr[lay[1]: lay[1] + mt]
    sup = arr[lay[4]: lay[4] + 2 * max(mt - 1, 0)].view(np.complex128)
    t = np.zeros((mt, mt), dtype=np.complex128)
    t[np.arange(mt), np.arange(mt)] = diag
    for k in range(mt - 1):
        t[k + 1, k] = sub[k]
        t[k, k + 1] = sup[k]
    return t


def _adjoint_sweep_batched(pulse: _Pulse, checkpoints: dict, psi_final: torch.Tensor,
                           lam: torch.Tensor, tol: float, max_dim: int, want_u: bool):
    """The reverse sweep with both backward recurrences -- the state psi_{k+1}
    and the adjoint lambda_{k+1} under B = +i dt_k H_k -- as two registers of
    one batch engine that pauses after every step: their Lanczos iterations
    share each launch.  At every pause the Frechet term of that step is
    contracted from the two finished decompositions, the state is reset to
    the forward checkpoint where one exists, and both registers resume."""
    import ctypes

    from emulators_amd.emu_sv.batch import PHASES, _Reg

    s, n = pulse.steps, pulse.n
    dev = lam.device
    dtype = lam.dtype
    om_r = np.ascontiguousarray(pulse.om.numpy()[::-1])
    de_r = np.ascontiguousarray(pulse.de.numpy()[::-1])
    ph_r = np.ascontiguousarray(pulse.ph.numpy()[::-1])
    blocks = np.zeros((s, native.EMU_SV_HAM_DOUBLES))
    dp = ctypes.POINTER(ctypes.c_double)
    um = np.ascontiguousarray(pulse.u.numpy())
    native.check(native.LIB.emu_sv_pack_hamiltonian_steps(
        n, s, om_r.ctypes.data_as(dp), de_r.ctypes.data_as(dp), ph_r.ctypes.data_as(dp),
        um.ctypes.data_as(dp), blocks.ctypes.data_as(dp)), "pack_hamiltonian_steps")
    cplx = bool(np.any(ph_r != 0.0))
    ham = torch.from_numpy(blocks).to(dev)
    # sigma = -i dt per engine step: negative dt gives B = +i dt H
    dts = torch.as_tensor(-np.asarray(pulse.dts[::-1], dtype=np.float64)).to(dev)
    bufs = []
    regs = (_Reg * 2)()
    for i, start in enumerate((psi_final, lam)):
        psi = start.detach().clone().contiguous()
        alt = torch.empty_like(psi)
        rows = torch.empty((max_dim + 1) * psi.numel(), dtype=dtype, device=dev)
        bufs.append((psi, alt, rows))
        regs[i] = _Reg(n, s, native.EMU_SV_FLAG_COMPLEX if cplx else 0, 1, ham.data_ptr(),
                       dts.data_ptr(), psi.data_ptr(), alt.data_ptr(), rows.data_ptr())
    code = native.dtype_code(dtype)
    nbytes = native.LIB.emu_sv_batch_workspace_bytes(2, code, regs, 0, 0)
    ws = torch.zeros(nbytes, dtype=torch.uint8, device=dev)
    stream = native._stream(dev)
    handle = ctypes.c_void_p()
    native.check(native.LIB.emu_sv_batch_create(2, code, regs, max_dim, float(tol), 0, 0,
                                                native._ptr(ws), nbytes, stream, ctypes.byref(handle)),
                 "batch_create")
    lay = (ctypes.c_int * 6)()
    native.check(native.LIB.emu_sv_batch_array_layout(lay), "batch_array_layout")
    lay = list(lay)
    g_om = np.zeros((s, n))
    g_de = np.zeros((s, n))
    g_ph = np.zeros((s, n))
    g_u = np.zeros((n, n)) if want_u else None
    d = 2**n
    info = (ctypes.c_int * 5)()
    scal = (ctypes.c_double * 1)()
    arrs = [np.zeros(lay[5]), np.zeros(lay[5])]
    cur = ctypes.c_void_p()
    prev = ctypes.c_void_p()
    rowp = ctypes.c_void_p()
    guess = 4
    try:
        remaining = ctypes.c_int(1)
        for step in range(s):
            k = s - 1 - step  # forward step index
            native.check(native.LIB.emu_sv_batch_run(handle, max(guess - 1, 1), max(guess - 1, 1),
                                                     stream, ctypes.byref(remaining)), "batch_run")
            while remaining.value > 0:
                native.check(native.LIB.emu_sv_batch_run(handle, 1, 1, stream, ctypes.byref(remaining)),
                             "batch_run")
            dec = []
            for i in range(2):
                a = arrs[i]
                native.check(native.LIB.emu_sv_batch_step_info(
                    handle, i, info, scal, a.ctypes.data_as(dp), ctypes.byref(cur), ctypes.byref(prev),
                    ctypes.byref(rowp), stream), "batch_step_info")
                phase = PHASES.get(info[0])
                if phase not in ("paused", "done"):
                    raise RecursionError("adjoint Lanczos did not converge to precision "
                                         "in allotted number of steps.")
                m, happy = int(info[2]), bool(info[3])
                dec.append((m, happy, float(scal[0]), a.copy(), cur.value, prev.value))
            guess = max(dec[0][0], dec[1][0])
            (ml, hl, nl, al, _, prev_l), (mp, hp, npsi, ap, cur_p, prev_p) = dec[1], dec[0]
            tl_, tp_ = _step_matrix(al, lay, ml, hl), _step_matrix(ap, lay, mp, hp)
            ml, mp = tl_.shape[0], tp_.shape[0]
            dt = pulse.dts[k]
            x = frechet_weights_eig(tl_, 1j * dt, tp_, 1j * dt, False)
            sl = al[lay[2]: lay[2] + ml]
            sp = ap[lay[2]: lay[2] + mp]
            coef = (sl[:, None] * x * sp[None, :]) * (nl * npsi)
            cdev = torch.as_tensor(coef, device=dev).to(dtype)
            psi_t, alt_t, rows_p = bufs[0]
            lam_t, lalt_t, rows_l = bufs[1]
            prev_p_t = psi_t if prev_p == psi_t.data_ptr() else alt_t
            prev_l_t = lam_t if prev_l == lam_t.data_ptr() else lalt_t
            z = cdev[:, :1] * prev_p_t[None, :]
            if mp > 1:
                z = z + cdev[:, 1:] @ rows_p[: (mp - 1) * d].view(mp - 1, d)
            tr0 = native.parameter_traces(n, prev_l_t[None, :], z[:1], pairs=want_u)
            gk = _traces_to_grads(tr0, -1j * dt, pulse.om[k].numpy(), pulse.ph[k].numpy())
            if ml > 1:
                tr1 = native.parameter_traces(n, rows_l[: (ml - 1) * d].view(ml - 1, d), z[1:],
                                              pairs=want_u)
                g1 = _traces_to_grads(tr1, -1j * dt, pulse.om[k].numpy(), pulse.ph[k].numpy())
                gk = tuple(None if a_ is None else a_ + b_ for a_, b_ in zip(gk, g1))
            g_om[k], g_de[k], g_ph[k] = gk[0], gk[1], gk[2]
            if want_u:
                g_u += gk[3]
            if step < s - 1:
                if k in checkpoints:
                    tgt = psi_t if cur_p == psi_t.data_ptr() else alt_t
                    tgt.copy_(checkpoints[k])
                native.check(native.LIB.emu_sv_batch_resume(handle, stream), "batch_resume")
        native.check(native.LIB.emu_sv_batch_publish(handle, stream), "batch_publish")
        out_lam = bufs[1][0]
    finally:
        native.LIB.emu_sv_batch_destroy(handle)
    return g_om, g_de, g_ph, g_u, out_lam
