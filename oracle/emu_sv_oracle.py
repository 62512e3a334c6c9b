"""CPU oracle for the emu_sv state-vector Krylov time evolution.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this module:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker (or as the timed CPU baseline),
never as a code path of the product.

It is a torch-CPU, complex128 restatement of the reference algorithm.  Every
function cites the reference file:line it follows (paths relative to the
reference checkout).  The restatement keeps the reference's control flow on
purpose (per-qubit ``index_add_`` H.psi, Lanczos with two-vector modified
Gram-Schmidt, Expokit extended-T error estimate, CPU ``matrix_exp``), because
it is also the CPU timing baseline.

How each part is pinned (tests/test_oracle.py, tests/test_golden.py,
tests/test_noise_host.py, tests/test_lindblad_host.py):
  * H.psi and the diagonal: dense brute force (``dense_rydberg_hamiltonian``)
    at atol 1e-12, the closed-form energy of (|rrr>+|rgg>)/sqrt2, and the
    reference's seeded energy-variance / second-moment constants.
  * krylov_exp: dense matrix_exp, and the reference's exact iteration counts
    (1, 1, 3, 7, 7, 6) on its seeded matrices.
  * step accuracy: dense expm(-i dt H) at atol = tolerance.
  * double_krylov / step backward: dense block-expm Frechet derivative and
    torch autograd through the dense matrix_exp.
  * time loop: single-atom Rabi closed form, the reference's SequenceData
    pi-pulse table (occupation 0 -> 1), norm and energy conservation.
  * PCHIP: the reference's PCHIP formula tests restated as closed forms.
  * noise term: dense brute force H + sum_q K_q, and the reference's check
    that K = diag(0, -i/2) leaves Re<psi|H|psi> unchanged.
  * Brent root finder: the reference's known answers (roots and exact
    evaluation counts 12, 13, 18) and a dense-expm jump-time root.
  * expect_batch / apply_1q: dense Kronecker-product operators.
  * Monte-Carlo step: a single decaying atom, whose jump time has the closed
    form exp(-gamma t) = threshold; and the reference's seeded known answer
    of ten two-atom emission trajectories (mean occupations [0.5, 0.2],
    test_end_to_end.py:458-545), which fixes the draw order.
  * Lindbladian: the textbook generator -i[H, rho] + sum (A rho A^H -
    1/2 {A^H A, rho}) as a dense 4^N superoperator, the reference's no-jump
    and Tr(H rho) checks, and the e^{-gamma t} decay of one atom.
  * inverse-CDF sampling: a brute-force scan of the cumulative weights, a
    point mass, and zero-weight entries that are never drawn.
"""

from __future__ import annotations

import cmath
import math
import random
from dataclasses import dataclass, field
from typing import Callable, Sequence

import torch

DTYPE = torch.complex128
DEFAULT_MAX_KRYLOV_DIM = 100  # emu_base/math/krylov_exp.py:4
TIME_CONVERSION_COEFF = 0.001  # emu_sv/sv_backend_impl.py:29 (rad/us with dt in ns)


# ---------------------------------------------------------------------------
# Rydberg Hamiltonian  (emu_sv/hamiltonian.py)
# ---------------------------------------------------------------------------
class OracleHamiltonian:
    """Restates ``RydbergHamiltonian`` (emu_sv/hamiltonian.py:5-172).

    H = sum_j (O_j/2)[cos(phi_j) X_j + sin(phi_j) Y_j] - sum_j d_j n_j
        + sum_{i<j} U_ij n_i n_j,  qubit 0 = most significant bit,
    plus, with ``noise`` = K (2x2), K on every qubit (hamiltonian.py:61-67,
    147-151).
    """

    def __init__(self, omegas, deltas, phis, interaction_matrix, noise=None):
        # hamiltonian.py:51-67
        self.nqubits = len(omegas)
        self.omegas = torch.as_tensor(omegas).detach().to(DTYPE) / 2.0
        self.deltas = torch.as_tensor(deltas).detach().to(DTYPE)
        self.phis = torch.as_tensor(phis).detach().to(DTYPE)
        self.interaction_matrix = torch.as_tensor(interaction_matrix).detach()
        self.noise = None if noise is None else torch.as_tensor(noise, dtype=DTYPE)
        self.diag = self._create_diagonal()
        self.inds = torch.tensor([1, 0])
        self.complex = bool(self.phis.any()) or self.noise is not None
        if self.complex:
            oc = self.omegas * torch.exp(1.0j * self.phis)
            self.omegas_c = torch.stack((oc, oc.conj()), dim=1)
            if self.noise is not None:
                self.omegas_c[:, 0] += self.noise[1, 0]
                self.omegas_c[:, 1] += self.noise[0, 1]

    def _create_diagonal(self) -> torch.Tensor:
        # hamiltonian.py:139-162: -sum_i d_i n_i + sum_{i<j} U_ij n_i n_j,
        # upper triangle of U only; the noise diagonal first (:147-151).
        n = self.nqubits
        diag = torch.zeros(2**n, dtype=DTYPE)
        if self.noise is not None:
            for i in range(n):
                diag = diag.view(2**i, 2, -1)
                diag[:, 0, :] += self.noise[0, 0]
                diag[:, 1, :] += self.noise[1, 1]
        for i in range(n):
            diag = diag.view(2**i, 2, -1)
            i_fixed = diag[:, 1, :]
            i_fixed -= self.deltas[i]
            for j in range(i + 1, n):
                i_fixed = i_fixed.view(2**i, 2 ** (j - i - 1), 2, -1)
                i_j_fixed = i_fixed[:, :, 1, :]
                i_j_fixed += self.interaction_matrix[i, j].item()
        return diag.view(-1)

    def __mul__(self, vec: torch.Tensor) -> torch.Tensor:
        # hamiltonian.py:69-91 and :93-137 (index_add_ per qubit)
        result = self.diag * vec
        n = self.nqubits
        for q in range(n):
            shape = (2**q, 2, 2 ** (n - q - 1))
            v = vec.view(shape)
            r = result.view(shape)
            if self.complex:
                c0, c1 = self.omegas_c[q, 0].item(), self.omegas_c[q, 1].item()
                r.index_add_(1, self.inds[0:1], v[:, 0, :].unsqueeze(1), alpha=c0)
                r.index_add_(1, self.inds[1:2], v[:, 1, :].unsqueeze(1), alpha=c1)
            else:
                r.index_add_(1, self.inds, v, alpha=self.omegas[q].item())
        return result

    def expect(self, state: torch.Tensor) -> torch.Tensor:
        # hamiltonian.py:164-172
        return torch.vdot(state, self * state).real


def noise_from_lindbladians(lindbladians) -> torch.Tensor:
    """K = -i/2 sum_l L_l^+ L_l (emu_base/jump_lindblad_operators.py:80-93)."""
    acc = torch.zeros(2, 2, dtype=DTYPE)
    for L in lindbladians:
        L = torch.as_tensor(L, dtype=DTYPE)
        acc = acc + L.conj().T @ L
    return -0.5j * acc


# ---------------------------------------------------------------------------
# Dense brute force  (test/utils_testing/utils_dense_hamiltonians.py:8-67)
# ---------------------------------------------------------------------------
def _single_gate(i: int, n: int, g: torch.Tensor) -> torch.Tensor:
    out = torch.ones(1, 1, dtype=DTYPE)
    eye = torch.eye(2, dtype=DTYPE)
    for k in range(n):
        out = torch.kron(out, g if k == i else eye)
    return out


def dense_rydberg_hamiltonian(omegas, deltas, phis, interaction_matrix) -> torch.Tensor:
    """Kronecker-product H, restating utils_dense_hamiltonians.py:37-67."""
    n = len(omegas)
    sx = torch.tensor([[0.0, 1.0], [1.0, 0.0]], dtype=DTYPE)
    sy = torch.tensor([[0.0, -1j], [1j, 0.0]], dtype=DTYPE)
    nop = torch.tensor([[0.0, 0.0], [0.0, 1.0]], dtype=DTYPE)
    om = torch.as_tensor(omegas).detach().to(DTYPE)
    de = torch.as_tensor(deltas).detach().to(DTYPE)
    ph = torch.as_tensor(phis).detach().to(DTYPE)
    um = torch.as_tensor(interaction_matrix).detach().to(DTYPE)
    h = torch.zeros(2**n, 2**n, dtype=DTYPE)
    for i in range(n):
        h += om[i] * torch.cos(ph[i]) * _single_gate(i, n, sx) / 2
        h += om[i] * torch.sin(ph[i]) * _single_gate(i, n, sy) / 2
        h -= de[i] * _single_gate(i, n, nop)
        for j in range(i + 1, n):
            h += um[i, j] * (_single_gate(i, n, nop) @ _single_gate(j, n, nop))
    return h


def dense_noisy_hamiltonian(omegas, deltas, phis, interaction_matrix, noise) -> torch.Tensor:
    """Kronecker-product H + sum_q K_q: the brute-force form of the noisy
    Hamiltonian of hamiltonian.py:61-67, 147-151 (K on every qubit)."""
    n = len(omegas)
    h = dense_rydberg_hamiltonian(omegas, deltas, phis, interaction_matrix)
    k = torch.as_tensor(noise, dtype=DTYPE)
    for q in range(n):
        h = h + _single_gate(q, n, k)
    return h


def dense_rydberg_hamiltonian_autograd(omegas, deltas, phis, interaction_matrix) -> torch.Tensor:
    """The same Kronecker-product H as dense_rydberg_hamiltonian, built from
    the parameter tensors without detaching them, so torch autograd through
    a dense matrix_exp gives exact parameter gradients at small N (the
    reference's check of its backward, test_time_evolution.py:79-103)."""
    n = len(omegas)
    sx = torch.tensor([[0.0, 1.0], [1.0, 0.0]], dtype=DTYPE)
    sy = torch.tensor([[0.0, -1j], [1j, 0.0]], dtype=DTYPE)
    nop = torch.tensor([[0.0, 0.0], [0.0, 1.0]], dtype=DTYPE)
    h = torch.zeros(2**n, 2**n, dtype=DTYPE)
    for i in range(n):
        h = h + omegas[i] * torch.cos(phis[i]) * _single_gate(i, n, sx) / 2
        h = h + omegas[i] * torch.sin(phis[i]) * _single_gate(i, n, sy) / 2
        h = h - deltas[i] * _single_gate(i, n, nop)
        for j in range(i + 1, n):
            h = h + interaction_matrix[i, j] * (_single_gate(i, n, nop) @ _single_gate(j, n, nop))
    return h


def dense_step(dt, omegas, deltas, phis, interaction_matrix, state) -> torch.Tensor:
    """expm(-i dt H) psi, as test/emu_sv/test_time_evolution.py:13-24."""
    h = dense_rydberg_hamiltonian(omegas, deltas, phis, interaction_matrix)
    return torch.linalg.matrix_exp(-1j * dt * h) @ state


# ---------------------------------------------------------------------------
# Krylov exponential  (emu_base/math/krylov_exp.py)
# ---------------------------------------------------------------------------
@dataclass
class KrylovExpResult:
    """emu_base/math/krylov_exp.py:7-20"""

    result: torch.Tensor
    converged: bool
    happy_breakdown: bool
    iteration_count: int
    T: torch.Tensor | None = None


def krylov_exp_impl(
    op: Callable[[torch.Tensor], torch.Tensor],
    v: torch.Tensor,
    is_hermitian: bool,
    exp_tolerance: float,
    norm_tolerance: float,
    max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM,
) -> KrylovExpResult:
    """Restates krylov_exp_impl (emu_base/math/krylov_exp.py:23-106).

    ``v`` is normalised in place (:43-44).  T is a CPU tensor (:47).
    """
    initial_norm = v.norm()
    v /= initial_norm
    lanczos_vectors = [v]
    T = torch.zeros(max_krylov_dim + 2, max_krylov_dim + 2, dtype=v.dtype)

    def combine(coeffs):
        return initial_norm * sum(a * b for a, b in zip(coeffs, lanczos_vectors))

    expd = None
    for j in range(max_krylov_dim):
        w = op(lanczos_vectors[-1])  # :50
        n = w.norm()  # :52, before orthogonalisation
        k_start = max(0, j - 1) if is_hermitian else 0  # :54
        for k in range(k_start, j + 1):
            overlap = torch.tensordot(lanczos_vectors[k].conj(), w, dims=w.dim())
            T[k, j] = overlap
            w -= overlap * lanczos_vectors[k]
        n2 = w.norm()  # :60
        T[j + 1, j] = n2
        if n2 < norm_tolerance:  # :63 happy breakdown
            expd = torch.linalg.matrix_exp(T[: j + 1, : j + 1])
            return KrylovExpResult(combine(expd[:, 0]), True, True, j + 1, T[: j + 2, : j + 1].clone())
        w /= n2  # :73
        lanczos_vectors.append(w)
        T[j + 2, j + 1] = 1  # :77 extended T (Sidje 1998)
        expd = torch.linalg.matrix_exp(T[: j + 3, : j + 3])
        err1 = abs(expd[j + 1, 0])  # :81-84
        err2 = abs(expd[j + 2, 0] * n)
        err = err1 if err1 < err2 else (err1 * err2 / (err1 - err2))
        if err < exp_tolerance:  # :86
            res = combine(expd[: len(lanczos_vectors), 0])
            return KrylovExpResult(res, True, False, j + 1, T[: j + 2, : j + 1].clone())
    res = combine(expd[: len(lanczos_vectors), 0])
    return KrylovExpResult(res, False, False, max_krylov_dim, None)


def krylov_exp(op, v, exp_tolerance, norm_tolerance, is_hermitian=True,
               max_krylov_dim=DEFAULT_MAX_KRYLOV_DIM) -> torch.Tensor:
    """emu_base/math/krylov_exp.py:109-131 (RecursionError on non-convergence)."""
    r = krylov_exp_impl(op, v, is_hermitian=is_hermitian, exp_tolerance=exp_tolerance,
                        norm_tolerance=norm_tolerance, max_krylov_dim=max_krylov_dim)
    if not r.converged:
        raise RecursionError(
            "exponentiation algorithm did not converge to precision in allotted number of steps."
        )
    return r.result


# ---------------------------------------------------------------------------
# Double Krylov / Frechet derivative  (emu_base/math/double_krylov.py)
# ---------------------------------------------------------------------------
def lanczos(op, v, tolerance, max_krylov_dim):
    """Restates double_krylov.py:61-113 (Lanczos basis and T, no combination)."""
    converged = False
    vecs = [v / v.norm()]
    T = torch.zeros(max_krylov_dim + 2, max_krylov_dim + 2, dtype=v.dtype)
    for j in range(max_krylov_dim):
        w = op(vecs[-1])
        n = w.norm()
        for k in range(max(0, j - 1), j + 1):
            overlap = torch.tensordot(vecs[k].conj(), w, dims=w.dim())
            T[k, j] = overlap
            w -= overlap * vecs[k]
        n2 = w.norm()
        T[j + 1, j] = n2
        if n2 < tolerance:
            converged = True
            break
        vecs.append(w / n2)
        T[j + 2, j + 1] = 1
        expd = torch.linalg.matrix_exp(T[: j + 3, : j + 3])
        err1 = abs(expd[j + 1, 0])
        err2 = abs(expd[j + 2, 0] * n)
        err = err1 if err1 < err2 else (err1 * err2 / (err1 - err2))
        if err < tolerance:
            converged = True
            break
    if not converged:
        raise RecursionError(
            "Lanczos iteration did not converge to precision in allotted number of steps."
        )
    size = len(vecs)
    return vecs, T[:size, :size]


def double_krylov(op, state, grad, tolerance, max_krylov_dim):
    """Restates double_krylov.py:5-58: dU(op, |state><grad|) = Vs^T dS Vg^*."""
    vs, ts = lanczos(op, state, tolerance, max_krylov_dim)
    vg, tg = lanczos(op, grad, tolerance, max_krylov_dim)
    size_s = len(vs)
    big = torch.block_diag(ts, tg)
    big[0, size_s] = state.norm() * grad.norm()
    ds = torch.matrix_exp(big)[:size_s, size_s:]
    return vs, ds, vg


def frechet_exp(a: torch.Tensor, e: torch.Tensor) -> torch.Tensor:
    """Dense Frechet derivative, test/emu_base/math/test_double_krylov.py:10-20."""
    big = torch.block_diag(a, a)
    n = a.shape[0]
    big[:n, n:] = e
    return torch.linalg.matrix_exp(big)[:n, n:]


# ---------------------------------------------------------------------------
# One time step with the reference's custom backward
# (emu_sv/time_evolution/torch.py:122-378)
# ---------------------------------------------------------------------------
def _flip_apply(vec2d: torch.Tensor, q: int, n: int, alpha: complex, complex_form: bool):
    """alpha sigma+_q + conj(alpha) sigma-_q (or alpha X_q) on a batch (torch.py:8-33)."""
    v = vec2d.view(vec2d.shape[0], 2**q, 2, 2 ** (n - q - 1))
    out = torch.zeros_like(v)
    if complex_form:
        out[:, :, 1, :] += alpha * v[:, :, 0, :]
        out[:, :, 0, :] += alpha.conjugate() * v[:, :, 1, :]
    else:
        out[:, :, 1, :] += alpha * v[:, :, 0, :]
        out[:, :, 0, :] += alpha * v[:, :, 1, :]
    return out.view(vec2d.shape[0], -1)


def step_forward(dt, omegas, deltas, phis, interaction_matrix, state, tol, max_krylov_dim):
    """torch.py:142-172: exp(-i dt H) psi with the Lanczos exponential."""
    ham = OracleHamiltonian(omegas, deltas, phis, interaction_matrix)
    res = krylov_exp(lambda x: -1j * dt * (ham * x), state.detach().clone(), tol, tol,
                     is_hermitian=True, max_krylov_dim=max_krylov_dim)
    return res, ham


def parameter_traces(bra: torch.Tensor, ket: torch.Tensor, omegas, phis, fac: complex):
    """Re(fac * Tr(bra^dag dH/dtheta ket)) for every qubit parameter: the
    DHDOmegaSparse / DHDDeltaSparse / DHDPhiSparse / DHDUSparse traces of
    emu_sv/time_evolution/torch.py:8-119 as used at :318-351.  bra, ket:
    (rows, 2**n).  Returns (g_omegas, g_deltas, g_phis, g_U upper)."""
    n = len(omegas)
    om = torch.as_tensor(omegas).detach()
    ph = torch.as_tensor(phis).detach()

    def trace(v):
        return (fac * torch.tensordot(bra.conj(), v)).real

    e_l = ket
    g_om = torch.zeros(n, dtype=torch.float64)
    g_de = torch.zeros(n, dtype=torch.float64)
    g_ph = torch.zeros(n, dtype=torch.float64)
    g_u = torch.zeros(n, n, dtype=torch.float64)
    for i in range(n):
        phi_i = complex(ph[i]).real
        alpha = 0.5 * cmath.exp(1j * phi_i)  # torch.py:48, in complex128
        g_om[i] = trace(_flip_apply(e_l, i, n, alpha, phi_i != 0))  # DHDOmegaSparse
        v = e_l.clone().view(e_l.shape[0], 2**i, 2, -1)
        v[:, :, 0] = 0.0
        g_de[i] = trace(-v.view(e_l.shape[0], -1))  # DHDDeltaSparse torch.py:87-101
        alpha_p = 0.5 * complex(om[i]) * cmath.exp(1j * (phi_i + math.pi / 2))  # torch.py:77
        g_ph[i] = trace(_flip_apply(e_l, i, n, alpha_p, True))  # DHDPhiSparse
        for j in range(i + 1, n):  # DHDUSparse torch.py:104-119
            u = e_l.clone().view(e_l.shape[0], 2**i, 2, 2 ** (j - i - 1), 2, -1)
            u[:, :, 0] = 0.0
            u[:, :, 1, :, 0] = 0.0
            g_u[i, j] = trace(u.view(e_l.shape[0], -1))
    return g_om, g_de, g_ph, g_u


def step_backward(dt, omegas, deltas, phis, interaction_matrix, state, grad_state_out,
                  tol, max_krylov_dim):
    """Restates DifferentiableStateVectorEvolution.backward (torch.py:221-378).

    Returns (g_omegas, g_deltas, g_phis, g_interaction, g_state_in) in torch's
    complex-gradient convention (the same convention the reference's backward
    receives from autograd).
    """
    n = len(omegas)
    ham = OracleHamiltonian(omegas, deltas, phis, interaction_matrix)

    def op(x):
        return -1j * dt * (ham * x)

    vs, ds, vg = double_krylov(op, state.detach(), grad_state_out.detach(), tol, max_krylov_dim)
    Vs = torch.stack(vs)
    Vg = torch.stack(vg)
    e_l = ds.mT @ Vs  # torch.py:318
    g_om, g_de, g_ph, g_u = parameter_traces(Vg, e_l, omegas, phis, -1j * dt)
    g_state = krylov_exp(lambda x: (1j * dt) * (ham * x), grad_state_out.detach().clone(),
                         tol, tol, max_krylov_dim=max_krylov_dim)  # torch.py:355-366
    return g_om, g_de, g_ph, g_u, g_state


class ReferenceStep(torch.autograd.Function):
    """Autograd wrapper over step_forward/step_backward, so a loss over a whole
    pulse can be differentiated exactly as the reference does it (one custom
    Function per time step, torch.py:122-219)."""

    @staticmethod
    def forward(ctx, dt, omegas, deltas, phis, interaction_matrix, state, tol, max_krylov_dim):
        res, _ = step_forward(dt, omegas, deltas, phis, interaction_matrix, state, tol,
                              max_krylov_dim)
        ctx.save_for_backward(omegas, deltas, phis, interaction_matrix, state)
        ctx.dt, ctx.tol, ctx.mk = dt, tol, max_krylov_dim
        return res

    @staticmethod
    def backward(ctx, grad_out):
        om, de, ph, um, st = ctx.saved_tensors
        g = step_backward(ctx.dt, om, de, ph, um, st, grad_out, ctx.tol, ctx.mk)
        return (None, g[0].to(om.dtype) if ctx.needs_input_grad[1] else None,
                g[1].to(de.dtype) if ctx.needs_input_grad[2] else None,
                g[2].to(ph.dtype) if ctx.needs_input_grad[3] else None,
                g[3].to(um.dtype) if ctx.needs_input_grad[4] else None,
                g[4] if ctx.needs_input_grad[5] else None, None, None)


# ---------------------------------------------------------------------------
# Observables  (emu_sv/custom_callback_implementations.py, dense_state.py)
# ---------------------------------------------------------------------------
def occupation(state: torch.Tensor) -> torch.Tensor:
    """custom_callback_implementations.py:30-47"""
    n = int(math.log2(state.numel()))
    out = torch.zeros(n, dtype=torch.float64)
    for i in range(n):
        out[i] = torch.linalg.vector_norm(state.view(2**i, 2, -1)[:, 1]) ** 2
    return out


def correlation_matrix(state: torch.Tensor) -> torch.Tensor:
    """custom_callback_implementations.py:77-102"""
    n = int(math.log2(state.numel()))
    corr = torch.zeros(n, n, dtype=torch.float64)
    for i in range(n):
        sel = state.view(2**i, 2, -1)[:, 1]
        corr[i, i] = torch.linalg.vector_norm(sel) ** 2
        for j in range(i + 1, n):
            sij = sel.reshape(2**i, 2 ** (j - i - 1), 2, -1)[:, :, 1, :]
            corr[i, j] = torch.linalg.vector_norm(sij) ** 2
            corr[j, i] = corr[i, j]
    return corr


def energy_second_moment(ham: OracleHamiltonian, state: torch.Tensor) -> torch.Tensor:
    """custom_callback_implementations.py:170-185"""
    hs = ham * state
    return torch.vdot(hs, hs).real


def energy_variance(ham: OracleHamiltonian, state: torch.Tensor) -> torch.Tensor:
    """custom_callback_implementations.py:131-147"""
    hs = ham * state
    return torch.vdot(hs, hs).real - torch.vdot(state, hs).real ** 2


def fidelity(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """dense_state.py:136-139 for two vectors: |<a|b>|^2."""
    return torch.abs(torch.vdot(a, b)) ** 2


# ---------------------------------------------------------------------------
# Time loop  (emu_sv/sv_backend_impl.py:186-262)
# ---------------------------------------------------------------------------
@dataclass
class OracleRun:
    states: dict = field(default_factory=dict)      # step index -> state (eval steps)
    energy: dict = field(default_factory=dict)
    occupation: dict = field(default_factory=dict)
    iterations: list = field(default_factory=list)  # Lanczos iterations per step


def run_sequence(omega, delta, phi, interaction_matrix: Callable[[float], torch.Tensor],
                 target_times: Sequence[float], krylov_tolerance: float = 1e-10,
                 max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM,
                 eval_steps: Sequence[int] | None = None, initial_state=None,
                 n_steps: int | None = None) -> OracleRun:
    """Restates SVBackendImpl._run / step / _evolve_step / _apply_observables
    (sv_backend_impl.py:186-262) for the noiseless state-vector case.

    Observables at step k use the Hamiltonian of the step that ended there;
    at k=0 the Hamiltonian of omega[0] with U at the first midpoint
    (sv_backend_impl.py:224-234).  ``eval_steps`` are indices into
    target_times (default: only the last).
    """
    nsteps = omega.shape[0] if n_steps is None else n_steps
    nq = omega.shape[1]
    if eval_steps is None:
        eval_steps = [omega.shape[0]]
    eval_steps = set(eval_steps)
    if initial_state is None:
        psi = torch.zeros(2**nq, dtype=DTYPE)
        psi[0] = 1.0
    else:
        psi = initial_state.clone().to(DTYPE)
    out = OracleRun()

    def observe(k, ham):
        norm = torch.linalg.vector_norm(psi)
        st = psi if torch.allclose(norm, torch.tensor(1.0, dtype=torch.float64)) else psi / norm
        out.states[k] = st.clone()
        out.energy[k] = float(ham.expect(st))
        out.occupation[k] = occupation(st)

    if 0 in eval_steps:
        t_mid = 0.5 * (target_times[0] + target_times[1])
        observe(0, OracleHamiltonian(omega[0], delta[0], phi[0], interaction_matrix(t_mid)[0]))
    for i in range(nsteps):
        dt = (target_times[i + 1] - target_times[i]) * TIME_CONVERSION_COEFF
        ham = OracleHamiltonian(omega[i], delta[i], phi[i], interaction_matrix(target_times[i])[0])
        r = krylov_exp_impl(lambda x: -1j * dt * (ham * x), psi, is_hermitian=True,
                            exp_tolerance=krylov_tolerance, norm_tolerance=krylov_tolerance,
                            max_krylov_dim=max_krylov_dim)
        if not r.converged:
            raise RecursionError(
                "exponentiation algorithm did not converge to precision in allotted number of steps."
            )
        psi = r.result
        out.iterations.append(r.iteration_count)
        if i + 1 in eval_steps:
            observe(i + 1, ham)
    return out


# ---------------------------------------------------------------------------
# PCHIP midpoint sampling  (emu_base/math/pchip_torch.py, pulser_adapter.py)
# ---------------------------------------------------------------------------
def pchip_eval(x: torch.Tensor, y: torch.Tensor, xq: torch.Tensor) -> torch.Tensor:
    """Restates PCHIP1D (emu_base/math/pchip_torch.py:97-269), float64."""
    x = x.to(torch.float64)
    y = y.to(torch.float64)
    h = x[1:] - x[:-1]
    delta = (y[1:] - y[:-1]) / h
    npts = h.numel() + 1
    d = torch.zeros(npts, dtype=torch.float64)
    if npts == 2:
        d.fill_(delta[0])
    else:
        dl, dr, hl, hr = delta[:-1], delta[1:], h[:-1], h[1:]
        wl, wr = hl + 2.0 * hr, 2.0 * hl + hr
        hm = (wl + wr) / (wl / dl + wr / dr)
        d[1:-1] = torch.where(dl * dr > 0, hm, torch.zeros_like(hm))

        def endpoint(s_l, s_r, h_l, h_r):
            de = ((2.0 * h_l + h_r) * s_l - h_l * s_r) / (h_l + h_r)
            de = torch.where(de * s_l < 0, torch.zeros_like(de), de)
            cap = (s_l * s_r < 0) & (torch.abs(de) > 3.0 * torch.abs(s_l))
            return torch.where(cap, 3.0 * s_l, de)

        d[0] = endpoint(delta[0], delta[1], h[0], h[1])
        d[-1] = endpoint(delta[-1], delta[-2], h[-1], h[-2])
    p0, p1 = y[:-1], d[:-1]
    p2 = (3.0 * delta - 2.0 * d[:-1] - d[1:]) / h
    p3 = (d[:-1] + d[1:] - 2.0 * delta) / (h * h)
    xq = torch.as_tensor(xq, dtype=torch.float64)
    i = (torch.searchsorted(x, xq, right=True) - 1).clamp(0, x.numel() - 2)
    t = xq - x[i]
    return p0[i] + t * (p1[i] + t * (p2[i] + t * p3[i]))


def midpoint_samples(samples_1ns: torch.Tensor, target_times: Sequence[float],
                     is_amplitude: bool) -> torch.Tensor:
    """Per-step midpoint values from 1 ns samples (pulser_adapter.py:193-226):
    PCHIP on t = 0..T-1, evaluated at (t_k + t_{k+1})/2, last amplitude
    value clamped at >= 0."""
    tt = torch.as_tensor(target_times, dtype=torch.float64)
    t_mid = 0.5 * (tt[:-1] + tt[1:])
    t_grid = torch.arange(target_times[-1], dtype=torch.float64)
    out = pchip_eval(t_grid, torch.as_tensor(samples_1ns, dtype=torch.float64), t_mid)
    if is_amplitude:
        out[-1] = torch.where(out[-1] > 0, out[-1], torch.zeros_like(out[-1]))
    return out


# ---------------------------------------------------------------------------
# Quantum jumps  (emu_base/math/brents_root_finding.py, emu_sv/algebra.py,
# emu_sv/time_evolution/implementations.py:123-250)
# ---------------------------------------------------------------------------
class OracleBrent:
    """Brent's bracketing root finder, driven one abscissa at a time
    (brents_root_finding.py:5-110): inverse quadratic interpolation through
    (a, b, c), secant when two ordinates nearly coincide, bisection when the
    step leaves the bracket or does not halve."""

    def __init__(self, start, end, f_start, f_end, epsilon=1e-6):
        assert start <= end and f_start * f_end < 0
        self.eps = epsilon
        self.a, self.b, self.fa, self.fb = start, end, f_start, f_end
        if abs(self.fa) < abs(self.fb):  # b is the better end (:32-35)
            self.a, self.b, self.fa, self.fb = self.b, self.a, self.fb, self.fa
        self.c, self.d, self.fc = self.a, self.a, self.fa
        self.bisect = True
        self.pending = None

    def next_abscissa(self) -> float:
        # :41-78
        a, b, c, d, fa, fb, fc = self.a, self.b, self.c, self.d, self.fa, self.fb, self.fc
        if abs(fc - fa) < self.eps or abs(fc - fb) < self.eps:
            dx = fb * (b - a) / (fa - fb)
        else:
            s_, r_, t_ = fb / fa, fb / fc, fa / fc
            dx = s_ * (t_ * (r_ - t_) * (c - b) + (r_ - 1) * (b - a)) / ((t_ - 1) * (s_ - 1) * (r_ - 1))
        ref = abs(b - c) if self.bisect else abs(c - d)
        if (abs(dx) >= abs(0.75 * (a - b)) or dx * (a - b) < 0 or abs(dx) >= ref / 2
                or ref < abs(2 * self.eps * b)):
            dx, self.bisect = (a - b) / 2, True
        else:
            self.bisect = False
        self.pending = b + dx
        self.d, self.c, self.fc = c, b, fb
        return self.pending

    def provide(self, x: float, fx: float) -> None:
        # :80-97
        assert x == self.pending
        if self.fa * fx < 0:
            self.b, self.fb = x, fx
        else:
            self.a, self.fa = x, fx
        if abs(self.fa) < abs(self.fb):
            self.a, self.b, self.fa, self.fb = self.b, self.a, self.fb, self.fa

    def converged(self, tol: float) -> bool:
        return abs(self.b - self.a) < tol


def brent_root(f, start, end, tolerance=1e-6, epsilon=1e-6):
    """find_root_brents (brents_root_finding.py:113-150); returns (root, calls)."""
    calls = [0]

    def g(x):
        calls[0] += 1
        return f(x)

    r = OracleBrent(start, end, g(start), g(end), epsilon)
    while not r.converged(tolerance):
        x = r.next_abscissa()
        r.provide(x, g(x))
    return r.b, calls[0]


def expect_batch(state: torch.Tensor, ops: torch.Tensor, n: int) -> torch.Tensor:
    """T[q, i] = Tr(ops_i rho_q) (emu_sv/algebra.py:4-25)."""
    rows = []
    for q in range(n):
        v = state.view(2**q, 2, -1)
        rho = torch.tensordot(v, v.conj(), dims=([0, 2], [0, 2]))
        rows.append(torch.stack([torch.trace(op @ rho) for op in ops]))
    return torch.stack(rows)


def apply_1q(state: torch.Tensor, q: int, op: torch.Tensor) -> torch.Tensor:
    """op on qubit q (emu_sv/algebra.py:28-37)."""
    return (torch.as_tensor(op, dtype=DTYPE) @ state.view(2**q, 2, -1)).reshape(-1)


class OracleMonteCarlo:
    """EvolveMonteCarlo (implementations.py:123-250) with the CPU Hamiltonian
    and the oracle's Arnoldi krylov_exp; draws from Python's ``random`` in
    the same order (one threshold at construction, random.choices and a new
    threshold per jump)."""

    JUMP_TIME_TOLERANCE_FRACTION = 0.1  # implementations.py:13

    def __init__(self, max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM):
        self.max_krylov_dim = max_krylov_dim
        self.jump_threshold = random.uniform(0.0, 1.0)
        self.jumps: list = []

    def evolve(self, dt, ham, state, tol):
        return krylov_exp(lambda x: -1j * dt * (ham * x), state, exp_tolerance=tol,
                          norm_tolerance=tol, is_hermitian=False,
                          max_krylov_dim=self.max_krylov_dim)

    def jump(self, state, lindblads, aggregated, n):
        w = expect_batch(state, aggregated, n).real
        q, op = random.choices([(q, L) for q in range(n) for L in lindblads],
                               weights=w.reshape(-1).tolist())[0]
        self.jumps.append(q)
        state = apply_1q(state, q, op)
        state = state / state.norm()
        self.jump_threshold = random.uniform(0.0, state.norm().item() ** 2)
        return state

    def step(self, dt, omegas, deltas, phis, interaction_matrix, state, tol, lindblads):
        ham = OracleHamiltonian(omegas, deltas, phis, interaction_matrix,
                                noise=noise_from_lindbladians(lindblads))
        n = len(omegas)
        ftol = dt * self.JUMP_TIME_TOLERANCE_FRACTION
        stacked = torch.stack([torch.as_tensor(L, dtype=DTYPE) for L in lindblads])
        aggregated = stacked.conj().transpose(1, 2) @ stacked
        t = 0.0
        gap = state.norm().item() ** 2 - self.jump_threshold
        while abs(dt - t) > 1e-5:
            old = gap
            state = self.evolve(dt - t, ham, state.clone(), tol)
            gap = state.norm().item() ** 2 - self.jump_threshold
            if gap > 0.0:
                break
            r = OracleBrent(t, dt, old, gap, ftol)
            t = dt
            while not r.converged(ftol):
                x = r.next_abscissa()
                state = self.evolve(x - t, ham, state.clone(), tol)
                t = x
                gap = state.norm().item() ** 2 - self.jump_threshold
                r.provide(t, gap)
            state = self.jump(state, lindblads, aggregated, n)
            gap = state.norm().item() ** 2 - self.jump_threshold
        return state


# ---------------------------------------------------------------------------
# Lindblad density-matrix path  (emu_sv/lindblad_operator.py,
# emu_sv/time_evolution/implementations.py:63-120)
# ---------------------------------------------------------------------------
class OracleLindbladian:
    """Restates RydbergLindbladian (lindblad_operator.py:11-223) on CPU
    torch: local 2x2 terms applied through (2^q, 2, rest) views of rho (rows)
    and (2^(q+N), 2, rest) views (columns), the interaction diagonal on the
    rows, then the anti-Hermitian parts of :196-216."""

    def __init__(self, omegas, deltas, phis, pulser_lindblads, interaction_matrix):
        self.nqubits = len(omegas)
        self.omegas = torch.as_tensor(omegas).detach().to(DTYPE) / 2.0
        self.deltas = torch.as_tensor(deltas).detach().to(DTYPE)
        self.phis = torch.as_tensor(phis).detach().to(DTYPE)
        self.lindblads = [torch.as_tensor(L, dtype=DTYPE) for L in pulser_lindblads]
        self.u = torch.as_tensor(interaction_matrix).detach()
        self.complex = bool(self.phis.any())
        n = self.nqubits
        diag = torch.zeros(2**n, dtype=DTYPE)  # :66-83, interaction only
        for i in range(n):
            diag = diag.view(2**i, 2, -1)
            i_fixed = diag[:, 1, :]
            for j in range(i + 1, n):
                i_fixed = i_fixed.view(2**i, 2 ** (j - i - 1), 2, -1)
                i_fixed[:, :, 1, :] += self.u[i, j].item()
        self.diag = diag.view(-1)

    def left(self, rho, op, q):  # :94-113
        return (op @ rho.reshape(2**q, 2, -1)).reshape(rho.shape)

    def right_dag(self, rho, op, q):  # :115-135, rho @ op^H on qubit q
        return (op.conj() @ rho.reshape(2 ** (q + self.nqubits), 2, -1)).reshape(rho.shape)

    def local_h(self, q, k):  # :158-180
        sx = torch.tensor([[0.0, 1.0], [1.0, 0.0]], dtype=DTYPE)
        sy = torch.tensor([[0.0, -1j], [1j, 0.0]], dtype=DTYPE)
        nop = torch.tensor([[0.0, 0.0], [0.0, 1.0]], dtype=DTYPE)
        drive = torch.cos(self.phis[q]) * sx + torch.sin(self.phis[q]) * sy if self.complex else sx
        return self.omegas[q] * drive - self.deltas[q] * nop + k

    def h_eff(self, rho, k=None):  # :137-156
        k = torch.zeros(2, 2, dtype=DTYPE) if k is None else k
        out = torch.zeros_like(rho)
        for q in range(self.nqubits):
            out += self.left(rho, self.local_h(q, k), q)
        return out + self.diag.view(-1, 1) * rho

    def __matmul__(self, rho):  # :182-216
        k = noise_from_lindbladians(self.lindblads)
        h = self.h_eff(rho, k)
        h = h - h.conj().T
        jumps = torch.zeros_like(rho)
        for q in range(self.nqubits):
            for L in self.lindblads:
                jumps += self.right_dag(self.left(rho, L, q), L, q)
        res = h + 1.0j * jumps
        return 0.5 * (res - res.conj().T)

    def expect(self, rho):  # :218-223
        return self.h_eff(rho).trace().real


def dense_lindblad_superoperator(omegas, deltas, phis, interaction_matrix, lindblads):
    """The Liouvillian as a 4^N x 4^N matrix on row-major vec(rho):
    d rho/dt = -i [H, rho] + sum_k (A rho A^H - 1/2 {A^H A, rho}), the
    textbook Lindblad form that -i * (L @ rho) must equal."""
    h = dense_rydberg_hamiltonian(omegas, deltas, phis, interaction_matrix)
    n = len(omegas)
    d = 2**n
    eye = torch.eye(d, dtype=DTYPE)
    # row-major vec: vec(A X B) = (A kron B^T) vec(X)
    sup = -1j * (torch.kron(h, eye) - torch.kron(eye, h.T.contiguous()))
    for L in lindblads:
        L = torch.as_tensor(L, dtype=DTYPE)
        for q in range(n):
            a = _single_gate(q, n, L)
            ada = a.conj().T @ a
            sup += torch.kron(a, a.conj()) - 0.5 * (torch.kron(ada, eye)
                                                   + torch.kron(eye, ada.T.contiguous()))
    return sup


# ---------------------------------------------------------------------------
# Sampling  (emu_sv/utils.py:17-49, emu_sv/dense_state.py:40-95)
# ---------------------------------------------------------------------------
def inverse_cdf_indices(probabilities: torch.Tensor, uniforms: torch.Tensor) -> torch.Tensor:
    """The inverse-CDF branch of sample_indices (emu_sv/utils.py:41-49) with
    the draws given: cdf = cumsum in float64, cdf /= cdf[-1], then
    searchsorted(cdf, u, right=True) clamped to the last index."""
    cdf = torch.cumsum(probabilities.to(torch.float64), dim=0)
    total = cdf[-1].item()
    if not total > 0:
        raise RuntimeError("invalid multinomial distribution (sum of probabilities <= 0)")
    cdf /= total
    idx = torch.searchsorted(cdf, uniforms.to(torch.float64), right=True)
    return idx.clamp_(max=probabilities.numel() - 1)
