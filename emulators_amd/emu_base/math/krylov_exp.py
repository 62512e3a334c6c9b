"""Krylov exponential exp(op) v by Lanczos / Arnoldi with the Expokit
extended-T error estimate.

Same contract as the reference (emu_base/math/krylov_exp.py:23-131):
``krylov_exp(op, v, exp_tolerance, norm_tolerance, is_hermitian=True,
max_krylov_dim=100)`` returns exp(op) v and raises RecursionError when the
iteration does not converge within ``max_krylov_dim`` steps; ``v`` is
invalid after the call.

Two implementations sit behind it:

* ``op`` a ``ScaledHamiltonian`` (op(x) = sigma H x for a
  ``RydbergHamiltonian``) with a device vector and ``is_hermitian``: the
  fused engine.  Each Lanczos iteration is one H.psi launch sequence
  (tiled passes, the diagonal generated on the fly): its first pass
  orthogonalises the previous iteration's vector,
  z = w - T[j-1,j] v_{j-1} - T[j,j] v_j, reducing ||z||^2, <z|v_j> and
  <z|w>, and its last pass reduces ||Hz||^2 and <z|Hz>.  With H Hermitian,
  T[j,j+1] = <v_j|sigma H v_{j+1}> = sigma conj(<v_{j+1}|sigma H v_j>)/conj(sigma),
  so the previous vector is not read again; one read-back per iteration.  Basis vectors are stored unnormalised with a host-side scale
  (the normalisation is folded into the next launch).  T is complex128 on
  the host; its exponential and the Expokit error terms come from the
  symmetric tridiagonal form (emu_sv_lanczos_exp: T = W^-1 sigma J W with J
  real symmetric, O(m^2) per iteration instead of a dense (j+3)^2 expm), and
  the result is one linear-combination launch.  The arithmetic is the
  reference's (two-vector modified Gram-Schmidt, T[j-1,j] taken first and
  T[j,j] = <v_j | w - T[j-1,j] v_{j-1}>).  Vectors may be complex64: the
  Lanczos scalars, reductions and T stay FP64.

* ``op`` a ``ScaledHamiltonian`` and ``is_hermitian=False`` (the
  Monte-Carlo effective Hamiltonian with its -i/2 sum L^+L term): the device
  Arnoldi engine.  Per iteration one H.psi launch sequence, then the full
  modified Gram-Schmidt of the reference (h_k = <v_k|w>, w -= h_k v_k for
  k = 0..j in order) as a chain of launches that keeps every h_k on the
  device (emu_sv_arnoldi_orthogonalize), one read-back of the j+2 records,
  and the first column of the (j+3)^2 Hessenberg matrix's exponential on
  the host (the only column the estimate and the combination read).

* any other callable: the same algorithm on torch tensors (Lanczos or full
  Gram-Schmidt Arnoldi), with the small exponential from the native expm.
"""

from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Callable

import numpy as np
import torch

from emulators_amd import _native as native
from emulators_amd import autotune

DEFAULT_MAX_KRYLOV_DIM: int = 100


@dataclass
class KrylovExpResult:
    """Outcome of one Krylov exponential (reference krylov_exp.py:7-20)."""

    result: torch.Tensor
    converged: bool
    happy_breakdown: bool
    iteration_count: int

    def __post_init__(self):
        assert (not self.happy_breakdown) or self.converged


class ScaledHamiltonian:
    """op(x) = sigma * (H x) for a RydbergHamiltonian H (e.g. sigma = -i dt).

    Callable like the reference's ``op`` closures; ``krylov_exp`` detects it
    and runs the fused engine."""

    def __init__(self, hamiltonian, sigma: complex):
        self.hamiltonian = hamiltonian
        self.sigma = complex(sigma)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.hamiltonian.apply(x, sigma=self.sigma)


class KrylovStats:
    """Counters of the fused engine (matvecs and Krylov calls), for reporting."""

    matvecs = 0
    calls = 0

    @classmethod
    def reset(cls):
        cls.matvecs = 0
        cls.calls = 0


def _error_estimate(expd: np.ndarray, j: int, n: float) -> float:
    """Expokit local truncation error from the extended-T exponential
    (krylov_exp.py:81-84)."""
    err1 = abs(expd[j + 1, 0])
    err2 = abs(expd[j + 2, 0] * n)
    return err1 if err1 < err2 else (err1 * err2 / (err1 - err2))


@dataclass
class LanczosRun:
    """Lanczos decomposition of one fused run: op(V) ~ V T, v = initial_norm * V e_0.

    ``basis[k] * scales[k]`` is the k-th orthonormal vector;
    ``T`` the m x m Lanczos matrix (complex128, op = sigma H) and ``n2`` the
    norm of the next vector (0 after a happy breakdown);
    ``coeffs`` = exp(T_ext) e_0 over the m (+1) vectors of the corrected
    Krylov approximation, exp(op) v = initial_norm * sum_k coeffs_k V_k."""

    basis: list
    scales: list
    T: np.ndarray
    coeffs: np.ndarray
    initial_norm: float
    converged: bool
    happy_breakdown: bool
    iteration_count: int
    n2: float
    result: torch.Tensor | None = None  # exp(op) v when the engine formed it already

    @property
    def m(self) -> int:
        return self.T.shape[0]

    def extended(self) -> np.ndarray:
        """The matrix whose exponential gives ``coeffs``: T with the row of the
        next vector appended ((m+1) x (m+1), last column zero) unless the
        iteration broke down."""
        if self.happy_breakdown:
            return self.T.copy()
        m = self.m
        t = np.zeros((m + 1, m + 1), dtype=np.complex128)
        t[:m, :m] = self.T
        t[m, m - 1] = self.n2
        return t

    def combine(self, coeffs: np.ndarray | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
        """initial_norm * sum_k coeffs_k V_k (default coeffs: exp(T) e_0)."""
        if coeffs is None and self.result is not None:
            if out is None:
                return self.result
            return out.copy_(self.result)
        c = self.coeffs if coeffs is None else np.asarray(coeffs)
        k = len(c)
        n = native.nqubits_of(self.basis[0])
        cc = self.initial_norm * c * np.asarray(self.scales[:k])
        if out is None:
            out = torch.empty_like(self.basis[0])
        return native.lincomb(n, self.basis[:k], cc, out)


def _fused_lanczos(op: ScaledHamiltonian, v: torch.Tensor, exp_tolerance: float,
                   norm_tolerance: float, max_krylov_dim: int,
                   rows: torch.Tensor | None = None) -> LanczosRun:
    """Lanczos of krylov_exp.py:23-106 with one fused launch sequence and one
    read-back per iteration.

    Call j computes, in its first H.psi pass, the orthogonalised vector
    z_j = w'_j - T[j-1,j] v_{j-1} - T[j,j] v_j (stored as the next basis
    vector, unnormalised) and then r = sigma H z_j, with ||z_j||^2 and the
    dots of r that give T[j,j+1] and T[j+1,j+1].  The convergence test of
    iteration j therefore lands after the matvec of iteration j+1: one matvec
    per Krylov step is computed and not used.

    ``rows``: optional (max_krylov_dim + 2, 2**n) tensor that receives the
    basis (row 0 a copy of v), so that the basis can enter a GEMM."""
    ham = op.hamiltonian
    n = native.nqubits_of(v)
    if n != ham.nqubits:
        raise ValueError(f"state has {n} qubits, the Hamiltonian {ham.nqubits}")
    params = ham.device_params(v.device)
    sched = autotune.best_schedule(n, v.dtype, v.device, params, ham.complex)
    kw = sched.kwargs()
    KrylovStats.calls += 1
    if rows is not None:
        if rows.dim() != 2 or rows.shape[0] < max_krylov_dim + 2 or rows.shape[1] != v.numel() \
                or rows.dtype != v.dtype or not rows.is_contiguous():
            raise ValueError("rows must be a contiguous (max_krylov_dim + 2, 2**n) tensor of v's dtype")
        rows[0].copy_(v)
        v = rows[0]

    def new_vec(k: int) -> torch.Tensor:
        return rows[k] if rows is not None else torch.empty_like(v)

    initial_norm = math.sqrt(native.dot(n, v)[0].item())
    basis = [v]
    scales = [1.0 / initial_norm]
    T = np.zeros((max_krylov_dim + 2, max_krylov_dim + 2), dtype=np.complex128)

    def finish(m: int, n2: float, happy: bool, converged: bool) -> LanczosRun:
        # the result combines m+1 vectors with the extended-matrix coefficients
        # (krylov_exp.py:88-90), or m after a happy breakdown (:66)
        _, _, c = native.lanczos_exp(T, m, op.sigma, n2, True)
        c = np.asarray(c[:m] if happy else c[: m + 1]).copy()
        return LanczosRun(basis, scales, T[:m, :m].copy(), c, initial_norm, converged, happy, m,
                          0.0 if happy else n2)

    # w'_0 = op(v_0)
    wbuf = new_vec(1)
    d = native.hpsi(n, params, op.sigma, v, scales[0], wbuf, dots=True, complex_coef=ham.complex,
                    **kw).cpu().numpy()
    KrylovStats.matvecs += 1
    w_scale = 1.0
    ww, vw, pw, vp = d[0], complex(d[1], d[2]), 0j, 0j
    for j in range(max_krylov_dim):
        nrm = math.sqrt(max(ww, 0.0))  # ||op(v_j)|| before orthogonalisation (:52)
        if j > 0:
            a = pw
            T[j - 1, j] = a
            b = vw - a * vp
        else:
            a, b = 0j, vw
        T[j, j] = b
        r = new_vec(j + 2)
        o = native.lanczos_matvec(n, params, op.sigma, wbuf, w_scale,
                                  basis[j - 1] if j > 0 else None,
                                  scales[j - 1] if j > 0 else 0.0, a, basis[j], scales[j], b, r,
                                  complex_coef=ham.complex, keep_diag=True, **kw).cpu().numpy()
        KrylovStats.matvecs += 1
        n2 = math.sqrt(max(o[0], 0.0))
        T[j + 1, j] = n2
        if n2 < norm_tolerance:  # happy breakdown (:63-71)
            return finish(j + 1, 0.0, True, True)
        basis.append(wbuf)  # z_j, normalised by its scale 1/n2
        scales.append(1.0 / n2)
        T[j + 2, j + 1] = 1.0
        # exp of the extended matrix: the two Expokit terms (:77-84)
        e1, e2, _ = native.lanczos_exp(T, j + 1, op.sigma, n2, False)
        err2 = e2 * nrm
        err = e1 if e1 < err2 else (e1 * err2 / (e1 - err2))
        if err < exp_tolerance:
            return finish(j + 1, n2, False, True)
        # coefficients of iteration j+1: w'_{j+1} = r/n2, v_{j+1} = z/n2,
        # <v_j|w'_{j+1}> = sigma conj(<v_{j+1}|w'_j>) / conj(sigma) (H Hermitian)
        wbuf, w_scale = r, 1.0 / n2
        vp = complex(o[1], o[2]) / n2
        pw = op.sigma * complex(o[3], o[4]).conjugate() / (n2 * op.sigma.conjugate())
        ww = o[5] / (n2 * n2)
        vw = complex(o[6], o[7]) / (n2 * n2)
    return finish(max_krylov_dim, T[max_krylov_dim, max_krylov_dim - 1].real, False, False)


class _PipelineBuffers:
    """Per-(device, stream) records of the pipelined engine: for iteration j,
    device row j holds [16:22] the coefficients it reads; the pinned host row
    j receives, from the iteration's last kernel, [0:12] the reductions and
    [24:30] the coefficients' side outputs (n2, b, a, ww of j+1)."""

    _cache: dict = {}
    WIDTH = 32

    def __init__(self, device: torch.device, rows: int):
        self.dev = torch.zeros((rows, self.WIDTH), dtype=torch.float64, device=device)
        self.host = torch.zeros((rows, self.WIDTH), dtype=torch.float64, pin_memory=True)
        self.rows = rows

    @classmethod
    def get(cls, device: torch.device, rows: int) -> "_PipelineBuffers":
        key = native.stream_key(device)
        b = cls._cache.get(key)
        if b is None or b.rows < rows:
            b = _PipelineBuffers(device, rows)
            cls._cache[key] = b
        return b


def _pipelined_lanczos(op: ScaledHamiltonian, v: torch.Tensor, exp_tolerance: float,
                       norm_tolerance: float, max_krylov_dim: int, sched,
                       rows: torch.Tensor | None = None) -> LanczosRun:
    """The arithmetic of _fused_lanczos with the host one iteration behind the
    device: the coefficients of iteration j+1 are formed on the device
    (emu_sv_lanczos_next_coef) right after iteration j, so iteration j+1 is
    enqueued before the host has read iteration j's reductions; the host
    checks convergence of iteration j while j+1 runs.  When the error
    sequence predicts that iteration j converges, j+1 is not enqueued ahead;
    otherwise the iteration in flight when convergence is found is discarded.
    T's entries come from the same device values the kernels used."""
    ham = op.hamiltonian
    n = native.nqubits_of(v)
    params = ham.device_params(v.device)
    KrylovStats.calls += 1
    if rows is not None:
        if rows.dim() != 2 or rows.shape[0] < max_krylov_dim + 2 or rows.shape[1] != v.numel() \
                or rows.dtype != v.dtype or not rows.is_contiguous():
            raise ValueError("rows must be a contiguous (max_krylov_dim + 2, 2**n) tensor of v's dtype")
        rows[0].copy_(v)
        v = rows[0]

    def new_vec(k: int) -> torch.Tensor:
        return rows[k] if rows is not None else torch.empty_like(v)

    buf = _PipelineBuffers.get(v.device, max_krylov_dim + 2)
    dev, host = buf.dev, buf.host
    stream = torch.cuda.current_stream(v.device)
    sig = op.sigma
    kw = dict(kernel=sched.kernel, tile_log=sched.tile_log, min_contig=sched.min_contig,
              partners=sched.partners, first_log=sched.first_log)

    initial_norm = math.sqrt(native.dot(n, v)[0].item())
    basis = [v]
    scales = [1.0 / initial_norm]
    T = np.zeros((max_krylov_dim + 2, max_krylov_dim + 2), dtype=np.complex128)

    def finish(m: int, n2: float, happy: bool, converged: bool) -> LanczosRun:
        _, _, c = native.lanczos_exp(T, m, sig, n2, True)
        c = np.asarray(c[:m] if happy else c[: m + 1]).copy()
        return LanczosRun(basis, scales, T[:m, :m].copy(), c, initial_norm, converged, happy, m,
                          0.0 if happy else n2)

    wbuf = new_vec(1)
    d = native.hpsi(n, params, sig, v, scales[0], wbuf, dots=True, complex_coef=ham.complex,
                    **kw).cpu().numpy()
    KrylovStats.matvecs += 1
    b0 = complex(d[1], d[2])
    T[0, 0] = b0
    ww = d[0]  # ||op(v_0)||^2 for iteration 0's error estimate
    # coefficients of iteration 0: a = 0, b = b0, pb = v (scale s0), w_scale = 1
    host[0, 16:22] = torch.tensor([0.0, 0.0, b0.real * scales[0], b0.imag * scales[0], scales[0], 1.0],
                                  dtype=torch.float64)
    dev[0, 16:22].copy_(host[0, 16:22], non_blocking=True)
    events = []
    outs = []  # r buffers

    def enqueue(j: int) -> None:
        r = new_vec(j + 2)
        outs.append(r)
        w = wbuf_of(j)
        # the first-pass diagonal table of the opening hpsi call serves every
        # iteration; the last reduction also forms iteration j+1's coefficients
        # the reductions and side outputs go straight into the pinned host
        # rows (no copy command between iterations); the coefficients the
        # next iteration reads stay on the device
        native.lanczos_iterate_dev(n, params, sig, w, basis_at(j - 1) if j > 0 else None, basis_at(j),
                                   dev[j, 16:22], r, host[j, 0:12], dev[j + 1, 16:22], host[j, 24:30],
                                   complex_coef=ham.complex, keep_diag=True, **kw)
        ev = torch.cuda.Event()
        ev.record(stream)
        events.append(ev)
        KrylovStats.matvecs += 1

    # buffers: iteration j reads w_j (basis j+1 before orthogonalisation), writes r_j
    def wbuf_of(j: int) -> torch.Tensor:
        return wbuf if j == 0 else outs[j - 1]

    def basis_at(k: int) -> torch.Tensor:
        return v if k == 0 else wbuf_of(k - 1)

    enqueue(0)
    errs: list = []
    for j in range(max_krylov_dim):
        # Speculate on iteration j+1 unless iteration j is likely to converge:
        # its error estimate, taken as half the last one, is below the
        # tolerance.  Near convergence the estimates fall by 0.4 per
        # iteration (median at N=26) but plateau for several iterations, so
        # a geometric extrapolation from the last two raises many false
        # alarms.  A false alarm costs one host round trip (~0.15 ms), a
        # missed convergence one discarded matvec.
        likely_done = len(errs) >= 1 and 0.5 * errs[-1] < exp_tolerance
        if j + 1 < max_krylov_dim and not likely_done:
            enqueue(j + 1)
        events[j].synchronize()
        rec = host[j].numpy()
        nrm = math.sqrt(max(ww, 0.0))
        n2 = float(rec[24])
        T[j + 1, j] = n2
        if n2 < norm_tolerance:
            return finish(j + 1, 0.0, True, True)
        basis.append(wbuf_of(j))
        scales.append(1.0 / n2)
        T[j + 2, j + 1] = 1.0
        e1, e2, _ = native.lanczos_exp(T, j + 1, sig, n2, False)
        err2 = e2 * nrm
        err = e1 if e1 < err2 else (e1 * err2 / (e1 - err2))
        if err < exp_tolerance:
            return finish(j + 1, n2, False, True)
        errs.append(err)
        T[j + 1, j + 1] = complex(rec[25], rec[26])
        T[j, j + 1] = complex(rec[27], rec[28])
        ww = float(rec[29])
        if j + 1 < max_krylov_dim and len(events) == j + 1:
            enqueue(j + 1)  # the speculation was skipped and j did not converge
    return finish(max_krylov_dim, T[max_krylov_dim, max_krylov_dim - 1].real, False, False)


PIPELINE = os.environ.get("EMU_SV_PIPELINE", "1") != "0"


# Up to this many qubits a Hermitian Krylov step that needs only its result
# runs in the library's driver (emu_sv_krylov_step: the loop of
# _pipelined_lanczos in native code), which holds a basis of max_dim + 1
# vectors for the call (101 x 64 MiB at 22 complex128 qubits).  Above it the
# Python engine allocates the m + 2 vectors it uses; there an iteration's GPU
# time (0.5 ms and more) already hides the interpreter's work.
NATIVE_LANCZOS_MAX_QUBITS = 22
_LDS_DEFAULT = autotune.Schedule("lds", 0, 0)


def _lanczos_schedule(op: ScaledHamiltonian, v: torch.Tensor):
    """(schedule, pipelined): the tuned H.psi schedule of this run and whether
    the device-coefficient (pipelined) iteration applies to it."""
    ham = op.hamiltonian
    n = native.nqubits_of(v)
    sched = autotune.best_schedule(n, v.dtype, v.device, ham.device_params(v.device), ham.complex)
    if sched.kernel == "btile":  # one big tile: 2^15 (c64) / 2^14 (c128) amplitudes
        tl = fl = 16 if v.dtype == torch.complex64 else 15
    else:
        tl = sched.tile_log or (13 if v.dtype == torch.complex64 else 12)
        fl = sched.first_log or tl
    ok = PIPELINE and sched.kernel in ("tile", "lds", "btile") and n >= max(tl, fl)
    return sched, ok


def _native_lanczos(op: ScaledHamiltonian, v: torch.Tensor, exp_tolerance: float,
                    norm_tolerance: float, max_krylov_dim: int, sched) -> KrylovExpResult:
    """The Hermitian branch in one library call (emu_sv_krylov_step: the loop
    of _pipelined_lanczos in native code)."""
    shape = v.shape
    v = v.contiguous().view(-1)
    n = native.nqubits_of(v)
    ham = op.hamiltonian
    out = torch.empty_like(v)
    KrylovStats.calls += 1
    iters, converged, happy, matvecs, _ = native.krylov_step(
        n, ham.device_params(v.device), op.sigma, v, out, max_krylov_dim, exp_tolerance,
        norm_tolerance, complex_coef=ham.complex, **sched.kwargs())
    KrylovStats.matvecs += matvecs
    return KrylovExpResult(out.view(shape), converged, happy, iters)


def _native_lanczos_run(op: ScaledHamiltonian, v: torch.Tensor, exp_tolerance: float,
                        norm_tolerance: float, max_krylov_dim: int, sched,
                        rows: torch.Tensor | None = None) -> LanczosRun:
    """The Lanczos decomposition of one step by the library's driver
    (emu_sv_krylov_step with its decomposition output), for callers that
    reuse the Krylov space (the adjoint).  With ``rows`` (a contiguous
    (max_krylov_dim + 2, 2**n) tensor) the basis lands in its rows, row 0 a
    copy of v; otherwise the used rows are kept in a compact tensor.  The
    combination exp(op) v is computed by the driver and cached on the run."""
    ham = op.hamiltonian
    v = v.contiguous().view(-1)
    n = native.nqubits_of(v)
    KrylovStats.calls += 1
    if rows is not None:
        if rows.dim() != 2 or rows.shape[0] < max_krylov_dim + 2 or rows.shape[1] != v.numel() \
                or rows.dtype != v.dtype or not rows.is_contiguous():
            raise ValueError("rows must be a contiguous (max_krylov_dim + 2, 2**n) tensor of v's dtype")
        rows[0].copy_(v)
        v = rows[0]
        buf = rows[1:]
    else:
        buf = torch.empty((max_krylov_dim + 1, v.numel()), dtype=v.dtype, device=v.device)
    out = torch.empty_like(v)
    iters, converged, happy, matvecs, dec = native.krylov_step(
        n, ham.device_params(v.device), op.sigma, v, out, max_krylov_dim, exp_tolerance,
        norm_tolerance, complex_coef=ham.complex, basis=buf, decomp=True, **sched.kwargs())
    KrylovStats.matvecs += matvecs
    big = max_krylov_dim + 2
    m = int(dec[0])
    count = m if happy else m + 1
    T = np.zeros((m, m), dtype=np.complex128)
    diag = dec[4: 4 + 2 * m].view(np.complex128)
    sub = dec[4 + 2 * big: 4 + 2 * big + m]
    sup = dec[4 + 3 * big: 4 + 3 * big + 2 * m].view(np.complex128)
    idx = np.arange(m)
    T[idx, idx] = diag
    T[idx[1:], idx[:-1]] = sub[: m - 1]
    T[idx[:-1], idx[1:]] = sup[: m - 1]
    scales = [float(x) for x in dec[4 + 5 * big: 4 + 5 * big + count]]
    coeffs = dec[4 + 6 * big: 4 + 6 * big + 2 * count].view(np.complex128).copy()
    if rows is None and count > 1:
        buf = buf[: count - 1].clone()  # only the rows this run used stay alive
    basis = [v] + [buf[k - 1] for k in range(1, count)]
    run = LanczosRun(basis, scales, T, coeffs, 1.0 / scales[0], converged, happy, m,
                     float(dec[1]))
    run.result = out
    return run


def fused_lanczos(op: ScaledHamiltonian, v: torch.Tensor, exp_tolerance: float,
                  norm_tolerance: float, max_krylov_dim: int,
                  rows: torch.Tensor | None = None) -> LanczosRun:
    """The fused engine: pipelined (device-side coefficients) when the tuned
    schedule is the tile or LDS kernel and N covers its tiles, else one
    read-back per iteration.  EMU_SV_PIPELINE=0 turns the pipeline off."""
    n = native.nqubits_of(v)
    if n == op.hamiltonian.nqubits:
        sched, pipelined = _lanczos_schedule(op, v)
        if PIPELINE and n <= NATIVE_LANCZOS_MAX_QUBITS and max_krylov_dim <= 256:
            return _native_lanczos_run(op, v, exp_tolerance, norm_tolerance, max_krylov_dim,
                                       sched if pipelined else _LDS_DEFAULT, rows)
        if pipelined:
            return _pipelined_lanczos(op, v, exp_tolerance, norm_tolerance, max_krylov_dim, sched,
                                      rows)
    return _fused_lanczos(op, v, exp_tolerance, norm_tolerance, max_krylov_dim, rows)


def _arnoldi_matvec(op, v: torch.Tensor, n: int):
    """dst = op(scale * src) on flat device vectors for the Arnoldi engines."""
    if not isinstance(op, ScaledHamiltonian):
        return op.device_matvec
    ham = op.hamiltonian
    if n != ham.nqubits:
        raise ValueError(f"vector has {n} qubits, the Hamiltonian {ham.nqubits}")
    params = ham.device_params(v.device)
    sched = ham._schedule(v)
    built = []

    def matvec(src, scale, dst):
        # the first call of this run builds the diagonal table; the later
        # ones keep it (the library rebuilds it if another block intervened)
        native.hpsi(n, params, op.sigma, src, scale, dst, complex_coef=ham.complex,
                    noise=ham.noisy, keep_diag=bool(built), **sched.kwargs())
        built.append(True)

    return matvec


class _ArnoldiBuffers:
    """Per-(device, stream) pinned host rows of the pipelined Arnoldi engine:
    iteration j's (j+2) x 4 Gram-Schmidt records land in host[j] as they
    complete, with events[j] recorded after the copy."""

    _cache: dict = {}

    def __init__(self, rows: int, device: torch.device):
        self.host = torch.zeros((rows, rows + 2, 4), dtype=torch.float64, pin_memory=True)
        self.events = [torch.cuda.Event() for _ in range(rows)]
        self.rows = rows
        self.ld = rows + 1  # Gram-form state: up to rows basis vectors plus the next
        self.gram_state = native.arnoldi_gram_state(self.ld, device)

    @classmethod
    def get(cls, device: torch.device, rows: int) -> "_ArnoldiBuffers":
        key = native.stream_key(device)
        b = cls._cache.get(key)
        if b is None or b.rows < rows:
            b = cls._cache[key] = cls(rows, device)
        return b


def _gram_schmidt_in_one_workgroup(n: int, max_krylov_dim: int) -> bool:
    """The one-workgroup modified Gram-Schmidt kernel serves 2^10..2^13
    elements (up to 128 basis vectors); elsewhere the pipelined engines use
    the Gram-matrix form (emu_sv_arnoldi_gram)."""
    return 10 <= n <= 13 and max_krylov_dim <= 128


def _pipelined_arnoldi(op, v: torch.Tensor, exp_tolerance: float, norm_tolerance: float,
                       max_krylov_dim: int) -> KrylovExpResult:
    """The Arnoldi branch with the host one iteration behind the device.
    Each new basis vector is normalised on the device by its own Gram-Schmidt
    launch (emu_sv_arnoldi_orthonormalize: v_{j+1} = w / T[j+1, j], the
    reference's ``w /= n2``), so iteration j+1 is enqueued before the host
    has read iteration j's records; the host then builds column j of T,
    evaluates the Expokit test and decides.  Iteration j+1 is not enqueued
    ahead when iteration j is likely to converge (half the last error
    estimate below the tolerance, as in the Lanczos pipeline); a discarded
    speculative iteration costs one matvec.  Same arithmetic as
    ``_fused_arnoldi`` apart from the rounding of the stored unit vectors.

    Outside the one-workgroup range the Gram-Schmidt step is the Gram-matrix
    form (emu_sv_arnoldi_gram): the same coefficients from one overlap sweep
    and one subtraction sweep, with the new basis vector left unnormalised
    and its scale kept on the device (the records then hold T[k, j]
    itself)."""
    shape = v.shape
    v = v.contiguous().view(-1)
    n = native.nqubits_of(v)
    matvec = _arnoldi_matvec(op, v, n)
    initial_norm = math.sqrt(float(native.dot(n, v)[0].item()))
    basis = [v]
    scales = [1.0 / initial_norm]  # later basis vectors: unit (MGS) or 1 / ||u|| (Gram form)
    T = np.zeros((max_krylov_dim + 2, max_krylov_dim + 2), dtype=np.complex128)
    bufs = _ArnoldiBuffers.get(v.device, max_krylov_dim)
    gram = not _gram_schmidt_in_one_workgroup(n, max_krylov_dim)

    def enqueue(j: int) -> None:
        w = torch.empty_like(v)
        matvec(basis[j], 1.0 if gram else scales[j], w)
        KrylovStats.matvecs += 1
        if gram:
            rec = native.arnoldi_gram(n, basis[: j + 1], w, scales[0], bufs.gram_state, bufs.ld)
        else:
            rec = native.arnoldi_orthogonalize(n, basis[: j + 1], scales[: j + 1], w, normalize=True)
        bufs.host[j, : j + 2].copy_(rec, non_blocking=True)
        bufs.events[j].record()
        basis.append(w)
        scales.append(1.0)

    def combine(coeffs) -> torch.Tensor:
        c = np.asarray(coeffs, dtype=np.complex128) * np.asarray(scales[: len(coeffs)]) * initial_norm
        return native.lincomb(n, basis[: len(coeffs)], c, torch.empty_like(v)).view(shape)

    enqueue(0)
    errs: list = []
    expd = None
    for j in range(max_krylov_dim):
        likely_done = len(errs) >= 1 and 0.5 * errs[-1] < exp_tolerance
        if j + 1 < max_krylov_dim and not likely_done:
            enqueue(j + 1)
        bufs.events[j].synchronize()
        rec = bufs.host[j, : j + 2].numpy()
        nrm = math.sqrt(max(rec[0, 0], 0.0))  # ||op(v_j)||, before orthogonalisation
        for k in range(j + 1):
            T[k, j] = (1.0 if gram else scales[k]) * complex(rec[k, 1], rec[k, 2])
        n2 = math.sqrt(max(rec[j + 1, 0], 0.0))
        T[j + 1, j] = n2
        if gram and n2 > 0.0:
            scales[j + 1] = 1.0 / n2
        if n2 < norm_tolerance:  # happy breakdown
            expd = native.expm_column(T[: j + 1, : j + 1])[:, None]
            return KrylovExpResult(combine(expd[:, 0]), True, True, j + 1)
        T[j + 2, j + 1] = 1.0
        expd = native.expm_column(T[: j + 3, : j + 3])[:, None]
        err = _error_estimate(expd, j, nrm)
        if err < exp_tolerance:
            return KrylovExpResult(combine(expd[: j + 2, 0]), True, False, j + 1)
        errs.append(err)
        if j + 1 < max_krylov_dim and len(basis) == j + 2:
            enqueue(j + 1)  # the speculation was skipped and j did not converge
    return KrylovExpResult(combine(expd[: max_krylov_dim + 1, 0]), False, False, max_krylov_dim)


# Up to this many qubits the Arnoldi iteration runs in the library's driver
# (emu_sv_arnoldi_exp, with a basis of max_krylov_dim vectors from torch's
# allocator for the call: 1.6 GiB at 20 complex128 qubits): there the
# per-iteration GPU work is short enough that the interpreter's work per
# iteration would show (a 50-step N=18 trajectory 0.32 -> 0.24 s; at N=20
# the two engines tie).
NATIVE_ARNOLDI_MAX_QUBITS = 20


def _native_arnoldi(op: ScaledHamiltonian, v: torch.Tensor, exp_tolerance: float,
                    norm_tolerance: float, max_krylov_dim: int) -> KrylovExpResult:
    """The Arnoldi branch in one library call (emu_sv_arnoldi_exp: the loop
    of _pipelined_arnoldi in native code)."""
    shape = v.shape
    v = v.contiguous().view(-1)
    n = native.nqubits_of(v)
    ham = op.hamiltonian
    if n != ham.nqubits:
        raise ValueError(f"vector has {n} qubits, the Hamiltonian {ham.nqubits}")
    out = torch.empty_like(v)
    iters, converged, happy, matvecs = native.arnoldi_exp(
        n, ham.device_params(v.device), op.sigma, v, out, max_krylov_dim, exp_tolerance,
        norm_tolerance, complex_coef=ham.complex, noise=ham.noisy, **ham._schedule(v).kwargs())
    KrylovStats.matvecs += matvecs
    return KrylovExpResult(out.view(shape), converged, happy, iters)


def _fused_arnoldi(op, v: torch.Tensor, exp_tolerance: float,
                   norm_tolerance: float, max_krylov_dim: int) -> KrylovExpResult:
    """Arnoldi branch of krylov_exp_impl (krylov_exp.py:44-107 with
    is_hermitian=False) on the device kernels.  ``op`` is a
    ``ScaledHamiltonian`` or an operator with ``device_matvec(src, scale,
    dst)`` (dst = op(scale * src) on flat device vectors, e.g. the scaled
    Lindbladian on a density matrix).  Basis vectors are stored unnormalised
    with a host-side scale; v itself is basis vector 0.  One host round trip
    per iteration (EMU_SV_PIPELINE=0; the default is _pipelined_arnoldi)."""
    if PIPELINE:
        if (isinstance(op, ScaledHamiltonian) and native.nqubits_of(v) <= NATIVE_ARNOLDI_MAX_QUBITS
                and max_krylov_dim <= 256):
            return _native_arnoldi(op, v, exp_tolerance, norm_tolerance, max_krylov_dim)
        return _pipelined_arnoldi(op, v, exp_tolerance, norm_tolerance, max_krylov_dim)
    shape = v.shape
    v = v.contiguous().view(-1)
    n = native.nqubits_of(v)
    matvec = _arnoldi_matvec(op, v, n)
    initial_norm = math.sqrt(float(native.dot(n, v)[0].item()))
    basis = [v]
    scales = [1.0 / initial_norm]
    T = np.zeros((max_krylov_dim + 2, max_krylov_dim + 2), dtype=np.complex128)

    def combine(coeffs) -> torch.Tensor:
        c = np.asarray(coeffs, dtype=np.complex128) * np.asarray(scales[: len(coeffs)]) * initial_norm
        return native.lincomb(n, basis[: len(coeffs)], c, torch.empty_like(v)).view(shape)

    expd = None
    for j in range(max_krylov_dim):
        w = torch.empty_like(v)
        matvec(basis[j], scales[j], w)
        KrylovStats.matvecs += 1
        rec = native.arnoldi_orthogonalize(n, basis, scales, w).cpu().numpy()
        nrm = math.sqrt(max(rec[0, 0], 0.0))  # ||op(v_j)||, before orthogonalisation
        for k in range(j + 1):
            T[k, j] = scales[k] * complex(rec[k, 1], rec[k, 2])
        n2 = math.sqrt(max(rec[j + 1, 0], 0.0))
        T[j + 1, j] = n2
        if n2 < norm_tolerance:  # happy breakdown
            expd = native.expm_column(T[: j + 1, : j + 1])[:, None]
            return KrylovExpResult(combine(expd[:, 0]), True, True, j + 1)
        basis.append(w)
        scales.append(1.0 / n2)
        T[j + 2, j + 1] = 1.0
        expd = native.expm_column(T[: j + 3, : j + 3])[:, None]
        if _error_estimate(expd, j, nrm) < exp_tolerance:
            return KrylovExpResult(combine(expd[: len(basis), 0]), True, False, j + 1)
    return KrylovExpResult(combine(expd[: len(basis), 0]), False, False, max_krylov_dim)


def _generic_krylov(op: Callable, v: torch.Tensor, is_hermitian: bool, exp_tolerance: float,
                    norm_tolerance: float, max_krylov_dim: int) -> KrylovExpResult:
    initial_norm = v.norm()
    v /= initial_norm
    vecs = [v]
    T = np.zeros((max_krylov_dim + 2, max_krylov_dim + 2), dtype=np.complex128)

    def combine(coeffs):
        acc = None
        for c, x in zip(coeffs, vecs):
            term = complex(c) * x
            acc = term if acc is None else acc + term
        return initial_norm * acc

    expd = None
    for j in range(max_krylov_dim):
        w = op(vecs[-1])
        nrm = float(w.norm())
        k_start = max(0, j - 1) if is_hermitian else 0
        for k in range(k_start, j + 1):
            overlap = torch.tensordot(vecs[k].conj(), w, dims=w.dim())
            T[k, j] = complex(overlap)
            w -= overlap * vecs[k]
        n2 = float(w.norm())
        T[j + 1, j] = n2
        if n2 < norm_tolerance:
            expd = native.expm(T[: j + 1, : j + 1])
            return KrylovExpResult(combine(expd[:, 0]), True, True, j + 1)
        w /= n2
        vecs.append(w)
        T[j + 2, j + 1] = 1.0
        expd = native.expm(T[: j + 3, : j + 3])
        if _error_estimate(expd, j, nrm) < exp_tolerance:
            return KrylovExpResult(combine(expd[: len(vecs), 0]), True, False, j + 1)
    return KrylovExpResult(combine(expd[: len(vecs), 0]), False, False, max_krylov_dim)


def krylov_exp_impl(op: Callable[[torch.Tensor], torch.Tensor], v: torch.Tensor,
                    is_hermitian: bool, exp_tolerance: float, norm_tolerance: float,
                    max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM) -> KrylovExpResult:
    """exp(op) v; see the module docstring for the two engines."""
    if isinstance(op, ScaledHamiltonian):
        if not v.is_cuda:
            raise ValueError("the fused Krylov engine needs the state on the GPU")
        if not is_hermitian:
            return _fused_arnoldi(op, v, exp_tolerance, norm_tolerance, max_krylov_dim)
        if getattr(op.hamiltonian, "noisy", False):
            raise ValueError("the noise term makes H non-Hermitian: use is_hermitian=False")
        n = native.nqubits_of(v)
        if PIPELINE and n == op.hamiltonian.nqubits and n <= NATIVE_LANCZOS_MAX_QUBITS \
                and max_krylov_dim <= 256:
            sched, pipelined = _lanczos_schedule(op, v)
            # below the tuned sizes (or tiles larger than the register) the
            # LDS kernel with its default tiles, one pass for small registers
            return _native_lanczos(op, v, exp_tolerance, norm_tolerance, max_krylov_dim,
                                   sched if pipelined else _LDS_DEFAULT)
        run = fused_lanczos(op, v, exp_tolerance, norm_tolerance, max_krylov_dim)
        return KrylovExpResult(run.combine(), run.converged, run.happy_breakdown,
                               run.iteration_count)
    if hasattr(op, "device_matvec"):  # e.g. the scaled Lindbladian
        if not v.is_cuda:
            raise ValueError("the fused Krylov engine needs the state on the GPU")
        if is_hermitian:
            raise ValueError("this operator runs on the Arnoldi engine: use is_hermitian=False")
        return _fused_arnoldi(op, v, exp_tolerance, norm_tolerance, max_krylov_dim)
    return _generic_krylov(op, v, is_hermitian, exp_tolerance, norm_tolerance, max_krylov_dim)


def krylov_exp(op: Callable[[torch.Tensor], torch.Tensor], v: torch.Tensor,
               exp_tolerance: float, norm_tolerance: float, is_hermitian: bool = True,
               max_krylov_dim: int = DEFAULT_MAX_KRYLOV_DIM) -> torch.Tensor:
    """exp(op) v, RecursionError if not converged (krylov_exp.py:109-131)."""
    res = krylov_exp_impl(op, v, is_hermitian=is_hermitian, exp_tolerance=exp_tolerance,
                          norm_tolerance=norm_tolerance, max_krylov_dim=max_krylov_dim)
    if not res.converged:
        raise RecursionError(
            "exponentiation algorithm did not converge to precision in allotted number of steps."
        )
    return res.result
