"""Many registers evolved together by the device-driven batch engine.

Each register has its own atom count N, its own per-step Hamiltonian
parameters (Omega, delta, phi per qubit and step, interaction matrix), time
steps and initial state.  All of them run the Krylov time stepping of the
single-register engine (emu_base/math/krylov_exp.py:23-131 per step, as
emu_sv/sv_backend_impl.py:186-209 loops over steps), but the Lanczos
coefficients, the Expokit convergence test and the step-to-step transitions
happen on the device: one batched iteration is a fixed kernel sequence,
captured once in a hipGraph and replayed until every register is done
(emu_sv_batch_create / emu_sv_batch_run).  No counterpart in the reference,
which evolves one register per call.
"""

from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from emulators_amd import _native as native

PHASES = {0: "start", 1: "iterating", 2: "combining", 3: "done", 4: "failed", 5: "paused"}


@dataclass
class BatchRegister:
    """omegas, deltas, phis: (steps, N); interaction_matrix: (N, N) (upper
    triangle read); dts: (steps,) in us; psi0: optional (2**N,) state
    (default |g...g>)."""

    omegas: np.ndarray
    deltas: np.ndarray
    phis: np.ndarray
    interaction_matrix: np.ndarray
    dts: np.ndarray
    psi0: torch.Tensor | None = None
    tile_log: int = 0    # this register's H.psi schedule (0: the batch's)
    min_contig: int = 0

    @property
    def nqubits(self) -> int:
        return int(np.asarray(self.omegas).shape[1])

    @property
    def steps(self) -> int:
        return int(np.asarray(self.omegas).shape[0])


@dataclass
class BatchResult:
    states: list
    matvecs: list
    phases: list
    seconds: float
    iterations: int
    failed: list = field(default_factory=list)


class _Reg(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("steps", ctypes.c_int), ("flags", ctypes.c_int),
                ("pause_each_step", ctypes.c_int), ("ham", ctypes.c_void_p), ("dt", ctypes.c_void_p), ("psi", ctypes.c_void_p),
                ("psi_alt", ctypes.c_void_p), ("rows", ctypes.c_void_p)]


def _pack_steps(reg: BatchRegister) -> tuple[np.ndarray, bool]:
    n, s = reg.nqubits, reg.steps
    om = np.ascontiguousarray(np.real(np.asarray(reg.omegas, dtype=np.complex128)), dtype=np.float64)
    de = np.ascontiguousarray(np.real(np.asarray(reg.deltas, dtype=np.complex128)), dtype=np.float64)
    ph = np.ascontiguousarray(np.real(np.asarray(reg.phis, dtype=np.complex128)), dtype=np.float64)
    um = np.ascontiguousarray(np.asarray(reg.interaction_matrix, dtype=np.float64))
    if om.shape != (s, n) or de.shape != (s, n) or ph.shape != (s, n) or um.shape != (n, n):
        raise ValueError("batch register: omegas/deltas/phis must be (steps, N), U (N, N)")
    blocks = np.zeros((s, native.EMU_SV_HAM_DOUBLES), dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    native.check(native.LIB.emu_sv_pack_hamiltonian_steps(
        n, s, om.ctypes.data_as(dp), de.ctypes.data_as(dp), ph.ctypes.data_as(dp),
        um.ctypes.data_as(dp), blocks.ctypes.data_as(dp)), "pack_hamiltonian_steps")
    return blocks, bool(np.any(ph != 0.0))


def evolve_batch(registers: list[BatchRegister], dtype: torch.dtype = torch.complex64,
                 tolerance: float = 1e-10, max_krylov_dim: int = 60, tile_log: int = 0,
                 min_contig: int = 0, check_every: int = 32, max_iterations: int = 1_000_000,
                 device: torch.device | str = "cuda", tiles: str | None = None) -> BatchResult:
    """Evolve every register through all its steps; returns the final states
    (device tensors of ``dtype``) and per-register matvec counts.

    ``tiles="auto"``: each register without a schedule of its own takes the
    one the autotuner picked for its atom count on this register set
    (autotune.with_batch_schedules)."""
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError("the batch engine runs on the GPU")
    if tiles == "auto":
        from emulators_amd import autotune
        registers = autotune.with_batch_schedules(registers, dtype, device, tolerance,
                                                  max_krylov_dim)
    elif tiles is not None:
        raise ValueError("evolve_batch: tiles must be None or 'auto'")
    code = native.dtype_code(dtype)
    count = len(registers)
    if count < 1:
        raise ValueError("evolve_batch needs at least one register")
    keep = []  # device tensors referenced by the descriptors
    regs = (_Reg * count)()
    states = []
    for i, reg in enumerate(registers):
        n = reg.nqubits
        blocks, cplx = _pack_steps(reg)
        ham = torch.from_numpy(blocks).to(device)
        dts = torch.as_tensor(np.asarray(reg.dts, dtype=np.float64)).to(device)
        if dts.numel() != reg.steps:
            raise ValueError("batch register: dts must have one entry per step")
        if reg.psi0 is None:
            psi = torch.zeros(2**n, dtype=dtype, device=device)
            psi[0] = 1.0
        else:
            psi = reg.psi0.to(device=device, dtype=dtype).clone().contiguous()
            if psi.numel() != 2**n:
                raise ValueError("batch register: psi0 must have 2**N elements")
        alt = torch.empty_like(psi)
        rows = torch.empty((max_krylov_dim + 1) * 2**n, dtype=dtype, device=device)
        keep += [ham, dts, alt, rows]
        states.append(psi)
        if not (0 <= reg.tile_log <= 15 and 0 <= reg.min_contig <= 15):
            raise ValueError("batch register: tile_log and min_contig must be in [0, 15]")
        flags = (native.EMU_SV_FLAG_COMPLEX if cplx else 0) | (reg.tile_log << 16) | (reg.min_contig << 20)
        regs[i] = _Reg(n, reg.steps, flags, 0, ham.data_ptr(), dts.data_ptr(), psi.data_ptr(),
                       alt.data_ptr(), rows.data_ptr())
    nbytes = native.LIB.emu_sv_batch_workspace_bytes(count, code, regs, tile_log, min_contig)
    if nbytes == 0:
        raise ValueError("evolve_batch: invalid register set or tile schedule")
    ws = torch.zeros(nbytes, dtype=torch.uint8, device=device)
    stream = native._stream(device)
    handle = ctypes.c_void_p()
    native.check(native.LIB.emu_sv_batch_create(count, code, regs, max_krylov_dim, float(tolerance),
                                                tile_log, min_contig, native._ptr(ws), nbytes, stream,
                                                ctypes.byref(handle)), "batch_create")
    try:
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        remaining = ctypes.c_int(1)
        done_iters = 0
        chunk = max(check_every, 1)
        while remaining.value > 0 and done_iters < max_iterations:
            todo = min(chunk, max_iterations - done_iters)
            native.check(native.LIB.emu_sv_batch_run(handle, todo, check_every, stream,
                                                     ctypes.byref(remaining)), "batch_run")
            done_iters += todo
        native.check(native.LIB.emu_sv_batch_publish(handle, stream), "batch_publish")
        torch.cuda.synchronize(device)
        seconds = time.perf_counter() - t0
        phases = (ctypes.c_int * count)()
        mv = (ctypes.c_int * count)()
        native.check(native.LIB.emu_sv_batch_status(handle, phases, None, mv, stream), "batch_status")
    finally:
        native.LIB.emu_sv_batch_destroy(handle)
    phase_list = [PHASES.get(int(p), str(int(p))) for p in phases]
    failed = [i for i, p in enumerate(phase_list) if p != "done"]
    del keep
    return BatchResult(states=states, matvecs=[int(x) for x in mv], phases=phase_list,
                       seconds=seconds, iterations=done_iters, failed=failed)


def registers_from_sequences(seqs) -> list[BatchRegister]:
    """BatchRegister per SequenceData (midpoint tables, U at t = 0, dt from
    the target times, sv_backend_impl.py:194-209)."""
    out = []
    for seq in seqs:
        tt = np.asarray(seq.target_times, dtype=np.float64)
        dts = (tt[1:] - tt[:-1]) * 1e-3
        u = seq.interaction_matrix(0.0)[0]
        out.append(BatchRegister(seq.omega.numpy(), seq.delta.numpy(), seq.phi.numpy(),
                                 np.asarray(u.numpy() if hasattr(u, "numpy") else u), dts))
    return out
