"""Schedule autotuner for the H.psi passes.

A schedule is (kernel kind, tile_log, min_contig):

* "wave": a wavefront owns 2^tile_log amplitudes (tile_log = 6 lane bits +
  3..5 register bits); flips are register or cross-lane operations.
* "lds":  a 256-thread workgroup stages 2^tile_log amplitudes in LDS.
* "btile": 1024-thread workgroups holding 2^14 (c128) / 2^15 (c64)
  amplitudes in registers (wave-bit flips through LDS): a contiguous first
  pass and later passes of ``min_contig`` low bits plus high bits; two passes
  cover n = 26 (``tile_log`` is ignored).
* "tile": a contiguous first pass (orthogonalisation + diagonal) and later
  passes of ``min_contig`` contiguous low bits plus high bits, the last of
  which reads the flip partners of up to ``partners`` top bits directly
  (16-byte packets, compile-time tiles, 2^11..2^14 elements).

``tile_log`` and ``min_contig`` fix the qubit split: pass 1 covers bits
[0, tile_log); each later pass covers the next tile_log - c high bits plus c
contiguous low bits (c >= min_contig) for coalesced segments.

The first use of a (device, n, dtype, coefficient kind) with n >= 16 times
every candidate on one Lanczos iteration (emu_sv_lanczos_matvec: two
untimed and three timed calls each, HIP events on the current stream) and
keeps the fastest for the rest of the process.  Set EMU_SV_AUTOTUNE=0 to use the library default, or
EMU_SV_SCHEDULE="c128=tile:12:4:6,c64=lds:11:3" (kind:tile_log:min_contig[:partners[:first_log]])
to fix the schedule per type.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import torch

from emulators_amd import _native as native

MIN_TUNED_QUBITS = 16


@dataclass(frozen=True)
class Schedule:
    kernel: str = "default"
    tile_log: int = 0
    min_contig: int = 0
    partners: int = 0
    first_log: int = 0

    def kwargs(self) -> dict:
        return {"kernel": self.kernel, "tile_log": self.tile_log, "min_contig": self.min_contig,
                "partners": self.partners, "first_log": self.first_log}


DEFAULT = Schedule()

CANDIDATES = {
    torch.complex64: [Schedule("lds", 11, 3), Schedule("wave", 11, 3),
                      Schedule("tile", 12, 5, 0, 11), Schedule("tile", 12, 5, 0, 12),
                      Schedule("tile", 13, 5, 0, 11), Schedule("tile", 13, 6, 0, 12),
                      Schedule("tile", 13, 4, 4, 13), Schedule("tile", 12, 4, 6, 12),
                      Schedule("tile", 14, 5, 4, 13), Schedule("tile", 14, 5, 6, 12),
                      Schedule("tile", 14, 6, 6, 12), Schedule("tile", 13, 5, 6, 12),
                      Schedule("btile", 0, 4), Schedule("btile", 0, 5)],
    torch.complex128: [Schedule("lds", 10, 2), Schedule("lds", 11, 2), Schedule("wave", 10, 2),
                       Schedule("tile", 12, 4, 0, 10), Schedule("tile", 12, 4, 0, 11),
                       Schedule("tile", 12, 5, 0, 10), Schedule("tile", 13, 5, 0, 10),
                       Schedule("tile", 12, 4, 6, 12), Schedule("tile", 13, 4, 4, 13),
                       Schedule("tile", 13, 4, 6, 12), Schedule("tile", 13, 5, 6, 12),
                       Schedule("btile", 0, 2), Schedule("btile", 0, 4)],
}

_cache: dict = {}
_timings: dict = {}


def _enabled() -> bool:
    return os.environ.get("EMU_SV_AUTOTUNE", "1") != "0"


def time_schedule(n: int, dtype: torch.dtype, params: torch.Tensor, sched: Schedule,
                  complex_coef: bool, vecs, warmup: int = 2, reps: int = 3) -> float:
    """Average ms of one Lanczos iteration (emu_sv_lanczos_matvec with zero
    orthogonalisation coefficients: the traffic of a real iteration)."""
    w, pa, pb, r = vecs

    def call():
        native.lanczos_matvec(n, params, -0.01j, w, 1.0, pa, 1.0, 0j, pb, 1.0, 0j, r,
                              complex_coef=complex_coef, **sched.kwargs())

    for _ in range(warmup):
        call()
    stream = torch.cuda.current_stream(w.device)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        call()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def tune(n: int, dtype: torch.dtype, device: torch.device, params: torch.Tensor,
         complex_coef: bool, candidates=None) -> tuple[Schedule, dict]:
    """Time every candidate; returns (best, {schedule: ms})."""
    g = torch.Generator(device=device).manual_seed(1234)
    vecs = tuple(torch.randn(2**n, dtype=dtype, device=device, generator=g) for _ in range(4))
    times = {}
    for sched in candidates or CANDIDATES[dtype]:
        times[sched] = time_schedule(n, dtype, params, sched, complex_coef, vecs)
    best = min(times, key=times.get)
    del vecs
    return best, times


def forced_schedule(dtype: torch.dtype) -> Schedule | None:
    """The schedule EMU_SV_SCHEDULE fixes for ``dtype``, if any."""
    spec = os.environ.get("EMU_SV_SCHEDULE", "")
    key = "c128" if dtype == torch.complex128 else "c64"
    for item in filter(None, (x.strip() for x in spec.split(","))):
        name, _, val = item.partition("=")
        if name.strip() == key:
            fields = val.split(":")
            kind, tl, cmin = fields[:3]
            partners = int(fields[3]) if len(fields) > 3 else 0
            first = int(fields[4]) if len(fields) > 4 else 0
            return Schedule(kind.strip(), int(tl), int(cmin), partners, first)
    return None


def best_schedule(n: int, dtype: torch.dtype, device: torch.device, params: torch.Tensor,
                  complex_coef: bool) -> Schedule:
    """The tuned schedule for this shape (DEFAULT below MIN_TUNED_QUBITS or when disabled)."""
    forced = forced_schedule(dtype)
    if forced is not None:
        return forced
    if n < MIN_TUNED_QUBITS or not _enabled():
        return DEFAULT
    idx = device.index if device.index is not None else torch.cuda.current_device()
    key = (idx, n, dtype, bool(complex_coef))
    sched = _cache.get(key)
    if sched is None:
        sched, times = tune(n, dtype, device, params, complex_coef)
        _cache[key] = sched
        _timings[key] = times
    return sched


# Noisy H.psi (the Monte-Carlo effective Hamiltonian, EMU_SV_FLAG_NOISE): the
# LDS and big-tile kernels carry the noise term.  The Arnoldi iteration calls
# bare H.psi (no orthogonalisation), so the candidates are timed on that.
NOISE_CANDIDATES = {
    torch.complex64: [Schedule("lds", 0, 0), Schedule("lds", 9, 2), Schedule("lds", 10, 3),
                      Schedule("lds", 11, 3), Schedule("btile", 0, 4), Schedule("btile", 0, 5)],
    torch.complex128: [Schedule("lds", 0, 0), Schedule("lds", 9, 2), Schedule("lds", 10, 2),
                       Schedule("lds", 11, 2), Schedule("btile", 0, 2), Schedule("btile", 0, 4)],
}
NOISE_DEFAULT = Schedule("lds", 0, 0)
_noise_cache: dict = {}


def time_noise_schedule(n: int, params: torch.Tensor, sched: Schedule, vecs, warmup: int = 2,
                        reps: int = 3) -> float:
    """Average ms of one noisy bare H.psi (emu_sv_hpsi with EMU_SV_FLAG_NOISE)."""
    v, w = vecs

    def call():
        native.hpsi(n, params, -0.01j, v, 1.0, w, complex_coef=True, noise=True, **sched.kwargs())

    for _ in range(warmup):
        call()
    stream = torch.cuda.current_stream(v.device)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(reps):
        call()
    stop.record(stream)
    stop.synchronize()
    return start.elapsed_time(stop) / reps


def best_noise_schedule(n: int, dtype: torch.dtype, device: torch.device,
                        params: torch.Tensor) -> Schedule:
    """The fastest noisy H.psi schedule for this shape (NOISE_DEFAULT below
    MIN_TUNED_QUBITS or when tuning is disabled; EMU_SV_SCHEDULE applies when
    it names a kernel that carries the noise term)."""
    forced = forced_schedule(dtype)
    if forced is not None and forced.kernel in ("lds", "btile"):
        return forced
    if n < MIN_TUNED_QUBITS or not _enabled():
        return NOISE_DEFAULT
    idx = device.index if device.index is not None else torch.cuda.current_device()
    key = (idx, n, dtype, "noise")
    sched = _noise_cache.get(key)
    if sched is None:
        g = torch.Generator(device=device).manual_seed(1234)
        vecs = tuple(torch.randn(2**n, dtype=dtype, device=device, generator=g) for _ in range(2))
        times = {c: time_noise_schedule(n, params, c, vecs) for c in NOISE_CANDIDATES[dtype]}
        del vecs
        sched = _noise_cache[key] = min(times, key=times.get)
        _timings[key] = times
    return sched


# Batch engine: (tile_log, min_contig) candidates of the LDS pass kernel per
# element type, and the tuned choice per (device, dtype, atom count).
BATCH_CANDIDATES = {
    torch.complex64: [(11, 3), (12, 4), (13, 4), (13, 5), (14, 5)],
    torch.complex128: [(10, 2), (11, 2), (12, 3), (13, 4)],
}
_batch_cache: dict = {}
_batch_timings: dict = {}


def _truncated(registers, steps: int):
    from emulators_amd.emu_sv.batch import BatchRegister

    return [BatchRegister(r.omegas[:steps], r.deltas[:steps], r.phis[:steps], r.interaction_matrix,
                          r.dts[:steps], r.psi0) for r in registers]


def _time_batch(regs, picks: dict, dtype, device, tolerance, max_krylov_dim, reps: int = 2) -> float:
    """Best device seconds of ``reps`` batched runs of ``regs`` with schedule
    picks[n] per atom count (set-up and graph capture are not timed)."""
    from dataclasses import replace

    from emulators_amd.emu_sv.batch import evolve_batch

    sched = [replace(r, tile_log=picks[r.nqubits][0], min_contig=picks[r.nqubits][1]) for r in regs]
    return min(evolve_batch(sched, dtype=dtype, tolerance=tolerance, max_krylov_dim=max_krylov_dim,
                            device=device).seconds for _ in range(reps))


def tune_batch(registers, dtype: torch.dtype, device: torch.device, tolerance: float,
               max_krylov_dim: int, steps: int = 4, gain: float = 0.02, refine: int = 4) -> dict:
    """(tile_log, min_contig) per atom count for this set of registers, timed
    on the set itself: every register truncated to its first ``steps``
    steps, run as one batch.  First the best uniform schedule, then, for the
    atom counts in decreasing order (the largest carry most of the work),
    each alternative schedule kept when it saves more than ``gain`` of the
    batch time (the ``refine`` largest counts only: the work per register
    grows as 2^N).  Atom counts of at most 10 (2^10 amplitudes: one tile of
    any candidate) keep the library default."""
    sizes = sorted({r.nqubits for r in registers}, reverse=True)
    regs = _truncated(registers, steps)
    cands = BATCH_CANDIDATES[dtype]
    small = {n: (0, 0) for n in sizes if n <= 10}
    best_t, best = None, None
    table = {}
    for tl, cmin in cands:
        picks = {**{n: (tl, cmin) for n in sizes}, **small}
        t = _time_batch(regs, picks, dtype, device, tolerance, max_krylov_dim)
        table[("uniform", tl, cmin)] = t
        if best_t is None or t < best_t:
            best_t, best = t, picks
    for n in [x for x in sizes if x > 10][:refine]:
        for alt in cands:
            if alt == best[n]:
                continue
            picks = {**best, n: alt}
            t = _time_batch(regs, picks, dtype, device, tolerance, max_krylov_dim)
            table[(n, alt[0], alt[1])] = t
            if t < best_t * (1.0 - gain):
                best_t, best = t, picks
    return {"picks": best, "seconds": best_t, "table": table}


def with_batch_schedules(registers, dtype: torch.dtype, device: torch.device, tolerance: float,
                         max_krylov_dim: int) -> list:
    """Copies of the registers, each without a schedule of its own given the
    tuned one for its atom count (cached per device, dtype and set of sizes)."""
    from dataclasses import replace

    if not _enabled():
        return list(registers)
    idx = device.index if device.index is not None else torch.cuda.current_device()
    key = (idx, dtype, tuple(sorted(r.nqubits for r in registers)))
    res = _batch_cache.get(key)
    if res is None:
        res = tune_batch(registers, dtype, device, tolerance, max_krylov_dim)
        _batch_cache[key] = res
        _batch_timings[key] = res["table"]
    picks = res["picks"]
    return [r if (r.tile_log or r.min_contig) else replace(r, tile_log=picks[r.nqubits][0],
                                                            min_contig=picks[r.nqubits][1])
            for r in registers]


def batch_tuned_table() -> dict:
    """{(device, dtype, sizes): {candidate: seconds}} of every batch tuning so far."""
    return dict(_batch_timings)


def tuned_table() -> dict:
    """{(device, n, dtype, complex or "noise"): {schedule: ms}} of every tuning done so far."""
    return dict(_timings)
