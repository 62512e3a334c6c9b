"""ctypes binding of ``libemu_sv_hip.so`` (the C ABI in include/emu_sv_hip.h).

The library is loaded from ``emulators_amd/_lib``.  When it is missing the
import of this module raises: there is no CPU or pure-torch substitute for
the state-vector kernels.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np
import torch

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libemu_sv_hip.so"

EMU_SV_OK = 0
EMU_SV_EINVAL = -1
EMU_SV_EHIP = -2
EMU_SV_ENOTCONVERGED = -3
EMU_SV_C128 = 0
EMU_SV_C64 = 1
EMU_SV_FLAG_COMPLEX = 1
EMU_SV_FLAG_NOISE = 2
EMU_SV_FLAG_KEEP_DIAG = 4
EMU_SV_KERNEL_DEFAULT = 0
EMU_SV_KERNEL_LDS = 1 << 8
EMU_SV_KERNEL_WAVE = 2 << 8
EMU_SV_KERNEL_TILE = 3 << 8
EMU_SV_KERNEL_BTILE = 4 << 8
KERNELS = {"default": EMU_SV_KERNEL_DEFAULT, "lds": EMU_SV_KERNEL_LDS, "wave": EMU_SV_KERNEL_WAVE,
           "tile": EMU_SV_KERNEL_TILE, "btile": EMU_SV_KERNEL_BTILE}


def kernel_flags(kernel: str, partners: int = 0, first_log: int = 0) -> int:
    """Kernel-kind bits of the flags word.  Tile kernel only: ``partners`` is
    the number of top bits the last pass may read as flip partners and
    ``first_log`` the tile log of the first (contiguous) pass (0: tile_log)."""
    if not 0 <= partners <= 15 or not 0 <= first_log <= 15:
        raise ValueError("partners and first_log must be in [0, 15]")
    return KERNELS[kernel] | (partners << 12) | (first_log << 16)
EMU_SV_MAX_QUBITS = 32
EMU_SV_HAM_DOUBLES = 1136

# Every symbol declared in include/emu_sv_hip.h, with (restype, argtypes).
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "emu_sv_last_error": (ctypes.c_char_p, []),
    "emu_sv_version": (ctypes.c_char_p, []),
    "emu_sv_pack_hamiltonian": (_i, [_i, _dp, _dp, _dp, _dp, _dp]),
    "emu_sv_plan_passes": (_i, [_i, _i, _i, _i, _i]),
    "emu_sv_workspace_bytes": (ctypes.c_size_t, [_i, _i]),
    "emu_sv_hpsi": (_i, [_i, _i, _vp, _d, _d, _vp, _d, _vp, _d, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "emu_sv_hpsi_static": (_i, [_i, _i, _vp, _d, _d, _vp, _d, _vp, _d, _vp, _vp, _i, _i, _i, _vp]),
    "emu_sv_lanczos_update": (_i, [_i, _i, _vp, _vp, _d, _d, _d, _vp, _d, _d, _d, _vp, _vp, _vp]),
    "emu_sv_lanczos_matvec": (_i, [_i, _i, _vp, _d, _d, _vp, _d, _vp, _d, _d, _d, _vp, _d, _d, _d,
                                   _vp, _vp, _vp, _vp, _i, _i, _i, _vp]),
    "emu_sv_lanczos_matvec_dev": (_i, [_i, _i, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                       _i, _i, _i, _vp]),
    "emu_sv_lanczos_iterate_dev": (_i, [_i, _i, _vp, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                        _vp, _vp, _i, _i, _i, _vp]),
    "emu_sv_lanczos_next_coef": (_i, [_vp, _vp, _d, _d, _vp, _vp, _vp, _vp]),
    "emu_sv_lincomb": (_i, [_i, _i, _i, ctypes.POINTER(_vp), _dp, _vp, _vp]),
    "emu_sv_dot": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp]),
    "emu_sv_scale": (_i, [_i, _i, _vp, _d, _vp]),
    "emu_sv_pack_noise": (_i, [_dp, _dp]),
    "emu_sv_arnoldi_orthogonalize": (_i, [_i, _i, _i, ctypes.POINTER(_vp), _dp, _vp, _vp, _vp,
                                          _vp]),
    "emu_sv_arnoldi_orthonormalize": (_i, [_i, _i, _i, ctypes.POINTER(_vp), _dp, _vp, _vp, _vp,
                                           _vp]),
    "emu_sv_krylov_step_records_doubles": (ctypes.c_size_t, [_i]),
    "emu_sv_krylov_step_decomp_doubles": (ctypes.c_size_t, [_i]),
    "emu_sv_krylov_step": (_i, [_i, _i, _vp, _d, _d, _vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _i, _i,
                                _i, _vp, ctypes.POINTER(_i), _vp]),
    "emu_sv_arnoldi_gram": (_i, [_i, _i, _i, ctypes.POINTER(_vp), _vp, _d, _vp, _i, _vp, _vp]),
    "emu_sv_arnoldi_gram_state_doubles": (ctypes.c_size_t, [_i]),
    "emu_sv_arnoldi_exp_records_doubles": (ctypes.c_size_t, [_i]),
    "emu_sv_arnoldi_exp": (_i, [_i, _i, _vp, _d, _d, _vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _i, _i,
                                _i, _vp, ctypes.POINTER(_i)]),
    "emu_sv_apply_1q": (_i, [_i, _i, _vp, _i, _dp, _vp]),
    "emu_sv_apply_pair": (_i, [_i, _i, _vp, _vp, _i, _i, _dp, _i, _vp]),
    "emu_sv_antihermitian": (_i, [_i, _i, _vp, _vp, _d, _d, _vp]),
    "emu_sv_apply_pairs_all": (_i, [_i, _i, _vp, _vp, _dp, _vp]),
    "emu_sv_sample_workspace_bytes": (ctypes.c_size_t, [_i]),
    "emu_sv_sample": (_i, [_i, _i, _vp, ctypes.c_uint64, _i, _vp, _i, _vp, _vp, ctypes.c_size_t,
                           _vp]),
    "emu_sv_convert": (_i, [_i, _i, _vp, _i, _vp, _vp]),
    "emu_sv_correlation": (_i, [_i, _i, _vp, _vp, _vp, _vp]),
    "emu_sv_parameter_traces": (_i, [_i, _i, _i, _vp, _vp, ctypes.c_uint64, _i, _vp, _vp, _vp]),
    "emu_sv_batch_workspace_bytes": (ctypes.c_size_t, [_i, _i, _vp, _i, _i]),
    "emu_sv_batch_create": (_i, [_i, _i, _vp, _i, _d, _i, _i, _vp, ctypes.c_size_t, _vp, _vp]),
    "emu_sv_batch_run": (_i, [_vp, _i, _i, _vp, _vp]),
    "emu_sv_batch_status": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "emu_sv_batch_destroy": (_i, [_vp]),
    "emu_sv_batch_resume": (_i, [_vp, _vp]),
    "emu_sv_batch_publish": (_i, [_vp, _vp]),
    "emu_sv_batch_step_info": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "emu_sv_batch_array_layout": (_i, [_vp]),
    "emu_sv_pack_hamiltonian_steps": (_i, [_i, _i, _dp, _dp, _dp, _dp, _dp]),
    "emu_sv_tridiag_eig": (_i, [_i, _dp, _dp, _dp, _dp]),
    "emu_sv_expm": (_i, [_i, _dp, _dp]),
    "emu_sv_expm_column": (_i, [_i, _dp, _dp]),
    "emu_sv_lanczos_exp": (_i, [_i, _dp, _dp, _d, _d, _d, _i, _dp]),
}


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m emulators_amd.build` "
            "(the emu_sv engine has no CPU fallback)"
        )
    lib = ctypes.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | ctypes.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


LIB = _load()


class NativeError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    """Map a C ABI status to the reference's exception types."""
    if rc == EMU_SV_OK:
        return
    msg = LIB.emu_sv_last_error().decode(errors="replace")
    if rc == EMU_SV_EINVAL:
        raise ValueError(f"{what}: {msg}")
    if rc == EMU_SV_ENOTCONVERGED:
        raise RecursionError(msg)
    raise NativeError(f"{what}: {msg}")


def version() -> str:
    return LIB.emu_sv_version().decode()


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.complex128:
        return EMU_SV_C128
    if dtype == torch.complex64:
        return EMU_SV_C64
    raise ValueError(f"state vectors must be complex64 or complex128, got {dtype}")


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


# torch's current-stream query through the Stream object costs ~5 us per call;
# the raw query returns the same handle (at::cuda::getCurrentCUDAStream).
_RAW_STREAM = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def _stream_handle(device: torch.device) -> int:
    if _RAW_STREAM is not None:
        idx = device.index if device.index is not None else torch.cuda.current_device()
        return int(_RAW_STREAM(idx))
    return int(torch.cuda.current_stream(device).cuda_stream)


def _stream(device: torch.device):
    return ctypes.c_void_p(_stream_handle(device))


def stream_key(device: torch.device) -> tuple[int, int]:
    """(device index, current stream handle): scratch that kernels update
    (reduction tickets, partial sums, result slots) is private to one stream,
    so calls enqueued concurrently on two streams of one device never share it."""
    idx = device.index if device.index is not None else torch.cuda.current_device()
    return idx, _stream_handle(device)


def nqubits_of(vec: torch.Tensor) -> int:
    d = vec.numel()
    n = d.bit_length() - 1
    if d < 2 or (1 << n) != d:
        raise ValueError(f"state length must be a power of two >= 2, got {d}")
    return n


def _require_device_vec(vec: torch.Tensor, what: str) -> None:
    if not vec.is_cuda:
        raise ValueError(f"{what}: the emu_sv engine needs device (GPU) tensors, got {vec.device}")
    if not vec.is_contiguous():
        raise ValueError(f"{what}: tensor must be contiguous")
    if vec.data_ptr() % 16:
        raise ValueError(f"{what}: vectors must be 16-byte aligned (the kernels move 16-byte packets)")


def _require_doubles(t: torch.Tensor, count: int, what: str) -> None:
    """A contiguous float64 device tensor of at least ``count`` elements."""
    if not t.is_cuda or t.dtype != torch.float64 or not t.is_contiguous() or t.numel() < count:
        raise ValueError(f"{what}: needs a contiguous float64 device tensor of >= {count} elements")


def _require_out_doubles(t: torch.Tensor, count: int, what: str) -> None:
    """An output the kernels write: a contiguous float64 device tensor, or one
    in pinned host memory (written there directly by the kernels)."""
    if not (t.is_cuda or t.is_pinned()) or t.dtype != torch.float64 or not t.is_contiguous() \
            or t.numel() < count:
        raise ValueError(f"{what}: needs a contiguous float64 device or pinned host tensor of "
                         f">= {count} elements")


def _require_iteration(n: int, ham_dev: torch.Tensor, w: torch.Tensor, pa: torch.Tensor | None,
                       pb: torch.Tensor, r: torch.Tensor, coef: torch.Tensor, rec: torch.Tensor,
                       what: str) -> None:
    """Shapes of one device-coefficient Lanczos iteration, checked before any launch."""
    _require_doubles(ham_dev, EMU_SV_HAM_DOUBLES, what + " (parameter block)")
    for t in (w, pb, r) + ((pa,) if pa is not None else ()):
        _require_device_vec(t, what)
        if t.dtype != w.dtype or t.numel() != (1 << n):
            raise ValueError(f"{what}: vectors must be 2**n elements of one dtype")
    _require_doubles(coef, 6, what + " (coefficients)")
    _require_out_doubles(rec, 12, what + " (reductions)")


def pack_hamiltonian(omega, delta, phi, interaction) -> tuple[np.ndarray, bool]:
    """Host packing of the parameter block (real float64 per qubit)."""
    om = np.ascontiguousarray(np.asarray(omega, dtype=np.float64))
    de = np.ascontiguousarray(np.asarray(delta, dtype=np.float64))
    ph = np.ascontiguousarray(np.asarray(phi, dtype=np.float64))
    um = np.ascontiguousarray(np.asarray(interaction, dtype=np.float64))
    n = om.shape[0]
    if de.shape != (n,) or ph.shape != (n,) or um.shape != (n, n):
        raise ValueError("pack_hamiltonian: inconsistent parameter shapes")
    block = np.zeros(EMU_SV_HAM_DOUBLES, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    check(LIB.emu_sv_pack_hamiltonian(n, om.ctypes.data_as(dp), de.ctypes.data_as(dp),
                                      ph.ctypes.data_as(dp), um.ctypes.data_as(dp),
                                      block.ctypes.data_as(dp)), "pack_hamiltonian")
    return block, bool(np.any(ph != 0.0))


def _complex_2x2(a) -> np.ndarray:
    m = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    if m.shape != (2, 2):
        raise ValueError("expected a 2x2 complex matrix")
    return np.ascontiguousarray(m.reshape(4)).view(np.float64)


def pack_noise(block: np.ndarray, noise) -> np.ndarray:
    """Add the single-qubit term K = -i/2 sum L^+L (2x2 complex) to a packed
    block (host, in place)."""
    k = _complex_2x2(noise)
    dp = ctypes.POINTER(ctypes.c_double)
    check(LIB.emu_sv_pack_noise(k.ctypes.data_as(dp), block.ctypes.data_as(dp)), "pack_noise")
    return block


def plan_passes(n: int, dtype: torch.dtype, tile_log: int = 0, min_contig: int = 0,
                kernel: str = "default", partners: int = 0, first_log: int = 0) -> int:
    rc = LIB.emu_sv_plan_passes(n, dtype_code(dtype), tile_log, min_contig,
                                kernel_flags(kernel, partners, first_log))
    if rc < 0:
        check(rc, "plan_passes")
    return rc


class Workspace:
    """Scratch of one (device, stream): partial sums, the reduction ticket,
    the diagonal table and small result slots.  Every stream that launches
    the kernels gets its own, so work enqueued concurrently on two streams of
    one device cannot race on the ticket or the partial rows."""

    _cache: dict = {}

    def __init__(self, device: torch.device, n: int):
        nbytes = LIB.emu_sv_workspace_bytes(n, 0)
        # zero-filled: holds the reduction ticket counter, which the library re-arms
        self.scratch = torch.zeros(nbytes, dtype=torch.uint8, device=device)
        self.results = torch.zeros(64, dtype=torch.float64, device=device)
        # the fill is complete before the first kernel of any stream reads it
        torch.cuda.current_stream(device).synchronize()
        self.n = n

    @classmethod
    def get(cls, device: torch.device, n: int) -> "Workspace":
        key = stream_key(device)
        ws = cls._cache.get(key)
        if ws is None or ws.n < n:
            ws = Workspace(device, max(n, 1))
            cls._cache[key] = ws
        return ws


def hpsi(n: int, ham_dev: torch.Tensor, sigma: complex, v: torch.Tensor, v_scale: float,
         w: torch.Tensor, prev: torch.Tensor | None = None, prev_scale: float = 1.0,
         dots: bool = False, complex_coef: bool = True, tile_log: int = 0,
         min_contig: int = 0, kernel: str = "default", partners: int = 0,
         first_log: int = 0, noise: bool = False, keep_diag: bool = False) -> torch.Tensor | None:
    """w = sigma * H (v_scale v); optionally returns the 7 fused dots (device).
    ``noise``: the block carries the non-Hermitian term (LDS or btile kernel).
    ``keep_diag``: reuse the workspace's diagonal table when it was built for
    this block, n and first-pass tile (EMU_SV_FLAG_KEEP_DIAG)."""
    _require_device_vec(v, "hpsi")
    _require_device_vec(w, "hpsi")
    if v.dtype != w.dtype or v.numel() != w.numel() or v.numel() != (1 << n):
        raise ValueError("hpsi: v and w must be 2**n-element vectors of one dtype")
    if prev is not None and (prev.dtype != v.dtype or prev.numel() != v.numel()):
        raise ValueError("hpsi: prev must match v")
    _require_doubles(ham_dev, EMU_SV_HAM_DOUBLES, "hpsi (parameter block)")
    ws = Workspace.get(v.device, n)
    out = ws.results[0:7] if dots else None
    flags = (EMU_SV_FLAG_COMPLEX if complex_coef else 0) | kernel_flags(kernel, partners, first_log)
    if noise:
        flags |= EMU_SV_FLAG_NOISE
    if keep_diag:
        flags |= EMU_SV_FLAG_KEEP_DIAG
    check(LIB.emu_sv_hpsi(n, dtype_code(v.dtype), _ptr(ham_dev), float(sigma.real),
                          float(sigma.imag), _ptr(v), float(v_scale), _ptr(prev),
                          float(prev_scale), _ptr(w), _ptr(out), _ptr(ws.scratch), tile_log,
                          min_contig, flags, _stream(v.device)), "hpsi")
    return out


def lanczos_matvec(n: int, ham_dev: torch.Tensor, sigma: complex, w: torch.Tensor, w_scale: float,
                   pa: torch.Tensor | None, pa_scale: float, a: complex, pb: torch.Tensor,
                   pb_scale: float, b: complex, r: torch.Tensor, complex_coef: bool = True,
                   tile_log: int = 0, min_contig: int = 0, kernel: str = "default",
                   partners: int = 0, first_log: int = 0, keep_diag: bool = False) -> torch.Tensor:
    """z = w_scale w - a pa_scale pa - b pb_scale pb (into w); r = sigma H z.
    Returns a device slice [||z||^2, Re/Im <z|pb_s>, Re/Im <z|w_s>, ||r||^2,
    Re/Im <z|r>] (pb_s = pb_scale pb, w_s = w_scale w).  ``keep_diag``: as
    for hpsi."""
    for t in (w, pb, r) + ((pa,) if pa is not None else ()):
        _require_device_vec(t, "lanczos_matvec")
        if t.dtype != w.dtype or t.numel() != (1 << n):
            raise ValueError("lanczos_matvec: vectors must be 2**n elements of one dtype")
    _require_doubles(ham_dev, EMU_SV_HAM_DOUBLES, "lanczos_matvec (parameter block)")
    ws = Workspace.get(w.device, n)
    out = ws.results[24:36]
    flags = (EMU_SV_FLAG_COMPLEX if complex_coef else 0) | kernel_flags(kernel, partners, first_log)
    if keep_diag:
        flags |= EMU_SV_FLAG_KEEP_DIAG
    check(LIB.emu_sv_lanczos_matvec(n, dtype_code(w.dtype), _ptr(ham_dev), float(sigma.real),
                                    float(sigma.imag), _ptr(w), float(w_scale), _ptr(pa),
                                    float(pa_scale), float(a.real), float(a.imag), _ptr(pb),
                                    float(pb_scale), float(b.real), float(b.imag), _ptr(r),
                                    _ptr(out[0:5]), _ptr(out[5:12]), _ptr(ws.scratch), tile_log,
                                    min_contig, flags, _stream(w.device)), "lanczos_matvec")
    return out


def lanczos_matvec_dev(n: int, ham_dev: torch.Tensor, sigma: complex, w: torch.Tensor,
                       pa: torch.Tensor | None, pb: torch.Tensor, coef: torch.Tensor, r: torch.Tensor,
                       rec: torch.Tensor, complex_coef: bool = True, tile_log: int = 0,
                       min_contig: int = 0, partners: int = 0, first_log: int = 0,
                       kernel: str = "tile", keep_diag: bool = False) -> None:
    """lanczos_matvec with device coefficients coef (6 doubles: a*pa_scale,
    b*pb_scale, pb_scale, w_scale) for the tile or LDS kernel; rec (device,
    >= 12 doubles) receives [||z||^2, <z|pb_s>, <z|w_s>, ||r||^2, <z|r>, ...].
    ``keep_diag``: as for hpsi."""
    _require_iteration(n, ham_dev, w, pa, pb, r, coef, rec, "lanczos_matvec_dev")
    flags = (EMU_SV_FLAG_COMPLEX if complex_coef else 0) | kernel_flags(kernel, partners, first_log)
    if keep_diag:
        flags |= EMU_SV_FLAG_KEEP_DIAG
    check(LIB.emu_sv_lanczos_matvec_dev(n, dtype_code(w.dtype), _ptr(ham_dev), float(sigma.real),
                                        float(sigma.imag), _ptr(w), _ptr(pa), _ptr(pb), _ptr(coef),
                                        _ptr(r), _ptr(rec[0:5]), _ptr(rec[5:12]), _ptr(
                                            Workspace.get(w.device, n).scratch), tile_log,
                                        min_contig, flags, _stream(w.device)), "lanczos_matvec_dev")


def lanczos_iterate_dev(n: int, ham_dev: torch.Tensor, sigma: complex, w: torch.Tensor,
                        pa: torch.Tensor | None, pb: torch.Tensor, coef: torch.Tensor,
                        r: torch.Tensor, rec: torch.Tensor, nxt: torch.Tensor, tout: torch.Tensor,
                        complex_coef: bool = True, tile_log: int = 0, min_contig: int = 0,
                        partners: int = 0, first_log: int = 0, kernel: str = "tile",
                        keep_diag: bool = False) -> None:
    """lanczos_matvec_dev followed by lanczos_next_coef in the same launch
    sequence: the last reduction also writes the next iteration's
    coefficients (nxt, 6 doubles) and tout (n2, b, a, ww)."""
    _require_iteration(n, ham_dev, w, pa, pb, r, coef, rec, "lanczos_iterate_dev")
    _require_doubles(nxt, 6, "lanczos_iterate_dev (next coefficients)")
    _require_out_doubles(tout, 6, "lanczos_iterate_dev (side outputs)")
    flags = (EMU_SV_FLAG_COMPLEX if complex_coef else 0) | kernel_flags(kernel, partners, first_log)
    if keep_diag:
        flags |= EMU_SV_FLAG_KEEP_DIAG
    check(LIB.emu_sv_lanczos_iterate_dev(n, dtype_code(w.dtype), _ptr(ham_dev), float(sigma.real),
                                         float(sigma.imag), _ptr(w), _ptr(pa), _ptr(pb), _ptr(coef),
                                         _ptr(r), _ptr(rec[0:5]), _ptr(rec[5:12]), _ptr(nxt),
                                         _ptr(tout), _ptr(Workspace.get(w.device, n).scratch),
                                         tile_log, min_contig, flags, _stream(w.device)),
          "lanczos_iterate_dev")


def lanczos_next_coef(rec: torch.Tensor, sigma: complex, prev: torch.Tensor, nxt: torch.Tensor,
                      tout: torch.Tensor) -> None:
    """Device coefficients of the next iteration from rec (see lanczos_matvec_dev)."""
    check(LIB.emu_sv_lanczos_next_coef(_ptr(rec[0:5]), _ptr(rec[5:8]), float(sigma.real),
                                       float(sigma.imag), _ptr(prev), _ptr(nxt), _ptr(tout),
                                       _stream(rec.device)), "lanczos_next_coef")


def lanczos_update(n: int, w: torch.Tensor, pa: torch.Tensor | None, pa_scale: float, a: complex,
                   pb: torch.Tensor | None, pb_scale: float, b: complex) -> torch.Tensor:
    """w -= a pa_scale pa + b pb_scale pb; returns a device slot holding ||w||^2."""
    _require_device_vec(w, "lanczos_update")
    ws = Workspace.get(w.device, n)
    out = ws.results[8:9]
    check(LIB.emu_sv_lanczos_update(n, dtype_code(w.dtype), _ptr(w), _ptr(pa), float(pa_scale),
                                    float(a.real), float(a.imag), _ptr(pb), float(pb_scale),
                                    float(b.real), float(b.imag), _ptr(out), _ptr(ws.scratch),
                                    _stream(w.device)), "lanczos_update")
    return out


def lincomb(n: int, vecs: list[torch.Tensor], coef: np.ndarray, out: torch.Tensor) -> torch.Tensor:
    _require_device_vec(out, "lincomb")
    count = len(vecs)
    if count != len(coef):
        raise ValueError("lincomb: one coefficient per vector")
    for v in vecs:
        _require_device_vec(v, "lincomb")
        if v.dtype != out.dtype or v.numel() != out.numel() or v.numel() != (1 << n):
            raise ValueError("lincomb: vectors must match the output (2**n elements, one dtype)")
    arr = (ctypes.c_void_p * count)(*[v.data_ptr() for v in vecs])
    c = np.ascontiguousarray(np.asarray(coef, dtype=np.complex128)).view(np.float64)
    check(LIB.emu_sv_lincomb(n, dtype_code(out.dtype), count, arr,
                             c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _ptr(out),
                             _stream(out.device)), "lincomb")
    return out


def dot(n: int, a: torch.Tensor, b: torch.Tensor | None = None) -> torch.Tensor:
    """Device slot [||a||^2, Re<a|b>, Im<a|b>]."""
    _require_device_vec(a, "dot")
    ws = Workspace.get(a.device, n)
    out = ws.results[16:19]
    check(LIB.emu_sv_dot(n, dtype_code(a.dtype), _ptr(a), _ptr(b), _ptr(out), _ptr(ws.scratch),
                         _stream(a.device)), "dot")
    return out


def arnoldi_orthogonalize(n: int, basis: list[torch.Tensor], scales: list[float],
                          w: torch.Tensor, normalize: bool = False) -> torch.Tensor:
    """Modified Gram-Schmidt of w against scales[k] * basis[k] on the device.
    Returns a device (count+1, 4) float64 tensor: row k < count =
    [||w||^2 before subtraction k, Re, Im <basis_k|w>] (unscaled), row count =
    [||w||^2 after the last subtraction].  ``normalize``: w is then divided
    by that norm on the device (the next orthonormal basis vector)."""
    _require_device_vec(w, "arnoldi_orthogonalize")
    count = len(basis)
    if count != len(scales) or count < 1:
        raise ValueError("arnoldi_orthogonalize: one scale per basis vector")
    size, dt = 1 << n, w.dtype
    if w.numel() != size:
        raise ValueError("arnoldi_orthogonalize: w must have 2**n elements")
    for v in basis:
        if not (v.is_cuda and v.dtype == dt and v.numel() == size and v.is_contiguous()):
            raise ValueError("arnoldi_orthogonalize: basis vectors must be contiguous device "
                             "tensors matching w")
    ws = Workspace.get(w.device, n)
    rec = torch.empty(count + 1, 4, dtype=torch.float64, device=w.device)
    arr = (ctypes.c_void_p * count)(*[v.data_ptr() for v in basis])
    sc = np.ascontiguousarray(np.asarray(scales, dtype=np.float64))
    fn = LIB.emu_sv_arnoldi_orthonormalize if normalize else LIB.emu_sv_arnoldi_orthogonalize
    check(fn(n, dtype_code(w.dtype), count, arr, sc.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
             _ptr(w), _ptr(rec), _ptr(ws.scratch), _stream(w.device)), "arnoldi_orthogonalize")
    return rec


def arnoldi_gram_state(ld: int, device) -> torch.Tensor:
    """A zeroed device state buffer for emu_sv_arnoldi_gram with leading
    dimension ``ld`` (more than the largest basis count it will serve)."""
    size = int(LIB.emu_sv_arnoldi_gram_state_doubles(ld))
    if size <= 0:
        raise ValueError("arnoldi_gram: ld must be at least 2")
    return torch.zeros(size, dtype=torch.float64, device=device)


def arnoldi_gram(n: int, basis: list[torch.Tensor], w: torch.Tensor, v0_scale: float,
                 state: torch.Tensor, ld: int, records: torch.Tensor | None = None) -> torch.Tensor:
    """Gram-Schmidt of w = op(u_j) (j = len(basis) - 1, u_j unscaled) against
    the unnormalised basis a_k = s_k u_k (s_0 = ``v0_scale``, later scales
    on the device in ``state``) by the Gram-matrix form of modified
    Gram-Schmidt; w becomes u_{j+1} = s_j w - sum_k h_k a_k, unnormalised.
    Returns the device (count+1, 4) records: row k < count = [., Re h_k,
    Im h_k, 0] (T[k, j]), row 0 [0] = ||s_j w||^2, row count [0] = ||u_{j+1}||^2
    (s_{j+1} = 1 / its root is stored in ``state``)."""
    _require_device_vec(w, "arnoldi_gram")
    count = len(basis)
    size, dt = 1 << n, w.dtype
    if count < 1 or w.numel() != size:
        raise ValueError("arnoldi_gram: w must have 2**n elements and the basis one vector at least")
    for v in basis:
        if not (v.is_cuda and v.dtype == dt and v.numel() == size and v.is_contiguous()):
            raise ValueError("arnoldi_gram: basis vectors must be contiguous device tensors matching w")
    if state.dtype != torch.float64 or state.device != w.device or \
            state.numel() < int(LIB.emu_sv_arnoldi_gram_state_doubles(ld)):
        raise ValueError("arnoldi_gram: state must be a float64 device buffer of the state size for ld")
    if records is None:
        records = torch.empty(count + 1, 4, dtype=torch.float64, device=w.device)
    arr = (ctypes.c_void_p * count)(*[v.data_ptr() for v in basis])
    check(LIB.emu_sv_arnoldi_gram(n, dtype_code(dt), count, arr, _ptr(w), float(v0_scale), _ptr(state),
                                  ld, _ptr(records), _stream(w.device)), "arnoldi_gram")
    return records


class _ArnoldiExpBuffers:
    """Per-(device, stream, max_dim) device records of emu_sv_arnoldi_exp and
    their pinned host mirror (the basis is allocated per call)."""

    _cache: dict = {}

    def __init__(self, device, max_dim: int):
        rows = int(LIB.emu_sv_arnoldi_exp_records_doubles(max_dim))
        if rows <= 0:
            raise ValueError("arnoldi_exp: max_dim must be in [1, 256]")
        self.records = torch.zeros(rows, dtype=torch.float64, device=device)
        self.host = torch.zeros(rows, dtype=torch.float64, pin_memory=True)

    @classmethod
    def get(cls, device, max_dim: int) -> "_ArnoldiExpBuffers":
        key = (stream_key(device), max_dim)
        b = cls._cache.get(key)
        if b is None:
            b = cls._cache[key] = cls(device, max_dim)
        return b


def arnoldi_exp(n: int, ham_dev: torch.Tensor, sigma: complex, v: torch.Tensor, out: torch.Tensor,
                max_dim: int, exp_tolerance: float, norm_tolerance: float, complex_coef: bool = True,
                noise: bool = False, tile_log: int = 0, min_contig: int = 0, kernel: str = "default",
                partners: int = 0, first_log: int = 0) -> tuple[int, bool, bool, int]:
    """out = exp(sigma H) v by the library's pipelined Arnoldi driver; returns
    (iterations, converged, happy breakdown, matvecs)."""
    _require_device_vec(v, "arnoldi_exp")
    _require_device_vec(out, "arnoldi_exp")
    if v.numel() != 1 << n or out.numel() != 1 << n or out.dtype != v.dtype:
        raise ValueError("arnoldi_exp: v and out must be 2**n elements of one dtype")
    bufs = _ArnoldiExpBuffers.get(v.device, max_dim)
    basis = torch.empty(max_dim << n, dtype=v.dtype, device=v.device)
    ws = Workspace.get(v.device, n)
    flags = (EMU_SV_FLAG_COMPLEX if complex_coef else 0) | kernel_flags(kernel, partners, first_log)
    if noise:
        flags |= EMU_SV_FLAG_NOISE
    res = (ctypes.c_int * 4)()
    check(LIB.emu_sv_arnoldi_exp(n, dtype_code(v.dtype), _ptr(ham_dev), float(sigma.real), float(sigma.imag),
                                 _ptr(v), _ptr(out), _ptr(basis), max_dim, float(exp_tolerance),
                                 float(norm_tolerance), _ptr(bufs.records), _ptr(bufs.host),
                                 _ptr(ws.scratch), tile_log, min_contig, flags, _stream(v.device), res),
          "arnoldi_exp")
    return int(res[0]), bool(res[1]), bool(res[2]), int(res[3])


class _KrylovStepBuffers:
    """Per-(device, stream, max_dim) device records and their pinned host
    mirror for emu_sv_krylov_step (the basis is allocated per call)."""

    _cache: dict = {}

    def __init__(self, device, max_dim: int):
        size = int(LIB.emu_sv_krylov_step_records_doubles(max_dim))
        if size <= 0:
            raise ValueError("krylov_step: max_dim must be in [1, 256]")
        self.records = torch.zeros(size, dtype=torch.float64, device=device)
        self.host = torch.zeros(size, dtype=torch.float64, pin_memory=True)

    @classmethod
    def get(cls, device, max_dim: int) -> "_KrylovStepBuffers":
        key = (stream_key(device), max_dim)
        b = cls._cache.get(key)
        if b is None:
            b = cls._cache[key] = cls(device, max_dim)
        return b


def krylov_step(n: int, ham_dev: torch.Tensor, sigma: complex, v: torch.Tensor, out: torch.Tensor,
                max_dim: int, exp_tolerance: float, norm_tolerance: float, complex_coef: bool = True,
                tile_log: int = 0, min_contig: int = 0, kernel: str = "lds", partners: int = 0,
                first_log: int = 0, basis: torch.Tensor | None = None, decomp: bool = False):
    """out = exp(sigma H) v by the library's pipelined Lanczos driver
    (emu_sv_krylov_step); returns (iterations, converged, happy breakdown,
    matvecs, decomposition or None).  ``basis``: a contiguous device buffer of
    at least (max_dim + 1) * 2**n elements that receives basis rows 1..
    (default: one from torch's allocator for the call).  ``decomp``: also
    return the Lanczos decomposition (numpy, the layout of
    emu_sv_krylov_step_decomp_doubles)."""
    _require_device_vec(v, "krylov_step")
    _require_device_vec(out, "krylov_step")
    if v.numel() != 1 << n or out.numel() != 1 << n or out.dtype != v.dtype:
        raise ValueError("krylov_step: v and out must be 2**n elements of one dtype")
    _require_doubles(ham_dev, EMU_SV_HAM_DOUBLES, "krylov_step (parameter block)")
    bufs = _KrylovStepBuffers.get(v.device, max_dim)
    if basis is None:
        basis = torch.empty((max_dim + 1) << n, dtype=v.dtype, device=v.device)
    elif not (basis.is_cuda and basis.device == v.device and basis.dtype == v.dtype
              and basis.is_contiguous() and basis.numel() >= (max_dim + 1) << n):
        raise ValueError("krylov_step: basis must be a contiguous device buffer of "
                         "(max_dim + 1) * 2**n elements of v's dtype")
    ws = Workspace.get(v.device, n)
    flags = (EMU_SV_FLAG_COMPLEX if complex_coef else 0) | kernel_flags(kernel, partners, first_log)
    res = (ctypes.c_int * 4)()
    dec = np.zeros(int(LIB.emu_sv_krylov_step_decomp_doubles(max_dim))) if decomp else None
    check(LIB.emu_sv_krylov_step(n, dtype_code(v.dtype), _ptr(ham_dev), float(sigma.real),
                                 float(sigma.imag), _ptr(v), _ptr(out), _ptr(basis), max_dim,
                                 float(exp_tolerance), float(norm_tolerance), _ptr(bufs.records),
                                 _ptr(bufs.host), _ptr(ws.scratch), tile_log, min_contig, flags,
                                 _stream(v.device), res,
                                 dec.ctypes.data_as(ctypes.c_void_p) if dec is not None else None),
          "krylov_step")
    return int(res[0]), bool(res[1]), bool(res[2]), int(res[3]), dec


def apply_1q(n: int, psi: torch.Tensor, qubit: int, op) -> torch.Tensor:
    """psi <- op acting on qubit ``qubit`` (bit n-1-qubit), in place."""
    _require_device_vec(psi, "apply_1q")
    if not 0 <= qubit < n:
        raise ValueError("apply_1q: qubit out of range")
    a = _complex_2x2(op.detach().cpu().numpy() if isinstance(op, torch.Tensor) else op)
    check(LIB.emu_sv_apply_1q(n, dtype_code(psi.dtype), _ptr(psi), n - 1 - qubit,
                              a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              _stream(psi.device)), "apply_1q")
    return psi


def apply_pair(n: int, src: torch.Tensor, dst: torch.Tensor, bit_a: int, bit_b: int, op,
               accumulate: bool = True) -> torch.Tensor:
    """dst (+)= S on the bit pair (bit_a, bit_b), S 4x4 complex over the local
    index 2*x_a + x_b."""
    _require_device_vec(src, "apply_pair")
    _require_device_vec(dst, "apply_pair")
    if src.dtype != dst.dtype or src.numel() != dst.numel() or src.numel() != (1 << n):
        raise ValueError("apply_pair: src and dst must be 2**n elements of one dtype")
    m = np.ascontiguousarray(np.asarray(op, dtype=np.complex128))
    if m.shape != (4, 4):
        raise ValueError("apply_pair: the operator must be 4x4")
    a = np.ascontiguousarray(m.reshape(16)).view(np.float64)
    check(LIB.emu_sv_apply_pair(n, dtype_code(src.dtype), _ptr(src), _ptr(dst), bit_a, bit_b,
                                a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                int(accumulate), _stream(src.device)), "apply_pair")
    return dst


def apply_pairs_all(n_side: int, src: torch.Tensor, dst: torch.Tensor, op: np.ndarray) -> None:
    """dst += the 4x4 op on every (row bit, column bit) pair of an
    n_side-qubit density matrix (flat 4^n_side device vectors, n_side <= 8);
    ``op`` is a contiguous complex128 (4, 4) array."""
    a = np.ascontiguousarray(op.reshape(16)).view(np.float64)
    check(LIB.emu_sv_apply_pairs_all(n_side, dtype_code(src.dtype), _ptr(src), _ptr(dst),
                                      a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      _stream(src.device)), "apply_pairs_all")


def antihermitian(m: torch.Tensor, out: torch.Tensor, s: complex = 1.0) -> torch.Tensor:
    """out = s (M - M^H) for a square 2^k x 2^k device matrix (out may be M)."""
    _require_device_vec(m, "antihermitian")
    _require_device_vec(out, "antihermitian")
    side = m.shape[0]
    k = side.bit_length() - 1
    if m.dim() != 2 or m.shape[1] != side or (1 << k) != side or out.shape != m.shape or \
            out.dtype != m.dtype:
        raise ValueError("antihermitian: M and out must be square 2^k x 2^k of one dtype")
    check(LIB.emu_sv_antihermitian(k, dtype_code(m.dtype), _ptr(m), _ptr(out), float(s.real),
                                   float(s.imag), _stream(m.device)), "antihermitian")
    return out


def sample(n: int, src: torch.Tensor, uniforms: torch.Tensor, density_matrix: bool = False
           ) -> torch.Tensor:
    """Basis indices drawn by inverse CDF from |psi|^2 (state vector) or
    |diag(rho)| (density matrix), one per uniform draw in [0, 1) (device
    float64).  Returns a device int64 tensor."""
    _require_device_vec(src, "sample")
    side = 1 << n
    if density_matrix:
        if src.numel() != side * side:
            raise ValueError("sample: the density matrix must be 2**n x 2**n")
        stride, mode = side + 1, 1
    else:
        if src.numel() != side:
            raise ValueError("sample: the state must have 2**n elements")
        stride, mode = 1, 0
    u = uniforms.to(device=src.device, dtype=torch.float64).contiguous()
    out = torch.empty(u.numel(), dtype=torch.int64, device=src.device)
    nbytes = LIB.emu_sv_sample_workspace_bytes(n)
    ws = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=src.device)
    check(LIB.emu_sv_sample(n, dtype_code(src.dtype), _ptr(src), ctypes.c_uint64(stride), mode,
                            _ptr(u), u.numel(), _ptr(out), _ptr(ws), ctypes.c_size_t(ws.numel()),
                            _stream(src.device)), "sample")
    return out


def scale(n: int, v: torch.Tensor, s: float) -> None:
    _require_device_vec(v, "scale")
    check(LIB.emu_sv_scale(n, dtype_code(v.dtype), _ptr(v), float(s), _stream(v.device)), "scale")


def convert(n: int, src: torch.Tensor, dst: torch.Tensor) -> torch.Tensor:
    _require_device_vec(src, "convert")
    _require_device_vec(dst, "convert")
    check(LIB.emu_sv_convert(n, dtype_code(src.dtype), _ptr(src), dtype_code(dst.dtype), _ptr(dst),
                             _stream(src.device)), "convert")
    return dst


def correlation(n: int, psi: torch.Tensor) -> torch.Tensor:
    """n x n float64 device tensor <n_i n_j> (qubit order)."""
    _require_device_vec(psi, "correlation")
    ws = Workspace.get(psi.device, n)
    out = torch.empty(n, n, dtype=torch.float64, device=psi.device)
    check(LIB.emu_sv_correlation(n, dtype_code(psi.dtype), _ptr(psi), _ptr(out), _ptr(ws.scratch),
                                 _stream(psi.device)), "correlation")
    return out


class Traces:
    """Per-qubit sums of <W|s+_q|Z>, <W|s-_q|Z>, <W|n_q|Z> and <W|n_q n_q'|Z>
    (complex128 numpy, QUBIT order; ``pair`` is upper-triangular or None)."""

    def __init__(self, up, dn, occ, pair):
        self.up, self.dn, self.occ, self.pair = up, dn, occ, pair


def parameter_traces(n: int, w: torch.Tensor, z: torch.Tensor, pairs: bool = False) -> Traces:
    """Traces of sum_l <W_l| . |Z_l> for rows W_l, Z_l of two (count, 2**n)
    (or (2**n,)) device tensors of one complex dtype (emu_sv_parameter_traces)."""
    w2 = w.reshape(-1, w.shape[-1]) if w.dim() > 1 else w.reshape(1, -1)
    z2 = z.reshape(-1, z.shape[-1]) if z.dim() > 1 else z.reshape(1, -1)
    if w2.shape != z2.shape or w2.dtype != z2.dtype or w2.shape[1] != (1 << n):
        raise ValueError("parameter_traces: W and Z must be (count, 2**n) of one dtype")
    if not (w2.is_cuda and z2.is_cuda):
        raise ValueError("parameter_traces: the emu_sv engine needs device (GPU) tensors")
    w2 = w2.contiguous()
    z2 = z2.contiguous()
    ws = Workspace.get(w2.device, n)
    nout = 6 * n + (2 * n * n if pairs else 0)
    out = torch.zeros(nout, dtype=torch.float64, device=w2.device)
    check(LIB.emu_sv_parameter_traces(n, dtype_code(w2.dtype), w2.shape[0], _ptr(w2), _ptr(z2),
                                      ctypes.c_uint64(1 << n), int(pairs), _ptr(out),
                                      _ptr(ws.scratch), _stream(w2.device)), "parameter_traces")
    o = out.cpu().numpy()
    c = o[: 6 * n].view(np.complex128)
    rev = slice(None, None, -1)  # bit position b <-> qubit n-1-b
    up, dn, occ = c[:n][rev].copy(), c[n: 2 * n][rev].copy(), c[2 * n: 3 * n][rev].copy()
    pair = None
    if pairs:
        pb = o[6 * n:].view(np.complex128).reshape(n, n)  # [b][b'] for b < b'
        pair = pb[rev][:, rev].T.copy()  # qubit (i, j) = bits (n-1-i, n-1-j): i < j  <=>  b > b'
    return Traces(up, dn, occ, pair)


def expm_column(a: np.ndarray) -> np.ndarray:
    """expm(a)[:, 0] of a small host complex128 matrix (Taylor substeps)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    m = a.shape[0]
    if a.shape != (m, m):
        raise ValueError("expm_column: square matrix expected")
    out = np.empty(m, dtype=np.complex128)
    dp = ctypes.POINTER(ctypes.c_double)
    check(LIB.emu_sv_expm_column(m, a.view(np.float64).ctypes.data_as(dp),
                                 out.view(np.float64).ctypes.data_as(dp)), "expm_column")
    return out


def expm(a: np.ndarray) -> np.ndarray:
    """Host complex128 matrix exponential (Pade 13, scaling and squaring)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    m = a.shape[0]
    if a.shape != (m, m):
        raise ValueError("expm: square matrix expected")
    e = np.empty_like(a)
    dp = ctypes.POINTER(ctypes.c_double)
    check(LIB.emu_sv_expm(m, a.view(np.float64).ctypes.data_as(dp),
                          e.view(np.float64).ctypes.data_as(dp)), "expm")
    return e


def tridiag_eig(d: np.ndarray, e: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(lambda, Z) of the real symmetric tridiagonal (d, e): J = Z diag(lambda) Z^T."""
    d = np.ascontiguousarray(np.asarray(d, dtype=np.float64))
    m = d.shape[0]
    e = np.ascontiguousarray(np.asarray(e, dtype=np.float64)) if m > 1 else np.zeros(1)
    lam = np.empty(m)
    z = np.empty((m, m))
    dp = ctypes.POINTER(ctypes.c_double)
    check(LIB.emu_sv_tridiag_eig(m, d.ctypes.data_as(dp), e.ctypes.data_as(dp), lam.ctypes.data_as(dp),
                                 z.ctypes.data_as(dp)), "tridiag_eig")
    return lam, z


def lanczos_exp(T: np.ndarray, m: int, sigma: complex, n2: float, want_coeffs: bool):
    """(|expd[m,0]|, |expd[m+1,0]|, coeffs or None) for the leading m x m block
    of a Hermitian Lanczos matrix T (see emu_sv_lanczos_exp)."""
    dp = ctypes.POINTER(ctypes.c_double)
    diag = np.ascontiguousarray(np.diagonal(T)[:m]).astype(np.complex128)
    sub = np.ascontiguousarray(np.real(np.diagonal(T, -1)[: max(m - 1, 0)])).astype(np.float64)
    if sub.size == 0:
        sub = np.zeros(1)
    out = np.empty(2 + 2 * (m + 1) if want_coeffs else 2, dtype=np.float64)
    check(LIB.emu_sv_lanczos_exp(m, diag.view(np.float64).ctypes.data_as(dp), sub.ctypes.data_as(dp),
                                 float(sigma.real), float(sigma.imag), float(n2), int(want_coeffs),
                                 out.ctypes.data_as(dp)), "lanczos_exp")
    coeffs = out[2:].view(np.complex128) if want_coeffs else None
    return out[0], out[1], coeffs
