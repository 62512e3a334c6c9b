"""The Lindblad superoperator on a density matrix, on the device kernels.

Interface of the reference ``RydbergLindbladian`` (emu_sv/lindblad_operator.py:11-227):
``L @ rho`` returns 1/2 (res - res^H) with
res = H_eff rho - (H_eff rho)^H + i sum_k A_k rho A_k^H and
H_eff = H - i/2 sum_k A_k^H A_k (the -i of the evolution is applied by the
caller), ``h_eff(rho)`` and ``expect(state)`` = Tr(H rho).

rho (2^N x 2^N, row-major) is handled as a 2N-bit vector whose high N bits
are the row index.  With P = H_eff rho and Q = sum_k A_k rho A_k^H,
1/2 (res - res^H) = M - M^H for M = P + (i/2) Q, so one application is:

* H rho: the H.psi kernel on the 2N-bit register with the Rydberg
  parameters on the row qubits only (zero on the column qubits);
* for every qubit q, S = K (x) I + (i/2) sum_k A_k (x) conj(A_k) on its
  (row bit, column bit) pair, K = -i/2 sum_k A_k^H A_k (emu_sv_apply_pair);
* M - M^H with the evolution's scale folded in (emu_sv_antihermitian).
"""

from __future__ import annotations

import numpy as np
import torch

from emulators_amd import _native as native
from emulators_amd import autotune
from emulators_amd.emu_base.jump_lindblad_operators import compute_noise_from_lindbladians

dtype = torch.complex128
_SCHED = autotune.Schedule("lds", 0, 0)


def _cpu_f64(x) -> torch.Tensor:
    t = torch.as_tensor(x).detach()
    if t.is_complex():
        t = t.real
    return t.to(device="cpu", dtype=torch.float64)


class RydbergLindbladian:
    def __init__(self, omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
                 pulser_lindblads: list, interaction_matrix: torch.Tensor,
                 device: torch.device | str):
        self.nqubits: int = len(omegas)
        self.omegas = omegas / 2.0
        self.deltas = deltas
        self.phis = phis
        self.interaction_matrix = interaction_matrix
        self.pulser_lindblads = pulser_lindblads
        self.device = torch.device(device)
        n = self.nqubits
        if 2 * n > native.EMU_SV_MAX_QUBITS:
            raise ValueError(f"a density matrix of {n} qubits exceeds the kernels' 32-bit index")
        om, de, ph = _cpu_f64(omegas), _cpu_f64(deltas), _cpu_f64(phis)
        self.complex = bool(torch.any(ph != 0))
        z = torch.zeros(n, dtype=torch.float64)
        u2 = torch.zeros(2 * n, 2 * n, dtype=torch.float64)
        u2[:n, :n] = _cpu_f64(interaction_matrix)
        # register qubit q (bit 2n-1-q) is row qubit q for q < n
        self._block, _ = native.pack_hamiltonian(torch.cat([om, z]).numpy(),
                                                 torch.cat([de, z]).numpy(),
                                                 torch.cat([ph, z]).numpy(), u2.numpy())
        self._s_full_cache = None  # built on first use, as the reference reads the operators
        self._dev: dict = {}
        self._diag_built: set = set()  # device blocks whose diagonal table was built once

    @property
    def _s_full(self) -> torch.Tensor:
        """S = K (x) I + (i/2) sum_k A_k (x) conj(A_k) on one qubit's (row, column) bits."""
        if self._s_full_cache is None:
            ops = [torch.as_tensor(L, dtype=dtype).cpu() for L in self.pulser_lindblads]
            k = compute_noise_from_lindbladians(ops)
            jump = sum((torch.kron(L, L.conj()) for L in ops), torch.zeros(4, 4, dtype=dtype))
            self._s_full_cache = torch.kron(k, torch.eye(2, dtype=dtype)) + 0.5j * jump
        return self._s_full_cache

    def _params(self, device: torch.device) -> torch.Tensor:
        key = device.index if device.index is not None else torch.cuda.current_device()
        t = self._dev.get(key)
        if t is None:
            t = torch.from_numpy(self._block).to(device)
            self._dev[key] = t
        return t

    def _flat(self, rho: torch.Tensor) -> torch.Tensor:
        side = 1 << self.nqubits
        if rho.numel() != side * side:
            raise ValueError(f"density matrix must be {side} x {side}")
        if not rho.is_cuda:
            raise ValueError("the emu_sv engine needs device (GPU) tensors")
        return rho.contiguous().view(-1)

    def _accumulate_m(self, src: torch.Tensor, scale: float, dst: torch.Tensor,
                      s_op: torch.Tensor | None) -> None:
        """dst = H (scale src) + sum_q S_q (scale src) (flat 2N-bit vectors)."""
        n = self.nqubits
        params = self._params(src.device)
        # this block's diagonal table, once built, is reused (the library
        # rebuilds it whenever another block used the workspace in between)
        key = params.data_ptr()
        native.hpsi(2 * n, params, 1.0, src, scale, dst, complex_coef=self.complex,
                    keep_diag=key in self._diag_built, **_SCHED.kwargs())
        self._diag_built.add(key)
        if s_op is None:
            return
        op = np.ascontiguousarray((s_op * scale).numpy())
        if n <= 8:  # every qubit's pair in one gathered launch
            native.apply_pairs_all(n, src, dst, op)
            return
        for q in range(n):
            native.apply_pair(2 * n, src, dst, 2 * n - 1 - q, n - 1 - q, op, accumulate=True)

    def device_matvec(self, src: torch.Tensor, scale: float, dst: torch.Tensor,
                      sigma: complex = 1.0) -> None:
        """dst = sigma * L(scale * src) on flat device vectors (Krylov engine)."""
        self._accumulate_m(src, scale, dst, self._s_full)
        side = 1 << self.nqubits
        m = dst.view(side, side)
        native.antihermitian(m, m, complex(sigma))

    def __matmul__(self, density_matrix: torch.Tensor) -> torch.Tensor:
        src = self._flat(density_matrix)
        out = torch.empty_like(src)
        self.device_matvec(src, 1.0, out)
        return out.view(density_matrix.shape)

    def apply_local_op_to_density_matrix(self, density_matrix: torch.Tensor,
                                         local_op: torch.Tensor, target_qubit: int) -> torch.Tensor:
        """L rho for a 2x2 L on one qubit (lindblad_operator.py:94-113)."""
        return self._pair_apply(density_matrix, torch.kron(
            torch.as_tensor(local_op, dtype=dtype).cpu(), torch.eye(2, dtype=dtype)), target_qubit)

    def apply_density_matrix_to_local_op_T(self, density_matrix: torch.Tensor,
                                           local_op: torch.Tensor, target_qubit: int) -> torch.Tensor:
        """rho L^H for a 2x2 L on one qubit (lindblad_operator.py:115-135)."""
        return self._pair_apply(density_matrix, torch.kron(
            torch.eye(2, dtype=dtype), torch.as_tensor(local_op, dtype=dtype).cpu().conj()),
            target_qubit)

    def _pair_apply(self, density_matrix, s_op, q):
        n = self.nqubits
        src = self._flat(density_matrix)
        out = torch.empty_like(src)
        native.apply_pair(2 * n, src, out, 2 * n - 1 - q, n - 1 - q, s_op.numpy(), accumulate=False)
        return out.view(density_matrix.shape)

    def h_eff(self, density_matrix: torch.Tensor,
              lindblad_ops: torch.Tensor = torch.zeros(2, 2, dtype=dtype)) -> torch.Tensor:
        """H rho + sum_q (K_q rho) for the single-qubit term K = lindblad_ops
        (zero by default), lindblad_operator.py:137-156."""
        src = self._flat(density_matrix)
        out = torch.empty_like(src)
        k = torch.as_tensor(lindblad_ops, dtype=dtype).cpu()
        s_op = torch.kron(k, torch.eye(2, dtype=dtype)) if torch.any(k != 0) else None
        self._accumulate_m(src, 1.0, out, s_op)
        return out.view(density_matrix.shape)

    def expect(self, state) -> torch.Tensor:
        """E = Tr(H rho) (lindblad_operator.py:218-223)."""
        data = state.data if hasattr(state, "data") else state
        en = self.h_eff(data).diagonal().sum()
        assert abs(float(en.imag)) <= 1e-8, "Tr(H rho) is not real"
        return en.real.to(torch.float64)


class ScaledLindbladian:
    """op(rho) = sigma * L(rho) for the Krylov engine (e.g. sigma = -i dt)."""

    def __init__(self, lindbladian: RydbergLindbladian, sigma: complex):
        self.lindbladian = lindbladian
        self.sigma = complex(sigma)

    def device_matvec(self, src: torch.Tensor, scale: float, dst: torch.Tensor) -> None:
        self.lindbladian.device_matvec(src, scale, dst, self.sigma)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        flat = self.lindbladian._flat(x)
        out = torch.empty_like(flat)
        self.device_matvec(flat, 1.0, out)
        return out.view(x.shape)
