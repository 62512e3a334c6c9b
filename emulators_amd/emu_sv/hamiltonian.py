"""The Rydberg Hamiltonian as a GPU operator.

    H = sum_j (O_j/2)[cos(phi_j) X_j + sin(phi_j) Y_j] - sum_j d_j n_j
        + sum_{i<j} U_ij n_i n_j

Interface of the reference ``RydbergHamiltonian`` (emu_sv/hamiltonian.py:5-172):
``H * vec``, ``H.expect(state)``, attributes ``omegas`` (= O/2), ``deltas``,
``phis``, ``interaction_matrix``, ``device``, ``noise``, ``nqubits``,
``complex``.  Unlike the reference, the diagonal is not stored: the
parameters are packed once into a 1136-double block on the device and the
H.psi kernels regenerate the diagonal per tile (``diag`` is computed on
request).

``noise`` is the single-qubit term K = -i/2 sum_l L_l^+ L_l of the Monte-Carlo
effective Hamiltonian (hamiltonian.py:61-67, 147-151): <1|H|0> += K10 and
<0|H|1> += K01 on every qubit, and K00 / K11 per ground / Rydberg qubit on
the diagonal.  H is then not Hermitian; its products run on the LDS kernel,
which carries the term (the tiled Lanczos kernels reject it).
"""

from __future__ import annotations

import torch

from emulators_amd import _native as native
from emulators_amd import autotune


def _real_params(x: torch.Tensor | float, what: str) -> torch.Tensor:
    t = torch.as_tensor(x).detach()
    if t.is_complex():
        if torch.any(t.imag != 0):
            raise ValueError(f"{what} must be real (the noise term is given separately)")
        t = t.real
    return t.to(device="cpu", dtype=torch.float64)


_LDS = autotune.Schedule("lds", 0, 0)


class RydbergHamiltonian:
    def __init__(self, omegas: torch.Tensor, deltas: torch.Tensor, phis: torch.Tensor,
                 interaction_matrix: torch.Tensor, device: torch.device | str,
                 noise: torch.Tensor | None = None):
        self.nqubits: int = len(omegas)
        self.omegas = omegas / 2.0
        self.deltas = deltas
        self.phis = phis
        self.interaction_matrix = interaction_matrix
        self.device = torch.device(device)
        self.noise = noise
        om = _real_params(omegas, "omegas")
        de = _real_params(deltas, "deltas")
        ph = _real_params(phis, "phis")
        um = _real_params(interaction_matrix, "interaction_matrix")
        if om.shape != (self.nqubits,) or de.shape != om.shape or ph.shape != om.shape:
            raise ValueError("omegas, deltas and phis must be 1-D of the same length")
        if um.shape != (self.nqubits, self.nqubits):
            raise ValueError(f"interaction matrix must be ({self.nqubits}, {self.nqubits})")
        self._block, self.complex = native.pack_hamiltonian(om.numpy(), de.numpy(), ph.numpy(),
                                                            um.numpy())
        self._noise_np = None
        if noise is not None:
            k = torch.as_tensor(noise).detach().to(device="cpu", dtype=torch.complex128)
            if k.shape != (2, 2):
                raise ValueError("noise must be a 2x2 single-qubit matrix")
            self._noise_np = k.numpy()
            native.pack_noise(self._block, self._noise_np)
            self.complex = True
        self._dev: dict = {}

    def device_params(self, device: torch.device | None = None) -> torch.Tensor:
        """The packed parameter block on ``device`` (uploaded once)."""
        device = torch.device(device) if device is not None else self.device
        if device.type != "cuda":
            raise ValueError("the emu_sv engine runs on the GPU; got device " + str(device))
        key = device.index if device.index is not None else torch.cuda.current_device()
        t = self._dev.get(key)
        if t is None:
            t = torch.from_numpy(self._block).to(device)
            self._dev[key] = t
        return t

    def apply(self, vec: torch.Tensor, sigma: complex = 1.0,
              out: torch.Tensor | None = None) -> torch.Tensor:
        """sigma * H vec (vec: 2^N complex64/complex128 device vector)."""
        vec = vec.contiguous()
        n = native.nqubits_of(vec)
        if n != self.nqubits:
            raise ValueError(f"vector has {n} qubits, the Hamiltonian {self.nqubits}")
        if out is None:
            out = torch.empty_like(vec)
        native.hpsi(n, self.device_params(vec.device), complex(sigma), vec, 1.0, out,
                    complex_coef=self.complex, noise=self.noisy, **self._schedule(vec).kwargs())
        return out

    @property
    def noisy(self) -> bool:
        return self._noise_np is not None

    def _schedule(self, vec: torch.Tensor) -> autotune.Schedule:
        if self.noisy:
            return autotune.best_noise_schedule(self.nqubits, vec.dtype, vec.device,
                                                self.device_params(vec.device))
        return autotune.best_schedule(self.nqubits, vec.dtype, vec.device,
                                      self.device_params(vec.device), self.complex)

    def __mul__(self, vec: torch.Tensor) -> torch.Tensor:
        return self.apply(vec)

    def moments(self, vec: torch.Tensor) -> tuple[float, float, float]:
        """(<v|v>, Re<v|H v>, <Hv|Hv>) from one fused H.psi launch."""
        vec = vec.contiguous()
        n = native.nqubits_of(vec)
        out = torch.empty_like(vec)
        d = native.hpsi(n, self.device_params(vec.device), 1.0, vec, 1.0, out, dots=True,
                        complex_coef=self.complex, noise=self.noisy,
                        **self._schedule(vec).kwargs()).cpu()
        return float(native.dot(n, vec)[0].item()), float(d[1]), float(d[0])

    @property
    def diag(self) -> torch.Tensor:
        """The diagonal of H (complex128, device), evaluated by the kernels."""
        n = self.nqubits
        zeros = torch.zeros(n, dtype=torch.float64)
        block, _ = native.pack_hamiltonian(zeros.numpy(), _real_params(self.deltas, "deltas").numpy(),
                                           zeros.numpy(),
                                           _real_params(self.interaction_matrix, "U").numpy())
        if self.noisy:  # the diagonal entries of K only (its off-diagonal ones are flips)
            native.pack_noise(block, [[self._noise_np[0, 0], 0.0], [0.0, self._noise_np[1, 1]]])
        dev = self.device if self.device.type == "cuda" else torch.device("cuda")
        params = torch.from_numpy(block).to(dev)
        ones = torch.ones(2**n, dtype=torch.complex128, device=dev)
        out = torch.empty_like(ones)
        native.hpsi(n, params, 1.0, ones, 1.0, out, complex_coef=self.noisy, noise=self.noisy,
                    **(_LDS.kwargs() if self.noisy else {}))
        return out

    def expect(self, state) -> torch.Tensor:
        """Energy <psi|H|psi> (real part), hamiltonian.py:164-172."""
        data = state.data if hasattr(state, "data") else state
        data = data.contiguous()
        n = native.nqubits_of(data)
        out = torch.empty_like(data)
        d = native.hpsi(n, self.device_params(data.device), 1.0, data, 1.0, out, dots=True,
                        complex_coef=self.complex, noise=self.noisy,
                        **self._schedule(data).kwargs()).cpu()
        en = complex(d[1], d[2])
        if not self.noisy:  # with the noise term <psi|H|psi> has an imaginary part
            assert abs(en.imag) <= 1e-8 * max(1.0, abs(en.real)), "<psi|H|psi> is not real"
        return torch.tensor(en.real, dtype=torch.float64)
