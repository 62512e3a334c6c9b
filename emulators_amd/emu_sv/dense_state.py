"""Dense states (reference emu_sv/dense_state.py:17-240, state-vector part)."""

from __future__ import annotations

import math
from collections import Counter
from typing import Mapping, Sequence, Type, TypeVar

import torch

from emulators_amd.emu_base import DEVICE_COUNT, apply_measurement_errors
from emulators_amd.emu_sv.utils import index_to_bitstring, sample_indices

dtype = torch.complex128
DenseStateType = TypeVar("DenseStateType", bound="DenseState")


SAMPLE_MAX_CATEGORIES = 2**24  # torch.multinomial limit (emu_sv/utils.py:38)


class DenseState:
    """A complex128 vector over the qubit basis; index bit N-1-q is qubit q,
    bit value 1 = "r" (dense_state.py:183-185)."""

    def __init__(self, data: torch.Tensor, *, gpu: bool = True,
                 eigenstates: Sequence[str] = ("r", "g")):
        assert data.dim() == 1 or data.dim() == 2
        self.eigenstates = tuple(eigenstates)
        device = "cuda" if gpu and DEVICE_COUNT > 0 else "cpu"
        self.data = data.to(dtype=dtype, device=device)

    def __repr__(self) -> str:
        return repr(self.data)

    @property
    def n_qudits(self) -> int:
        return int(math.log2(self.data.shape[0]))

    def sample(self, num_shots: int = 1000, one_state=None, p_false_pos: float = 0.0,
               p_false_neg: float = 0.0) -> Counter:
        """Bitstring counts drawn from |psi|^2 or |diag(rho)| (dense_state.py:40-95).

        Up to 2^24 categories the indices come from torch.multinomial, as in
        the reference (emu_sv/utils.py:38-39), so seeded counts are the
        reference's.  Beyond, on the GPU, the native inverse-CDF sampler
        (emu_sv_sample) turns the reference's torch.rand draws into indices
        with the reference's test, searchsorted(cdf / total, u, right=True)
        (utils.py:41-49); on the CPU ``sample_indices`` does."""
        density = self.data.dim() == 2
        categories = self.data.shape[0]
        if categories <= SAMPLE_MAX_CATEGORIES or not self.data.is_cuda:
            probabilities = torch.abs(self.data.diagonal()) if density else torch.abs(self.data) ** 2
            # two arguments, as the reference calls it (its default limit is 2^24)
            outcomes = sample_indices(probabilities, num_shots).cpu()
        else:
            from emulators_amd import _native as native
            data = self.data.contiguous()
            n = self.n_qudits
            total = (float(torch.abs(data.diagonal()).sum()) if density
                     else float(native.dot(n, data.view(-1))[0].item()))
            if not total > 0:
                # same exception and message as torch.multinomial
                raise RuntimeError("invalid multinomial distribution (sum of probabilities <= 0)")
            u = torch.rand(num_shots, dtype=torch.float64, device=data.device)
            outcomes = native.sample(n, data, u, density_matrix=density).cpu()
        counts = Counter(index_to_bitstring(self.n_qudits, int(o)) for o in outcomes.tolist())
        if p_false_neg > 0 or p_false_pos > 0:
            counts = apply_measurement_errors(counts, p_false_pos=p_false_pos,
                                              p_false_neg=p_false_neg)
        return counts

    def overlap(self, other: "DenseState") -> torch.Tensor:
        """|<a|b>|^2 for two vectors, Tr(a^dag b) with density matrices
        (dense_state.py:97-140)."""
        assert isinstance(other, DenseState), "Only overlaps between emu-sv states are implemented."
        l_data, r_data = self.data, other.data
        if l_data.dim() < r_data.dim():
            l_data = torch.outer(l_data, l_data.conj())
        if r_data.dim() < l_data.dim():
            r_data = torch.outer(r_data, r_data.conj())
        assert l_data.shape == r_data.shape, "States do not have the same shape."
        ov = torch.vdot(l_data.flatten(), r_data.to(self.data.device).flatten())
        if l_data.dim() == 1:
            ov = torch.abs(ov) ** 2
        return ov

    @staticmethod
    def _build_state_vector(n_qudits: int, amplitudes: Mapping[str, complex]) -> torch.Tensor:
        acc = torch.zeros(2**n_qudits, dtype=torch.complex128,
                          device="cuda" if torch.cuda.is_available() else "cpu")
        for bits, amp in amplitudes.items():
            acc[int(bits.replace("r", "1").replace("g", "0"), 2)] = amp
        acc /= acc.norm()
        return acc

    @classmethod
    def _from_state_amplitudes(cls: Type[DenseStateType], *, eigenstates: Sequence[str],
                               n_qudits: int, amplitudes: Mapping[str, complex]):
        basis = set(eigenstates)
        if basis == {"0", "1"}:
            raise NotImplementedError(
                "{'0','1'} basis is related to XY Hamiltonian, which is not implemented"
            )
        if basis != {"r", "g"}:
            raise ValueError("Unsupported basis provided")
        sv = cls._build_state_vector(n_qudits=n_qudits, amplitudes=amplitudes)
        return cls(sv, gpu=sv.is_cuda, eigenstates=eigenstates), amplitudes

    @classmethod
    def from_state_amplitudes(cls: Type[DenseStateType], *, eigenstates: Sequence[str],
                              amplitudes: Mapping[str, complex]) -> DenseStateType:
        """State from a {bitstring: amplitude} mapping, normalised.  The
        bitstrings must all have one length and use only the eigenstates'
        labels (the check pulser's State.from_state_amplitudes makes before
        dense_state.py:192-240 runs): ValueError otherwise."""
        if not amplitudes:
            raise ValueError("at least one amplitude is needed")
        lengths = {len(k) for k in amplitudes}
        if len(lengths) != 1:
            raise ValueError(f"every bitstring must have the same length, got {sorted(lengths)}")
        labels = set(eigenstates)
        bad = sorted({c for k in amplitudes for c in k} - labels)
        if bad:
            raise ValueError(f"bitstrings use labels {bad} outside the eigenstates {tuple(eigenstates)}")
        n = lengths.pop()
        state, amps = cls._from_state_amplitudes(eigenstates=eigenstates, n_qudits=n,
                                                 amplitudes=amplitudes)
        state._amplitudes = dict(amps)  # kept for the serialised form
        return state

    def _to_abstract_repr(self) -> dict:
        """{"eigenstates": ..., "amplitudes": {bitstring: amplitude}}: the
        amplitudes the state was built from, else its nonzero amplitudes
        (state vectors only)."""
        amps = getattr(self, "_amplitudes", None)
        if amps is None:
            data = self.data.detach().cpu()
            if data.dim() != 1:
                raise ValueError("only state vectors can be serialised")
            n = self.n_qudits  # bit value 1 = "r", 0 = "g", as _build_state_vector reads them
            amps = {format(i, f"0{n}b").replace("1", "r").replace("0", "g"): complex(v)
                    for i, v in enumerate(data.tolist()) if v != 0}
        return {"eigenstates": tuple(self.eigenstates), "amplitudes": dict(amps)}
