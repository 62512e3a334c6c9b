"""Time loop of the state-vector backend (reference emu_sv/sv_backend_impl.py:74-262).

Per step: dt from the target times, one Krylov step exp(-i dt H(t_mid)) psi
on the GPU, then the observables whose evaluation time it is.  Parity
semantics kept from the reference:

* observables at t_k see the Hamiltonian of the step that ended at t_k; at
  t = 0 the one built from omega[0], delta[0], phi[0] and U at the first
  midpoint (sv_backend_impl.py:224-234);
* observables see a renormalised copy when ||psi|| is not close to 1
  (torch.allclose defaults, sv_backend_impl.py:213-218);
* dt in ns, Omega/delta in rad/us: the step uses dt * 1e-3 (:29, :201).

With ``precision="complex64"`` the working vector is complex64 and is
promoted to a complex128 StateVector only when an observable needs it.
"""

from __future__ import annotations

import os
import time

import torch

from emulators_amd import _native as native
from emulators_amd.emu_base import SequenceData
from emulators_amd.emu_sv.observables import Results, Statistics
from emulators_amd.emu_sv.state_vector import StateVector
from emulators_amd.emu_sv.sv_config import SVConfig
from emulators_amd.emu_sv.sv_config import Solver
from emulators_amd.emu_sv.time_evolution import (EvolveDensityMatrix, EvolveMonteCarlo,
                                                 EvolveStateVector)
from emulators_amd.emu_sv.dense_state import DenseState

_TIME_CONVERSION_COEFF = 0.001  # Omega and delta in rad/us, dt in ns


def select_stepper(config: SVConfig, data):
    """The stepper choice of sv_backend_impl.py:81-95: Lindblad noise runs as
    Monte-Carlo trajectories when the noise is stochastic (or asked for),
    else as a density matrix under the Lindbladian.  Returns
    (stepper, evolves_a_density_matrix)."""
    if data.lindblad_ops:
        solver = getattr(config, "solver", Solver.DEFAULT)
        if (solver == Solver.DEFAULT and data.has_stochastic_noise) or \
                solver == Solver.MONTECARLO:
            return EvolveMonteCarlo(config.max_krylov_dim), False
        return EvolveDensityMatrix(config.max_krylov_dim), True
    return EvolveStateVector(config.max_krylov_dim), False


class SVBackendImpl:
    def __init__(self, config: SVConfig, data: SequenceData):
        self.pulser_lindblads = list(data.lindblad_ops)
        self.stepper, self.density_matrix = select_stepper(config, data)
        self._config = config
        self._data = data
        self.target_times = data.target_times
        self.omega = data.omega
        self.delta = data.delta
        self.phi = data.phi
        self.nsteps = data.omega.shape[0]
        self.nqubits = data.omega.shape[1]
        self.interaction_matrix = data.interaction_matrix
        requested_gpu = True if config.gpu is None else config.gpu
        if not requested_gpu:
            raise ValueError("emulators_amd evolves the state on the GPU: gpu=False is not supported")
        if not torch.cuda.is_available():
            raise RuntimeError("emulators_amd needs a GPU: torch.cuda.is_available() is False")
        self.resolved_gpu = True
        self._work_dtype = (torch.complex64 if getattr(config, "precision", "complex128")
                            == "complex64" else torch.complex128)
        if config.initial_state is not None and config.initial_state.n_qudits != self.nqubits:
            raise ValueError(
                "Mismatch in number of atoms: initial state has "
                f"{config.initial_state.n_qudits} and the sequence has {self.nqubits}"
            )
        if config.initial_state is not None and data.state_prep_error > 0.0:
            raise NotImplementedError("Initial state and state preparation error can not be together.")
        self._init_state()
        self.time = time.time()
        self.results = Results(atom_order=data.qubit_ids, total_duration=int(self.target_times[-1]))
        self.statistics = Statistics(
            evaluation_times=[t / self.target_times[-1] for t in self.target_times], data=[],
            timestep_count=self.nsteps)
        self._current_H = None
        self._init_dark_qubits()
        self.iterations: list[int] = []

    # -- state ---------------------------------------------------------------
    def _init_state(self) -> None:
        # sv_backend_impl.py:142-160: a state vector is promoted to |psi><psi|
        # for the density-matrix solver
        cfg_state = self._config.initial_state
        if cfg_state is not None:
            if not isinstance(cfg_state, DenseState):
                raise ValueError("initial_state must be a dense state")
            data = cfg_state.data.detach().clone().to("cuda", torch.complex128)
            if data.dim() == 2 and not self.density_matrix:
                raise ValueError("a density-matrix initial state needs the Lindblad solver")
            if self.density_matrix and data.dim() == 1:
                data = torch.outer(data, data.conj())
        else:
            side = 2**self.nqubits
            if self.density_matrix:
                data = torch.zeros(side, side, dtype=torch.complex128, device="cuda")
                data[0, 0] = 1.0
            else:
                data = torch.zeros(side, dtype=torch.complex128, device="cuda")
                data[0] = 1.0
        if self.density_matrix and self._work_dtype != torch.complex128:
            raise ValueError("the Lindblad solver runs in complex128")
        self._work = data.to(self._work_dtype).contiguous()

    @property
    def state(self):
        """The current state as a complex128 StateVector, or a DensityMatrix
        for the Lindblad solver."""
        if self.density_matrix:
            from emulators_amd.emu_sv.density_matrix_state import DensityMatrix
            return DensityMatrix(self._work, gpu=True)
        data = self._work
        if data.dtype != torch.complex128:
            data = native.convert(native.nqubits_of(data), data,
                                  torch.empty(data.numel(), dtype=torch.complex128,
                                              device=data.device))
        return StateVector(data, gpu=True)

    def _init_dark_qubits(self) -> None:
        if self._data.state_prep_error > 0.0:
            bad = torch.tensor(self._data.bad_atoms)
            indices = torch.where(bad)[0]
            original = self.interaction_matrix

            def interaction_matrix(t: float) -> torch.Tensor:
                mat = original(t).clone()
                mat[:, indices] = 0.0
                mat[:, :, indices] = 0.0
                return mat

            self.interaction_matrix = interaction_matrix
            # zeroed on copies: the sequence data's tables may be shared with
            # other runs (the same SequenceData with other bad atoms)
            self.omega = self.omega.clone()
            self.delta = self.delta.clone()
            self.phi = self.phi.clone()
            self.omega[:, bad] = 0.0
            self.delta[:, bad] = 0.0
            self.phi[:, bad] = 0.0

    # -- loop ----------------------------------------------------------------
    def step(self, step_idx: int) -> None:
        dt = self._compute_dt(step_idx)
        self._evolve_step(dt, step_idx)
        step_idx += 1
        self._apply_observables(step_idx)
        self._save_statistics(step_idx)

    def _compute_dt(self, step_idx: int) -> float:
        return self.target_times[step_idx + 1] - self.target_times[step_idx]

    def _evolve_step(self, dt: float, step_idx: int) -> None:
        self._current_H = None
        out, self._current_H = self.stepper.apply(
            dt * _TIME_CONVERSION_COEFF, self.omega[step_idx], self.delta[step_idx],
            self.phi[step_idx], self.interaction_matrix(self.target_times[step_idx])[0],
            self._work, self._config.krylov_tolerance, self.pulser_lindblads)
        self._work = out

    def _callbacks_at(self, norm_time: float) -> list:
        from emulators_amd.emu_base.pulser_adapter import is_evaluation_time

        return [cb for cb in self._config.observables
                if is_evaluation_time(self._config, cb, norm_time)]

    def _apply_observables(self, step_idx: int) -> None:
        norm_time = self.target_times[step_idx] / self.target_times[-1]
        callbacks = self._callbacks_at(norm_time)
        if not callbacks:
            return
        state = self.state
        norm = state.norm()
        if not torch.allclose(norm, torch.tensor(1.0, dtype=torch.float64)):
            state = (1 / norm.item()) * state
        if self._current_H is None:
            nxt = min(step_idx + 1, len(self.target_times) - 1)
            self._current_H = self.stepper.get_hamiltonian(
                omegas=self.omega[0], deltas=self.delta[0], phis=self.phi[0],
                pulser_lindblads=self.pulser_lindblads,
                interaction_matrix=self.interaction_matrix(
                    0.5 * (self.target_times[step_idx] + self.target_times[nxt]))[0],
                device=state.data.device)
        for cb in callbacks:
            cb(self._config, norm_time, state, self._current_H, self.results)

    def _save_statistics(self, step_idx: int) -> None:
        norm_time = self.target_times[step_idx] / self.target_times[-1]
        self.statistics.data.append(time.time() - self.time)
        self.statistics(self._config, norm_time, _StatsView(self._work), self._current_H,
                        self.results)
        self.time = time.time()

    # -- device-driven segments (EMU_SV_SEGMENTS=1) ------------------------------
    # Between two observable evaluations (and while U(t) stays the same), the
    # steps of a small noiseless register run as one register of the batch
    # engine: the Lanczos coefficients, the convergence test and the step
    # transitions stay on the device, with no host round trip per iteration.
    # The arithmetic is the per-step engine's (same iteration counts); a
    # segment that needs more than 62 Krylov vectors is redone step by step.
    # Off by default: the per-step loop, whose Krylov steps run in the
    # library's driver (emu_sv_krylov_step) up to 22 qubits, is faster
    # (chain10 0.093 s against 0.116 s, square20 0.384 s against 0.572 s).
    SEGMENT_MAX_QUBITS = 20
    SEGMENT_MAX_STEPS = 512

    def _use_segments(self) -> bool:
        if os.environ.get("EMU_SV_SEGMENTS", "0") != "1":
            return False
        if not isinstance(self.stepper, EvolveStateVector) or self.nqubits > self.SEGMENT_MAX_QUBITS:
            return False
        if torch.is_grad_enabled() and any(getattr(t, "requires_grad", False)
                                           for t in (self.omega, self.delta, self.phi)):
            return False
        return True

    def _segment_end(self, k0: int) -> int:
        u0 = self.interaction_matrix(self.target_times[k0])
        k1 = k0 + 1
        while k1 < self.nsteps and k1 - k0 < self.SEGMENT_MAX_STEPS:
            if self._callbacks_at(self.target_times[k1] / self.target_times[-1]):
                break
            if not torch.equal(self.interaction_matrix(self.target_times[k1]), u0):
                break
            k1 += 1
        return k1

    def _run_segment(self, k0: int, k1: int) -> bool:
        import numpy as np

        from emulators_amd.emu_sv.batch import BatchRegister, evolve_batch

        tt = self.target_times
        u = self.interaction_matrix(tt[k0])[0]
        reg = BatchRegister(
            omegas=self.omega[k0:k1].detach().cpu().numpy(),
            deltas=self.delta[k0:k1].detach().cpu().numpy(),
            phis=self.phi[k0:k1].detach().cpu().numpy(),
            interaction_matrix=np.asarray(u.detach().cpu().numpy()),
            dts=np.asarray([(tt[k + 1] - tt[k]) * _TIME_CONVERSION_COEFF for k in range(k0, k1)]),
            psi0=self._work)
        res = evolve_batch([reg], dtype=self._work_dtype, tolerance=self._config.krylov_tolerance,
                           max_krylov_dim=min(self._config.max_krylov_dim, 62),
                           device=self._work.device)
        if res.failed:
            return False
        self._work = res.states[0]
        self._current_H = self.stepper.get_hamiltonian(
            omegas=self.omega[k1 - 1], deltas=self.delta[k1 - 1], phis=self.phi[k1 - 1],
            pulser_lindblads=self.pulser_lindblads, interaction_matrix=u,
            device=self._work.device)
        per_step = (time.time() - self.time) / (k1 - k0)
        for k in range(k0 + 1, k1 + 1):
            if k == k1:
                self._apply_observables(k)
            self.statistics.data.append(per_step)
            self.statistics(self._config, tt[k] / tt[-1], _StatsView(self._work),
                            self._current_H, self.results)
        self.time = time.time()
        return True

    def _run(self) -> Results:
        self._apply_observables(0)
        segments = self._use_segments()
        step = 0
        while step < self.nsteps:
            end = self._segment_end(step) if segments else step + 1
            if end - step > 1 and self._run_segment(step, end):
                step = end
            else:
                self.step(step)
                step += 1
        return self.results


class _StatsView:
    """Minimal stand-in passed to the statistics observable (reads .data.is_cuda)."""

    def __init__(self, data: torch.Tensor):
        self.data = data


def run_sequence_data(sequence_data: SequenceData, config: SVConfig) -> Results:
    return SVBackendImpl(config, sequence_data)._run()


__all__ = ["SVBackendImpl", "run_sequence_data", "select_stepper"]
