"""SVConfig (reference emu_sv/sv_config.py:33-155).

Same keyword arguments and defaults as the reference: ``dt=10``,
``max_krylov_dim=100``, ``krylov_tolerance=1e-10``, ``gpu=None``,
``interaction_cutoff=0``, ``log_level=INFO``, ``log_file=None``,
``solver=Solver.DEFAULT``, plus the emulation-config options the
state-vector path reads (``observables``, ``default_evaluation_times``,
``initial_state``, ``interaction_matrix``, ``with_modulation``,
``noise_model``).

One option is new: ``precision`` = "complex128" (default, the reference's
arithmetic) or "complex64" (state and Krylov vectors stored in complex64;
Lanczos coefficients, reductions and the small exponential stay FP64).
"""

from __future__ import annotations

import logging
import pathlib
import warnings
from enum import Enum
from typing import Any, Sequence

from emulators_amd.emu_base import init_logging

PRECISIONS = ("complex128", "complex64")


class Solver(str, Enum):
    """Solver choice for Lindblad noise (reference emu_sv/solver.py)."""

    DEFAULT = "default"
    MONTECARLO = "montecarlo"
    LINDBLAD = "lindblad"


class SVConfig:
    _expected = {
        "dt", "max_krylov_dim", "krylov_tolerance", "gpu", "interaction_cutoff", "log_level",
        "log_file", "solver", "precision", "observables", "default_evaluation_times",
        "initial_state", "interaction_matrix", "with_modulation", "noise_model",
        "prefer_device_noise_model", "n_trajectories",
    }

    def __init__(self, *, dt: float = 10.0, max_krylov_dim: int = 100,
                 krylov_tolerance: float = 1e-10, gpu: bool | None = None,
                 interaction_cutoff: float = 0.0, log_level: int = logging.INFO,
                 log_file: pathlib.Path | None = None, solver: Solver = Solver.DEFAULT,
                 precision: str = "complex128", observables: Sequence = (),
                 default_evaluation_times: Sequence[float] | str = (1.0,),
                 initial_state=None, interaction_matrix=None, with_modulation: bool = False,
                 noise_model=None, **kwargs: Any):
        unexpected = set(kwargs) - self._expected
        if unexpected:
            warnings.warn(f"SVConfig: unexpected keyword arguments {sorted(unexpected)}",
                          stacklevel=2)
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}, got {precision!r}")
        if dt <= 0:
            raise ValueError("dt must be positive")
        if max_krylov_dim < 1:
            raise ValueError("max_krylov_dim must be >= 1")
        self.dt = dt
        self.max_krylov_dim = int(max_krylov_dim)
        self.krylov_tolerance = float(krylov_tolerance)
        self.gpu = gpu
        self.interaction_cutoff = float(interaction_cutoff)
        self.log_level = log_level
        self.log_file = log_file
        self.solver = solver
        self.precision = precision
        self.observables = tuple(observables)
        self.default_evaluation_times = (default_evaluation_times
                                         if isinstance(default_evaluation_times, str)
                                         else tuple(float(t) for t in default_evaluation_times))
        self.initial_state = initial_state
        self.interaction_matrix = interaction_matrix
        self.with_modulation = with_modulation
        self.noise_model = noise_model
        self.p_false_pos = 0.0
        self.p_false_neg = 0.0
        self._extra = dict(kwargs)
        tags = [o.tag for o in self.observables]
        if len(tags) != len(set(tags)):
            raise ValueError(f"observable tags must be unique, got {tags}")
        init_logging(log_level, log_file)

    def _expected_kwargs(self) -> set[str]:
        return set(self._expected)

    def is_evaluation_time(self, t: float, tol: float = 1e-10) -> bool:
        times = self.default_evaluation_times
        return isinstance(times, str) or any(abs(t - x) <= tol for x in times)

    def is_time_in_evaluation_times(self, t: float, times: Sequence[float], tol: float = 1e-10) -> bool:
        return any(abs(t - float(x)) <= tol for x in times)
