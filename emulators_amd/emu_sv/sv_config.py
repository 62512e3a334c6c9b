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

``to_abstract_repr`` / ``from_abstract_repr`` serialise a config as JSON
(sv_config.py and pulser's EmulationConfig): the keyword arguments, the
observables (type, tag suffix, evaluation times and their own options; an
``Expectation``'s operator as its class and operator representation) and
an initial state as eigenstates plus {bitstring: amplitude}.  A real
amplitude is written as a plain number, so the parsed JSON compares equal
to the amplitudes given (test/emu_sv/test_svconfig.py:76-98); a complex
one as {"real": re, "imag": im}.  The reader also takes [re, im] pairs.
pulser's own schema is not available here, so the rest of the wire format
matches this package's reader only.
"""

from __future__ import annotations

import json
import logging
import pathlib
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
        if unexpected:  # the message of pulser's EmulationConfig
            raise ValueError(
                f"'{type(self).__name__}' received unexpected keyword arguments: {unexpected}; "
                f"only the following keyword arguments are expected: {self._expected}")
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


    # -- serialised form ------------------------------------------------------
    _SCALARS = ("dt", "max_krylov_dim", "krylov_tolerance", "gpu", "interaction_cutoff",
                "log_level", "precision", "with_modulation")

    def to_abstract_repr(self) -> str:
        """The config as a JSON string (see the module docstring)."""
        d: dict = {k: getattr(self, k) for k in self._SCALARS}
        d["log_file"] = None if self.log_file is None else str(self.log_file)
        d["solver"] = Solver(self.solver).value
        d["default_evaluation_times"] = (self.default_evaluation_times
                                         if isinstance(self.default_evaluation_times, str)
                                         else list(self.default_evaluation_times))
        d["observables"] = [_observable_repr(o) for o in self.observables]
        d["initial_state"] = None if self.initial_state is None else _state_repr(self.initial_state)
        return json.dumps(d)

    @classmethod
    def from_abstract_repr(cls, text: str) -> "SVConfig":
        d = json.loads(text)
        kw = {k: d[k] for k in cls._SCALARS if k in d}
        kw["log_file"] = None if d.get("log_file") is None else pathlib.Path(d["log_file"])
        kw["solver"] = Solver(d.get("solver", "default"))
        kw["default_evaluation_times"] = d.get("default_evaluation_times", (1.0,))
        kw["observables"] = [_observable_from_repr(o) for o in d.get("observables", [])]
        if d.get("initial_state") is not None:
            kw["initial_state"] = _state_from_repr(d["initial_state"])
        return cls(**kw)


def _state_repr(state) -> dict:
    d = state._to_abstract_repr()
    return {"eigenstates": list(d["eigenstates"]),
            "amplitudes": {k: _amplitude_repr(v) for k, v in d["amplitudes"].items()}}


def _amplitude_repr(v):
    v = complex(v)
    return v.real if v.imag == 0.0 else {"real": v.real, "imag": v.imag}


def _amplitude_from_repr(v) -> complex:
    if isinstance(v, dict):
        return complex(v["real"], v["imag"])
    if isinstance(v, (list, tuple)):
        return complex(v[0], v[1])
    return complex(v)


def _state_from_repr(d: dict):
    from emulators_amd.emu_sv.state_vector import StateVector
    amps = {k: _amplitude_from_repr(v) for k, v in d["amplitudes"].items()}
    return StateVector.from_state_amplitudes(eigenstates=tuple(d["eigenstates"]), amplitudes=amps)


def _observable_repr(obs) -> dict:
    d = {"type": type(obs).__name__, "tag_suffix": obs.tag_suffix,
         "evaluation_times": None if obs.evaluation_times is None
         else [float(t) for t in obs.evaluation_times]}
    if hasattr(obs, "num_shots"):
        d["num_shots"] = obs.num_shots
    if type(obs).__name__ == "Fidelity":
        d["state"] = _state_repr(obs.state)
    if type(obs).__name__ == "Expectation":
        d["operator"] = _operator_repr(obs.operator)
    return d


def _operator_repr(op) -> dict:
    """An operator built by from_operator_repr: its class and representation
    (complex weights as plain numbers or real/imag objects)."""
    r = op._to_abstract_repr()
    terms = [[_amplitude_repr(coeff),
              [[{k: _amplitude_repr(w) for k, w in qop.items()}, sorted(int(t) for t in targets)]
               for qop, targets in factors]]
             for coeff, factors in r["operations"]]
    return {"type": type(op).__name__, "eigenstates": r["eigenstates"],
            "n_qudits": r["n_qudits"], "operations": terms}


def _operator_from_repr(d: dict):
    from emulators_amd.emu_sv import matrix_operator
    cls = getattr(matrix_operator, d["type"])
    ops = [(_amplitude_from_repr(coeff),
            [({k: _amplitude_from_repr(w) for k, w in qop.items()}, list(targets))
             for qop, targets in factors])
           for coeff, factors in d["operations"]]
    return cls.from_operator_repr(eigenstates=tuple(d["eigenstates"]),
                                  n_qudits=int(d["n_qudits"]), operations=ops)


def _observable_from_repr(d: dict):
    from emulators_amd.emu_sv import observables as obs_mod
    cls = getattr(obs_mod, d["type"])
    kw = {"evaluation_times": d.get("evaluation_times"), "tag_suffix": d.get("tag_suffix")}
    if "num_shots" in d:
        kw["num_shots"] = d["num_shots"]
    if d["type"] == "Fidelity":
        return cls(_state_from_repr(d["state"]), **kw)
    if d["type"] == "Expectation":
        return cls(_operator_from_repr(d["operator"]), **kw)
    return cls(**kw)
