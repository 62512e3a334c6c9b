"""SVConfig keyword handling and its JSON round trip, after the reference's
test/emu_sv/test_svconfig.py (pulser's own schema is not available, so the
round trip is checked against this package's reader)."""

import json
import logging
import re

import pytest

from emulators_amd.emu_sv import BitStrings, Occupation, Solver, StateVector, SVConfig

SV_PARAMS = {"dt": 1, "max_krylov_dim": 15, "krylov_tolerance": 1e-1, "gpu": False,
             "interaction_cutoff": 1.1, "log_level": logging.ERROR, "log_file": None,
             "solver": Solver.LINDBLAD}


def test_default_round_trip():
    cfg = SVConfig()
    assert len(cfg.observables) == 0
    back = SVConfig.from_abstract_repr(cfg.to_abstract_repr())
    for attr in SV_PARAMS:
        assert getattr(back, attr) == getattr(cfg, attr), attr


def test_round_trip_with_parameters_and_observables():
    obs = [BitStrings(evaluation_times=[1.0], num_shots=77),
           Occupation(evaluation_times=[0.5, 1.0], tag_suffix="x")]
    cfg = SVConfig(**SV_PARAMS, observables=obs)
    back = SVConfig.from_abstract_repr(cfg.to_abstract_repr())
    for attr, val in SV_PARAMS.items():
        assert getattr(back, attr) == getattr(cfg, attr) == val, attr
    assert [o.tag for o in back.observables] == ["bitstrings", "occupation_x"]
    assert back.observables[0].num_shots == 77
    assert list(back.observables[1].evaluation_times) == [0.5, 1.0]


def test_unexpected_kwargs_message():
    msg = ("'SVConfig' received unexpected keyword arguments: {'blabla'}; "
           "only the following keyword arguments are expected:")
    with pytest.raises(ValueError, match="^" + re.escape(msg)):
        SVConfig(blabla=10)


def test_serialization_with_state():
    basis = ["r", "g"]
    amp = {"gg": (0.7071 + 0.0j), "rr": (0.7071 + 0.0j)}
    state = StateVector.from_state_amplitudes(eigenstates=basis, amplitudes=amp)
    text = SVConfig(initial_state=state).to_abstract_repr()
    d = json.loads(text)
    assert d["initial_state"]["eigenstates"] == basis
    assert d["initial_state"]["amplitudes"] == amp  # as the reference compares them
    back = SVConfig.from_abstract_repr(text)
    assert back.initial_state.data.cpu().allclose(state.data.cpu())


def test_serialization_with_complex_state():
    """Complex amplitudes as {"real", "imag"}; the reader also takes [re, im]."""
    basis = ["r", "g"]
    amp = {"gg": 0.6 + 0.0j, "gr": 0.0 + 0.8j}
    state = StateVector.from_state_amplitudes(eigenstates=basis, amplitudes=amp)
    text = SVConfig(initial_state=state).to_abstract_repr()
    d = json.loads(text)
    assert d["initial_state"]["amplitudes"] == {"gg": 0.6, "gr": {"real": 0.0, "imag": 0.8}}
    back = SVConfig.from_abstract_repr(text)
    assert back.initial_state.data.cpu().allclose(state.data.cpu())
    d["initial_state"]["amplitudes"] = {"gg": [0.6, 0.0], "gr": [0.0, 0.8]}
    again = SVConfig.from_abstract_repr(json.dumps(d))
    assert again.initial_state.data.cpu().allclose(state.data.cpu())
