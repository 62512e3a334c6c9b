"""Backend set-up after the reference's test/emu_sv/test_sv_backend.py:
the stepper each solver and noise combination selects (host), the type of
the state built from an initial state, and SVConfig.max_krylov_dim reaching
every Krylov call of a time step (GPU)."""

import importlib
import inspect
from types import SimpleNamespace

import pytest
import torch

from emulators_amd.emu_sv import DensityMatrix, Solver, StateVector, SVConfig
from emulators_amd.emu_sv.dense_state import DenseState
from emulators_amd.emu_sv.sv_backend_impl import select_stepper
from emulators_amd.emu_sv.time_evolution import EvolveDensityMatrix, EvolveMonteCarlo, EvolveStateVector

EYE = torch.eye(2, dtype=torch.complex128)


@pytest.mark.parametrize("solver,lindblad,stochastic,expected", [
    (Solver.DEFAULT, [1.0, 2.0], False, EvolveDensityMatrix),
    (Solver.DEFAULT, [], False, EvolveStateVector),
    (Solver.DEFAULT, [1.0, 2.0], True, EvolveMonteCarlo),
    (Solver.LINDBLAD, [1.0, 2.0], True, EvolveDensityMatrix),
    (Solver.MONTECARLO, [1.0, 2.0], False, EvolveMonteCarlo),
    (Solver.MONTECARLO, [], False, EvolveStateVector),
    (Solver.LINDBLAD, [], False, EvolveStateVector),
])
def test_stepper_selection(solver, lindblad, stochastic, expected):
    cfg = SimpleNamespace(solver=solver, max_krylov_dim=100)
    data = SimpleNamespace(lindblad_ops=lindblad, has_stochastic_noise=stochastic)
    stepper, density = select_stepper(cfg, data)
    assert type(stepper) is expected
    assert density == (expected is EvolveDensityMatrix)


def _one_step_data(n, lindblad=(), stochastic=False):
    return SimpleNamespace(
        omega=torch.ones(1, n, dtype=torch.float64), delta=torch.ones(1, n, dtype=torch.float64),
        phi=torch.zeros(1, n, dtype=torch.float64),
        interaction_matrix=lambda t: torch.zeros((1, n, n), dtype=torch.float64),
        state_prep_error=0.0, target_times=[0.0, 1.0], qubit_ids=tuple(f"q{i}" for i in range(n)),
        bad_atoms=(), lindblad_ops=list(lindblad), has_stochastic_noise=stochastic)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,stochastic,lindblad", [
    (DenseState, False, False), (StateVector, False, False), (DenseState, True, False),
    (StateVector, True, False), (DensityMatrix, False, True)])
def test_initial_state_type(gpu_device, kind, stochastic, lindblad):
    from emulators_amd.emu_sv.sv_backend_impl import SVBackendImpl
    state = kind.from_state_amplitudes(eigenstates=("r", "g"), amplitudes={"r": 1.0})
    cfg = SVConfig(initial_state=state, log_level=40)
    impl = SVBackendImpl(cfg, _one_step_data(1, [EYE] if (lindblad or stochastic) else [], stochastic))
    expected = DensityMatrix if (lindblad and not stochastic) else StateVector
    assert type(impl.state) is expected


@pytest.mark.gpu
def test_initial_state_with_the_wrong_atom_count(gpu_device):
    from emulators_amd.emu_sv.sv_backend_impl import SVBackendImpl
    state = StateVector.from_state_amplitudes(eigenstates=("r", "g"), amplitudes={"rr": 1.0})
    with pytest.raises(ValueError, match="number of atoms"):
        SVBackendImpl(SVConfig(initial_state=state, log_level=40), _one_step_data(3))


MAX_DIM = 17  # not the default of 100, and enough for a 2-qubit density matrix (16)


@pytest.mark.gpu
@pytest.mark.usefixtures("restore_rng_state")
@pytest.mark.parametrize("solver,lindblad,stepper", [
    (Solver.DEFAULT, [], EvolveStateVector),
    (Solver.LINDBLAD, [EYE], EvolveDensityMatrix),
    (Solver.MONTECARLO, [EYE], EvolveMonteCarlo)])
def test_max_krylov_dim_reaches_every_krylov_call(gpu_device, monkeypatch, solver, lindblad, stepper):
    from emulators_amd.emu_base.math.krylov_exp import krylov_exp
    from emulators_amd.emu_sv.sv_backend_impl import SVBackendImpl
    te_torch = importlib.import_module("emulators_amd.emu_sv.time_evolution.torch")
    te_impl = importlib.import_module("emulators_amd.emu_sv.time_evolution.implementations")
    dims = []
    sig = inspect.signature(krylov_exp)

    def spy(*args, **kwargs):
        bound = sig.bind(*args, **kwargs)
        bound.apply_defaults()
        dims.append(bound.arguments["max_krylov_dim"])
        return krylov_exp(*args, **kwargs)

    monkeypatch.setattr(te_torch, "krylov_exp", spy)
    monkeypatch.setattr(te_impl, "krylov_exp", spy)
    cfg = SVConfig(max_krylov_dim=MAX_DIM, solver=solver, log_level=40)
    impl = SVBackendImpl(cfg, _one_step_data(2, lindblad))
    assert type(impl.stepper) is stepper
    impl.step(0)
    assert dims and dims == [MAX_DIM] * len(dims)
