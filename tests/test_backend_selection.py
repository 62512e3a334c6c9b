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


@pytest.mark.gpu
@pytest.mark.usefixtures("restore_rng_state")
def test_max_krylov_dim_reaches_backward(gpu_device, monkeypatch):
    """The reference's test_max_krylov_dim_reaches_backward (test/emu_sv/
    test_sv_backend.py:274-307): the per-step gradient uses the stepper's
    max_krylov_dim in the forward krylov_exp, the backward krylov_exp of the
    input state and the double_krylov of the pulse parameters."""
    te_torch = importlib.import_module("emulators_amd.emu_sv.time_evolution.torch")
    calls = {"krylov_exp": [], "double_krylov": []}
    real_exp, real_double = te_torch.krylov_exp, te_torch.double_krylov
    sig_exp, sig_double = inspect.signature(real_exp), inspect.signature(real_double)

    def spy_exp(*args, **kwargs):
        b = sig_exp.bind(*args, **kwargs)
        b.apply_defaults()
        calls["krylov_exp"].append(b.arguments["max_krylov_dim"])
        return real_exp(*args, **kwargs)

    def spy_double(*args, **kwargs):
        b = sig_double.bind(*args, **kwargs)
        calls["double_krylov"].append(b.arguments["max_krylov_dim"])
        return real_double(*args, **kwargs)

    monkeypatch.setattr(te_torch, "krylov_exp", spy_exp)
    monkeypatch.setattr(te_torch, "double_krylov", spy_double)
    torch.manual_seed(1337)
    n = 2
    om, de, ph = (torch.rand(n, dtype=torch.float64, requires_grad=True) for _ in range(3))
    u = torch.triu(torch.rand(n, n, dtype=torch.float64), 1)
    u = (u + u.T).requires_grad_(True)
    psi = torch.randn(2**n, dtype=torch.complex128)
    psi = (psi / psi.norm()).to(gpu_device).requires_grad_(True)
    out, _ = EvolveStateVector(MAX_DIM).apply(1.0, om, de, ph, u, psi, 1e-10, None)
    torch.autograd.grad(out.real.sum(), (om, de, ph, u, psi))
    assert calls["krylov_exp"] == [MAX_DIM] * 2  # forward, then the input state's gradient
    assert calls["double_krylov"] == [MAX_DIM]  # the pulse parameters' gradient


@pytest.mark.gpu
@pytest.mark.parametrize("state_requires_grad", [True, False])
def test_forward_output_requires_grad(gpu_device, state_requires_grad):
    """The reference's test_forward_with_requires_grad (test/emu_sv/
    test_time_evolution.py:47-55): the evolved state takes part in autograd
    when either the state or the Hamiltonian parameters require grad."""
    g = torch.Generator().manual_seed(4)
    om, de, ph = (torch.rand(1, dtype=torch.float64, generator=g) for _ in range(3))
    if not state_requires_grad:
        om.requires_grad_(True)
    u = torch.zeros(1, 1, dtype=torch.float64)
    psi = torch.randn(2, dtype=torch.complex128, generator=g)
    psi = (psi / psi.norm()).to(gpu_device).requires_grad_(state_requires_grad)
    out, _ = EvolveStateVector().apply(0.3, om, de, ph, u, psi, 1e-5, None)
    assert out.requires_grad
