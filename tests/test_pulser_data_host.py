"""PulserData (the SequenceData producer from a pulser Sequence) on stand-ins
for pulser's objects, with the scenarios and expected values of the
reference's own tests (test/emu_base/test_adapter.py:690-903): one XY
trajectory with an interaction-matrix override, the cutoff, the SLM mask and
the jump operators; two ising trajectories with 3 + 2 repetitions.  pulser
itself is not installed, so what pulser computes (HamiltonianData's noisy
samples, NoiseModel's noise types) is supplied by the stand-ins, and the
producer's own steps are what these tests pin."""

import math
from types import SimpleNamespace

import pytest
import torch

from emulators_amd.emu_base import pulser_adapter as pa
from emulators_amd.emu_base.jump_lindblad_operators import get_all_lindblad_noise_operators

QUBITS = ["test_qubit_0", "test_qubit_1", "test_qubit_2"]


class StandInNoiseModel:
    """The NoiseModel attributes the producer reads (pulser's defaults)."""

    def __init__(self, **kw):
        self.relaxation_rate = kw.get("relaxation_rate", 0.0)
        self.dephasing_rate = kw.get("dephasing_rate", 0.0)
        self.hyperfine_dephasing_rate = 0.0
        self.depolarizing_rate = kw.get("depolarizing_rate", 0.0)
        self.eff_noise_rates = kw.get("eff_noise_rates", [])
        self.eff_noise_opers = kw.get("eff_noise_opers", [])
        self.state_prep_error = kw.get("state_prep_error", 0.0)
        types = [name for name, rate in (("dephasing", self.dephasing_rate),
                                         ("depolarizing", self.depolarizing_rate),
                                         ("relaxation", self.relaxation_rate)) if rate]
        if self.eff_noise_rates:
            types.append("eff_noise")
        self.noise_types = tuple(kw.get("noise_types", types))

    def __bool__(self):
        return bool(self.noise_types)


class StandInSamples:
    def __init__(self, basis, per_qubit, duration):
        self._d = {"Local": {basis: per_qubit}}
        self.max_duration = duration

    def to_nested_dict(self, all_local, samples_type):
        assert all_local and samples_type == "tensor"
        return self._d


class StandInRegister:
    qubit_ids = QUBITS

    def find_indices(self, ids):
        return [QUBITS.index(q) if isinstance(q, str) else int(q) for q in ids]


def _sequence(duration, slm_time=(), slm_targets=()):
    return SimpleNamespace(register=StandInRegister(), _slm_mask_time=list(slm_time),
                           _slm_mask_targets=list(slm_targets), _in_ising=True,
                           get_duration=lambda include_fall_time=False: duration,
                           device=SimpleNamespace(default_noise_model=None))


def _hamiltonian_data(trajectories, interaction_type, dim=2):
    class StandInHamiltonianData:
        calls = []

        @classmethod
        def from_sequence(cls, sequence, with_modulation, noise_model, n_trajectories):
            cls.calls.append((with_modulation, n_trajectories))
            return SimpleNamespace(
                basis_data=SimpleNamespace(dim=dim, interaction_type=interaction_type,
                                           eigenbasis=["r", "g"]),
                noisy_samples=trajectories)

    return StandInHamiltonianData


@pytest.fixture
def stand_ins(monkeypatch):
    def install(hd):
        monkeypatch.setattr(pa, "HamiltonianData", hd)
        monkeypatch.setattr(pa, "NoiseModel", StandInNoiseModel)
        monkeypatch.setattr(pa, "has_shot_to_shot_except_spam", lambda model: False)
    return install


def _config(**kw):
    from emulators_amd.emu_sv import SVConfig
    return SVConfig(log_level=40, **kw)


def test_one_xy_trajectory(stand_ins):
    """test_adapter.py:690-850: dt 2 over 10 ns, an XY sample table for every
    qubit, the config's 2-stack interaction matrix cut at 0.15, then the SLM
    mask over [1, 10] ns on qubit 1, and the noise model's jump operators
    (the effective-noise operator unflipped in the XY basis)."""
    amp = torch.tensor([10.0, 8.88888889, 7.77777778, 6.66666667, 5.55555556, 4.44444444,
                        3.33333333, 2.22222222, 1.11111111, 0.0], dtype=torch.complex128)
    det = torch.tensor([-10.0, -7.77777778, -5.55555556, -3.33333333, -1.11111111, 1.11111111,
                        3.33333333, 5.55555556, 7.77777778, 10.0], dtype=torch.complex128)
    phase = torch.full((10,), 0.2, dtype=torch.complex128)
    per_qubit = {q: {"amp": amp, "det": det, "phase": phase} for q in QUBITS}
    samples = StandInSamples("XY", per_qubit, 10)
    traj = SimpleNamespace(bad_atoms={q: False for q in QUBITS})
    hd = _hamiltonian_data([SimpleNamespace(trajectory=traj, samples=samples, reps=1)], "XY")
    stand_ins(hd)
    m = [[0.0, 0.0929, -0.4], [0.0929, 0.0, 0.1067], [-0.4, 0.1067, 0.0]]
    torch.manual_seed(3)
    collapse = torch.rand(2, 2, dtype=torch.complex128)
    noise = StandInNoiseModel(depolarizing_rate=0.16, dephasing_rate=0.005, eff_noise_rates=[0.0036],
                              eff_noise_opers=[collapse])
    cfg = _config(noise_model=noise, interaction_matrix=[m, m], interaction_cutoff=0.15)
    target = [0.0, 2.0, 4.0, 6.0, 8.0, 10.0]
    cut = torch.tensor([[[0.0, 0.0, -0.4], [0.0, 0.0, 0.0], [-0.4, 0.0, 0.0]]] * 2, dtype=torch.float64)
    ops = get_all_lindblad_noise_operators(noise, interact_type="XY")
    for slm_time, slm_targets in (((), ()), ((1.0, 10.0), (1,))):
        data = pa.PulserData(sequence=_sequence(10, slm_time, slm_targets), config=cfg, dt=2.0)
        assert data.target_times == target
        out = list(data.get_sequences())
        assert len(out) == 1
        sd = out[0]
        om, de, ph = pa.extract_omega_delta_phi(per_qubit, QUBITS, target)
        assert torch.equal(sd.omega, om) and torch.equal(sd.delta, de) and torch.equal(sd.phi, ph)
        assert sd.hamiltonian_type == pa.HamiltonianType.XY
        assert torch.allclose(sd.interaction_matrix(10.0), cut)
        masked = cut.clone()
        if slm_targets:
            masked[:, 1] = 0.0
            masked[:, :, 1] = 0.0
        assert torch.allclose(sd.interaction_matrix(0.0), masked)
        assert torch.allclose(sd.interaction_matrix(9.99), masked)
        assert len(sd.lindblad_ops) == len(ops) == 5
        for a, b in zip(ops, sd.lindblad_ops):
            assert torch.equal(a, b)
        assert torch.allclose(ops[-1], math.sqrt(0.0036) * collapse)
        assert sd.qubit_ids == tuple(QUBITS) and sd.bad_atoms == (False, False, False)
        assert not sd.has_stochastic_noise and sd.state_prep_error == 0.0


def test_two_trajectories_with_repetitions(stand_ins):
    """test_adapter.py:853-903: two ising trajectories of 3 and 2 repetitions
    give 5 SequenceData with the same midpoint tables; 13 ns at dt 2 ends on
    the duration; the trajectory's own interaction matrix is cut at 0.15;
    bad atoms are carried through."""
    torch.manual_seed(5)
    per_qubit = {q: {"amp": torch.rand(13), "det": torch.randn(13), "phase": torch.zeros(13)}
                 for q in QUBITS[1:]}
    samples = StandInSamples("ground-rydberg", per_qubit, 13)
    mat = torch.randn(3, 3, dtype=torch.float64)
    mat = (mat + mat.T).fill_diagonal_(0).reshape(1, 3, 3)
    bad = {"test_qubit_0": True, "test_qubit_1": False, "test_qubit_2": False}
    traj = SimpleNamespace(bad_atoms=bad, interaction_matrix=SimpleNamespace(as_tensor=lambda: mat))
    hd = _hamiltonian_data([SimpleNamespace(trajectory=traj, samples=samples, reps=3),
                            SimpleNamespace(trajectory=traj, samples=samples, reps=2)], "ising")
    stand_ins(hd)
    cfg = _config(noise_model=None, interaction_cutoff=0.15, n_trajectories=5)
    data = pa.PulserData(sequence=_sequence(13), config=cfg, dt=2)
    assert data.target_times == [0.0, 2.0, 4.0, 6.0, 8.0, 10.0, 12.0, 13.0]
    assert hd.calls[-1] == (False, 5)
    out = list(data.get_sequences())
    assert len(out) == 5
    om, de, ph = pa.extract_omega_delta_phi(per_qubit, QUBITS, data.target_times)
    cut = mat.clone()
    cut[torch.abs(cut) < 0.15] = 0.0
    for sd in out:
        assert torch.equal(sd.omega, om) and torch.equal(sd.delta, de) and torch.equal(sd.phi, ph)
        assert torch.all(sd.omega[:, 0] == 0)  # a qubit without samples stays at zero
        assert torch.equal(sd.interaction_matrix(0.0), cut)
        assert sd.bad_atoms == (True, False, False)
        assert sd.hamiltonian_type == pa.HamiltonianType.Rydberg and sd.lindblad_ops == []


def test_producer_errors(stand_ins):
    """The reference's refusals: an unsupported basis, an interaction matrix
    of the wrong size or stack, samples not ending at the last target time,
    and more than one addressed basis."""
    per_qubit = {q: {"amp": torch.zeros(10), "det": torch.zeros(10), "phase": torch.zeros(10)}
                 for q in QUBITS}
    samples = StandInSamples("ground-rydberg", per_qubit, 10)
    zero = torch.zeros(1, 3, 3, dtype=torch.float64)
    traj = SimpleNamespace(bad_atoms={q: False for q in QUBITS},
                           interaction_matrix=SimpleNamespace(as_tensor=lambda: zero))
    stand_ins(_hamiltonian_data([SimpleNamespace(trajectory=traj, samples=samples, reps=1)], "digital"))
    with pytest.raises(ValueError, match="Unsupported basis"):
        pa.PulserData(sequence=_sequence(10), config=_config(), dt=2.0)
    stand_ins(_hamiltonian_data([SimpleNamespace(trajectory=traj, samples=samples, reps=1)], "ising"))
    with pytest.raises(ValueError, match="number of qubits"):
        pa.PulserData(sequence=_sequence(10), config=_config(interaction_matrix=[[0.0, 1.0], [1.0, 0.0]]),
                      dt=2.0)
    m = torch.zeros(2, 3, 3).tolist()
    with pytest.raises(ValueError, match="interaction matrices in the stack"):
        pa.PulserData(sequence=_sequence(10), config=_config(interaction_matrix=m), dt=2.0)
    short = StandInSamples("ground-rydberg", per_qubit, 12)
    stand_ins(_hamiltonian_data([SimpleNamespace(trajectory=traj, samples=short, reps=1)], "ising"))
    with pytest.raises(ValueError, match="last target time"):
        list(pa.PulserData(sequence=_sequence(10), config=_config(), dt=2.0).get_sequences())
    two = StandInSamples("ground-rydberg", per_qubit, 10)
    two._d["Local"]["XY"] = per_qubit
    with pytest.raises(ValueError, match="single interaction type"):
        pa.local_samples(two)


def test_without_pulser_the_producer_says_so(monkeypatch):
    """With pulser absent and no stand-ins, building PulserData names the
    pulser-free entry points."""
    try:
        import pulser  # noqa: F401
        pytest.skip("pulser is installed")
    except ImportError:
        pass
    monkeypatch.setattr(pa, "HamiltonianData", None)
    with pytest.raises(ImportError, match="needs pulser"):
        pa.PulserData(sequence=_sequence(10), config=_config(), dt=2.0)


@pytest.mark.gpu
def test_backend_run_on_a_sequence(stand_ins, gpu_device):
    """SVBackend(sequence).run() builds PulserData and runs each SequenceData:
    the same occupations as running that SequenceData directly, and as the
    oracle's dense evolution of the same tables."""
    from emulators_amd.emu_sv import Occupation, SVBackend, SVConfig
    from oracle import emu_sv_oracle as orc

    t = torch.arange(40, dtype=torch.float64)
    per_qubit = {q: {"amp": 4.0 * torch.sin(math.pi * t / 40) ** 2, "det": -5.0 + 0.25 * t,
                     "phase": torch.zeros(40)} for q in QUBITS}
    samples = StandInSamples("ground-rydberg", per_qubit, 40)
    u = torch.tensor([[[0.0, 1.5, 0.2], [1.5, 0.0, 1.5], [0.2, 1.5, 0.0]]], dtype=torch.float64)
    traj = SimpleNamespace(bad_atoms={q: False for q in QUBITS},
                           interaction_matrix=SimpleNamespace(as_tensor=lambda: u))
    stand_ins(_hamiltonian_data([SimpleNamespace(trajectory=traj, samples=samples, reps=1)], "ising"))
    cfg = SVConfig(observables=[Occupation(evaluation_times=[1.0])], dt=5, log_level=40)
    res = SVBackend(_sequence(40), config=cfg).run()
    (sd,) = list(pa.PulserData(sequence=_sequence(40), config=cfg, dt=5).get_sequences())
    direct = SVBackend._run_from_sequence_data(sd, cfg)
    assert torch.allclose(res.occupation[-1], direct.occupation[-1], atol=1e-12, rtol=0)
    psi = torch.zeros(8, dtype=torch.complex128)
    psi[0] = 1.0
    for k in range(len(sd.target_times) - 1):
        dt = (sd.target_times[k + 1] - sd.target_times[k]) * 1e-3
        psi = orc.dense_step(dt, sd.omega[k].real, sd.delta[k].real, sd.phi[k].real,
                             sd.interaction_matrix(sd.target_times[k])[0], psi)
    occ = torch.tensor([float(sum(abs(psi[x]) ** 2 for x in range(8) if (x >> (2 - q)) & 1))
                        for q in range(3)], dtype=torch.float64)
    assert torch.allclose(res.occupation[-1].cpu().to(torch.float64), occ, atol=1e-9, rtol=0)
