"""SVBackend (reference emu_sv/sv_backend.py:7-42)."""

from __future__ import annotations

from emulators_amd.emu_base import PulserData, SequenceData
from emulators_amd.emu_sv.observables import BitStrings, Results
from emulators_amd.emu_sv.sv_backend_impl import SVBackendImpl
from emulators_amd.emu_sv.sv_config import SVConfig


class SVBackend:
    """Emulates a pulser Sequence (or SequenceData) with GPU state vectors."""

    default_config = None

    def __init__(self, sequence=None, *, config: SVConfig | None = None, mimic_qpu: bool = False):
        self._sequence = sequence
        self._config = config if config is not None else SVConfig(
            observables=[BitStrings(evaluation_times=[1.0])])
        self._mimic_qpu = mimic_qpu

    def run(self) -> Results:
        assert isinstance(self._config, SVConfig)
        if isinstance(self._sequence, SequenceData):
            return self._run_from_sequence_data(self._sequence, self._config)
        assert getattr(self._sequence, "_in_ising", True), "Only Rydberg interaction is currently supported"
        pulser_data = PulserData(sequence=self._sequence, config=self._config, dt=self._config.dt)
        return Results.aggregate([self._run_from_sequence_data(sd, self._config)
                                  for sd in pulser_data.get_sequences()])

    @staticmethod
    def _run_from_sequence_data(sequence_data: SequenceData, config: SVConfig) -> Results:
        return SVBackendImpl(config, sequence_data)._run()
