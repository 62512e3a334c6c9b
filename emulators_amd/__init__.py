"""emulators_amd: MI355X-native state-vector Krylov time evolution.

Drop-in for the emu_sv state-vector path of the reference emulators
(``emu_sv.SVBackend`` / ``SVConfig``): the Rydberg Hamiltonian applied to a
2^N complex state inside Lanczos Krylov exponential steps, computed by
hand-written gfx950 HIP kernels behind the C ABI in ``include/emu_sv_hip.h``.

Sub-packages mirror the reference layout:
  emulators_amd.emu_base   SequenceData, krylov_exp, PCHIP, logging helpers
  emulators_amd.emu_sv     SVBackend, SVConfig, StateVector, observables
"""

__version__ = "0.1.0"
