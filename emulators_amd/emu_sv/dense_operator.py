"""DenseOperator (reference emu_sv/dense_operator.py); defined with the
other matrix operators."""

from emulators_amd.emu_sv.matrix_operator import DenseOperator

__all__ = ["DenseOperator"]
