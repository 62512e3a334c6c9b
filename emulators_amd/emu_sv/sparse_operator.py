"""SparseOperator and its COO helpers (reference emu_sv/sparse_operator.py);
defined with the other matrix operators."""

from emulators_amd.emu_sv.matrix_operator import SparseOperator, sparse_add, sparse_kron

__all__ = ["SparseOperator", "sparse_add", "sparse_kron"]
