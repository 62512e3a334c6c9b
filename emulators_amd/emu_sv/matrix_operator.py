"""Operators for custom observables (reference emu_sv/matrix_operator.py,
dense_operator.py, sparse_operator.py).

An operator is a 2^n x 2^n complex128 matrix in the state's basis (index
bit n-1-q is qubit q, 0 = |g>, 1 = |r>), held dense (``DenseOperator``) or
as a CSR tensor (``SparseOperator``).  They are user-supplied matrices off
the time-stepping path, so they are applied with the library matrix-vector
products (rocBLAS for dense, hipSPARSE for CSR, through torch).

``from_operator_repr`` takes pulser's operator representation: a list of
terms ``(coeff, [(qudit_op, targets), ...])`` where ``qudit_op`` maps
two-letter keys ``"ab"`` (|a><b| over the eigenstates) to complex weights
and ``targets`` lists the qubits it acts on; qubits not named in a term
carry the identity.  pulser's ``Operator`` base class, which validates the
representation and keeps it for serialisation, is not a dependency here:
``Operator`` below does both, with pulser's error message for a bad key.
"""

from __future__ import annotations

import copy
from typing import Any, Sequence

import torch

from emulators_amd.emu_base import DEVICE_COUNT
from emulators_amd.emu_sv.state_vector import StateVector

dtype = torch.complex128

# above this many qubits a dense 2^n x 2^n matrix (4^n * 16 bytes) is not built
MAX_DENSE_QUBITS = 14

_INDEX = {"g": 0, "r": 1}


class Operator:
    """What pulser's ``Operator`` provides to the emu-sv operators: the
    validated ``from_operator_repr`` constructor and the stored
    representation (``_eigenstates``, ``_n_qudits``, ``_operations``)."""

    _eigenstates: tuple | None = None
    _n_qudits: int | None = None
    _operations: list | None = None

    @classmethod
    def from_operator_repr(cls, *, eigenstates: Sequence[str], n_qudits: int,
                           operations: Sequence) -> "Operator":
        eigenstates = tuple(eigenstates)
        _validate_operator_repr(eigenstates, n_qudits, operations)
        op, ops = cls._from_operator_repr(eigenstates=eigenstates, n_qudits=n_qudits,
                                          operations=operations)
        op._eigenstates = eigenstates
        op._n_qudits = int(n_qudits)
        op._operations = list(ops)
        return op

    @classmethod
    def _from_operator_repr(cls, *, eigenstates, n_qudits, operations):
        raise NotImplementedError

    def _to_abstract_repr(self) -> dict:
        if self._operations is None:
            raise ValueError("only operators built by from_operator_repr can be serialised")
        return {"eigenstates": list(self._eigenstates), "n_qudits": self._n_qudits,
                "operations": self._operations}


def _validate_operator_repr(eigenstates: tuple, n_qudits: int, operations: Sequence) -> None:
    """Key, target and eigenstate checks of pulser's operator representation."""
    if len(set(eigenstates)) != len(eigenstates):
        raise ValueError(f"eigenstates must be distinct, got {eigenstates}")
    for term in operations:
        coeff, factors = term
        seen: set[int] = set()
        for qudit_op, targets in factors:
            if not isinstance(qudit_op, torch.Tensor):
                for key in qudit_op:
                    if len(key) != 2 or any(s not in eigenstates for s in key):
                        raise ValueError(
                            "Every QuditOp key must be made up of two eigenstates among "
                            f"{eigenstates}; instead, got '{key}'.")
            targets = set(int(t) for t in targets)
            if any(t < 0 or t >= n_qudits for t in targets):
                raise ValueError(f"targets {sorted(targets)} out of range for {n_qudits} qudits")
            if seen & targets:
                raise ValueError(f"qudits {sorted(seen & targets)} are targeted twice in one term")
            seen |= targets


def _factor_matrix(qudit_op: Any) -> torch.Tensor:
    """The 2x2 matrix of one qudit operator: sum of w |a><b| over its keys."""
    if isinstance(qudit_op, torch.Tensor):
        return qudit_op.to(dtype=dtype, device="cpu")
    m = torch.zeros(2, 2, dtype=dtype)
    for key, w in qudit_op.items():
        m[_INDEX[key[0]], _INDEX[key[1]]] += complex(w)
    return m


def _term_factors(n_qudits: int, factors: Sequence) -> list[torch.Tensor | None]:
    """Per qubit, its 2x2 factor in one term (None = identity)."""
    per_qubit: list[torch.Tensor | None] = [None] * n_qudits
    for qudit_op, targets in factors:
        m = _factor_matrix(qudit_op)
        for t in targets:
            per_qubit[int(t)] = m
    return per_qubit


def _check_basis(eigenstates: Sequence[str]) -> None:
    if set(eigenstates) == {"r", "g"}:
        return
    if set(eigenstates) == {"0", "1"}:
        raise NotImplementedError(
            "the {'0','1'} basis belongs to the XY Hamiltonian, which is not implemented")
    raise ValueError(f"unsupported eigenstate basis {tuple(eigenstates)}: only ('r', 'g')")


def _dense_from_repr(n_qudits: int, operations: Sequence) -> torch.Tensor:
    """sum_terms coeff * (F_0 kron F_1 kron ... kron F_{n-1}), qubit 0 the
    most significant factor, accumulated densely on the CPU."""
    dim = 1 << n_qudits
    out = torch.zeros(dim, dim, dtype=dtype)
    eye = torch.eye(2, dtype=dtype)
    for coeff, factors in operations:
        term = torch.ones(1, 1, dtype=dtype)
        for f in _term_factors(n_qudits, factors):
            term = torch.kron(term, eye if f is None else f)
        out += complex(coeff) * term
    return out


def _sparse_from_repr(n_qudits: int, operations: Sequence) -> torch.Tensor:
    """The same sum built from the factors' nonzero entries: walking the
    qubits from the most significant, each nonzero (row, col, value) of the
    partial product branches into the nonzeros of the next factor (the
    identity's two diagonal entries for qubits outside the term).  Returns
    a coalesced CSR tensor."""
    rows_all, cols_all, vals_all = [], [], []
    eye = torch.eye(2, dtype=dtype)
    for coeff, factors in operations:
        rows = torch.zeros(1, dtype=torch.int64)
        cols = torch.zeros(1, dtype=torch.int64)
        vals = torch.full((1,), complex(coeff), dtype=dtype)
        for f in _term_factors(n_qudits, factors):
            m = eye if f is None else f
            nz = torch.nonzero(m, as_tuple=False)  # (k, 2): (row, col)
            fr, fc, fv = nz[:, 0], nz[:, 1], m[nz[:, 0], nz[:, 1]]
            rows = (rows[:, None] * 2 + fr[None, :]).reshape(-1)
            cols = (cols[:, None] * 2 + fc[None, :]).reshape(-1)
            vals = (vals[:, None] * fv[None, :]).reshape(-1)
        rows_all.append(rows)
        cols_all.append(cols)
        vals_all.append(vals)
    dim = 1 << n_qudits
    idx = torch.stack([torch.cat(rows_all), torch.cat(cols_all)]) if rows_all else \
        torch.zeros(2, 0, dtype=torch.int64)
    val = torch.cat(vals_all) if vals_all else torch.zeros(0, dtype=dtype)
    return torch.sparse_coo_tensor(idx, val, (dim, dim)).coalesce().to_sparse_csr()


def sparse_kron(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Kronecker product of two sparse COO matrices (reference
    sparse_operator.py:18-28): entry (ra, ca, va) of a and (rb, cb, vb) of b
    give (ra * rows(b) + rb, ca * cols(b) + cb, va * vb).  The result is
    coalesced: with both inputs coalesced (row-major sorted, unique), the
    a-major, b-minor enumeration is sorted and unique too."""
    a, b = a.coalesce(), b.coalesce()
    (ra, ca), (rb, cb) = a.indices(), b.indices()
    hb, wb = b.shape
    rows = (ra[:, None] * hb + rb[None, :]).reshape(-1)
    cols = (ca[:, None] * wb + cb[None, :]).reshape(-1)
    vals = (a.values()[:, None] * b.values()[None, :]).reshape(-1)
    shape = (a.shape[0] * hb, a.shape[1] * wb)
    return torch.sparse_coo_tensor(torch.stack([rows, cols]), vals, shape, is_coalesced=True)


def sparse_add(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """Sum of two sparse COO matrices of one shape, coalesced (reference
    sparse_operator.py:10-15): the entries of both, duplicates summed."""
    a, b = a.coalesce(), b.coalesce()
    return torch.sparse_coo_tensor(torch.cat([a.indices(), b.indices()], dim=1),
                                   torch.cat([a.values(), b.values()]), size=a.shape).coalesce()


class MatrixOperator(Operator):
    """A 2^n x 2^n complex128 operator, dense or CSR (matrix_operator.py:25-279).

    Args:
        matrix: square complex tensor (dense or sparse CSR).
        gpu: place it on the GPU when one is present (default True).
    """

    def __init__(self, matrix: torch.Tensor, *, gpu: bool = True):
        device = "cuda" if gpu and DEVICE_COUNT > 0 else "cpu"
        self.data = matrix.to(dtype=dtype, device=device)

    def __repr__(self) -> str:
        return repr(self.data)

    def _new(self, matrix: torch.Tensor) -> "MatrixOperator":
        return type(self)(matrix, gpu=self.data.is_cuda)

    def __rmul__(self, scalar: complex) -> "MatrixOperator":
        return self._new(scalar * self.data)

    def __add__(self, other: "MatrixOperator") -> "MatrixOperator":
        t = type(self)
        assert isinstance(other, t), \
            f"Operator of type {t} can only be added to another of the same type."
        if self.data.is_sparse_csr:
            s = sparse_add(self.data.to_sparse_coo(), other.data.to(self.data.device).to_sparse_coo())
            return self._new(s.to_sparse_csr())
        return self._new(self.data + other.data.to(self.data.device))

    def __matmul__(self, other: "MatrixOperator") -> "MatrixOperator":
        raise NotImplementedError("operator products are defined for DenseOperator only")

    def apply_to(self, other: StateVector) -> StateVector:
        """O|psi> as a new StateVector (on the operator's device)."""
        assert isinstance(other, StateVector), \
            "emu-sv operators can only be applied to a StateVector."
        vec = other.data.to(device=self.data.device, dtype=dtype)
        return StateVector(self.data @ vec, gpu=self.data.is_cuda)

    def expect(self, state: StateVector) -> torch.Tensor:
        """<psi|O|psi> as a complex128 scalar tensor on the CPU."""
        assert isinstance(state, StateVector), \
            "Only expectation values of StateVectors are supported."
        vec = state.data.to(device=self.data.device, dtype=dtype)
        return torch.vdot(vec, self.data @ vec).cpu()

    @classmethod
    def _build(cls, *, eigenstates: Sequence[str], n_qudits: int, operations: Sequence,
               sparse: bool) -> torch.Tensor:
        assert len(set(eigenstates)) == 2, "Only qubits are supported in EMU-SV."
        _check_basis(eigenstates)
        if sparse:
            return _sparse_from_repr(n_qudits, operations)
        return _dense_from_repr(n_qudits, operations)

    @classmethod
    def _from_operator_repr(cls, *, eigenstates, n_qudits, operations):
        """Dense up to MAX_DENSE_QUBITS qubits, CSR beyond."""
        matrix = cls._build(eigenstates=eigenstates, n_qudits=n_qudits, operations=operations,
                            sparse=n_qudits > MAX_DENSE_QUBITS)
        return MatrixOperator(matrix), operations

    def __deepcopy__(self, memo: dict) -> "MatrixOperator":
        out = type(self)(self.data.clone(), gpu=self.data.is_cuda)
        out._eigenstates = self._eigenstates
        out._n_qudits = self._n_qudits
        out._operations = copy.deepcopy(self._operations, memo)
        memo[id(self)] = out
        return out


class DenseOperator(MatrixOperator):
    """An operator held as a dense matrix (reference emu_sv/dense_operator.py:16-69)."""

    def __matmul__(self, other: "DenseOperator") -> "DenseOperator":
        assert isinstance(other, DenseOperator), \
            "DenseOperator can only be multiplied with a DenseOperator."
        return self._new(self.data @ other.data.to(self.data.device))

    @classmethod
    def _from_operator_repr(cls, *, eigenstates, n_qudits, operations):
        matrix = cls._build(eigenstates=eigenstates, n_qudits=n_qudits, operations=operations,
                            sparse=False)
        return DenseOperator(matrix), operations


class SparseOperator(MatrixOperator):
    """An operator held as a CSR matrix (reference emu_sv/sparse_operator.py:31-71)."""

    @classmethod
    def _from_operator_repr(cls, *, eigenstates, n_qudits, operations):
        matrix = cls._build(eigenstates=eigenstates, n_qudits=n_qudits, operations=operations,
                            sparse=True)
        return SparseOperator(matrix), operations
