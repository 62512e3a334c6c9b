"""Operators for custom observables, after the reference's
test/emu_sv/test_dense_operator.py and test_sparse_operator.py (the same
operator representations against torch.kron of the Pauli matrices), plus
closed forms for the CSR path above 14 qubits and the serialised
Expectation observable."""

from functools import reduce

import pytest
import torch

from emulators_amd.emu_sv import (DenseOperator, Expectation, SparseOperator, StateVector,
                                  SVConfig)
from emulators_amd.emu_sv.matrix_operator import MatrixOperator
from emulators_amd.emu_sv.sparse_operator import sparse_add, sparse_kron

c128 = torch.complex128
X = torch.tensor([[0, 1], [1, 0]], dtype=c128)
Y = torch.tensor([[0, -1j], [1j, 0]], dtype=c128)
Z = torch.tensor([[1, 0], [0, -1]], dtype=c128)
Id = torch.eye(2, dtype=c128)

OPERATORS = (DenseOperator, SparseOperator)


def _dense(op):
    d = op.data.cpu()
    return d.to_dense() if d.is_sparse_csr else d


@pytest.mark.parametrize("cls", OPERATORS)
def test_from_operator_repr_and_rmul(cls):
    """test_dense_operator.py:16-40: {"gr": 1, "rg": -1} is iY on qubit 0
    (|g><r| - |r><g| with g = index 0), Z on qubit 1, then 3 * op."""
    ops = [(1.0, [({"gr": 1.0, "rg": -1.0}, {0}), ({"gg": 1.0, "rr": -1.0}, {1})])]
    op = cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=2, operations=ops)
    assert type(op) is cls
    assert torch.allclose(_dense(op), torch.kron(1j * Y, Z))
    op3 = 3 * op
    assert type(op3) is cls
    assert torch.allclose(_dense(op3), 3 * torch.kron(1j * Y, Z))
    assert op._n_qudits == 2 and op._eigenstates == ("r", "g")


@pytest.mark.parametrize("cls", OPERATORS)
def test_from_operator_repr_errors(cls):
    """test_dense_operator.py:43-76 and :164-181: the XY basis is not
    implemented, an unknown basis and a bad key raise ValueError (pulser's
    message for the key)."""
    ops01 = [(1.0, [({"01": 1.0, "10": -1.0}, {0}), ({"00": 1.0, "11": -1.0}, {1})])]
    with pytest.raises(NotImplementedError):
        cls.from_operator_repr(eigenstates=("0", "1"), n_qudits=2, operations=ops01)
    opsab = [(1.0, [({"ab": 1.0, "ba": -1.0}, {0}), ({"aa": 1.0, "bb": -1.0}, {1})])]
    with pytest.raises(ValueError):
        cls.from_operator_repr(eigenstates=("a", "b"), n_qudits=2, operations=opsab)
    bad = [(1.0, [({"X": 2.0}, [0, 2]), ({"Z": 3.0}, [1])])]
    with pytest.raises(ValueError) as ve:
        cls.from_operator_repr(eigenstates=("g", "1"), n_qudits=3, operations=bad)
    assert str(ve.value) == ("Every QuditOp key must be made up of two eigenstates among "
                             "('g', '1'); instead, got 'X'.")
    with pytest.raises(ValueError):
        cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=2,
                               operations=[(1.0, [({"rr": 1.0}, [2])])])
    with pytest.raises(ValueError):
        cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=2,
                               operations=[(1.0, [({"rr": 1.0}, [0]), ({"gg": 1.0}, [0])])])


@pytest.mark.parametrize("cls", OPERATORS)
def test_add_and_matmul(cls):
    """test_dense_operator.py:79-126 / test_sparse_operator.py:42-77:
    2 X_0 + Z_1; the product only for dense operators."""
    o1 = cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=2,
                                operations=[(2.0, [({"gr": 1.0, "rg": 1.0}, [0])])])
    o2 = cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=2,
                                operations=[(1.0, [({"gg": 1.0, "rr": -1.0}, [1])])])
    s = o1 + o2
    assert isinstance(s, cls)
    assert torch.allclose(_dense(s), torch.kron(2 * X, Id) + torch.kron(Id, Z))
    if cls is DenseOperator:
        assert torch.allclose(_dense(o1 @ o2), torch.kron(2 * X, Id) @ torch.kron(Id, Z))
    else:
        with pytest.raises(NotImplementedError):
            o1 @ o2
    with pytest.raises(AssertionError):
        o1 + (DenseOperator if cls is SparseOperator else SparseOperator).from_operator_repr(
            eigenstates=("r", "g"), n_qudits=2, operations=[(1.0, [({"rr": 1.0}, [0])])])


@pytest.mark.parametrize("cls", OPERATORS)
def test_apply_to_and_expect(cls):
    """test_dense_operator.py:129-161: X_0 Z_2 on (|ggg> - |rrr>)/sqrt 2."""
    ops = [(1.0, [({"rg": 1.0, "gr": 1.0}, {0}), ({"gg": 1.0, "rr": -1.0}, {2})])]
    op = cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=3, operations=ops)
    state = StateVector.from_state_amplitudes(eigenstates=("r", "g"),
                                              amplitudes={"rrr": -1.0, "ggg": 1.0})
    psi = torch.zeros(8, dtype=c128)
    psi[0], psi[7] = 1.0, -1.0
    psi /= psi.norm()
    want = reduce(torch.kron, [X, Id, Z]) @ psi
    got = op.apply_to(state)
    assert isinstance(got, StateVector)
    assert torch.allclose(got.data.cpu(), want)
    e = op.expect(state)
    assert e.dtype == c128 and e.device.type == "cpu"
    assert torch.isclose(e, torch.vdot(psi, want))
    with pytest.raises(AssertionError):
        op.apply_to(psi)


def test_sparse_kron_and_add():
    """test_sparse_operator.py:144-164: against torch.kron and + on random
    sparse matrices (repeated positions overwrite in the dense fill)."""
    g = torch.Generator().manual_seed(5)
    size, density = 100, 4
    a = torch.zeros(size, size)
    b = torch.zeros(size, size)
    inds = torch.randint(0, size, (2, density * size, 2), generator=g)
    vals = torch.randn((2, density * size), generator=g)
    for i in range(density * size):
        a[inds[0, i, 0], inds[0, i, 1]] = vals[0, i]
        b[inds[1, i, 0], inds[1, i, 1]] = vals[1, i]
    k = sparse_kron(a.to_sparse_coo(), b.to_sparse_coo())
    assert k.layout == torch.sparse_coo and k.is_coalesced()
    assert torch.allclose(k.to_dense(), torch.kron(a, b))
    s = sparse_add(a.to_sparse_coo(), b.to_sparse_coo())
    assert s.layout == torch.sparse_coo and s.is_coalesced()
    assert torch.allclose(s.to_dense(), a + b)


def test_large_operator_is_sparse_with_closed_forms():
    """Above 14 qubits the base class builds CSR (matrix_operator.py:251-258).
    Closed forms at n = 15: sum_i n_i |x> = popcount(x) |x>, and
    sum_i X_i |g..g> = sum_i |r_i>; 2^15 - 1 nonzeros (none at x = 0) and
    15 * 2^15 (distinct positions per term)."""
    n = 15
    occ = MatrixOperator.from_operator_repr(eigenstates=("r", "g"), n_qudits=n,
                                            operations=[(1.0, [({"rr": 1.0}, [i])])
                                                        for i in range(n)])
    assert occ.data.is_sparse_csr
    assert occ.data.values().numel() == (1 << n) - 1
    g = torch.Generator().manual_seed(11)
    psi = torch.randn(1 << n, dtype=c128, generator=g)
    pc = torch.tensor([bin(x).count("1") for x in range(1 << n)], dtype=torch.float64)
    assert torch.allclose(occ.apply_to(StateVector(psi, gpu=False)).data.cpu(),
                          pc.to(c128) * psi)
    xs = SparseOperator.from_operator_repr(eigenstates=("r", "g"), n_qudits=n,
                                           operations=[(1.0, [({"rg": 1.0, "gr": 1.0}, [i])])
                                                       for i in range(n)])
    assert xs.data.values().numel() == n << n
    out = xs.apply_to(StateVector.make(n, gpu=False)).data.cpu()
    want = torch.zeros(1 << n, dtype=c128)
    for i in range(n):
        want[1 << (n - 1 - i)] = 1.0  # qubit i is bit n-1-i
    assert torch.equal(out, want)


def test_expectation_serialises_with_its_operator():
    """An Expectation observable survives to_abstract_repr / from_abstract_repr
    with its operator's class, representation and tag suffix."""
    ops = [(0.5 + 0.25j, [({"rg": 1.0, "gr": 1.0}, [0, 1])]), (2.0, [({"rr": 1.0}, [2])])]
    for cls in OPERATORS:
        op = cls.from_operator_repr(eigenstates=("r", "g"), n_qudits=3, operations=ops)
        cfg = SVConfig(observables=[Expectation(op, evaluation_times=[0.5, 1.0],
                                                tag_suffix="x")], gpu=False)
        back = SVConfig.from_abstract_repr(cfg.to_abstract_repr())
        (obs,) = back.observables
        assert isinstance(obs, Expectation) and obs.tag == "expectation_x"
        assert list(obs.evaluation_times) == [0.5, 1.0]
        assert type(obs.operator) is cls
        assert torch.allclose(_dense(obs.operator), _dense(op))
