"""Host-side pieces of the whole-pulse adjoint (no GPU needed)."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg


def _random_lanczos_like(m: int, sigma: complex, rng) -> np.ndarray:
    """sigma * J with J real symmetric tridiagonal (the form of a Hermitian
    Lanczos matrix of op = sigma H)."""
    d = rng.normal(size=m)
    e = np.abs(rng.normal(size=m - 1)) + 0.1
    j = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    return sigma * j


def _quadrature(tb, tk, ket_forward):
    x, w = np.polynomial.legendre.leggauss(80)
    u = 0.5 * (x + 1.0)
    w = 0.5 * w
    e0b = np.zeros(tb.shape[0]); e0b[0] = 1.0
    e0k = np.zeros(tk.shape[0]); e0k[0] = 1.0
    acc = np.zeros((tb.shape[0], tk.shape[0]), dtype=np.complex128)
    for ui, wi in zip(u, w):
        a = scipy.linalg.expm(ui * tb) @ e0b
        b = scipy.linalg.expm(((1.0 - ui) if ket_forward else ui) * tk) @ e0k
        acc += wi * np.outer(np.conj(a), b)
    return acc


@pytest.mark.parametrize("ket_forward", [False, True])
@pytest.mark.parametrize("mb,mk", [(1, 1), (3, 5), (7, 4)])
def test_frechet_weights_vs_quadrature(ket_forward, mb, mk):
    """X[l,k] = int_0^1 conj([e^{u Tb} e0]_l) [e^{g(u) Tk} e0]_k du against
    80-point Gauss-Legendre quadrature of the same integrand."""
    from emulators_amd.emu_sv.adjoint import frechet_weights

    rng = np.random.default_rng(mb * 10 + mk)
    tb = _random_lanczos_like(mb, 0.7j, rng)
    tk = _random_lanczos_like(mk, -0.7j if ket_forward else 0.7j, rng)
    x = frechet_weights(tb, tk, ket_forward=ket_forward)
    ref = _quadrature(tb, tk, ket_forward)
    np.testing.assert_allclose(x, ref, atol=1e-13, rtol=0)


def test_frechet_weights_scalar_closed_form():
    """1x1: int_0^1 conj(e^{u a}) e^{u b} du = (e^{conj(a)+b} - 1)/(conj(a)+b)."""
    from emulators_amd.emu_sv.adjoint import frechet_weights

    a, b = 0.3 + 1.1j, -0.2 + 0.4j
    x = frechet_weights(np.array([[a]]), np.array([[b]]))
    s = np.conj(a) + b
    assert abs(x[0, 0] - (np.exp(s) - 1.0) / s) < 1e-14
