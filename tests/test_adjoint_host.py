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


def _lanczos_form(m: int, sigma: complex, rng) -> np.ndarray:
    """W^-1 (sigma J) W with W = diag(w^k), w = sigma/|sigma|: the structure of
    a Hermitian Lanczos matrix of op = sigma H (real positive subdiagonal)."""
    j = np.real(_random_lanczos_like(m, 1.0, rng))
    w = sigma / abs(sigma)
    wd = w ** np.arange(m)
    return (np.diag(1.0 / wd) @ (sigma * j) @ np.diag(wd))


@pytest.mark.parametrize("ket_forward", [False, True])
@pytest.mark.parametrize("mb,mk", [(1, 1), (4, 6), (9, 3), (25, 30)])
def test_frechet_weights_eig_matches_block_exponential(ket_forward, mb, mk):
    """The eigen-decomposition weights equal the Van Loan block-exponential
    weights for Lanczos-structured matrices (including complex sigma)."""
    from emulators_amd.emu_sv.adjoint import frechet_weights, frechet_weights_eig

    rng = np.random.default_rng(7 * mb + mk)
    sb = 0.3j if ket_forward else 0.25j + 0.05
    sk = -0.3j if ket_forward else 0.25j + 0.05
    tb = _lanczos_form(mb, sb, rng)
    tk = _lanczos_form(mk, sk, rng)
    ref = frechet_weights(tb, tk, ket_forward=ket_forward)
    got = frechet_weights_eig(tb, sb, tk, sk, ket_forward=ket_forward)
    np.testing.assert_allclose(got, ref, atol=1e-12 * max(1.0, np.abs(ref).max()), rtol=0)


def test_lanczos_trajectory_reproduces_expm():
    from emulators_amd.emu_sv.adjoint import lanczos_trajectory

    rng = np.random.default_rng(2)
    sigma = -0.4j
    t = _lanczos_form(12, sigma, rng)
    a, r = lanczos_trajectory(t, sigma)
    for u in (0.0, 0.3, 1.0):
        ref = scipy.linalg.expm(u * t)[:, 0]
        np.testing.assert_allclose(a @ np.exp(u * r), ref, atol=1e-13, rtol=0)
