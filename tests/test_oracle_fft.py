"""Pins for the oracle's FFT (FFTW is not available): numpy's pocketfft in
FP64, a brute-force O(N^2) DFT, Parseval, plane waves and the c2r Hermitian
semantics."""
import numpy as np
import pytest

from helpers import random_real_grid


def brute_dft(x, sign):
    n = x.size
    j = np.arange(n)
    W = np.exp(sign * 2j * np.pi * np.outer(j, j) / n)
    return W @ x


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 9, 12, 16, 18, 27, 48, 96, 100])
@pytest.mark.parametrize("sign", [-1, 1])
def test_dft1d_vs_brute(orc, n, sign):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    got = orc.dft1d(x, sign)
    ref = brute_dft(x, sign)
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [4, 6, 8, 12, 16, 24, 48])
def test_r2c_vs_numpy(orc, n):
    a = random_real_grid(n, n)
    ref = np.fft.rfftn(a[:, :, :n]) / np.sqrt(n ** 3)
    orc.fft_r2c(a)
    assert np.max(np.abs(orc.kview(a) - ref)) < 1e-13


@pytest.mark.parametrize("n", [6, 12])
def test_r2c_vs_brute_3d(orc, n):
    a = random_real_grid(n, 7)
    x = a[:, :, :n]
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    ref = np.einsum("ia,jb,kc,abc->ijk", W, W, W[: n // 2 + 1], x) / np.sqrt(n ** 3)
    orc.fft_r2c(a)
    assert np.max(np.abs(orc.kview(a) - ref)) < 1e-12


@pytest.mark.parametrize("n", [8, 18, 32])
def test_roundtrip_and_parseval(orc, n):
    a = random_real_grid(n, 3)
    x = a[:, :, :n].copy()
    orc.fft_r2c(a)
    k = orc.kview(a)
    w = np.full(k.shape, 2.0)
    w[:, :, 0] = 1.0
    w[:, :, -1] = 1.0
    assert abs(np.sum(w * np.abs(k) ** 2) - np.sum(x ** 2)) < 1e-10 * np.sum(x ** 2)
    orc.fft_c2r(a)
    assert np.max(np.abs(a[:, :, :n] - x)) < 1e-13


def test_plane_wave(orc):
    n = 16
    g = np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = np.cos(2 * np.pi * (2 * X + 3 * Y + 5 * Z) / n)
    orc.fft_r2c(a)
    k = orc.kview(a)
    assert abs(k[2, 3, 5] - np.sqrt(n ** 3) / 2) < 1e-12
    k[2, 3, 5] = 0
    assert np.max(np.abs(k)) < 1e-12


def test_c2r_ignores_imaginary_dc_nyquist(orc):
    """FFTW c2r semantics: the z-line c2r drops Im X[0] and Im X[n/2], i.e. the
    result equals the Hermitian-symmetrised input"""
    n = 8
    rng = np.random.default_rng(0)
    a = rng.standard_normal((n, n, n + 2))
    b = a.copy()
    orc.fft_c2r(a)
    kb = orc.kview(b)
    # symmetrise the kz=0 and kz=n/2 planes explicitly
    for kz in (0, n // 2):
        P = kb[:, :, kz]
        idx = (-np.arange(n)) % n
        kb[:, :, kz] = 0.5 * (P + np.conj(P[idx][:, idx]))
    full = np.zeros((n, n, n), dtype=complex)
    full[:, :, : n // 2 + 1] = kb
    for kz in range(1, n // 2):
        full[:, :, n - kz] = np.conj(kb[(-np.arange(n)) % n][:, (-np.arange(n)) % n, kz])
    ref = np.fft.ifftn(full).real * n ** 3 / np.sqrt(n ** 3)
    assert np.max(np.abs(a[:, :, :n] - ref)) < 1e-12
