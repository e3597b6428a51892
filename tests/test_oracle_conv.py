"""Pins for the oracle's Orszag convolver and LPT sequence against closed forms
and an independent dealiased product (numpy FFT at twice the resolution)."""
import numpy as np
import pytest

from conftest import rel_max

BOX = 100.0


def bandlimited(n, seed):
    """real field with all Nyquist planes zero (exactly representable on the grid)"""
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, n, n))
    k = np.fft.rfftn(x)
    k[n // 2, :, :] = 0
    k[:, n // 2, :] = 0
    k[:, :, n // 2] = 0
    return np.fft.irfftn(k, s=(n, n, n))


def to_kgrid(orc, x):
    n = x.shape[0]
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = x
    orc.fft_r2c(a)
    return a


def exact_product_spectrum(x, y):
    """Fourier coefficients of the continuous product of two band-limited
    fields, from a 2n-point evaluation (no aliasing), in the n-grid unitary
    normalisation, truncated to |k_d| < n/2 and Nyquist zeroed"""
    n = x.shape[0]
    m = 2 * n

    def up(f):
        F = np.fft.fftn(f)
        G = np.zeros((m, m, m), dtype=complex)
        idx = np.r_[0:n // 2, m - n // 2 + 1:m]
        src = np.r_[0:n // 2, n - n // 2 + 1:n]
        G[np.ix_(idx, idx, idx)] = F[np.ix_(src, src, src)]
        return np.fft.ifftn(G).real * (m / n) ** 3

    P = np.fft.fftn(up(x) * up(y)) / m ** 3  # continuous coefficients P_k
    out = np.zeros((n, n, n // 2 + 1), dtype=complex)
    ki = np.r_[0:n // 2 + 1, -(n // 2) + 1:0]
    for a, kx in enumerate(ki):
        for b, ky in enumerate(ki):
            out[a, b, :] = P[kx % m, ky % m, : n // 2 + 1] * np.sqrt(n ** 3)
    out[n // 2, :, :] = 0
    out[:, n // 2, :] = 0
    out[:, :, n // 2] = 0
    return out


@pytest.mark.parametrize("n", [8, 12, 16])
def test_multiply_field_is_dealiased_product(orc, n):
    x, y = bandlimited(n, 1), bandlimited(n, 2)
    A, B = to_kgrid(orc, x), to_kgrid(orc, y)
    out = np.zeros_like(A)
    orc.convolve2(A, orc.mult("identity"), B, orc.mult("identity"), out, orc.ORC_OP_ASSIGN, 0, BOX)
    ref = exact_product_spectrum(x, y)
    assert np.max(np.abs(orc.kview(out) - ref)) < 1e-12 * np.max(np.abs(ref))


def test_hessian_product_plane_waves(orc):
    """phi = cos(kx x) + cos(ky y): phi_xx * phi_yy = kx^2 ky^2 cos(kx x) cos(ky y)"""
    n = 16
    g = np.arange(n) * BOX / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    k0 = 2 * np.pi / BOX
    kx, ky = 2 * k0, 3 * k0
    phi = to_kgrid(orc, np.cos(kx * X) + np.cos(ky * Y))
    out = np.zeros_like(phi)
    orc.convolve2(phi, orc.mult("hessian", 0, 0), phi, orc.mult("hessian", 1, 1), out, orc.ORC_OP_ASSIGN, 0, BOX)
    orc.fft_c2r(out)
    ref = kx ** 2 * ky ** 2 * np.cos(kx * X) * np.cos(ky * Y)
    assert np.max(np.abs(out[:, :, :n] - ref)) < 1e-12 * np.max(np.abs(ref))


def test_2lpt_of_single_plane_wave_vanishes(orc):
    """for a 1D wave every 2LPT term phi_ii phi_jj - phi_ij^2 is zero"""
    n = 16
    g = np.arange(n) * BOX / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    k0 = 2 * np.pi / BOX
    phi = to_kgrid(orc, np.cos(k0 * (3 * X + 2 * Y + Z)))
    phi2, _, _ = orc.lpt_terms(phi, BOX, 2)
    assert np.max(np.abs(phi2)) < 1e-12


def test_2lpt_two_orthogonal_waves(orc):
    """phi = A cos(kx x) + B cos(ky y): the 2LPT source is phi_xx phi_yy and
    phi2 = -source/|k|^2 mode by mode"""
    n = 16
    g = np.arange(n) * BOX / n
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    k0 = 2 * np.pi / BOX
    kx, ky = 2 * k0, 3 * k0
    phi = to_kgrid(orc, 0.7 * np.cos(kx * X) + 1.3 * np.cos(ky * Y))
    phi2, _, _ = orc.lpt_terms(phi, BOX, 2)
    orc.fft_c2r(phi2)
    # source = 0.7*1.3 kx^2 ky^2 cos cos = (c/2)[cos(kx x + ky y) + cos(kx x - ky y)], |k|^2 = kx^2+ky^2
    ref = -0.7 * 1.3 * kx ** 2 * ky ** 2 * np.cos(kx * X) * np.cos(ky * Y) / (kx ** 2 + ky ** 2)
    assert np.max(np.abs(phi2[:, :, :n] - ref)) < 1e-12 * np.max(np.abs(ref))


def test_convolve3_is_two_truncated_products(orc):
    """convolve3 = convolve2(trunc(a*b), c) -- the intermediate is dealiased to
    the n-grid (convolution.hh:542-549), not a direct triple product"""
    n = 12
    x, y, z = bandlimited(n, 4), bandlimited(n, 5), bandlimited(n, 6)
    A, B, C = (to_kgrid(orc, v) for v in (x, y, z))
    out = np.zeros_like(A)
    I = orc.mult("identity")
    orc.convolve3(A, I, B, I, C, I, out, orc.ORC_OP_ASSIGN, 0, BOX)
    ab = exact_product_spectrum(x, y)
    abk = np.zeros_like(A)
    orc.kview(abk)[...] = ab
    orc.fft_c2r(abk)
    ref = exact_product_spectrum(abk[:, :, :n], z)
    assert rel_max(orc.kview(out), ref) < 1e-12
