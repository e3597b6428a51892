"""Pins for the oracle's k-space operators and power spectrum: identities of
the mathematics, written independently of the oracle's code."""
import numpy as np
import pytest

from helpers import random_kspace_field, random_real_grid

BOX = 100.0


def kvec(n, boxlen):
    """get_k per axis (grid_fft.hh:282-296): (i - [i > n/2] n) kfac, +n/2 kept"""
    kfac = 2 * np.pi / boxlen
    i = np.arange(n)
    kx = (i - (i > n // 2) * n) * kfac
    kz = np.arange(n // 2 + 1) * kfac
    return kx, kz


def k2_grid(n, boxlen):
    kx, kz = kvec(n, boxlen)
    return kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2


@pytest.mark.parametrize("n", [8, 16, 24])
def test_inverse_laplacian_inverts_laplacian(orc, n):
    a = random_kspace_field(orc, n, 4)
    f = orc.kview(a).copy()
    r = a.copy()
    orc.inverse_laplacian(r, BOX)
    g = orc.kview(r)
    lap = -k2_grid(n, BOX) * g  # Laplacian of the result
    f[0, 0, 0] = 0
    assert g[0, 0, 0] == 0
    assert np.max(np.abs(lap - f)) < 1e-12 * np.max(np.abs(f))


def test_noise_normalise_fixing_and_inversion(orc):
    n = 16
    a = random_real_grid(n, 5)
    orc.fft_r2c(a)
    vol = 0.37
    for fixing in (False, True):
        for inversion in (False, True):
            r = a.copy()
            orc.noise_normalise(r, vol, fixing, inversion)
            x, y = orc.kview(a), orc.kview(r)
            expect = x / vol
            if fixing:
                nz = np.abs(x) > 0
                expect = np.where(nz, x / np.where(nz, np.abs(x), 1) / vol, 0)
            if inversion:
                expect = -expect
            assert np.allclose(y, expect, rtol=1e-14, atol=1e-300)


def test_phi1_is_minus_amplitude_over_k2(orc):
    n = 16
    a = random_real_grid(n, 6)
    orc.fft_r2c(a)
    cos = orc.Cosmology()
    table = cos.amplitude_table(n, BOX)
    r = np.zeros_like(a)
    orc.phi1(r, a, BOX, table)
    k2 = k2_grid(n, BOX)
    kmag = np.sqrt(k2)
    amp = np.array([cos.amplitude(k) for k in kmag.ravel()]).reshape(kmag.shape)
    k2[0, 0, 0] = 1
    expect = -orc.kview(a) * amp / k2
    expect[0, 0, 0] = 0
    got = orc.kview(r)
    assert np.max(np.abs(got - expect)) < 1e-12 * np.max(np.abs(expect))


def test_power_spectrum_of_white_noise(orc):
    """unit-variance real white noise under the unitary transform has
    E|X_k|^2 = 1 for k != 0, so P(k) = L^3/(2 pi)^3 / n^3 in every bin"""
    n = 64
    Ps = []
    for seed in range(4):
        a = random_real_grid(n, 100 + seed)
        orc.fft_r2c(a)
        k, P, eP, c = orc.power_spectrum(a, BOX)
        Ps.append(P)
        dk = 2 * np.pi / BOX
        sel = c > 0
        b = np.arange(len(k))[sel]
        # mean |k| of each bin lies inside the bin
        assert np.all(k[sel] >= dk * b - 1e-12) and np.all(k[sel] < dk * (b + 1))
    P = np.mean(Ps, axis=0)
    expect = BOX ** 3 / (2 * np.pi) ** 3 / n ** 3
    good = c > 200
    assert abs(np.mean(P[good]) / expect - 1) < 0.01
    assert np.max(np.abs(P[good] / expect - 1)) < 0.2
    # counts: every mode with dk <= |k| < kmax counted once, kz > 0 twice
    kx, kz = kvec(n, BOX)
    kk = np.sqrt(k2_grid(n, BOX))
    w = np.where(np.arange(n // 2 + 1) == 0, 1, 2)[None, None, :] * np.ones_like(kk)
    dk = 2 * np.pi / BOX
    m = (kk >= dk) & (kk < dk * (n // 2))
    idx = (kk[m] / dk).astype(int)
    cnt = np.bincount(idx, weights=w[m], minlength=len(c))[:len(c)]
    assert np.array_equal(cnt, c)


def band_limited(shape, waves, seed=3):
    """a real field on the unit cube sampled on `shape`, sum of cosines of the
    integer wave vectors `waves`, and a function evaluating it on any grid"""
    rng = np.random.default_rng(seed)
    amp = rng.standard_normal(len(waves))
    pha = rng.uniform(0, 2 * np.pi, len(waves))

    def on(sh):
        x = np.meshgrid(*[np.arange(m) / m for m in sh], indexing="ij")
        return sum(a * np.cos(2 * np.pi * (w[0] * x[0] + w[1] * x[1] + w[2] * x[2]) + p)
                   for a, w, p in zip(amp, waves, pha))
    return on(shape), on


@pytest.mark.parametrize("fshape,tshape", [((16, 16, 16), (24, 24, 24)), ((24, 24, 24), (16, 16, 16)),
                                           ((16, 20, 12), (8, 30, 16)), ((8, 8, 8), (32, 32, 32))])
def test_fourier_copy_band_limited(orc, fshape, tshape):
    """FourierInterpolateCopyTo keeps every mode below both grids' Nyquist
    modes with its unitary amplitude, so a band-limited field comes out as the
    same function sampled on the new grid, times sqrt(N_from / N_to); numpy's
    FFT transforms both ends.  The last case more than doubles each side."""
    lim = [min(a, b) // 2 - 1 for a, b in zip(fshape, tshape)]
    waves = [(0, 0, 0), (lim[0], 0, 1), (-lim[0], lim[1], lim[2]), (1, -lim[1], 2), (2, 1, -lim[2])]
    f, on = band_limited(fshape, waves)
    Nf, Nt = np.prod(fshape), np.prod(tshape)
    F = np.fft.rfftn(f) / np.sqrt(Nf)
    T = orc.fourier_copy(F, tshape)
    t = np.fft.irfftn(T, s=tshape) * np.sqrt(Nt)
    np.testing.assert_allclose(t, on(tshape) * np.sqrt(Nf / Nt), rtol=0, atol=1e-12)
    # Nyquist planes of the destination stay zero
    assert np.all(T[tshape[0] // 2] == 0) and np.all(T[:, tshape[1] // 2] == 0) and np.all(T[:, :, -1] == 0)
