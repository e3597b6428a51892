"""k-space operators of Grid_FFT / ic_generator::run and Compute_PowerSpectrum
on the GPU against the oracle (grid_fft.hh:392-466,844-887; grid_fft.cc:1133-1210;
ic_generator.cc:341-346,450-457)."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms
from helpers import random_kspace_field, random_real_grid

pytestmark = pytest.mark.gpu

BOX = 100.0


def _kgrid(M, ctx, n, prec, a):
    g = M.Grid_FFT(ctx, n, BOX, prec, initialspace=M.kspace_id)
    g.upload(a)
    return g


@pytest.mark.parametrize("n,prec,tol", [(32, "double", 1e-14), (48, "double", 1e-14), (32, "float", 1e-6)])
def test_inverse_laplacian_matches_oracle(ctx, orc, n, prec, tol):
    import monofonic_amd as M
    print(f"inverse laplacian n={n} {prec}", flush=True)
    a = random_kspace_field(orc, n, 21)
    g = _kgrid(M, ctx, n, prec, a)
    g.apply_InverseLaplacian()
    r = a.copy()
    orc.inverse_laplacian(r, BOX)
    assert rel_rms(g.kdata(), orc.kview(r)) < tol
    assert rel_max(g.kdata(), orc.kview(r)) < 10 * tol


@pytest.mark.parametrize("n", [16, 24])
def test_zero_dc_scale_stats(ctx, orc, n):
    import monofonic_amd as M
    a = random_real_grid(n, 3)
    a[:, :, :n] += 0.25
    g = M.Grid_FFT(ctx, n, BOX, "double")
    g.upload(a)
    x = a[:, :, :n]
    m, s = g.mean(), g.std()
    assert abs(m - x.mean()) < 1e-13
    assert abs(s - x.std()) < 1e-12
    g *= 3.0
    assert np.allclose(g.rdata(), 3.0 * x, rtol=0, atol=1e-13)
    g /= 3.0
    g.FourierTransformForward()
    k = g.kdata()
    assert abs(k[0, 0, 0]) > 1.0
    g.zero_DC_mode()
    k2 = g.kdata()
    assert k2[0, 0, 0] == 0
    assert np.array_equal(k2.ravel()[1:], k.ravel()[1:])


@pytest.mark.parametrize("fixing,inversion", [(False, False), (True, False), (False, True), (True, True)])
def test_noise_normalise_matches_oracle(ctx, orc, fixing, inversion):
    import monofonic_amd as M
    n = 32
    a = random_real_grid(n, 8)
    orc.fft_r2c(a)
    volfac = (BOX / n / 2.0 / np.pi) ** 1.5
    g = _kgrid(M, ctx, n, "double", a)
    M.noise_normalise(g, volfac, fixing, inversion)
    r = a.copy()
    orc.noise_normalise(r, volfac, fixing, inversion)
    assert rel_max(g.kdata(), orc.kview(r)) < 1e-14


@pytest.mark.parametrize("n,prec,tol", [(32, "double", 1e-14), (64, "float", 1e-6)])
def test_phi1_matches_oracle(ctx, orc, n, prec, tol):
    import monofonic_amd as M
    a = random_real_grid(n, 9)
    orc.fft_r2c(a)
    table = orc.Cosmology().amplitude_table(n, BOX)
    wn = _kgrid(M, ctx, n, prec, a)
    phi = M.Grid_FFT(ctx, n, BOX, prec, initialspace=M.kspace_id)
    M.phi1(phi, wn, table)
    r = np.zeros_like(a)
    orc.phi1(r, a, BOX, table)
    assert rel_rms(phi.kdata(), orc.kview(r)) < tol
    assert rel_max(phi.kdata(), orc.kview(r)) < 10 * tol


@pytest.mark.parametrize("n,prec,tol", [(32, "double", 1e-13), (64, "double", 1e-13), (64, "float", 1e-5)])
def test_power_spectrum_matches_oracle(ctx, orc, n, prec, tol):
    import monofonic_amd as M
    print(f"power spectrum n={n} {prec}", flush=True)
    a = random_kspace_field(orc, n, 13, slope=-1.0)
    g = _kgrid(M, ctx, n, prec, a)
    k, P, eP, c = g.Compute_PowerSpectrum()
    rk, rP, reP, rc = orc.power_spectrum(a, BOX)
    assert len(k) == len(rk)
    assert np.array_equal(c, rc)
    sel = rc > 0
    assert np.max(np.abs(k[sel] / rk[sel] - 1)) < tol
    assert np.max(np.abs(P[sel] / rP[sel] - 1)) < tol
    assert np.max(np.abs(eP[sel] - reP[sel]) / rP[sel]) < 100 * tol


@pytest.mark.parametrize("prec", ["double", "float"])
def test_stats_in_k_space(ctx, prec):
    """std() of a k-space grid over its n0 n1 (n2/2+1) modes -- real parts
    and |z|^2 -- and mean() over relem(i,j,k), k < n2/2+1, the first n2/2+1
    reals of each padded row (grid_fft.hh:507-590)"""
    import monofonic_amd as M
    n = 24
    rng = np.random.default_rng(11)
    g = M.Grid_FFT(ctx, n, 1.0, prec, initialspace=M.kspace_id)
    raw = rng.standard_normal(n * n * (n + 2)) + 0.3
    g.upload(raw)
    x = raw.reshape(n, n, n + 2)
    z = x[:, :, 0::2] + 1j * x[:, :, 1::2]
    s1, s2, cnt = np.sum(z.real), np.sum(np.abs(z) ** 2), z.size
    want_sd = np.sqrt(s2 / cnt - (s1 / cnt) ** 2)
    want_mean = np.sum(x[:, :, : n // 2 + 1]) / (n * n * (n // 2 + 1))
    tol = 1e-12 if prec == "double" else 1e-6
    assert abs(g.std() - want_sd) < tol * want_sd
    assert abs(g.mean() - want_mean) < tol * max(1.0, abs(want_mean))


@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("float", 2e-6)])
def test_shift_field(ctx, orc, prec, tol):
    """Grid_FFT::shift_field (grid_fft.hh:832-842): a real field shifted by
    half cells against the oracle; a whole-cell shift is the cyclic roll
    f(x + s dx); on a k-space grid of unequal sides (transform_back false) the
    modes times exp(2 pi i sum_d s_d k_d / n_d), k_d signed as get_k forms it"""
    import monofonic_amd as M
    n, L = 32, 100.0
    print(f"shift_field {prec}", flush=True)
    rng = np.random.default_rng(5)
    f = np.zeros((n, n, n + 2))
    f[:, :, :n] = rng.standard_normal((n, n, n))
    s = (0.5, 0.25, -0.75)
    ref = orc.shift_field(f.copy(), L, s)
    g = M.Grid_FFT(ctx, n, L, prec)
    g.upload(f)
    g.shift_field(s)
    assert g.space_ == M.rspace_id
    got = g.download()
    assert rel_rms(got[:, :, :n], ref[:, :, :n]) < tol
    assert rel_max(got[:, :, :n], ref[:, :, :n]) < 10 * tol
    # whole cells: a cyclic roll
    g.upload(f)
    g.shift_field((1, 0, -2))
    want = np.roll(f[:, :, :n], shift=(-1, 0, 2), axis=(0, 1, 2))
    assert np.max(np.abs(g.download()[:, :, :n] - want)) < 20 * tol
    # k space, unequal sides
    shape = (16, 24, 32)
    h = M.Grid_FFT(ctx, shape, (1.0, 2.0, 3.0), prec)
    x = np.zeros((16, 24, 34))
    x[:, :, :32] = rng.standard_normal(shape)
    h.upload(x)
    h.FourierTransformForward()
    F = h.kdata().astype(np.complex128)
    s = (0.3, -1.5, 2.25)
    h.shift_field(s, transform_back=False)
    assert h.space_ == M.kspace_id
    idx = [np.arange(m) - (np.arange(m) > m // 2) * m for m in shape]
    ph = sum(s[d] * idx[d].reshape([-1 if e == d else 1 for e in range(3)]) / shape[d] for d in range(3))
    want = F * np.exp(2j * np.pi * ph[:, :, : shape[2] // 2 + 1])
    assert np.max(np.abs(h.kdata() - want)) < tol * np.max(np.abs(F)) * 10


@pytest.mark.parametrize("prec", ["double", "float"])
def test_fourier_interpolate_copy(ctx, orc, prec):
    """Grid_FFT::FourierInterpolateCopyTo (grid_fft.cc:274-333) between grids
    of other sides: the destination's modes equal the oracle's copy of the
    source's modes exactly (values are copied), the destination is zeroed
    first, equal sides copy the buffer, and a band-limited field arrives as the
    same function on the new grid times sqrt(N_from / N_to)"""
    import monofonic_amd as M
    from test_oracle_ops import band_limited
    print(f"fourier interpolate copy {prec}", flush=True)
    rng = np.random.default_rng(9)
    for fs, ts in (((16, 16, 16), (24, 24, 24)), ((24, 20, 16), (16, 30, 12)), ((8, 8, 8), (32, 32, 32))):
        a = M.Grid_FFT(ctx, fs, 1.0, prec)
        x = np.zeros((fs[0], fs[1], fs[2] + 2))
        x[:, :, :fs[2]] = rng.standard_normal(fs)
        a.upload(x)
        b = M.Grid_FFT(ctx, ts, 1.0, prec)
        b.upload(np.full((ts[0], ts[1], ts[2] + 2), 5.0))  # stale contents must not survive
        a.FourierInterpolateCopyTo(b)
        assert a.space_ == M.kspace_id and b.space_ == M.kspace_id
        want = orc.fourier_copy(a.kdata(), ts)
        assert np.array_equal(b.kdata().astype(np.complex128), want), (fs, ts)
        # band-limited field: the same function on the new grid
        lim = [min(p, q) // 2 - 1 for p, q in zip(fs, ts)]
        f, on = band_limited(fs, [(0, 0, 0), (lim[0], 1, lim[2]), (-1, -lim[1], 2)])
        x[:, :, :fs[2]] = f
        a.FourierTransformBackward(False)  # the upload below is a real-space field
        a.upload(x)
        a.FourierInterpolateCopyTo(b)
        b.FourierTransformBackward()
        tol = 1e-12 if prec == "double" else 2e-6
        assert np.max(np.abs(b.rdata() - on(ts) * np.sqrt(np.prod(fs) / np.prod(ts)))) < tol * np.max(np.abs(f))
    # equal sides: the buffer is copied
    c = M.Grid_FFT(ctx, (8, 8, 8), 1.0, prec)
    a.FourierInterpolateCopyTo(c)
    assert np.array_equal(c.kdata(), a.kdata())
