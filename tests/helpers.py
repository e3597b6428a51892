"""shared helpers for the parity tests"""
import numpy as np


def random_real_grid(n, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    a = np.zeros((n, n, n + 2), dtype=np.float64)
    a[:, :, :n] = rng.standard_normal((n, n, n))
    return a.astype(dtype)


def random_kspace_field(O, n, seed, boxlen=100.0, slope=-2.0):
    """a smooth random k-space field (r2c of white noise shaped by |k|^slope)"""
    a = random_real_grid(n, seed)
    O.fft_r2c(a)
    k = O.kview(a)
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    K2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + kz[None, None, :] ** 2
    K2[0, 0, 0] = 1.0
    k *= K2 ** (slope / 2)
    k[0, 0, 0] = 0
    return a


def real_part(a, n):
    return a[:, :, :n]
