"""3D r2c/c2r on the GPU against the FP64 oracle transform (grid_fft.cc:209-270)."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms
from helpers import random_real_grid

pytestmark = pytest.mark.gpu

SIZES = [16, 32, 48, 64, 96, 128, 144, 192, 256]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("float", 3e-6)])
def test_r2c_matches_oracle(ctx, orc, n, prec, tol):
    import monofonic_amd as M
    print(f"r2c n={n} {prec}", flush=True)
    a = random_real_grid(n, 1000 + n)
    g = M.Grid_FFT(ctx, n, 100.0, prec)
    g.upload(a)
    g.FourierTransformForward()
    got = g.kdata()
    ref = a.copy()
    orc.fft_r2c(ref)
    ref = orc.kview(ref)
    assert rel_rms(got, ref) < tol
    assert rel_max(got, ref) < 20 * tol
    # round trip
    g.FourierTransformBackward()
    back = g.rdata()
    assert rel_max(back, a[:, :, :n]) < 20 * tol


@pytest.mark.parametrize("n", [32, 64, 96])
def test_c2r_hermitian_semantics(ctx, orc, n):
    """c2r of a non-Hermitian kz=0 plane: imaginary parts at kz=0 and kz=n/2 ignored,
    exactly as the oracle's Hermitian extension"""
    import monofonic_amd as M
    rng = np.random.default_rng(n)
    k = np.zeros((n, n, n + 2))
    k[...] = rng.standard_normal(k.shape)
    g = M.Grid_FFT(ctx, n, 100.0, "double", initialspace=M.kspace_id)
    g.upload(k)
    g.FourierTransformBackward()
    got = g.rdata()
    ref = k.copy()
    orc.fft_c2r(ref)
    assert rel_max(got, ref[:, :, :n]) < 1e-12


def test_plane_wave_unitary(ctx):
    """cos(2 pi (a x + b y + c z)/n) -> two modes of amplitude sqrt(n^3)/2"""
    import monofonic_amd as M
    n = 64
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    a, b, c = 3, 5, 7
    f = np.zeros((n, n, n + 2))
    f[:, :, :n] = np.cos(2 * np.pi * (a * X + b * Y + c * Z) / n)
    g = M.Grid_FFT(ctx, n, 1.0, "double")
    g.upload(f)
    g.FourierTransformForward()
    k = g.kdata()
    amp = np.sqrt(n ** 3) / 2
    assert abs(k[a, b, c] - amp) < 1e-9
    k[a, b, c] = 0
    assert np.max(np.abs(k)) < 1e-9


@pytest.mark.parametrize("shape", [(1536, 16, 1536), (16, 1536, 1536), (1024, 16, 1024), (288, 24, 384),
                                   (768, 16, 768), (16, 768, 768), (512, 16, 512), (384, 16, 384),
                                   (16, 384, 512), (16, 512, 384)])
@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("float", 3e-6)])
def test_r2c_long_axes_vs_numpy(ctx, shape, prec, tol):
    """column lengths 1536, 1024, 768, 512, 384 and rows of 768, 512, 384,
    256 and 192 complex (the 1024^3, 512^3 and 256^3 configurations and their
    padded grids) on thin non-cubic grids, against numpy's rfftn (pocketfft)
    with the unitary normalisation"""
    import monofonic_amd as M
    n0, n1, n2 = shape
    print(f"r2c {shape} {prec}", flush=True)
    rng = np.random.default_rng(n0 + n1 + n2)
    x = rng.standard_normal((n0, n1, n2))
    a = np.zeros((n0, n1, n2 + 2))
    a[:, :, :n2] = x
    g = M.Grid_FFT(ctx, shape, (100.0, 100.0, 100.0), prec)
    g.upload(a)
    g.FourierTransformForward()
    ref = np.fft.rfftn(x) / np.sqrt(n0 * n1 * n2)
    got = g.kdata()
    assert rel_rms(got, ref) < tol
    assert rel_max(got, ref) < 20 * tol
    g.FourierTransformBackward()
    assert rel_max(g.rdata(), x) < 20 * tol


@pytest.mark.parametrize("n", [384, 512])
@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("float", 3e-6)])
def test_r2c_cubic_large_vs_numpy(ctx, n, prec, tol):
    """the cubic 384^3 and 512^3 transforms of the 256^3 and 512^3
    configurations against numpy's rfftn"""
    import monofonic_amd as M
    print(f"r2c {n}^3 {prec}", flush=True)
    rng = np.random.default_rng(n)
    x = rng.standard_normal((n, n, n))
    a = np.zeros((n, n, n + 2))
    a[:, :, :n] = x
    g = M.Grid_FFT(ctx, n, 100.0, prec)
    g.upload(a)
    del a
    g.FourierTransformForward()
    ref = np.fft.rfftn(x) / np.sqrt(float(n) ** 3)
    got = g.kdata()
    assert rel_rms(got, ref) < tol
    assert rel_max(got, ref) < 20 * tol
    del got, ref
    g.FourierTransformBackward()
    assert rel_max(g.rdata(), x) < 20 * tol
