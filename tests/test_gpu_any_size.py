"""Grid sizes without tuned kernels (the reference plans any size through
FFTW, src/grid_fft.cc:43-112): the generic runtime-plan transforms
(fft_generic.hpp) against numpy's rfftn, and whole IC realisations at such
sizes against the oracle."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(20, 20, 20), (40, 40, 40), (80, 80, 80), (100, 100, 100), (160, 160, 160),
                                   (200, 10, 120), (134, 8, 22), (56, 14, 70), (250, 6, 50), (98, 26, 242),
                                   (3, 5, 14), (320, 12, 640)])
@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("float", 3e-6)])
def test_r2c_any_length_vs_numpy(ctx, shape, prec, tol):
    """lengths with factors 5, 7, 11, 13, 67 and odd column lengths"""
    import monofonic_amd as M
    n0, n1, n2 = shape
    print(f"r2c {shape} {prec}", flush=True)
    rng = np.random.default_rng(n0 * 7 + n1 * 3 + n2)
    x = rng.standard_normal(shape)
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


@pytest.mark.parametrize("n,order,prec,rms_tol", [(160, 3, "double", 1e-11), (200, 3, "double", 1e-11),
                                                  (80, 2, "float", 1e-5), (320, 2, "float", 1e-5)])
def test_icgen_any_size_vs_oracle(ctx, orc, n, order, prec, rms_tol):
    """NGENIC realisations at grid sizes 160, 200 (padded 240, 300), 80 and
    320 (padded 120, 480): the convolve sequence on the generic transforms"""
    import monofonic_amd as M
    from test_gpu_configs import compare_fields
    from test_gpu_lpt import oracle_pipeline
    box, seed = 2.0 * n, 99
    print(f"icgen {order}LPT {n}^3 NGENIC {prec}", flush=True)
    ic = M.ICGenerator(ctx, n, box, 24.0, LPTorder=order, precision=prec, generator="NGENIC", seed=seed)
    ic.run()
    ref, _, _ = oracle_pipeline(orc, n, box, 24.0, order, seed, orc.Cosmology(), generator="NGENIC")
    compare_fields(ic, ref, n, rms_tol, 100 * rms_tol, f"{n}^3 {prec}")


def test_convolver_rejects_sizes_not_divisible_by_4(ctx):
    import monofonic_amd as M
    with pytest.raises(M.MfxError, match="divisible by 4"):
        M.OrszagConvolver(ctx, 90, 100.0, "double")
