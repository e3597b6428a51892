"""mfx_rng_music1_fill_noalloc, the allocation-free variant of the MUSIC1
fill, through the C ABI against mfx_rng_music1_fill: bit-equal output on two
calls in a row (the second on a smaller grid), both precisions, and the size
limit of its fixed scratch reported as an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def fill(lib, g, seed, cubesize, noalloc):
    name = "mfx_rng_music1_fill_noalloc" if noalloc else "mfx_rng_music1_fill"
    st = getattr(lib, name)(g.handle, ctypes.c_long(seed), ctypes.c_int(cubesize))
    assert st == 0, lib.mfx_last_error().decode()


@pytest.mark.parametrize("prec", ["double", "float"])
def test_noalloc_music1_equals_allocating_fill(ctx, prec):
    import monofonic_amd as M
    lib = M.load()
    print(f"allocation-free MUSIC1 fill {prec}", flush=True)
    for n, cs, seed in ((64, 32, 4711), (32, 16, 99)):
        ref = M.Grid_FFT(ctx, n, 100.0, prec)
        got = M.Grid_FFT(ctx, n, 100.0, prec)
        got.upload(np.full((n, n, n + 2), 7.0))  # stale contents must not matter
        fill(lib, ref, seed, cs, noalloc=False)
        fill(lib, got, seed, cs, noalloc=True)
        ctx.synchronize()
        a, b = got.download(), ref.download()
        assert np.array_equal(a[:, :, :n], b[:, :, :n]), (n, cs)
        # zero mean up to the rounding of the stored values
        assert abs(float(np.mean(a[:, :, :n].astype(np.float64)))) < (1e-12 if prec == "double" else 1e-6)


def test_noalloc_music1_rejects_oversized_scratch(ctx):
    """512^3 with cube size 8 needs 64^3 = 262144 per-cube sums"""
    import monofonic_amd as M
    lib = M.load()
    g = M.Grid_FFT(ctx, 512, 100.0, "float")
    st = lib.mfx_rng_music1_fill_noalloc(g.handle, ctypes.c_long(5), ctypes.c_int(8))
    assert st != 0
    assert b"at most" in lib.mfx_last_error()
