"""Device white-noise generators against the oracle streams.

Integer streams (MT19937 words, RANLUX ranlxd1 doubles, Threefry counters ->
uniforms) must be bit-exact.  Gaussian transforms that call log/sqrt/cos/sin
are compared to 1e-14 relative (the device libm and glibc may differ in the
last ulp)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mt19937_stream_bit_exact(ctx, orc):
    import monofonic_amd as M
    seeds = [1, 5489, 4357, 0, 123456789, 2 ** 32 + 17]
    got = M.mt19937_raw(ctx, seeds, 2000)
    for s, row in zip(seeds, got):
        ref = orc.mt19937(s, 2000)
        assert np.array_equal(row.astype(np.uint64), ref), s


def test_ranlxd1_stream_bit_exact(ctx, orc):
    import monofonic_amd as M
    seeds = [1, 2, 12345, 0x7ffffffe, 4000000000]
    got = M.ranlxd1_raw(ctx, seeds, 500)
    for s, row in zip(seeds, got):
        assert np.array_equal(row, orc.ranlxd1(s, 500)), s


def test_threefry_uniforms_bit_exact(ctx, orc):
    import monofonic_amd as M
    # the last three cells hit the Gaussian-tail replacement branch for u[0]
    cells = [(10, 0, 0, 0), (10, 1, 2, 3), (7, 127, 5, 99), (10, 786, 427, 0), (10, 259, 344, 2),
             (10, 437, 550, 2)]
    for seed in (0, 7, 2 ** 32 - 1):
        got = M.threefry_uniforms(ctx, np.array(cells), seed, 16)
        for c, row in zip(cells, got):
            ref = orc.panphasia_uniforms(*c, seed, 16)
            assert np.array_equal(row, ref), (c, seed)


# cube sizes below, equal to and above the grid size, and sizes that do not
# divide the grid (cubes repeat; random_music_wnoise_generator.hh:195-222)
@pytest.mark.parametrize("n,cs", [(32, 32), (64, 32), (64, 16), (16, 32), (64, 48), (64, 24), (32, 12), (128, 100)])
def test_music1_fill(ctx, orc, n, cs):
    import monofonic_amd as M
    g = M.Grid_FFT(ctx, n, 100.0, "double")
    M.RNG_music(4711, cs).Fill_Grid(g)
    got = g.download()
    ref = orc.music1_fill(n, 4711, cs)
    assert np.array_equal(got[:, :, :n], ref[:, :, :n])
    gf = M.Grid_FFT(ctx, n, 100.0, "float")
    M.RNG_music(4711, cs).Fill_Grid(gf)
    assert np.array_equal(gf.download()[:, :, :n], orc.music1_fill(n, 4711, cs, f32=True)[:, :, :n])


@pytest.mark.parametrize("n", [16, 48])
def test_ngenic_fill(ctx, orc, n):
    import monofonic_amd as M
    g = M.Grid_FFT(ctx, n, 100.0, "double")
    M.RNG_ngenic(12345).Fill_Grid(g)
    got = g.kdata()
    ref = orc.kview(orc.ngenic_fill(n, orc.ngenic_seedtable(n, 12345)))
    assert np.max(np.abs(got - ref)) < 1e-14 * np.max(np.abs(ref))
    assert np.array_equal(got == 0, ref == 0)


def test_threefry_fill(ctx, orc):
    import monofonic_amd as M
    n = 32
    g = M.Grid_FFT(ctx, n, 100.0, "double")
    M.RNG_threefry(3).Fill_Grid(g)
    got = g.rdata()
    ref = orc.threefry_fill(n, 3)[:, :, :n]
    assert np.max(np.abs(got - ref)) < 1e-14 * np.max(np.abs(ref))
