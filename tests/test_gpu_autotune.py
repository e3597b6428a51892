"""Column-pass tile autotuning: the tuner records a decision per length and
direction, and the transforms are bit-identical whichever tile width runs."""
import numpy as np
import pytest

from helpers import random_real_grid

pytestmark = pytest.mark.gpu


PIPE = 0x100  # MFX_COL_PIPELINED
ALT = 0x200  # MFX_COL_PLAN_ALT


def set_all_tiles(ctx, L, prec, flags):
    """the widest accepted tile of length L in both directions, or'ed with flags"""
    import monofonic_amd as M
    for inv in (0, 1):
        for w in (16, 8, 4, 2, 1):
            try:
                ctx.set_column_tile(L, prec, inv, w | flags)
                break
            except M.MfxError:
                continue
        else:
            raise AssertionError(f"no tile accepted for L={L} {prec} flags {flags:#x}")


@pytest.mark.parametrize("n,prec,wide,half,double", [(96, "float", 16, 8, 0), (192, "double", 16, 8, 0),
                                                     (256, "float", 16, 8, 0), (768, "float", 8, 4, 16),
                                                     (768, "double", 4, 2, 8)])
def test_tile_width_does_not_change_results(n, prec, wide, half, double):
    """tile widths (including the double-width tile in dynamic LDS beyond
    64 KiB) and the persistent pipelined kernel give identical transforms"""
    import monofonic_amd as M
    ctx = M.Context(0)
    a = random_real_grid(n, 77) if n <= 256 else None
    if a is None:  # a thin 768 x 16 x 768 grid: columns of length 768
        rng = np.random.default_rng(5)
        a = np.zeros((n, 16, n + 2))
        a[:, :, :n] = rng.standard_normal((n, 16, n))
    out = []
    tiles = [wide, half, wide | PIPE, half | PIPE] + ([double, double | PIPE] if double else [])
    for tile in tiles:
        for inv in (0, 1):
            ctx.set_column_tile(n, prec, inv, tile)
        g = M.Grid_FFT(ctx, a.shape[:2] + (n,), 100.0, prec)
        g.upload(a)
        g.FourierTransformForward()
        k = g.download().copy()
        g.FourierTransformBackward()
        out.append((k, g.download().copy()))
        del g
    for q in range(1, len(tiles)):
        assert np.array_equal(out[0][0], out[q][0]), q
        assert np.array_equal(out[0][1], out[q][1]), q
    ctx.close()


@pytest.mark.parametrize("path", ["stored", "streamed"])
def test_pipelined_columns_in_lpt_engine_vs_oracle(orc, path):
    """every column pass of the fused LPT engine (x/y stages, truncating
    forward passes, projection turns) on the persistent pipelined kernel,
    against the oracle"""
    import monofonic_amd as M
    from conftest import rel_max, rel_rms
    from helpers import random_kspace_field
    n, box = 64, 100.0
    ctx = M.Context(0)
    for L in (n, 3 * n // 2):
        wide = 16
        for inv in (0, 1):
            ctx.set_column_tile(L, "double", inv, (wide // 2) | PIPE)
    phi_h = random_kspace_field(orc, n, 8, slope=-3.0)
    grids = [M.Grid_FFT(ctx, n, box, "double", initialspace=M.kspace_id) for _ in range(6)]
    grids[0].upload(phi_h)
    conv = M.OrszagConvolver(ctx, n, box, "double")
    conv.set_lpt_path(path)
    conv.lpt_sources(grids[0], grids[1], grids[2], grids[3:], 3)
    r2, r3, rA = orc.lpt_terms(phi_h, box, 3)
    for q, r in enumerate([r2, r3] + list(rA)):
        got = grids[1 + q].kdata()
        assert rel_rms(got, orc.kview(r)) < 1e-12
        assert rel_max(got, orc.kview(r)) < 1e-10
    del conv, grids
    ctx.close()


@pytest.mark.parametrize("shape", [(384, 16, 384), (768, 16, 768), (16, 1024, 512), (1536, 8, 768)])
@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("float", 3e-6)])
@pytest.mark.parametrize("pipe", [0, PIPE])
def test_second_radix_plan_vs_numpy(shape, prec, tol, pipe):
    """the second radix plans of the column lengths 384 (12.8.4), 768
    (12.8.8), 1024 (16.16.4) and 1536 (12.8.8.2), plain and pipelined,
    against numpy's rfftn, forward and back"""
    import monofonic_amd as M
    from conftest import rel_max, rel_rms
    n0, n1, n2 = shape
    L = max(n0, n1)
    print(f"alt plan L={L} {shape} {prec} pipe={bool(pipe)}", flush=True)
    ctx = M.Context(0)
    set_all_tiles(ctx, L, prec, ALT | pipe)
    rng = np.random.default_rng(L)
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
    del g
    ctx.close()


def test_second_radix_plan_in_lpt_engine_vs_oracle(orc):
    """every column pass of the fused LPT engine at length 384 (n = 256) on
    the second radix plan, against the oracle"""
    import monofonic_amd as M
    from conftest import rel_max, rel_rms
    from helpers import random_kspace_field
    n, box = 256, 500.0
    ctx = M.Context(0)
    set_all_tiles(ctx, 3 * n // 2, "double", ALT)
    phi_h = random_kspace_field(orc, n, 9, slope=-3.0)
    grids = [M.Grid_FFT(ctx, n, box, "double", initialspace=M.kspace_id) for _ in range(6)]
    grids[0].upload(phi_h)
    conv = M.OrszagConvolver(ctx, n, box, "double")
    conv.lpt_sources(grids[0], grids[1], grids[2], grids[3:], 3)
    r2, r3, rA = orc.lpt_terms(phi_h, box, 3)
    for q, r in enumerate([r2, r3] + list(rA)):
        got = grids[1 + q].kdata()
        assert rel_rms(got, orc.kview(r)) < 1e-12
        assert rel_max(got, orc.kview(r)) < 1e-10
    del conv, grids
    ctx.close()


def test_autotune_records_decisions():
    import monofonic_amd as M
    ctx = M.Context(0)
    ctx.autotune(128, "float")
    rep = ctx.autotune_report()
    print(rep, flush=True)
    lines = [ln for ln in rep.splitlines() if ln]
    # lengths 128 and 192, forward and inverse, row-strided and plane-strided,
    # and the turn and truncating forward passes of 192 (a padded length 3n/2)
    assert len(lines) == 12
    assert sum("plane-strided" in ln for ln in lines) == 6
    assert sum(ln.startswith("L=192 f32 turn") for ln in lines) == 2
    assert sum(ln.startswith("L=192 f32 truncating forward") for ln in lines) == 2
    for ln in lines:
        assert "-> C=" in ln
        assert ln.count("ms") == 4  # two tile widths, plain and pipelined
    with pytest.raises(M.MfxError):
        ctx.set_column_tile(128, "float", 0, 3)
    with pytest.raises(M.MfxError):  # 128 has one radix plan
        ctx.set_column_tile(128, "float", 0, 16 | ALT)
    ctx.autotune(512, "float")  # lengths 512 and 768: twelve candidates at 768
    lines = [ln for ln in ctx.autotune_report().splitlines() if ln.startswith("L=768")]
    assert len(lines) == 8
    for ln in lines:  # widths 8, 4, 16; plain, pipelined; two radix plans (one for the truncating passes)
        assert ln.count("ms") == (6 if "truncating" in ln else 12)
    with pytest.raises(M.MfxError):  # no double-width tile at 512 (16 columns already)
        ctx.set_column_tile(512, "float", 0, 32)
    ctx.set_column_tile(768, "float", 0, 16 | PIPE | ALT)
    ctx.close()
