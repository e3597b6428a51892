"""Column-pass tile autotuning: the tuner records a decision per length and
direction, and the transforms are bit-identical whichever tile width runs."""
import numpy as np
import pytest

from helpers import random_real_grid

pytestmark = pytest.mark.gpu


PIPE = 0x100  # MFX_COL_PIPELINED


@pytest.mark.parametrize("n,prec,wide,half", [(96, "float", 16, 8), (192, "double", 16, 8), (256, "float", 16, 8),
                                              (768, "float", 8, 4)])
def test_tile_width_does_not_change_results(n, prec, wide, half):
    """tile widths and the persistent pipelined kernel give identical transforms"""
    import monofonic_amd as M
    ctx = M.Context(0)
    a = random_real_grid(n, 77) if n <= 256 else None
    if a is None:  # a thin 768 x 16 x 768 grid: columns of length 768
        rng = np.random.default_rng(5)
        a = np.zeros((n, 16, n + 2))
        a[:, :, :n] = rng.standard_normal((n, 16, n))
    out = []
    for tile in (wide, half, wide | PIPE, half | PIPE):
        for inv in (0, 1):
            ctx.set_column_tile(n, prec, inv, tile)
        g = M.Grid_FFT(ctx, a.shape[:2] + (n,), 100.0, prec)
        g.upload(a)
        g.FourierTransformForward()
        k = g.download().copy()
        g.FourierTransformBackward()
        out.append((k, g.download().copy()))
        del g
    for q in range(1, 4):
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


def test_autotune_records_decisions():
    import monofonic_amd as M
    ctx = M.Context(0)
    ctx.autotune(128, "float")
    rep = ctx.autotune_report()
    print(rep, flush=True)
    lines = [ln for ln in rep.splitlines() if ln]
    # lengths 128 and 192, forward and inverse
    assert len(lines) == 4
    for ln in lines:
        assert "-> C=" in ln
        assert ln.count("ms") == 4  # two tile widths, plain and pipelined
    with pytest.raises(M.MfxError):
        ctx.set_column_tile(128, "float", 0, 3)
    ctx.close()
