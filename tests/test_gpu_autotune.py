"""Column-pass tile autotuning: the tuner records a decision per length and
direction, and the transforms are bit-identical whichever tile width runs."""
import numpy as np
import pytest

from helpers import random_real_grid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,prec,wide,half", [(96, "float", 16, 8), (192, "double", 16, 8), (256, "float", 16, 8)])
def test_tile_width_does_not_change_results(n, prec, wide, half):
    import monofonic_amd as M
    ctx = M.Context(0)
    a = random_real_grid(n, 77)
    out = []
    for tile in (wide, half):
        for inv in (0, 1):
            ctx.set_column_tile(n, prec, inv, tile)
        g = M.Grid_FFT(ctx, n, 100.0, prec)
        g.upload(a)
        g.FourierTransformForward()
        k = g.download().copy()
        g.FourierTransformBackward()
        out.append((k, g.download().copy()))
        del g
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
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
    with pytest.raises(M.MfxError):
        ctx.set_column_tile(128, "float", 0, 3)
    ctx.close()
