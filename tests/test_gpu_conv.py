"""Orszag 3/2 convolutions on the GPU against the literal oracle restatement of
OrszagConvolver::convolve2/convolve3 (convolution.hh:512-669)."""
import numpy as np
import pytest

from conftest import rel_max, rel_rms
from helpers import random_kspace_field

pytestmark = pytest.mark.gpu

BOX = 100.0


def _grid(M, ctx, data, prec):
    g = M.Grid_FFT(ctx, data.shape[0], BOX, prec, initialspace=M.kspace_id)
    g.upload(data)
    return g


@pytest.mark.parametrize("n", [16, 32, 48])
@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("float", 2e-5)])
def test_convolve_hessians_all_ops(ctx, orc, n, prec, tol):
    import monofonic_amd as M
    print(f"conv n={n} {prec}", flush=True)
    A = random_kspace_field(orc, n, 11)
    B = random_kspace_field(orc, n, 12)
    O0 = random_kspace_field(orc, n, 13)
    ga, gb = _grid(M, ctx, A, prec), _grid(M, ctx, B, prec)
    conv = M.OrszagConvolver(ctx, n, BOX, prec)
    for (opname, code, fac) in [("assign", 0, 0.0), ("add", 1, 0.0), ("sub", 2, 0.0), ("muladd", 3, -5.0 / 7.0)]:
        for (d1, d2) in [((0, 0), (1, 1)), ((0, 1), (0, 2)), ((1, 2), (2, 2))]:
            go = _grid(M, ctx, O0, prec)
            o = {"assign": M.op.assign_to(go), "add": M.op.add_to(go), "sub": M.op.subtract_from(go),
                 "muladd": M.op.multiply_add_to(go, fac)}[opname]
            conv.convolve_Hessians(ga, d1, gb, d2, o)
            got = go.kdata()
            ref = O0.copy()
            orc.convolve2(A, orc.mult("hessian", *d1), B, orc.mult("hessian", *d2), ref, code, fac, BOX)
            assert rel_rms(got, orc.kview(ref)) < tol, (opname, d1, d2)
            assert rel_max(got, orc.kview(ref)) < 50 * tol, (opname, d1, d2)


@pytest.mark.parametrize("n", [32])
def test_convolver_variants(ctx, orc, n):
    import monofonic_amd as M
    A = random_kspace_field(orc, n, 21)
    B = random_kspace_field(orc, n, 22)
    C = random_kspace_field(orc, n, 23)
    ga, gb, gc = (_grid(M, ctx, X, "double") for X in (A, B, C))
    conv = M.OrszagConvolver(ctx, n, BOX, "double")
    cases = [
        ("sum", lambda o: conv.convolve_SumOfHessians(ga, (0, 0), gb, (1, 1), (2, 2), o),
         lambda r: orc.convolve2(A, orc.mult("hessian", 0, 0), B, orc.mult("sum", 1, 1, 2, 2), r, 0, 0, BOX)),
        ("diff", lambda o: conv.convolve_DifferenceOfHessians(ga, (1, 2), gb, (1, 1), (2, 2), o),
         lambda r: orc.convolve2(A, orc.mult("hessian", 1, 2), B, orc.mult("diff", 1, 1, 2, 2), r, 0, 0, BOX)),
        ("grad", lambda o: conv.convolve_Gradients(ga, (0,), gb, (2,), o),
         lambda r: orc.convolve2(A, orc.mult("grad", 0), B, orc.mult("grad", 2), r, 0, 0, BOX)),
        ("gradhess", lambda o: conv.convolve_Gradient_and_Hessian(ga, (1,), gb, (0, 2), o),
         lambda r: orc.convolve2(A, orc.mult("getk_grad", 1), B, orc.mult("getk_hess", 0, 2), r, 0, 0, BOX)),
        ("mult", lambda o: conv.multiply_field(ga, gb, o),
         lambda r: orc.convolve2(A, orc.mult("identity"), B, orc.mult("identity"), r, 0, 0, BOX)),
        ("hess3", lambda o: conv.convolve_Hessians(ga, (0, 1), gb, (0, 2), gc, (1, 2), o),
         lambda r: orc.convolve3(A, orc.mult("hessian", 0, 1), B, orc.mult("hessian", 0, 2), C,
                                 orc.mult("hessian", 1, 2), r, 0, 0, BOX)),
    ]
    for name, run, ref_fn in cases:
        go = M.Grid_FFT(ctx, n, BOX, "double", initialspace=M.kspace_id)
        run(M.op.assign_to(go))
        ref = np.zeros_like(A)
        ref_fn(ref)
        assert rel_rms(go.kdata(), orc.kview(ref)) < 1e-12, name
