"""The oracle's Panphasia-HO restatement pinned by mathematics.

* Basis tables: degree order, orthonormal constraint rows per parity class,
  hand-derived entries (the parent's P_1, P_2 on a child:
  -sqrt(3)/2, 1/2, -sqrt(15)/4, 1/4), 13-significant-digit values.
* Moment propagation (the reference's own self-test idea,
  high_order_panphasia_routines.c:778-814): the children's piecewise
  Legendre-block field, projected on the parent's 84 orthonormal functions
  by 3D Gauss-Legendre quadrature with numpy's Legendre polynomials, returns
  the parent's coefficients.  A planted sign error of the parity transform
  fails it.
* Spherical Bessel weights against scipy.
* k-space synthesis against a direct sum: the coefficient fields' DFT times
  the Bessel weights, the phase and scale of pan_mpi_routines.c:372-383 and
  the copy into a smaller grid, with scipy's j_l.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npl
from scipy.special import spherical_jn

from oracle import oracle as orc

DEG = [(px, py, n - px - py) for n in range(7) for px in range(n + 1) for py in range(n - px + 1)]


def test_basis_order_and_rows():
    C, deg, ri = orc.pan_tables()
    assert [tuple(deg[:, l]) for l in range(84)] == DEG
    cls = [4 * (a % 2) + 2 * (b % 2) + c % 2 for a, b, c in DEG]
    assert [cls[q] for q in ri] == sorted(cls[q] for q in ri)
    assert sorted(ri) == list(range(84))
    # rows of one class are orthonormal; rows touch only child functions of
    # degree <= the parent's per axis
    start = 0
    for c in range(8):
        rows = [r for r in range(84) if cls[ri[r]] == c]
        Cb = C[rows]
        np.testing.assert_allclose(Cb @ Cb.T, np.eye(len(rows)), atol=1e-12)
        for r in rows:
            a = DEG[ri[r]]
            for l in range(84):
                if any(DEG[l][d] > a[d] for d in range(3)):
                    assert C[r, l] == 0.0
        start += len(rows)
    # hand-derived 1D projections (parent P_1, P_2 on the upper child)
    row = {DEG[ri[r]]: r for r in range(84)}
    i000, i001, i002 = DEG.index((0, 0, 0)), DEG.index((0, 0, 1)), DEG.index((0, 0, 2))
    assert C[row[(0, 0, 1)], i000] == pytest.approx(-np.sqrt(3) / 2, abs=1e-13)
    assert C[row[(0, 0, 1)], i001] == pytest.approx(0.5, abs=1e-13)
    assert C[row[(0, 0, 2)], i001] == pytest.approx(-np.sqrt(15) / 4, abs=1e-13)
    assert C[row[(0, 0, 2)], i002] == pytest.approx(0.25, abs=1e-13)
    # printed with 13 significant digits
    nz = C[C != 0]
    assert np.all(np.array([float(f"{v:.12e}") for v in nz]) == nz)


def _phi(a, t):
    """orthonormal Legendre on [0, 1]"""
    c = np.zeros(a + 1)
    c[a] = 1
    return np.sqrt(2 * a + 1) * npl.legval(2 * t - 1, c)


def _project_children(children):
    """parent coefficients of the children's field: sum over children of
    int_child phi_parent(x) sum_l' w_l' psi_child,l'(x) dx"""
    xg, wg = npl.leggauss(10)
    t = 0.5 * (xg + 1)
    w = 0.5 * wg
    out = np.zeros(84)
    for c in range(8):
        ix, iy, iz = (c >> 2) & 1, (c >> 1) & 1, c & 1
        xs, ys, zs = (ix + t) / 2, (iy + t) / 2, (iz + t) / 2
        # child field at the child's quadrature points
        field = np.zeros((10, 10, 10))
        for l, (a, b, cc) in enumerate(DEG):
            field += children[c, l] * np.sqrt(8.0) * np.einsum("i,j,k->ijk", _phi(a, t), _phi(b, t), _phi(cc, t))
        wt = np.einsum("i,j,k->ijk", w, w, w) / 8.0  # child volume 1/8
        for l, (a, b, cc) in enumerate(DEG):
            basis = np.einsum("i,j,k->ijk", _phi(a, xs), _phi(b, ys), _phi(cc, zs))
            out[l] += np.sum(wt * field * basis)
    return out


def test_moment_propagation():
    rng = np.random.default_rng(3)
    for level, j in [(5, (3, 17, 30)), (40, (12345, 999, 7))]:
        parent = rng.standard_normal(84)
        g = orc.pan_cell_gauss(level, *j)
        ch = orc.pan_solve(parent, g, True)
        np.testing.assert_allclose(_project_children(ch), parent, atol=1e-11)


def test_moment_propagation_planted_sign_fails():
    parent = np.random.default_rng(4).standard_normal(84)
    g = orc.pan_cell_gauss(9, 1, 2, 3)
    orc.pan_set_mutation(1)
    try:
        ch = orc.pan_solve(parent, g, True)
    finally:
        orc.pan_set_mutation(0)
    assert np.max(np.abs(_project_children(ch) - parent)) > 1e-3


def test_unconstrained_children_are_orthonormal_transform():
    # without the parent constraint the solve is an orthogonal map of the
    # 672 deviates (Parseval)
    g = orc.pan_cell_gauss(3, 1, 0, 1)
    ch = orc.pan_solve(np.zeros(84), g, False)
    assert np.sum(ch ** 2) == pytest.approx(np.sum(g ** 2), rel=1e-13)
    root = orc.pan_root()
    np.testing.assert_array_equal(root, orc.pan_solve(np.zeros(84), orc.pan_cell_gauss(0, 2, 2, 2), False)[0])


def test_sph_bessel_vs_scipy():
    xs = np.linspace(0, np.pi, 257)
    for l in range(7):
        got = np.array([orc.sph_bessel(l, x) for x in xs])
        np.testing.assert_allclose(got, spherical_jn(l, xs), atol=1e-14, rtol=1e-12)


def _direct_noise(origin, level, size, kk, n):
    fdim = 2
    rel = 0
    while fdim * (size << rel) < n:
        rel += 1
    np_ = size << rel
    N0 = fdim * np_
    coef = orc.pan_coefficients(origin, level, rel, np_)  # (np, np, np, 84)
    x = np.arange(np_)
    k = np.arange(N0)
    E = np.exp(-2j * np.pi * np.outer(k, fdim * x) / N0)  # (N0, np)
    F = np.einsum("ax,by,cz,xyzl->abcl", E, E, E[: N0 // 2 + 1], coef)
    ks = np.where(k > N0 // 2, k - N0, k)

    def B(p, idx):
        kk_ = np.abs(ks[idx])
        v = np.sqrt(2 * p + 1) * (-1j) ** p * np.where(ks[idx] < 0, (-1.0) ** p, 1.0) * \
            spherical_jn(p, np.pi * fdim * kk_ / N0)
        return np.where(idx == N0 // 2, 1.0 if p == 0 else 0.0, v)

    ia, ib, ic = np.meshgrid(k, k, k[: N0 // 2 + 1], indexing="ij")
    acc = np.zeros(ia.shape, complex)
    W = np.zeros(ia.shape)
    for l, (a, b, c) in enumerate(DEG):
        w = B(a, ia) * B(b, ib) * B(c, ic)
        acc += w * F[..., l]
        W += np.abs(w) ** 2
    kx, ky, kz = ks[ia], ks[ib], ks[ic]
    nyq = (kx == N0 // 2) | (ky == N0 // 2) | (kz == N0 // 2)
    th = -fdim * np.pi * (kx + ky + kz) / np.sqrt(np.where(nyq, 1, W)) / N0
    f = np.where(nyq, 0, acc * np.sqrt(fdim ** 3) * np.exp(1j * th) / N0 ** 1.5)
    k2 = kx ** 2 + ky ** 2 + kz ** 2
    sel = (~nyq) & (k2 <= kk) & (k2 > 0) & (kk > 0)
    f = np.where(sel, f / np.abs(np.where(sel, f, 1)), f)
    out = np.zeros((n, n, n // 2 + 1), complex)
    for i in range(n):
        for j in range(n):
            if N0 == n:
                out[i, j] = f[i, j, : n // 2 + 1]
                continue
            if i == n // 2 or j == n // 2:
                continue
            is_ = i if i < n // 2 else i + N0 - n
            js = j if j < n // 2 else j + N0 - n
            out[i, j, : n // 2] = f[is_, js, : n // 2]
    return out


@pytest.mark.parametrize("origin,level,size,kk,n", [((5, 9, 2), 4, 1, 0, 8), ((100, 3, 77), 7, 3, 0, 8),
                                                    ((1, 2, 3), 5, 1, 3, 8)])
def test_synthesis_vs_direct_sum(origin, level, size, kk, n):
    got = orc.panphasia_noise(origin, level, size, kk, n)
    ref = _direct_noise(origin, level, size, kk, n)
    np.testing.assert_allclose(got, ref, atol=1e-12 * np.max(np.abs(ref)))
    assert np.max(np.abs(got)) > 0.1


def test_coefficients_depend_on_absolute_cell_only():
    # a cell's coefficients follow from its octree ancestry alone: two cubes
    # whose cells overlap -- one wrapping around the octree edge along x --
    # agree on the shared cells
    a = orc.pan_coefficients((15, 2, 3), 4, 1, 6)  # abs x at level 5: 30, 31, 0, 1, 2, 3
    b = orc.pan_coefficients((0, 2, 3), 4, 1, 6)   # abs x: 0 .. 5
    np.testing.assert_array_equal(a[2:6], b[0:4])
    c = orc.pan_coefficients((14, 2, 3), 4, 1, 6)  # abs x: 28 .. 33 -> 28..31, 0, 1
    np.testing.assert_array_equal(a[0:4], c[2:6])
    assert np.std(a) > 0.5
