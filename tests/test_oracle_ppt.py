"""The oracle's complex transform and semiclassical (PPT) Eulerian fields
(ic_generator.cc:674-790) pinned by a library transform and closed forms.

* fft_c2c against numpy.fft (unitary, both directions, a transposed axis
  order would fail).
* hbar against potentials whose largest neighbour difference is known: a ramp
  along x (the reduction stops one plane short, ic_generator.cc:703), along y
  and z (periodic wrap).
* delta and v for a plane-wave potential phi = A cos(2 pi l x_d / n) (and
  phi2 = B cos(...)) against the Jacobi-Anger series
  exp(i a cos t) = sum_m i^m J_m(a) exp(i m t): on the n-lattice harmonic m
  lands on DFT index p = m l mod n, whose drift phase uses get_k(p); the
  spectral gradient multiplies each term by i k_p.  A planted drift-sign
  mistake fails this check.
"""
import numpy as np
import pytest
from scipy.special import jv

from oracle import oracle as orc


def test_fft_c2c_matches_numpy():
    rng = np.random.default_rng(5)
    n = 12
    z = rng.standard_normal((n, n, n)) + 1j * rng.standard_normal((n, n, n))
    f = orc.fft_c2c(z.copy(), -1)
    np.testing.assert_allclose(f, np.fft.fftn(z, norm="ortho"), atol=1e-12)
    b = orc.fft_c2c(z.copy(), +1)
    np.testing.assert_allclose(b, np.fft.ifftn(z, norm="ortho"), atol=1e-12)


def _padded(f):
    n = f.shape[0]
    out = np.zeros((n, n, n + 2))
    out[:, :, :n] = f
    return out


@pytest.mark.parametrize("axis,expect", [(0, 1.0), (1, 15.0), (2, 15.0)])
def test_hbar_reduction_bounds(axis, expect):
    # ramp of slope 1 along one axis: x differences stop at i = n-2, y and z wrap
    n, D = 16, 0.7
    idx = np.arange(n, dtype=float)
    shape = [1, 1, 1]
    shape[axis] = n
    phi = np.broadcast_to(idx.reshape(shape), (n, n, n)).copy()
    _, _, hbar = orc.ppt_fields(_padded(phi), None, 100.0, D, 1.0)
    assert hbar == pytest.approx(expect / np.pi / D * 1.01, rel=1e-14)


def _getk(p, n, kfac):
    return np.where(p > n // 2, p - n, p) * kfac


def _closed_form(n, l, A, B, boxlen, D, vunit, axis):
    """delta and v_axis of the plane-wave potential, by Jacobi-Anger"""
    kfac = 2 * np.pi / boxlen
    x = np.arange(n)
    th = 2 * np.pi * l * x / n
    # hbar: max |A (cos(th(i+1)) - cos(th(i)))| = 2A |sin(pi l/n) sin(pi l (2i+1)/n)|
    ii = np.arange(n - 1) if axis == 0 else np.arange(n)
    maxd = np.max(np.abs(2 * A * np.sin(np.pi * l / n) * np.sin(np.pi * l * (2 * ii + 1) / n)))
    hbar = maxd / np.pi / D * 1.01
    a, b = (A + B) / (hbar * D), B / (hbar * D)
    M = 60
    m = np.arange(-M, M + 1)
    p = (m * l) % n
    kp = _getk(p, n, kfac)
    c = (1j ** m) * jv(m, a) * np.exp(-0.5j * hbar * kp ** 2 * D)
    dm = (1j ** m) * jv(m, b) if B != 0 else (m == 0).astype(complex)
    # psi and its spectral derivative: terms (m, m') at index (p(m) + p(m')) mod n
    P = (p[:, None] + p[None, :]) % n
    coef = c[:, None] * dm[None, :]
    kP = _getk(P, n, kfac)
    ph = np.exp(2j * np.pi * P[..., None] * x[None, None, :] / n)
    psi = np.sum(coef[..., None] * ph, axis=(0, 1))
    g = np.sum((coef * 1j * kP)[..., None] * ph, axis=(0, 1))
    delta = np.abs(psi) ** 2 - 1
    v = vunit * hbar * np.imag(np.conj(psi) * g) / (1 + delta)
    return delta, v, hbar


def _plane(n, l, amp, axis):
    x = np.arange(n)
    shape = [1, 1, 1]
    shape[axis] = n
    return np.broadcast_to((amp * np.cos(2 * np.pi * l * x / n)).reshape(shape), (n, n, n)).copy()


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("second", [False, True])
def test_ppt_plane_wave_closed_form(axis, second):
    n, l, boxlen, D, vunit = 16, 1, 100.0, 0.5, 1.3
    A, B = 60.0, (15.0 if second else 0.0)
    phi = _padded(_plane(n, l, A, axis))
    phi2 = _padded(_plane(n, l, B, axis)) if second else None
    delta, v, hbar = orc.ppt_fields(phi, phi2, boxlen, D, vunit)
    d_ref, v_ref, h_ref = _closed_form(n, l, A, B, boxlen, D, vunit, axis)
    assert hbar == pytest.approx(h_ref, rel=1e-12)
    sl = [0, 0, 0]
    sl[axis] = slice(None)
    assert np.max(np.abs(delta)) > 0.1  # the drift moved mass
    np.testing.assert_allclose(delta[tuple(sl)], d_ref, atol=1e-10)
    # the field is uniform across the wave
    assert np.allclose(delta, np.broadcast_to(d_ref.reshape([n if q == axis else 1 for q in range(3)]), delta.shape),
                       atol=1e-10)
    np.testing.assert_allclose(v[axis][tuple(sl)], v_ref, atol=1e-9 * np.max(np.abs(v_ref)))
    for q in range(3):
        if q != axis:
            assert np.max(np.abs(v[q])) < 1e-10 * np.max(np.abs(v_ref))


def test_ppt_planted_drift_sign_fails():
    n, l, boxlen, D, vunit = 16, 1, 100.0, 0.5, 1.0
    phi = _padded(_plane(n, l, 60.0, 0))
    d_ref, _, _ = _closed_form(n, l, 60.0, 0.0, boxlen, D, vunit, 0)
    orc.ppt_set_mutation(1)
    try:
        delta, _, _ = orc.ppt_fields(phi, None, boxlen, D, vunit)
    finally:
        orc.ppt_set_mutation(0)
    assert np.max(np.abs(delta[:, 0, 0] - d_ref)) > 1e-2


def test_ppt_aliased_harmonics_second_order():
    # l = 3 on n = 16: harmonics alias across the Nyquist index
    n, l, boxlen, D, vunit = 16, 3, 50.0, 1.2, 0.8
    A, B = 25.0, -6.0
    phi = _padded(_plane(n, l, A, 2))
    phi2 = _padded(_plane(n, l, B, 2))
    delta, v, hbar = orc.ppt_fields(phi, phi2, boxlen, D, vunit)
    d_ref, v_ref, h_ref = _closed_form(n, l, A, B, boxlen, D, vunit, 2)
    assert hbar == pytest.approx(h_ref, rel=1e-12)
    np.testing.assert_allclose(delta[0, 0, :], d_ref, atol=1e-10)
    np.testing.assert_allclose(v[2][0, 0, :], v_ref, atol=1e-9 * np.max(np.abs(v_ref)))
