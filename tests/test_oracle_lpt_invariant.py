"""The oracle's 3LPT terms pinned by the Lagrangian vorticity invariant
(src/testing.cc:102-241): see lpt_invariant.py."""
import numpy as np

from lpt_invariant import band_limited_phi, displacement_velocity, exact_product, grads, to_real

N, BOX = 32, 100.0


def invariant_rms(G, fields, D, with_a3):
    psi, vel = displacement_velocity(G, *fields, D, with_a3)
    out = []
    for i in range(3):
        j = (i + 1) % 3
        inv = G[j] * vel[i] - G[i] * vel[j]
        for k in range(3):
            inv = inv + exact_product(G[i] * psi[k], G[j] * vel[k], N) - exact_product(G[j] * psi[k], G[i] * vel[k], N)
        out.append(np.sqrt(np.mean(to_real(inv, N) ** 2)))
    return np.array(out)


def test_vorticity_invariant_cancels_through_third_order(orc):
    a = band_limited_phi(orc, N, 5, BOX, N // 8)
    r2, r3, rA = orc.lpt_terms(a, BOX, 3)
    fields = (orc.kview(a), orc.kview(r2), orc.kview(r3), [orc.kview(x) for x in rA])
    G = grads(N, BOX)
    with_a3 = [invariant_rms(G, fields, D, True) for D in (1.0, 0.5)]
    without = [invariant_rms(G, fields, D, False) for D in (1.0, 0.5)]
    assert np.all(np.abs(with_a3[0] / with_a3[1] - 16) < 0.5), with_a3   # D^4
    assert np.all(np.abs(without[0] / without[1] - 8) < 0.3), without    # D^3
    assert np.all(with_a3[1] < without[1] / 20)


def test_vorticity_invariant_detects_a_sign_error_in_A3(orc):
    """the same check with A3's sign flipped in the oracle output: the third
    order no longer cancels"""
    a = band_limited_phi(orc, N, 6, BOX, N // 8)
    r2, r3, rA = orc.lpt_terms(a, BOX, 3)
    fields = (orc.kview(a), orc.kview(r2), orc.kview(r3), [-orc.kview(x) for x in rA])
    G = grads(N, BOX)
    s = [invariant_rms(G, fields, D, True) for D in (1.0, 0.5)]
    assert np.all(np.abs(s[0] / s[1] - 8) < 0.5), s
