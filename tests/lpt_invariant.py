"""The Lagrangian vorticity invariant of src/testing.cc:102-241, a physical
check of the 3LPT source terms that does not depend on the grid size.

For x = q + Psi and the velocity v = sum_n n vfac Psi_n of growth-scaled LPT
potentials (D+ = D: g1 = -D, g2 = -3/7 D^2, g3 = D^3/3, g3c = D^3/7,
ic_generator.cc:224-227), the invariant

  I_ij = dv_i/dq_j - dv_j/dq_i + dPsi_k/dq_i dv_k/dq_j - dPsi_k/dq_j dv_k/dq_i

of a potential flow cancels through third order only if the transverse term A3
has the right sign and cyclic structure and phi2 the right normalisation: its
RMS then falls as D^4 (x16 when D is halved), and as D^3 (x8) with A3 left
out.  With phi1 band-limited to |k_d| <= n/8, every product up to third order
is exact under the 3/2 rule, so the cancellation holds to rounding."""
import numpy as np


def band_limited_phi(O, n, seed, boxlen, kmax, hess_rms=0.01):
    """a random k-space field (padded real layout) with modes |k_d| <= kmax,
    scaled so that the RMS of phi_xx is hess_rms"""
    from helpers import random_kspace_field
    a = random_kspace_field(O, n, seed, boxlen=boxlen, slope=-2.0)
    k = O.kview(a)
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    k[(np.abs(KX) > kmax) | (np.abs(KY) > kmax) | (KZ > kmax)] = 0
    G = grads(n, boxlen)
    k *= hess_rms / np.sqrt(np.mean(to_real(G[0] * G[0] * k, n) ** 2))
    return a


def grads(n, boxlen):
    """i k_d with the Nyquist index zeroed (Grid_FFT::gradient), d = 0, 1, 2"""
    kx = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1, dtype=float)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    out = []
    for K in (KX, KY, KZ):
        g = K.copy()
        g[np.abs(g) == n // 2] = 0
        out.append(1j * g * (2 * np.pi / boxlen))
    return out


def to_real(kc, n):
    """unitary c2r of a k-space array (n, n, n/2+1)"""
    return np.fft.irfftn(kc, s=(n, n, n), axes=(0, 1, 2)) * np.sqrt(float(n) ** 3)


def displacement_velocity(G, P1, P2, P3, A, D, with_a3):
    """k-space Psi_d and v_d (vfac = 1) of testing.cc:150-160"""
    g1, g2, g3, g3c = -D, -3.0 / 7.0 * D * D, D ** 3 / 3.0, D ** 3 / 7.0
    ptot = g1 * P1 + g2 * P2 + g3 * P3
    pv = g1 * P1 + 2 * g2 * P2 + 3 * g3 * P3
    AA = [g3c * a if with_a3 else 0 * a for a in A]
    psi, vel = [], []
    for d in range(3):
        p, pp = (d + 1) % 3, (d + 2) % 3
        curl = G[p] * AA[pp] - G[pp] * AA[p]
        psi.append(G[d] * ptot + curl)
        vel.append(G[d] * pv + 3 * curl)
    return psi, vel


def exact_product(fa, fb, n):
    """k-space product of two fields whose modes lie within |k_d| < n/2, on a
    2n grid (no aliasing), truncated back to the n-grid modes"""
    m = 2 * n
    lo = np.r_[0:n // 2, m - n // 2:m]
    src = np.r_[0:n // 2, n // 2:n]

    def up(f):
        F = np.zeros((m, m, m // 2 + 1), complex)
        F[np.ix_(lo, lo, np.arange(n // 2))] = f[np.ix_(src, src, np.arange(n // 2))]
        return np.fft.irfftn(F, s=(m, m, m), axes=(0, 1, 2)) * m ** 3 / np.sqrt(float(n) ** 3)

    P = np.fft.rfftn(up(fa) * up(fb), axes=(0, 1, 2)) / m ** 3 * np.sqrt(float(n) ** 3)
    out = np.zeros((n, n, n // 2 + 1), complex)
    out[:, :, :n // 2] = P[np.ix_(lo, lo, np.arange(n // 2))]
    return out
