"""Pins for the oracle's PLT restatement (include/particle_plt.hh): the eigen
solver against LAPACK, the physical limits and symmetries of the lattice
dynamical matrix, and the gradient/assembly algebra against closed forms."""
import numpy as np
import pytest

NG = 64


@pytest.fixture(scope="module")
def plt(orc):
    tabs, mu = orc.plt_tables(NG, with_mu=True)
    return tabs, mu


def test_sym3_eigen_matches_lapack(orc):
    rng = np.random.default_rng(1)
    mats = [rng.standard_normal((3, 3)) for _ in range(200)]
    mats = [m + m.T for m in mats]
    mats += [np.diag([3.0, 1.0, 2.0]), np.diag([1.0, 1.0, 0.5]),
             np.array([[1.0, 1e-9, 0], [1e-9, 1.0, 0], [0, 0, 0.2]])]
    for A in mats:
        ev, v = orc.sym3_eigen(A)
        w, V = np.linalg.eigh(A)
        assert np.allclose(ev, w, rtol=0, atol=1e-12 * max(1.0, np.abs(w).max()))
        if w[2] - w[1] > 1e-6 * max(1.0, abs(w[2])):
            assert abs(abs(np.dot(v, V[:, 2])) - 1) < 1e-10
            assert np.linalg.norm(A @ v - w[2] * v) < 1e-10 * max(1.0, abs(w[2]))


def test_small_k_limits(plt):
    """long-wavelength modes of the lattice follow the fluid: the longitudinal
    eigenvalue -> 1 (D_r = 1, D_theta, D_phi -> 0) and the growth-rate
    correction -> 1 (Joyce et al. 2005; particle_plt.hh:469-503)"""
    tabs, mu = plt
    assert np.all(tabs[0] == 1.0)
    for n in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (1, 1, 1), (2, 0, 0), (NG - 1, 1, 0)]:
        assert abs(mu[n] - 1) < 3e-3, n
        assert abs(tabs[1][n]) < 1e-3 and abs(tabs[2][n]) < 1e-3, n
        assert abs(tabs[3][n] - 1) < 2e-3, n
    assert tabs[3][0, 0, 0] == 1 and tabs[1][0, 0, 0] == 0 and tabs[2][0, 0, 0] == 0


def test_zone_corner_closed_form(plt):
    """at k = (pi,pi,pi) cubic symmetry makes D(k) a multiple of the identity;
    its trace is 1, so every eigenvalue is 1/3 and the growth-rate correction
    1/(0.25 (sqrt(1 + 24/3) - 1)) = 2"""
    tabs, mu = plt
    c = (NG // 2, NG // 2, NG // 2)
    assert abs(mu[c] - 1.0 / 3.0) < 1e-6
    assert abs(tabs[3][c] - 2.0) < 2e-5
    assert abs(tabs[1][c]) < 1e-12 and abs(tabs[2][c]) < 1e-12


def test_kohn_sum_rule(orc):
    """for a Coulomb lattice the three squared frequencies add up to the plasma
    frequency at every k: trace D(k) = 1"""
    D = orc.plt_dynamical_matrix(NG)
    tr = 0
    for c in (0, 3, 5):
        a = D[c].copy()
        orc.fft_r2c(a)
        tr = tr + orc.kview(a).real * NG ** 1.5
    tr[0, 0, 0] = 1.0
    assert np.max(np.abs(tr - 1)) < 1e-3


def test_cubic_symmetry(plt):
    """the sc lattice is invariant under axis permutations, which the
    reference's truncated Ewald sum (image cube [-3,3]^3 around every site)
    keeps exactly; reflections it keeps only up to that truncation, because
    the image cube around a site at x in [0,1) is not mirror-symmetric"""
    tabs, mu = plt
    h = NG // 2 + 1
    i = np.arange(NG)
    # x <-> y exchange within the stored half space
    assert np.max(np.abs(mu - mu.transpose(1, 0, 2))) < 1e-12
    assert np.max(np.abs(tabs[3] - tabs[3].transpose(1, 0, 2))) < 1e-12
    # y <-> z on modes with ky <= n/2
    m2 = mu[:, :h, :]
    assert np.max(np.abs(m2 - m2.transpose(0, 2, 1))) < 1e-12
    # x -> -x
    assert np.max(np.abs(mu - mu[(-i) % NG])) < 1e-4


def test_gradient_spherical_projection(orc):
    """with constant synthetic tables D_r = 1, D_theta = a, D_phi = b the PLT
    gradient is |k| (e_r + a e_theta + b e_phi), written here with the
    algebraic unit vectors instead of the reference's angles"""
    a, b = 0.3, -0.2
    tabs = np.zeros((4, NG, NG, NG // 2 + 1))
    tabs[0] = 1.0
    tabs[1] = a
    tabs[2] = b
    tabs[3] = 1.7
    n, L = 32, 250.0
    rng = np.random.default_rng(2)
    for _ in range(50):
        ijk = (int(rng.integers(0, n)), int(rng.integers(0, n)), int(rng.integers(0, n // 2 + 1)))
        g, vf = orc.plt_gradient(tabs, n, L, *ijk)
        kv = np.array([(x - n * (x > n // 2)) for x in ijk], float) * 2 * np.pi / L
        if ijk[2] > n // 2:
            continue
        kr = np.linalg.norm(kv)
        if kr == 0:
            assert np.all(g == 0)
            continue
        er = kv / kr
        rho = np.hypot(kv[0], kv[1])
        if rho > 0:
            ep = np.array([-kv[1], kv[0], 0.0]) / rho
            et = np.array([kv[2] * kv[0], kv[2] * kv[1], -rho * rho]) / (kr * rho)
        else:
            ep = np.array([0.0, 1.0, 0.0])
            et = np.array([np.sign(kv[2]), 0.0, 0.0])
        assert np.allclose(g, kr * (er + a * et + b * ep), rtol=0, atol=1e-12 * kr)
        assert abs(vf - 1.7) < 1e-14


def test_plt_assembly_with_fluid_tables_is_plain_assembly(orc):
    """D_theta = D_phi = 0 and a growth-rate factor of 1 make the PLT gradient
    i k: the PLT assembly must then equal the ordinary one on fields without
    Nyquist modes (where the ordinary gradient is zeroed)"""
    from helpers import random_kspace_field
    n, L = 16, 100.0
    tabs = np.zeros((4, NG, NG, NG // 2 + 1))
    tabs[0] = 1.0
    tabs[3] = 1.0
    f = [random_kspace_field(orc, n, s) for s in range(5)]
    for a in f:
        k = orc.kview(a)
        k[n // 2] = 0
        k[:, n // 2] = 0
        k[:, :, n // 2] = 0
    for kind in (0, 1):
        for idim in range(3):
            p = orc.assemble_component_plt(f[0], f[1], f[2], f[2:5], tabs, L, 3, kind, idim, 0.7, 1.3)
            q = orc.assemble_component(f[0], f[1], f[2], f[2:5], L, 3, kind, idim, 0.7, 1.3)
            assert np.max(np.abs(p - q)) < 1e-12 * np.max(np.abs(q))
