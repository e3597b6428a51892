"""The library's own host FP64 cosmology (mfx_cosmo_*, cosmology.cpp; no GPU
needed) against independent scipy computations, so that the amplitude, growth
and velocity factors of the pipeline are pinned without going through the
oracle's restatement of cosmology_calculator.hh:91-463 and
transfer_eisenstein.cc:58-220."""
import numpy as np
import pytest
import scipy.integrate as si

from test_oracle_pins import t0_closed_form


@pytest.fixture(scope="module")
def lib_cosmo():
    import monofonic_amd as M
    return M.CosmologyCalculator()


def growth_scipy(c, Om=0.3, H0=70.0):
    """D(a) and f(a) from D'' + (2 + dlnH/dlna) D' = 3/2 Om H0^2/(a^3 H^2) D
    in ln a, integrated by scipy's DOP853 on the library's H(a)"""
    def rhs(lna, y):
        a = np.exp(lna)
        H = c.H_of_a(a)
        eps = 1e-6
        dlnH = (np.log(c.H_of_a(a * (1 + eps))) - np.log(c.H_of_a(a * (1 - eps)))) / (2 * eps)
        return [y[1], -(2 + dlnH) * y[1] + 1.5 * Om * H0 ** 2 / (a ** 3 * H * H) * y[0]]

    a0 = 1e-6
    sol = si.solve_ivp(rhs, [np.log(a0), 0.0], [a0, a0], method="DOP853", rtol=1e-11, atol=1e-20, dense_output=True)
    D1 = sol.sol(0.0)[0]
    return (lambda a: sol.sol(np.log(a))[0] / D1), (lambda a: sol.sol(np.log(a))[1] / sol.sol(np.log(a))[0])


def test_library_growth_vs_scipy(lib_cosmo):
    D, f = growth_scipy(lib_cosmo)
    for a in (1 / 51, 1 / 25, 0.5, 1.0):
        assert abs(lib_cosmo.get_growth_factor(a) / D(a) - 1) < 2e-6, a
        assert abs(lib_cosmo.get_f(a) / f(a) - 1) < 1e-5, a


def test_library_vfact_relation(lib_cosmo):
    """vfact = f a H / h (cosmology_calculator.hh:360-363) with scipy's f"""
    _, f = growth_scipy(lib_cosmo)
    for a in (1 / 51, 1 / 25):
        assert abs(lib_cosmo.get_vfact(a) / (f(a) * a * lib_cosmo.H_of_a(a) / 0.7) - 1) < 1e-5


def test_library_sigma8_normalisation_vs_scipy(lib_cosmo):
    """pnorm = sigma8^2 / (4 pi int k^2 W^2(8k) P(k) dk) with the top-hat
    window, by scipy quad over the library's T(k)"""
    c = lib_cosmo

    def integrand(lk):
        k = np.exp(lk)
        x = 8 * k
        w = 1 - 0.1 * x * x if x < 1e-3 else 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        return k * k * w * w * k ** 0.96 * c.get_transfer(k) ** 2 * k

    pts = np.linspace(np.log(1e-4), np.log(1e4), 400)
    s = sum(si.quad(integrand, pts[i], pts[i + 1], epsabs=0, epsrel=1e-12, limit=200)[0] for i in range(len(pts) - 1))
    assert abs(c.pnorm / (0.81 ** 2 / (4 * np.pi * s)) - 1) < 1e-8
    # amplitude = k^(n_s/2) T(k) sqrt(pnorm)
    for k in (1e-3, 0.1, 3.0):
        assert abs(c.get_amplitude(k) / (k ** 0.48 * c.get_transfer(k) * np.sqrt(c.pnorm)) - 1) < 1e-14


def test_library_transfer_baryon_free_limit():
    import monofonic_amd as M
    c = M.CosmologyCalculator(Omega_b=1e-9 * 0.3)
    ks = np.logspace(-4, 1, 200)
    T = np.array([c.get_transfer(k) for k in ks])
    assert np.max(np.abs(T / t0_closed_form(ks, 0.3, 0.7) - 1)) < 1e-6


def test_library_growth_eds_limit():
    import monofonic_amd as M
    c = M.CosmologyCalculator(Omega_m=1.0, zero_radiation=True)
    for a in (0.01, 0.1, 0.5):
        assert abs(c.get_growth_factor(a) - a) < 1e-7 * a
        assert abs(c.get_f(a) - 1) < 1e-7
