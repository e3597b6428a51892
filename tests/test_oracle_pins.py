"""Checks that pin oracle functions which no reference fixture covers, each
by a property the mathematics fixes, and each with a mutation check showing
that a planted mistake in the oracle fails it:

  Leva ratio method (gsl_ran_ugaussian_ratio_method, used by
      random_music_wnoise_generator.cc:613): its quick accept/reject bounds
      must agree with the exact ratio-of-uniforms region v^2 <= -4 u^2 ln u
      on 10^7 MT19937 pairs, so the deviate stream equals the one the exact
      test selects
  NGenIC seed table (random_ngenic.cc:47-75): with the same seed the N and
      N/2 grids share every mode with |k_x|, |k_y|, k_z < N/4 -- the
      frequency mapping of the table rows and columns (the order of the
      eight blocks within a shell is pinned by reading only)
  Eisenstein & Hu (transfer_eisenstein.cc:58-201): the f_b -> 0 limit is
      the closed-form CDM shape T_0(q) with alpha_c = beta_c = 1; the sound
      horizon is within 2 % of E&H's fitting formula; T_b changes sign
      exactly where k s~(k) = n pi; the full T(k) follows the published
      no-wiggle shape
  PLT Ewald sum (particle_plt.hh:175-204): the real-space half is a true
      Ewald real-space kernel (with the textbook reciprocal sum its total does
      not depend on the split); the k-space half satisfies the separable
      Laplacian identity; and the reference's own combination of the two
      (period-1 images with k = 2 pi m / ng) does depend on the split, which
      is why parity follows the reference rather than a converged sum
"""
import numpy as np
import pytest


# ------------------------------------------------------------------ Leva
LEVA = (0.449871, -0.386595, 0.19600, 0.25472, 0.27597, 0.27846)


def leva_exact_stream(orc, seed, npairs):
    """the deviates the exact ratio-of-uniforms test keeps from the first
    npairs MT19937 pairs: u = 1 - U1, v = 1.7156 (U2 - 1/2), accept iff
    v^2 <= -4 u^2 ln u, deviate v/u; also the smallest distance of any
    pair from the boundary"""
    w = orc.mt_words(seed, 2 * npairs)
    u = 1.0 - w[0::2] / 4294967296.0
    v = (w[1::2] / 4294967296.0 - 0.5) * 1.7156
    g = v * v + 4.0 * u * u * np.log(u)
    return (v / u)[g <= 0], float(np.min(np.abs(g)))


def leva_check(orc, consts, seed=4357, npairs=10_000_000):
    ref, margin = leva_exact_stream(orc, seed, npairs)
    assert margin > 1e-12  # no pair so close to the boundary that rounding of log decides
    got = orc.gauss_ratio_bulk(seed, ref.size, consts)
    return np.array_equal(got, ref)


def test_leva_bounds_match_exact_region(orc):
    assert leva_check(orc, LEVA)
    assert leva_check(orc, LEVA, seed=12345, npairs=2_000_000)


@pytest.mark.parametrize("which,value", [(4, 0.2790), (0, 0.4505), (5, 0.2770)])
def test_leva_mutation_is_detected(orc, which, value):
    """a planted error in an inner (r1), centre (s) or outer (r2) bound"""
    c = list(LEVA)
    c[which] = value
    assert not leva_check(orc, c, npairs=2_000_000)


# ------------------------------------------------------------------ NGenIC
def ngenic_low_modes_agree(orc, n, seed, table_n=None):
    """N- and N/2-grid k-space noise on the modes |fx|, |fy|, kz < N/4"""
    big = orc.kview(orc.ngenic_fill(n, orc.ngenic_seedtable(n, seed) if table_n is None else table_n))
    m = n // 2
    small = orc.kview(orc.ngenic_fill(m, orc.ngenic_seedtable(m, seed)))
    f = np.arange(-(n // 4) + 1, n // 4)
    bi, si = f % n, f % m
    kz = np.arange(n // 4)
    a = big[np.ix_(bi, bi, kz)]
    b = small[np.ix_(si, si, kz)]
    assert np.count_nonzero(b) > 0.9 * b.size
    return np.array_equal(a, b)


@pytest.mark.parametrize("n,seed", [(32, 12345), (64, 7), (48, 99)])
def test_ngenic_nested_grids_share_low_modes(orc, n, seed):
    assert ngenic_low_modes_agree(orc, n, seed)


def test_ngenic_mutation_is_detected(orc):
    """negative-frequency rows of the seed table shifted by one (an off-by-one
    in the N-1-i mapping)"""
    n = 32
    t = orc.ngenic_seedtable(n, 12345).reshape(n, n).copy()
    t[n // 2:] = np.roll(t[n // 2:], 1, axis=0)
    assert not ngenic_low_modes_agree(orc, n, 12345, t.ravel())


# ------------------------------------------------------------------ Eisenstein & Hu
COSMOLOGIES = [(0.3, 0.045, 0.7), (0.25, 0.04, 0.73), (0.315, 0.049, 0.674), (0.5, 0.1, 0.6)]
TCMB = 2.7255


def t0_closed_form(k_hmpc, Om, h):
    """E&H 1998 eqs. (10), (19)-(20) with alpha_c = beta_c = 1"""
    theta = TCMB / 2.7
    keq = 0.0746 * Om * h * h / theta ** 2
    q = k_hmpc * h / (13.41 * keq)
    L = np.log(np.e + 1.8 * q)
    C = 14.2 + 386.0 / (1 + 69.9 * q ** 1.08)
    return L / (L + C * q * q)


def cdm_limit_error(c, Om, h):
    ks = np.logspace(-4, 1, 300)
    T = np.array([c.transfer(k) for k in ks])
    return float(np.max(np.abs(T / t0_closed_form(ks, Om, h) - 1)))


@pytest.mark.parametrize("Om,h", [(0.3, 0.7), (0.5, 0.6)])
def test_eh_baryon_free_limit(orc, Om, h):
    c = orc.Cosmology(H0=100 * h, Omega_m=Om, Omega_b=1e-9 * Om)
    assert cdm_limit_error(c, Om, h) < 1e-6


def test_eh_baryon_free_limit_mutation(orc):
    c = orc.Cosmology(Omega_b=1e-9 * 0.3)
    c.set_eh("keq", c.eh("keq") * 1.01)
    assert cdm_limit_error(c, 0.3, 0.7) > 1e-3


def s_fit(Om, Ob, h):
    """E&H 1998 eq. (26)"""
    return 44.5 * np.log(9.83 / (Om * h * h)) / np.sqrt(1 + 10 * (Ob * h * h) ** 0.75)


@pytest.mark.parametrize("Om,Ob,h", COSMOLOGIES)
def test_eh_sound_horizon_vs_fit(orc, Om, Ob, h):
    c = orc.Cosmology(H0=100 * h, Omega_m=Om, Omega_b=Ob)
    assert abs(c.eh("s") / s_fit(Om, Ob, h) - 1) < 0.02


def tb_zeros(c, kmin=0.005, kmax=2.0):
    """k (Mpc^-1) where T_b changes sign, bisected to double precision"""
    ks = np.logspace(np.log10(kmin), np.log10(kmax), 6000)
    tb = np.array([c.eh_parts(k)[0] for k in ks])
    out = []
    for i in np.nonzero(np.sign(tb[:-1]) != np.sign(tb[1:]))[0]:
        a, b = ks[i], ks[i + 1]
        fa = tb[i]
        for _ in range(200):
            m = 0.5 * (a + b)
            fm = c.eh_parts(m)[0]
            if np.sign(fm) == np.sign(fa):
                a, fa = m, fm
            else:
                b = m
            if b - a <= 4e-16 * b:
                break
        out.append(0.5 * (a + b))
    return np.array(out)


def tb_node_errors(c):
    """k_n s~(k_n) / (n pi) - 1 for consecutive zeros n = 1, 2, ...;
    s~ = s / (1 + (beta_node / (k s))^3)^(1/3) from the oracle's s, beta_node"""
    z = tb_zeros(c)
    s, bn = c.eh("s"), c.eh("beta_node")
    x = z * s / (1 + (bn / (z * s)) ** 3) ** (1 / 3)
    n = np.arange(1, len(z) + 1)
    return z, x / (n * np.pi) - 1


@pytest.mark.parametrize("Om,Ob,h", COSMOLOGIES[:2])
def test_eh_baryon_nodes(orc, Om, Ob, h):
    c = orc.Cosmology(H0=100 * h, Omega_m=Om, Omega_b=Ob)
    z, err = tb_node_errors(c)
    assert len(z) >= 8
    assert np.max(np.abs(err)) < 1e-9


def test_eh_baryon_nodes_mutation(orc):
    """a negative alpha_b makes the bracket of T_b change sign: extra zeros
    that are not nodes of j_0(k s~)"""
    c = orc.Cosmology()
    c.set_eh("alpha_b", -c.eh("alpha_b"))
    z, err = tb_node_errors(c)
    assert np.max(np.abs(err)) > 1e-3


def t_nowiggle(k_hmpc, Om, Ob, h):
    """E&H 1998 eqs. (29)-(31), the zero-baryon-oscillation shape"""
    omhh, fb, theta = Om * h * h, Ob / Om, TCMB / 2.7
    s = s_fit(Om, Ob, h)
    ag = 1 - 0.328 * np.log(431 * omhh) * fb + 0.38 * np.log(22.3 * omhh) * fb * fb
    k = k_hmpc * h
    geff = Om * h * (ag + (1 - ag) / (1 + (0.43 * k * s) ** 4))
    q = k_hmpc * theta ** 2 / geff
    L = np.log(2 * np.e + 1.8 * q)
    C = 14.2 + 731.0 / (1 + 62.5 * q)
    return L / (L + C * q * q)


def nowiggle_dev(c, Om, Ob, h):
    ks = np.logspace(-4, 1, 1500)
    r = np.array([c.transfer(k) for k in ks]) / t_nowiggle(ks, Om, Ob, h) - 1
    return float(np.max(np.abs(r))), float(np.max(np.abs(r[ks < 3e-3])))


@pytest.mark.parametrize("Om,Ob,h", COSMOLOGIES)
def test_eh_follows_nowiggle_shape(orc, Om, Ob, h):
    """the baryon oscillations stay within 5 % of the smooth shape and the
    two agree to 0.2 % on large scales"""
    c = orc.Cosmology(H0=100 * h, Omega_m=Om, Omega_b=Ob)
    full, large = nowiggle_dev(c, Om, Ob, h)
    assert full < 0.05 and large < 2e-3


@pytest.mark.parametrize("name,factor", [("alpha_b", 3.0), ("beta_node", 3.0), ("keq", 1.2)])
def test_eh_nowiggle_mutation(orc, name, factor):
    c = orc.Cosmology()
    c.set_eh(name, c.eh(name) * factor)
    full, large = nowiggle_dev(c, 0.3, 0.045, 0.7)
    assert full > 0.05 or large > 2e-3


# ------------------------------------------------------------------ PLT Ewald halves
def textbook_kspace(ar, eta, M=12):
    """the reciprocal half of the Ewald sum of the tidal tensor of a unit
    point mass in a periodic unit box: sum_{m != 0} k k exp(-k^2/4a^2)
    cos(k.r) / k^2 with k = 2 pi m (standard Ewald, e.g. Allen & Tildesley)"""
    alpha = 1 / np.sqrt(2) / eta
    m = np.arange(-M, M + 1)
    g = np.stack([x.ravel() for x in np.meshgrid(m, m, m, indexing="ij")], 1)
    k = 2 * np.pi * g[np.any(g != 0, axis=1)]
    k2 = np.sum(k * k, axis=1)
    w = np.exp(-k2 / (4 * alpha * alpha)) * np.cos(k @ np.asarray(ar)) / k2
    return np.einsum("i,ij,ik->jk", w, k, k)


POINTS = [np.array([0.1, 0.2, 0.3]), np.array([0.45, 0.05, 0.0]), np.array([0.5, 0.5, 0.5])]


def real_half_split_dependence(orc, ar):
    tot = [orc.plt_ewald(ar, eta)[0] + textbook_kspace(ar, eta) for eta in (0.15, 0.2, 0.25)]
    return max(np.max(np.abs(t - tot[1])) for t in tot) / np.max(np.abs(tot[1]))


@pytest.mark.parametrize("p", range(3))
def test_plt_real_half_is_an_ewald_kernel(orc, p):
    assert real_half_split_dependence(orc, POINTS[p]) < 1e-12


def kspace_laplacian_error(orc, ar, eta=4.0, nk=3, ng=64, h=3e-3):
    """finite-difference Laplacian of the k-space half against the separable
    closed form -1/ng^3 Re prod_d sum_m k_d^{p_d} exp(-k_d^2/4a^2 + i k_d r_d)"""
    lap = np.zeros((3, 3))
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        lap += (orc.plt_ewald(ar + e, eta)[1] - 2 * orc.plt_ewald(ar, eta)[1] + orc.plt_ewald(ar - e, eta)[1]) / h ** 2
    alpha = 1 / np.sqrt(2) / eta
    kk = 2 * np.pi * np.arange(-nk, nk + 1) / ng
    g = np.exp(-kk ** 2 / (4 * alpha * alpha))
    S = [[np.sum(kk ** q * g * np.exp(1j * kk * ar[d])) for q in range(3)] for d in range(3)]
    ref = np.zeros((3, 3))
    for mu in range(3):
        for nu in range(3):
            q = [(mu == d) + (nu == d) for d in range(3)]
            ref[mu, nu] = -np.real(S[0][q[0]] * S[1][q[1]] * S[2][q[2]]) / ng ** 3
    return np.max(np.abs(lap - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("p", range(2))
def test_plt_kspace_half_laplacian(orc, p):
    assert kspace_laplacian_error(orc, POINTS[p]) < 1e-6


def test_plt_ewald_mutations_are_detected(orc):
    try:
        orc.plt_set_mutation(1)
        assert real_half_split_dependence(orc, POINTS[0]) > 1e-3
        orc.plt_set_mutation(2)
        assert kspace_laplacian_error(orc, POINTS[0]) > 1e-2
    finally:
        orc.plt_set_mutation(0)


def test_plt_reference_split_is_not_converged(orc):
    """the reference's real-space images have period 1 while its k vectors
    are 2 pi m / ng: the sum of the two halves depends on eta (0.5 % at
    (10, 0, 5)/64), so D(k) is defined by eta = 4 and the 7^3 images and
    k vectors the reference uses, which the oracle and the library repeat"""
    ar = np.array([10, 0, 5]) / 64
    t = [sum(orc.plt_ewald(ar, eta)) for eta in (3.0, 4.0, 5.0)]
    assert np.max(np.abs(t[0] - t[2])) / np.max(np.abs(t[1])) > 1e-3
