"""ctypes bindings for the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` as the checker.  The shipped package never
imports this module.

Grids are numpy float64 arrays in the reference's in-place r2c layout
(src/grid_fft.cc:43-112): real space shape (n, n, n+2), k space viewed as
complex128 of shape (n, n, n//2+1).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ORC_OP_ASSIGN, ORC_OP_ADD, ORC_OP_SUB, ORC_OP_MULADD = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_threefry4x64_20.argtypes = [ctypes.POINTER(ctypes.c_uint64)] * 3
        L.orc_mt_next.restype = ctypes.c_uint32
        L.orc_gauss_ratio.restype = ctypes.c_double
        L.orc_ranlxd_uniform.restype = ctypes.c_double
        L.orc_get_k.restype = ctypes.c_double
        L.orc_grad_k.restype = ctypes.c_double
        for f in ("orc_cosmo_transfer", "orc_cosmo_Dplus", "orc_cosmo_f", "orc_cosmo_vfact",
                  "orc_cosmo_H", "orc_cosmo_amplitude"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_cosmo_sigma8_unnorm.restype = ctypes.c_double
        L.orc_cosmo_sigma8_unnorm.argtypes = [ctypes.c_void_p]
        L.orc_cosmo_init.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 6 + [ctypes.c_int]
        _LIB = L
    return _LIB


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def real_shape(n):
    return (n, n, n + 2)


def kview(a):
    """complex view (n, n, n//2+1) of a padded real-layout float64 array"""
    n = a.shape[0]
    return a.view(np.complex128).reshape(n, n, n // 2 + 1)


# ---------------------------------------------------------------- FFT
def fft_r2c(a):
    n = a.shape[0]
    lib().orc_fft_r2c(_p(a), ctypes.c_int(n))
    return a


def fft_c2r(a):
    n = a.shape[0]
    lib().orc_fft_c2r(_p(a), ctypes.c_int(n))
    return a


def fft_c2c(z, sign):
    """unitary transform of a complex128 (n, n, n) grid in place: sign -1
    forward, +1 backward (Grid_FFT<ccomplex_t>)"""
    assert z.dtype == np.complex128
    lib().orc_fft_c2c(_p(z), ctypes.c_int(z.shape[0]), ctypes.c_int(sign))
    return z


def dft1d(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    lib().orc_dft1d(_p(x), _p(out), ctypes.c_int(x.size), ctypes.c_int(sign))
    return out


# ---------------------------------------------------------------- multipliers
class Mult(ctypes.Structure):
    _fields_ = [("nterms", ctypes.c_int), ("coef_re", ctypes.c_double * 2),
                ("coef_im", ctypes.c_double * 2), ("pow", (ctypes.c_int * 3) * 2),
                ("nyq", ctypes.c_int * 2)]


def mult(kind, *dims):
    m = Mult()
    f = {"hessian": "orc_mult_hessian", "sum": "orc_mult_sum_hessians",
         "diff": "orc_mult_diff_hessians", "grad": "orc_mult_gradient",
         "getk_grad": "orc_mult_getk_gradient", "getk_hess": "orc_mult_getk_hessian",
         "identity": "orc_mult_identity"}[kind]
    getattr(lib(), f)(ctypes.byref(m), *[ctypes.c_int(d) for d in dims])
    return m


def convolve2(fa, ma, fb, mb, out, op, opfac, boxlen):
    n = fa.shape[0]
    lib().orc_convolve2(_p(fa), ctypes.byref(ma), _p(fb), ctypes.byref(mb), _p(out),
                        ctypes.c_int(op), ctypes.c_double(opfac), ctypes.c_int(n),
                        ctypes.c_double(boxlen))


def convolve3(fa, ma, fb, mb, fc, mc, out, op, opfac, boxlen):
    n = fa.shape[0]
    lib().orc_convolve3(_p(fa), ctypes.byref(ma), _p(fb), ctypes.byref(mb), _p(fc),
                        ctypes.byref(mc), _p(out), ctypes.c_int(op), ctypes.c_double(opfac),
                        ctypes.c_int(n), ctypes.c_double(boxlen))


# ---------------------------------------------------------------- operators
def noise_normalise(wn, volfac, fixing=False, inversion=False):
    lib().orc_noise_normalise(_p(wn), ctypes.c_int(wn.shape[0]), ctypes.c_double(volfac),
                              ctypes.c_int(int(fixing)), ctypes.c_int(int(inversion)))


def phi1(phi, wn, boxlen, amp_table):
    amp_table = np.ascontiguousarray(amp_table, dtype=np.float64)
    lib().orc_phi1(_p(phi), _p(wn), ctypes.c_int(phi.shape[0]), ctypes.c_double(boxlen), _p(amp_table))


def inverse_laplacian(f, boxlen):
    lib().orc_inverse_laplacian(_p(f), ctypes.c_int(f.shape[0]), ctypes.c_double(boxlen))


def scale(f, s):
    lib().orc_scale(_p(f), ctypes.c_int(f.shape[0]), ctypes.c_double(s))


def lpt_terms(phi, boxlen, order):
    n = phi.shape[0]
    z = lambda: np.zeros(real_shape(n))
    phi2 = z()
    phi3, ax, ay, az = (z(), z(), z(), z()) if order > 2 else (None, None, None, None)
    nul = ctypes.c_void_p(0)
    lib().orc_lpt_terms(_p(phi), _p(phi2), _p(phi3) if order > 2 else nul,
                        _p(ax) if order > 2 else nul, _p(ay) if order > 2 else nul,
                        _p(az) if order > 2 else nul, ctypes.c_int(n), ctypes.c_double(boxlen),
                        ctypes.c_int(order))
    return phi2, phi3, (ax, ay, az)


def assemble_component(phi, phi2, phi3, A3, boxlen, order, kind, idim, vfac, unit=1.0):
    n = phi.shape[0]
    out = np.zeros(real_shape(n))
    nul = ctypes.c_void_p(0)
    g = lambda a: _p(a) if a is not None else nul
    lib().orc_assemble_component(g(phi), g(phi2), g(phi3), g(A3[0]), g(A3[1]), g(A3[2]), _p(out),
                                 ctypes.c_int(n), ctypes.c_double(boxlen), ctypes.c_int(order),
                                 ctypes.c_int(kind), ctypes.c_int(idim), ctypes.c_double(vfac),
                                 ctypes.c_double(unit))
    return out


def set_positions(field, idim, lunit=1.0):
    n = field.shape[0]
    pos = np.zeros(n * n * n)
    lib().orc_set_positions(_p(field), _p(pos), ctypes.c_int(n), ctypes.c_int(idim),
                            ctypes.c_double(lunit))
    return pos


def set_positions_shifted(field, idim, s, lunit=1.0):
    """set_positions of a sublattice offset by s cells along idim"""
    n = field.shape[0]
    pos = np.zeros(n * n * n)
    lib().orc_set_positions_shifted(_p(field), _p(pos), ctypes.c_int(n), ctypes.c_int(idim),
                                    ctypes.c_double(lunit), ctypes.c_double(s))
    return pos


def fourier_copy(F, tshape):
    """Grid_FFT::FourierInterpolateCopyTo of a k-space array F (n0, n1, n2/2+1)
    complex into a grid of sides tshape; returns (t0, t1, t2/2+1) complex"""
    src = np.ascontiguousarray(F, dtype=np.complex128)
    f = (ctypes.c_int * 3)(src.shape[0], src.shape[1], 2 * (src.shape[2] - 1))
    t = (ctypes.c_int * 3)(*tshape)
    out = np.zeros((tshape[0], tshape[1], tshape[2] // 2 + 1), dtype=np.complex128)
    lib().orc_fourier_copy(_p(src), f, _p(out), t)
    return out


def shift_field(f, boxlen, s):
    """Grid_FFT::shift_field of a real padded field (in place, returned)"""
    n = f.shape[0]
    lib().orc_shift_field(_p(f), ctypes.c_int(n), ctypes.c_double(boxlen), ctypes.c_double(s[0]),
                          ctypes.c_double(s[1]), ctypes.c_double(s[2]))
    return f


# sublattice offsets of the Bravais loads in cells (particle_generator.hh:47-54)
LATTICE_SHIFTS = {
    "sc": [(0, 0, 0)],
    "bcc": [(0, 0, 0), (0.5, 0.5, 0.5)],
    "fcc": [(0, 0, 0), (0, 0.5, 0.5), (0.5, 0, 0.5), (0.5, 0.5, 0)],
    "rsc": [(0, 0, 0), (0, 0, 0.5), (0, 0.5, 0), (0, 0.5, 0.5), (0.5, 0, 0), (0.5, 0, 0.5), (0.5, 0.5, 0),
            (0.5, 0.5, 0.5)],
}


# offset of the secondary (baryon) lattice in cells (particle_generator.hh:56-63)
SECOND_LATTICE_SHIFT = {"sc": (0.5, 0.5, 0.5), "bcc": (0.5, 0.5, 0.0), "fcc": (0.5, 0.5, 0.5),
                        "rsc": (0.25, 0.25, 0.25)}


def lattice_values(field, boxlen, lattice, idim=None, lunit=1.0, secondary=False):
    """set_positions (idim given) or set_velocities (idim None) of a Bravais
    load as the reference runs them (particle_generator.hh:499-543,597-638):
    on the secondary lattice (lattice_index 1) the field is first shifted by
    its offset; for sublattice l it is shifted by the difference to the
    previous offset and read out; returns (sublattices, n, n, n)"""
    n = field.shape[0]
    f = np.array(field, dtype=np.float64, copy=True)
    sh = LATTICE_SHIFTS[lattice]
    o = SECOND_LATTICE_SHIFT[lattice] if secondary else (0.0, 0.0, 0.0)
    out = []
    for l, s in enumerate(sh):
        if l == 0 and secondary:
            shift_field(f, boxlen, o)
        if l > 0:
            shift_field(f, boxlen, [s[d] - sh[l - 1][d] for d in range(3)])
        if idim is None:
            out.append(f[:, :, :n].copy())
        else:
            out.append(set_positions_shifted(f, idim, s[idim] + o[idim], lunit).reshape(n, n, n))
    return np.stack(out)


def lattice_set_mutation(m):
    """plant a mistake in shift_field (0 = none, 1 = phase sign)"""
    ctypes.c_int.in_dll(lib(), "orc_lattice_mutation").value = int(m)


def ppt_fields(phi, phi2, boxlen, D, vunit):
    """semiclassical Eulerian fields (ic_generator.cc:674-790) from the
    growth-scaled real-space potentials phi, phi2 (padded (n, n, n+2) float64,
    phi2 None for LPTorder 1): (delta, v (3, n, n, n), hbar)"""
    n = phi.shape[0]
    delta = np.zeros((n, n, n))
    v = np.zeros((3, n, n, n))
    hb = ctypes.c_double()
    p2 = _p(phi2) if phi2 is not None else None
    lib().orc_ppt_fields(_p(phi), p2, ctypes.c_int(n), ctypes.c_double(boxlen), ctypes.c_double(D),
                         ctypes.c_double(vunit), _p(delta), _p(v), ctypes.byref(hb))
    return delta, v, hb.value


def ppt_set_mutation(m):
    """plant a mistake in ppt_fields (0 = none, 1 = drift phase sign)"""
    ctypes.c_int.in_dll(lib(), "orc_ppt_mutation").value = int(m)


# ---------------------------------------------------------------- Panphasia-HO
def pan_tables():
    """(C 84x84, degrees 3x84, row_index 84) of the order-6 basis"""
    C = np.zeros((84, 84))
    deg = np.zeros((3, 84), dtype=np.int32)
    ri = np.zeros(84, dtype=np.int32)
    lib().orc_pan_tables(_p(C), _p(deg), _p(ri))
    return C, deg, ri


def pan_solve(parent, gauss, constrained=True):
    """solve_panphasia_cell_: 8 x 84 child coefficients (child 4ix+2iy+iz)"""
    parent = np.ascontiguousarray(parent, dtype=np.float64)
    g = np.ascontiguousarray(gauss, dtype=np.float64).reshape(672)
    out = np.zeros(672)
    lib().orc_pan_solve(_p(parent), _p(g), _p(out), ctypes.c_int(1 if constrained else 0))
    return out.reshape(8, 84)


def pan_cell_gauss(level, j1, j2, j3):
    g = np.zeros(672)
    lib().orc_pan_cell_gauss(ctypes.c_size_t(level), ctypes.c_size_t(j1), ctypes.c_size_t(j2), ctypes.c_size_t(j3),
                             _p(g))
    return g


def pan_root():
    r = np.zeros(84)
    lib().orc_pan_root(_p(r))
    return r


def pan_coefficients(origin, base_level, rel, np_):
    """PANPHASIA_compute_coefficients_ of the cube: (np, np, np, 84)"""
    coef = np.zeros(np_ * np_ * (np_ + 2) * 84)
    lib().orc_pan_coefficients(ctypes.c_uint64(origin[0]), ctypes.c_uint64(origin[1]), ctypes.c_uint64(origin[2]),
                               ctypes.c_int(base_level), ctypes.c_int(rel), ctypes.c_size_t(np_), _p(coef))
    return coef.reshape(np_, np_, np_ + 2, 84)[:, :, :np_, :]


def sph_bessel(l, x):
    lib().orc_sph_bessel.restype = ctypes.c_double
    return lib().orc_sph_bessel(ctypes.c_int(l), ctypes.c_double(x))


def panphasia_noise(origin, base_level, size, kk, n):
    """k-space Panphasia-HO noise of an n-grid, complex (n, n, n//2+1)"""
    out = np.zeros(n * n * (n // 2 + 1) * 2)
    lib().orc_panphasia_noise(ctypes.c_uint64(origin[0]), ctypes.c_uint64(origin[1]), ctypes.c_uint64(origin[2]),
                              ctypes.c_int(base_level), ctypes.c_uint64(size), ctypes.c_uint64(kk), ctypes.c_int(n),
                              _p(out))
    return out.view(np.complex128).reshape(n, n, n // 2 + 1)


def pan_set_mutation(m):
    """plant a mistake in the cell solve (0 = none, 1 = parity transform sign)"""
    ctypes.c_int.in_dll(lib(), "orc_pan_mutation").value = int(m)


def power_spectrum(f, boxlen):
    n = f.shape[0]
    mb = n
    k, P, eP, c = (np.zeros(mb) for _ in range(4))
    nb = lib().orc_power_spectrum(_p(f), ctypes.c_int(n), ctypes.c_double(boxlen), _p(k), _p(P),
                                  _p(eP), _p(c), ctypes.c_int(mb))
    return k[:nb], P[:nb], eP[:nb], c[:nb]


# ---------------------------------------------------------------- RNG
class MT(ctypes.Structure):
    _fields_ = [("mt", ctypes.c_uint32 * 624), ("mti", ctypes.c_int)]


def mt19937(seed, count):
    s = MT()
    lib().orc_mt_seed(ctypes.byref(s), ctypes.c_ulong(seed))
    return np.array([lib().orc_mt_next(ctypes.byref(s)) for _ in range(count)], dtype=np.uint64)


def gauss_ratio(seed, count):
    s = MT()
    lib().orc_mt_seed(ctypes.byref(s), ctypes.c_ulong(seed))
    return np.array([lib().orc_gauss_ratio(ctypes.byref(s)) for _ in range(count)])


def gauss_ratio_bulk(seed, count, consts=None):
    """count deviates of the Leva ratio method on MT19937(seed); consts =
    (s, t, a, b, r1, r2) replaces the published constants"""
    out = np.zeros(count)
    c = None if consts is None else np.ascontiguousarray(consts, dtype=np.float64)
    lib().orc_gauss_ratio_bulk(ctypes.c_ulong(seed), ctypes.c_size_t(count),
                               None if c is None else _p(c), _p(out))
    return out


def mt_words(seed, count):
    out = np.zeros(count, dtype=np.uint32)
    lib().orc_mt_words(ctypes.c_ulong(seed), ctypes.c_size_t(count), _p(out))
    return out


def music1_fill(n, seed, cubesize=32, f32=False):
    if f32:
        a = np.zeros(real_shape(n), dtype=np.float32)
        lib().orc_music1_fill_f32(_p(a), ctypes.c_int(n), ctypes.c_long(seed), ctypes.c_int(cubesize))
    else:
        a = np.zeros(real_shape(n))
        lib().orc_music1_fill(_p(a), ctypes.c_int(n), ctypes.c_long(seed), ctypes.c_int(cubesize))
    return a


def threefry4x64_20(ctr, key):
    c = (ctypes.c_uint64 * 4)(*ctr)
    k = (ctypes.c_uint64 * 4)(*key)
    o = (ctypes.c_uint64 * 4)()
    lib().orc_threefry4x64_20(c, k, o)
    return [int(v) for v in o]


def panphasia_uniforms(l, j1, j2, j3, seed, count):
    u = np.zeros(count)
    lib().orc_panphasia_uniforms(ctypes.c_size_t(l), ctypes.c_size_t(j1), ctypes.c_size_t(j2),
                                 ctypes.c_size_t(j3), ctypes.c_size_t(seed), _p(u), ctypes.c_int(count))
    return u


def threefry_fill(n, seed):
    a = np.zeros(real_shape(n))
    lib().orc_threefry_fill(_p(a), ctypes.c_int(n), ctypes.c_size_t(seed))
    return a


def threefry_cells(n, seed, ijk):
    """threefry_fill's values at the cells ijk (m x 3) of an n^3 grid"""
    ijk = np.ascontiguousarray(ijk, dtype=np.int64).reshape(-1, 3)
    out = np.zeros(ijk.shape[0])
    lib().orc_threefry_cells(_p(ijk), ctypes.c_int(ijk.shape[0]), ctypes.c_int(n), ctypes.c_size_t(seed), _p(out))
    return out


class Ranlxd(ctypes.Structure):
    _fields_ = [("xdbl", ctypes.c_double * 12), ("carry", ctypes.c_double), ("ir", ctypes.c_uint),
                ("jr", ctypes.c_uint), ("is_", ctypes.c_uint), ("is_old", ctypes.c_uint),
                ("pr", ctypes.c_uint)]


def ranlxd1(seed, count):
    s = Ranlxd()
    lib().orc_ranlxd_seed(ctypes.byref(s), ctypes.c_ulong(seed), ctypes.c_int(202))
    return np.array([lib().orc_ranlxd_uniform(ctypes.byref(s)) for _ in range(count)])


def ngenic_seedtable(n, seed):
    t = np.zeros(n * n, dtype=np.uint32)
    lib().orc_ngenic_seedtable(_p(t), ctypes.c_int(n), ctypes.c_long(seed))
    return t


def ngenic_fill(n, table):
    a = np.zeros(real_shape(n))
    table = np.ascontiguousarray(table, dtype=np.uint32)
    lib().orc_ngenic_fill(_p(a), ctypes.c_int(n), _p(table))
    return a


# ---------------------------------------------------------------- cosmology
class Cosmo(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in
                ("h", "Omega_m", "Omega_b", "Omega_DE", "Omega_r", "Omega_k", "w_0", "w_a", "n_s",
                 "sigma_8", "Tcmb", "pnorm", "sqrtpnorm", "Dnow")] + \
               [("eh", ctypes.c_double * 24), ("ntab", ctypes.c_int)] + \
               [(k, ctypes.c_void_p) for k in ("la", "lD", "lf", "d2D", "d2f")]


class Cosmology:
    """FP64 cosmology calculator restated from include/cosmology_calculator.hh"""

    def __init__(self, H0=70.0, Omega_m=0.3, Omega_b=0.045, n_s=0.96, sigma_8=0.81, Tcmb=2.7255,
                 zero_radiation=False):
        self.c = Cosmo()
        lib().orc_cosmo_init(ctypes.byref(self.c), H0, Omega_m, Omega_b, n_s, sigma_8, Tcmb,
                             int(zero_radiation))

    def __del__(self):
        try:
            lib().orc_cosmo_free(ctypes.byref(self.c))
        except Exception:
            pass

    def _f(self, name, x):
        return getattr(lib(), name)(ctypes.byref(self.c), ctypes.c_double(x))

    def transfer(self, k):
        return self._f("orc_cosmo_transfer", k)

    EH_SLOTS = ("omhh", "obhh", "theta", "zeq", "keq", "zd", "Rd", "Req", "s", "ksilk", "alpha_c", "beta_c",
                "alpha_b", "beta_b", "beta_node", "h", "fb", "fc")

    def eh(self, name):
        """a derived Eisenstein & Hu quantity (Mpc units)"""
        return self.c.eh[self.EH_SLOTS.index(name)]

    def set_eh(self, name, value):
        """overwrite a derived quantity (plants a mistake for mutation checks)"""
        self.c.eh[self.EH_SLOTS.index(name)] = value

    def eh_parts(self, k_mpc):
        """(T_b, T_c) of TFfit_onek at k in Mpc^-1"""
        tb, tc = ctypes.c_double(), ctypes.c_double()
        lib().orc_cosmo_eh_parts(ctypes.byref(self.c), ctypes.c_double(k_mpc), ctypes.byref(tb), ctypes.byref(tc))
        return tb.value, tc.value

    def Dplus(self, a):
        return self._f("orc_cosmo_Dplus", a)

    def f(self, a):
        return self._f("orc_cosmo_f", a)

    def vfact(self, a):
        return self._f("orc_cosmo_vfact", a)

    def H(self, a):
        return self._f("orc_cosmo_H", a)

    def amplitude(self, k):
        return self._f("orc_cosmo_amplitude", k)

    @property
    def pnorm(self):
        return self.c.pnorm

    def amplitude_table(self, n, boxlen):
        """A(k) at k = kfac*sqrt(n2) for every integer n2 <= 3 (n/2)^2"""
        kfac = 2.0 * np.pi / boxlen
        nmax = 3 * (n // 2) ** 2 + 1
        t = np.zeros(nmax)
        for n2 in range(1, nmax):
            t[n2] = self.amplitude(kfac * np.sqrt(n2))
        return t


# ---------------------------------------------------------------- PLT (particle_plt.hh)
def plt_dynamical_matrix(ng):
    """6 padded real grids D_xx, D_xy, D_xz, D_yy, D_yz, D_zz on ng^3 sites"""
    out = np.zeros((6, ng, ng, ng + 2))
    lib().orc_plt_dynamical_matrix(ctypes.c_int(ng), _p(out))
    return out


def plt_ewald(ar, eta, nimg=3, nk=3, ng=64):
    """(real-space half, k-space half) of init_D's Ewald sum at ar, 3x3 each"""
    r = np.ascontiguousarray(ar, dtype=np.float64)
    Dr = np.zeros(9)
    Dk = np.zeros(9)
    lib().orc_plt_ewald_real(_p(r), ctypes.c_double(eta), ctypes.c_int(nimg), _p(Dr))
    lib().orc_plt_ewald_kspace(_p(r), ctypes.c_double(eta), ctypes.c_int(nk), ctypes.c_int(ng), _p(Dk))
    return Dr.reshape(3, 3), Dk.reshape(3, 3)


def plt_set_mutation(m):
    """plant a mistake in the Ewald halves (0 = none, 1 = K1 sign, 2 = k-space exponent)"""
    ctypes.c_int.in_dll(lib(), "orc_plt_mutation").value = int(m)


def sym3_eigen(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    ev, vec = np.zeros(3), np.zeros(3)
    lib().orc_sym3_eigen(_p(A), _p(ev), _p(vec))
    return ev, vec


def plt_tables(ng=64, with_mu=False):
    """(4, ng, ng, ng/2+1): D_r, D_theta, D_phi, growth-rate correction"""
    tabs = np.zeros((4, ng, ng, ng // 2 + 1))
    mu = np.zeros((ng, ng, ng // 2 + 1)) if with_mu else None
    lib().orc_plt_tables(ctypes.c_int(ng), _p(tabs), _p(mu) if with_mu else ctypes.c_void_p(0))
    return (tabs, mu) if with_mu else tabs


def plt_gradient(tabs, n, boxlen, i, j, k):
    ng = tabs.shape[1]
    g = np.zeros(3)
    vf = ctypes.c_double()
    lib().orc_plt_gradient(_p(tabs), ctypes.c_int(ng), ctypes.c_int(n), ctypes.c_double(boxlen),
                           ctypes.c_size_t(i), ctypes.c_size_t(j), ctypes.c_size_t(k), _p(g), ctypes.byref(vf))
    return g, vf.value


def assemble_component_plt(phi, phi2, phi3, A3, tabs, boxlen, order, kind, idim, vfac, unit=1.0):
    n = phi.shape[0]
    out = np.zeros(real_shape(n))
    nul = ctypes.c_void_p(0)
    g = lambda a: _p(a) if a is not None else nul
    lib().orc_assemble_component_plt(g(phi), g(phi2), g(phi3), g(A3[0]), g(A3[1]), g(A3[2]), _p(tabs),
                                     ctypes.c_int(tabs.shape[1]), _p(out), ctypes.c_int(n), ctypes.c_double(boxlen),
                                     ctypes.c_int(order), ctypes.c_int(kind), ctypes.c_int(idim),
                                     ctypes.c_double(vfac), ctypes.c_double(unit))
    return out
