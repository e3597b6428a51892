"""Python mirror of monofonIC's field-pipeline interfaces over the gfx950 C ABI.

Names, argument meaning and error behaviour follow the reference:
``Grid_FFT`` (include/grid_fft.hh), ``OrszagConvolver`` (include/convolution.hh),
the ``op::`` output operators (include/operators.hh), the RNG plugins
(include/random_plugin.hh) and ``cosmology::calculator``
(include/cosmology_calculator.hh).  All arithmetic runs on the GPU (or, for the
FP64 cosmology scalars, in the native host library); numpy arrays are only used
to move data in and out.
"""
import ctypes
from collections import namedtuple

import numpy as np

from . import _lib
from ._lib import call, load

kspace_id, rspace_id = _lib.MFX_KSPACE, _lib.MFX_RSPACE


def _prec(precision):
    if precision in ("double", "f64", np.float64, 8):
        return _lib.MFX_DOUBLE
    if precision in ("float", "f32", np.float32, 4):
        return _lib.MFX_FLOAT
    raise ValueError(f"unknown precision {precision!r}")


class Context:
    """One HIP stream on one device (replaces the FFTW/OpenMP set-up of src/main.cc)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        call("mfx_ctx_create", device, ctypes.byref(self._h))

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        call("mfx_ctx_synchronize", self._h)

    def record(self, slot):
        call("mfx_event_record", self._h, slot)

    def elapsed_ms(self, s0, s1):
        ms = ctypes.c_float()
        call("mfx_event_elapsed", self._h, s0, s1, ctypes.byref(ms))
        return ms.value

    def profile(self, on=True):
        call("mfx_profile_enable", self._h, int(on))

    def profile_reset(self):
        call("mfx_profile_reset", self._h)

    def profile_stats(self):
        n = ctypes.c_int()
        call("mfx_profile_count", self._h, ctypes.byref(n))
        out = {}
        buf = ctypes.create_string_buffer(128)
        for i in range(n.value):
            cnt, ms, by = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
            call("mfx_profile_get", self._h, i, buf, 128, ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(by))
            out[buf.value.decode()] = {"launches": cnt.value, "ms": ms.value, "bytes": by.value}
        return out

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        call("mfx_device_alloc", self._h, nbytes, ctypes.byref(p))
        return p

    def free(self, p):
        call("mfx_device_free", self._h, p)

    def autotune(self, n, precision="float"):
        """time the column-pass tile widths for grid size n (lengths n and 3n/2)"""
        call("mfx_autotune", self._h, int(n), _prec(precision))

    def set_column_tile(self, L, precision, inverse, tile):
        call("mfx_set_column_tile", self._h, int(L), _prec(precision), int(inverse), int(tile))

    def autotune_report(self):
        buf = ctypes.create_string_buffer(65536)
        call("mfx_autotune_report", self._h, buf, 65536)
        return buf.value.decode()

    def d2h(self, dev, nbytes, dtype):
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        call("mfx_memcpy_d2h", self._h, out.ctypes.data_as(ctypes.c_void_p), dev, nbytes)
        return out

    def close(self):
        if self._h:
            call("mfx_ctx_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid_FFT:
    """Grid_FFT<real_t,false> (include/grid_fft.hh:40-914) on one GPU."""

    def __init__(self, ctx, N, L, precision="double", initialspace=rspace_id):
        N = (N, N, N) if np.isscalar(N) else tuple(int(x) for x in N)
        L = (L, L, L) if np.isscalar(L) else tuple(float(x) for x in L)
        self.ctx = ctx
        self.n_ = N
        self.length_ = L
        self._h = ctypes.c_void_p()
        call("mfx_grid_create", ctx.handle, N[0], N[1], N[2], L[0], L[1], L[2], _prec(precision), initialspace,
             ctypes.byref(self._h))
        nt, sp, pr = ctypes.c_size_t(), ctypes.c_int(), ctypes.c_int()
        call("mfx_grid_info", self._h, ctypes.byref(nt), ctypes.byref(sp), ctypes.byref(pr))
        self.ntot_ = nt.value
        self.dtype = np.float64 if pr.value == 8 else np.float32
        self.kfac_ = tuple(2.0 * np.pi / x for x in L)

    @property
    def handle(self):
        return self._h

    @property
    def space_(self):
        sp = ctypes.c_int()
        call("mfx_grid_info", self._h, None, ctypes.byref(sp), None)
        return sp.value

    def reset(self):
        if self._h:
            call("mfx_grid_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.reset()
        except Exception:
            pass

    # --- data movement (host mirror) ---
    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.ntot_:
            raise ValueError("array size must equal ntot_")
        call("mfx_grid_upload", self._h, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)

    def download(self):
        a = np.empty(self.ntot_, dtype=self.dtype)
        call("mfx_grid_download", self._h, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)
        n0, n1, n2 = self.n_
        return a.reshape(n0, n1, n2 + 2)

    def kdata(self):
        """k-space view (n0, n1, n2/2+1) complex"""
        a = self.download()
        ct = np.complex128 if self.dtype == np.float64 else np.complex64
        return a.view(ct).reshape(self.n_[0], self.n_[1], self.n_[2] // 2 + 1)

    def rdata(self):
        return self.download()[:, :, :self.n_[2]]

    # --- reference methods ---
    def zero(self):
        call("mfx_grid_zero", self._h)

    def copy_from(self, g):
        call("mfx_grid_copy_from", self._h, g.handle)

    def FourierTransformForward(self, do_transform=True):
        call("mfx_grid_fft_forward", self._h, int(do_transform))

    def FourierTransformBackward(self, do_transform=True):
        call("mfx_grid_fft_backward", self._h, int(do_transform))

    def __imul__(self, x):
        call("mfx_grid_scale", self._h, float(x))
        return self

    def __itruediv__(self, x):
        call("mfx_grid_scale", self._h, 1.0 / float(x))
        return self

    def apply_InverseLaplacian(self):
        call("mfx_grid_inverse_laplacian", self._h)
        return self

    def zero_DC_mode(self):
        call("mfx_grid_zero_dc", self._h)

    def shift_field(self, s, transform_back=True):
        """Grid_FFT::shift_field (include/grid_fft.hh:832-842): multiply every
        mode by exp(i s.k dx), s in cells; ends in real space with
        transform_back, else in k space"""
        call("mfx_grid_shift_field", self._h, float(s[0]), float(s[1]), float(s[2]), int(bool(transform_back)))

    def FourierInterpolateCopyTo(self, grid_to):
        """Grid_FFT::FourierInterpolateCopyTo (src/grid_fft.cc:274-333): both
        grids end in k space; grid_to holds the modes of this grid that both
        represent below their Nyquist modes, zero elsewhere"""
        call("mfx_grid_fourier_interpolate_copy", self._h, grid_to.handle)

    def _stats(self):
        m, s = ctypes.c_double(), ctypes.c_double()
        call("mfx_grid_stats", self._h, ctypes.byref(m), ctypes.byref(s))
        return m.value, s.value

    def mean(self):
        return self._stats()[0]

    def std(self):
        return self._stats()[1]

    def Compute_PowerSpectrum(self):
        nb = self.n_[0]
        bk, bP, beP, bc = (np.zeros(nb) for _ in range(4))
        nbins = ctypes.c_int()
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        call("mfx_grid_power_spectrum", self._h, p(bk), p(bP), p(beP), p(bc), nb, ctypes.byref(nbins))
        k = nbins.value
        return bk[:k], bP[:k], beP[:k], bc[:k]


# ------------------------------------------------------------------ output operators (include/operators.hh)
Op = namedtuple("Op", "grid code fac")


class op:
    @staticmethod
    def assign_to(g):
        return Op(g, _lib.MFX_OP_ASSIGN, 0.0)

    @staticmethod
    def add_to(g):
        return Op(g, _lib.MFX_OP_ADD, 0.0)

    @staticmethod
    def subtract_from(g):
        return Op(g, _lib.MFX_OP_SUB, 0.0)

    @staticmethod
    def multiply_add_to(g, x):
        return Op(g, _lib.MFX_OP_MULADD, float(x))


def _i2(d):
    return (ctypes.c_int * 2)(int(d[0]), int(d[1]))


class OrszagConvolver:
    """OrszagConvolver<real_t> (include/convolution.hh:429-669) with the
    BaseConvolver convolve_* family (:74-292)."""

    def __init__(self, ctx, N, L, precision="double"):
        self.ctx = ctx
        self._h = ctypes.c_void_p()
        N = N[0] if not np.isscalar(N) else N
        L = L[0] if not np.isscalar(L) else L
        call("mfx_conv_create", ctx.handle, int(N), float(L), _prec(precision), ctypes.byref(self._h))

    def __del__(self):
        try:
            if self._h:
                call("mfx_conv_destroy", self._h)
                self._h = ctypes.c_void_p()
        except Exception:
            pass

    def convolve_Gradients(self, inl, d1l, inr, d1r, o):
        call("mfx_conv_gradients", self._h, inl.handle, int(d1l[0]), inr.handle, int(d1r[0]), o.grid.handle,
             o.code, o.fac)

    def convolve_Gradient_and_Hessian(self, inl, d1l, inr, d2r, o):
        call("mfx_conv_gradient_and_hessian", self._h, inl.handle, int(d1l[0]), inr.handle, _i2(d2r),
             o.grid.handle, o.code, o.fac)

    def convolve_Hessians(self, inl, d2l, inr, d2r, *rest):
        if len(rest) == 1:
            o = rest[0]
            call("mfx_conv_hessians", self._h, inl.handle, _i2(d2l), inr.handle, _i2(d2r), o.grid.handle, o.code,
                 o.fac)
        else:
            g3, d3, o = rest
            call("mfx_conv_hessians3", self._h, inl.handle, _i2(d2l), inr.handle, _i2(d2r), g3.handle, _i2(d3),
                 o.grid.handle, o.code, o.fac)

    def convolve_SumOfHessians(self, inl, d2l, inr, d2r1, d2r2, o):
        call("mfx_conv_sum_of_hessians", self._h, inl.handle, _i2(d2l), inr.handle, _i2(d2r1), _i2(d2r2),
             o.grid.handle, o.code, o.fac)

    def convolve_DifferenceOfHessians(self, inl, d2l, inr, d2r1, d2r2, o):
        call("mfx_conv_difference_of_hessians", self._h, inl.handle, _i2(d2l), inr.handle, _i2(d2r1), _i2(d2r2),
             o.grid.handle, o.code, o.fac)

    def multiply_field(self, inl, inr, o):
        call("mfx_conv_multiply_field", self._h, inl.handle, inr.handle, o.grid.handle, o.code, o.fac)

    def convolve2(self, a, ma, b, mb, o):
        call("mfx_conv2", self._h, a.handle, ctypes.byref(ma), b.handle, ctypes.byref(mb), o.grid.handle, o.code,
             o.fac)

    def lpt_sources(self, phi, phi2, phi3=None, A3=(None, None, None), order=2):
        h = lambda g: g.handle if g is not None else None
        call("mfx_lpt_sources", self._h, phi.handle, phi2.handle, h(phi3), h(A3[0]), h(A3[1]), h(A3[2]), order)

    LPT_PATHS = {"auto": 0, "streamed": 1, "stored": 2}

    def set_lpt_path(self, path):
        """3LPT evaluation path of lpt_sources: "streamed" (every stage
        transforms its Hessian rows), "stored" (phi's real-space Hessian rows
        are kept, half the padded transforms, larger workspace) or "auto"
        (stored when it fits in free device memory)."""
        call("mfx_conv_set_lpt_path", self._h, self.LPT_PATHS[path])

    @property
    def lpt_path(self):
        """path the 3LPT workspace was reserved for ("auto" before the first
        3LPT call)"""
        v = ctypes.c_int()
        call("mfx_conv_lpt_path", self._h, ctypes.byref(v))
        return {i: k for k, i in self.LPT_PATHS.items()}[v.value]


# ------------------------------------------------------------------ RNG plugins (include/random_plugin.hh)
class RNG_music:
    """MUSIC1 plugin (src/plugins/random_music.cc)"""

    def __init__(self, seed, cubesize=32):
        self.seed, self.cubesize = int(seed), int(cubesize)

    def isMultiscale(self):
        return True

    def Fill_Grid(self, g):
        call("mfx_rng_music1_fill", g.handle, self.seed, self.cubesize)


class RNG_ngenic:
    """NGENIC plugin (src/plugins/random_ngenic.cc)"""

    def __init__(self, seed):
        self.seed = int(seed)

    def isMultiscale(self):
        return False

    def Fill_Grid(self, g):
        call("mfx_rng_ngenic_fill", g.handle, self.seed)


class RNG_threefry:
    """Panphasia-HO Threefry-4x64-20 cell stream (first Box-Muller deviate per cell)"""

    def __init__(self, seed):
        self.seed = int(seed)

    def isMultiscale(self):
        return False

    def Fill_Grid(self, g):
        call("mfx_rng_threefry_fill", g.handle, self.seed)


class RNG_panphasia:
    """PANPHASIA plugin with a Panphasia-HO descriptor (src/plugins/random_panphasia.cc)"""

    def __init__(self, descriptor):
        self.descriptor = str(descriptor)

    def isMultiscale(self):
        return True

    def Fill_Grid(self, g):
        call("mfx_rng_panphasia_fill", g.handle, self.descriptor.encode())


def panphasia_info(descriptor, n=0):
    """parsed descriptor fields, its check digit and, for an n-grid, the
    relative level, coefficient cube and Fourier grid sizes"""
    out = (ctypes.c_longlong * 11)()
    call("mfx_panphasia_info", str(descriptor).encode(), int(n), out)
    keys = ["order", "level", "x", "y", "z", "size", "kk", "check", "rel_level", "np", "N0"]
    return dict(zip(keys, list(out)))


def panphasia_tables():
    """the library's order-6 basis tables: (C 84x84, degrees 3x84, row_index 84)"""
    C = np.zeros((84, 84))
    deg = np.zeros((3, 84), dtype=np.int32)
    ri = np.zeros(84, dtype=np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    call("mfx_panphasia_tables", p(C), p(deg), p(ri))
    return C, deg, ri


def select_RNG_plugin(name, seed, descriptor=None):
    """random_plugin.cc:136-191"""
    if name == "PANPHASIA":
        return RNG_panphasia(descriptor)
    table = {"MUSIC1": RNG_music, "NGENIC": RNG_ngenic, "PANPHASIA_THREEFRY": RNG_threefry}
    if name not in table:
        raise RuntimeError(f"Unknown random number generator plugin '{name}'")
    return table[name](seed)


def mt19937_raw(ctx, seeds, count):
    seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
    out = np.zeros(len(seeds) * count, dtype=np.uint32)
    call("mfx_rng_mt19937_raw", ctx.handle, seeds.ctypes.data_as(ctypes.c_void_p), len(seeds), count,
         out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(len(seeds), count)


def ranlxd1_raw(ctx, seeds, count):
    seeds = np.ascontiguousarray(seeds, dtype=np.uint32)
    out = np.zeros(len(seeds) * count)
    call("mfx_rng_ranlxd1_raw", ctx.handle, seeds.ctypes.data_as(ctypes.c_void_p), len(seeds), count,
         out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(len(seeds), count)


def threefry_uniforms(ctx, cells, seed, count):
    cells = np.ascontiguousarray(cells, dtype=np.uint64).reshape(-1, 4)
    out = np.zeros(len(cells) * count)
    call("mfx_rng_threefry_uniforms", ctx.handle, cells.ctypes.data_as(ctypes.c_void_p), len(cells), int(seed),
         count, out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(len(cells), count)


# ------------------------------------------------------------------ cosmology (host FP64)
class CosmologyCalculator:
    """cosmology::calculator with the Eisenstein & Hu transfer function"""
    TRANSFER, DPLUS, F, VFACT, H, AMPLITUDE, PNORM, OMEGA_DE = range(8)

    def __init__(self, H0=70.0, Omega_m=0.3, Omega_b=0.045, n_s=0.96, sigma_8=0.81, Tcmb=2.7255,
                 zero_radiation=False):
        load()
        self._h = ctypes.c_void_p()
        call("mfx_cosmo_create", H0, Omega_m, Omega_b, n_s, sigma_8, Tcmb, int(zero_radiation),
             ctypes.byref(self._h))

    def __del__(self):
        try:
            if self._h:
                call("mfx_cosmo_destroy", self._h)
        except Exception:
            pass

    def _q(self, what, x=0.0):
        out = ctypes.c_double()
        call("mfx_cosmo_query", self._h, what, float(x), ctypes.byref(out))
        return out.value

    def get_transfer(self, k):
        return self._q(self.TRANSFER, k)

    def get_growth_factor(self, a):
        return self._q(self.DPLUS, a)

    def get_f(self, a):
        return self._q(self.F, a)

    def get_vfact(self, a):
        return self._q(self.VFACT, a)

    def H_of_a(self, a):
        return self._q(self.H, a)

    def get_amplitude(self, k):
        return self._q(self.AMPLITUDE, k)

    @property
    def pnorm(self):
        return self._q(self.PNORM)

    def omega_de(self):
        return self._q(self.OMEGA_DE)

    def amplitude_table(self, n, boxlen):
        nt = 3 * (n // 2) ** 2 + 1
        out = np.zeros(nt)
        call("mfx_cosmo_amplitude_table", self._h, n, float(boxlen), out.ctypes.data_as(ctypes.c_void_p), nt)
        return out


# ------------------------------------------------------------------ k-space operators of ic_generator::run
def noise_normalise(wn, volfac, fixing=False, inversion=False):
    call("mfx_noise_normalise", wn.handle, float(volfac), int(fixing), int(inversion))


def phi1(phi, wn, amp_table):
    t = np.ascontiguousarray(amp_table, dtype=np.float64)
    call("mfx_phi1", phi.handle, wn.handle, t.ctypes.data_as(ctypes.c_void_p), t.size)


def assemble_kspace(tmp, phi, phi2, phi3, A3, order, kind, idim, vfac, unit=1.0, plt=None):
    """ic_generator.cc:821-919 for one component; with `plt` (a PLT object)
    the gradient is particle::lattice_gradient (ENABLE_PLT builds)"""
    h = lambda g: g.handle if g is not None else None
    if plt is None:
        call("mfx_assemble_kspace", tmp.handle, phi.handle, h(phi2), h(phi3), h(A3[0]), h(A3[1]), h(A3[2]), order,
             kind, idim, float(vfac), float(unit))
    else:
        call("mfx_assemble_kspace_plt", tmp.handle, phi.handle, h(phi2), h(phi3), h(A3[0]), h(A3[1]), h(A3[2]),
             order, kind, idim, float(vfac), float(unit), plt.handle)


class PLT:
    """particle::lattice_gradient (include/particle_plt.hh:48-585) for the sc
    lattice: init_D runs on the device at construction"""

    def __init__(self, ctx, ngrid=64):
        self.ctx = ctx
        self.ng = int(ngrid)
        self._h = ctypes.c_void_p()
        call("mfx_plt_create", ctx.handle, self.ng, ctypes.byref(self._h))

    @property
    def handle(self):
        return self._h

    def __del__(self):
        try:
            if self._h:
                call("mfx_plt_destroy", self._h)
        except Exception:
            pass

    def tables(self):
        """(4, ng, ng, ng/2+1): D_r, D_theta, D_phi, growth-rate factor"""
        nh = self.ng // 2 + 1
        out = np.zeros((4, self.ng, self.ng, nh))
        call("mfx_plt_tables", self._h, out.ctypes.data_as(ctypes.c_void_p), out.size)
        return out

    def dmatrix(self):
        """(6, ng, ng, ng/2+1) real parts of the unnormalised k-space D_xx..D_zz"""
        nh = self.ng // 2 + 1
        out = np.zeros((6, self.ng, self.ng, nh))
        call("mfx_plt_dmatrix", self._h, out.ctypes.data_as(ctypes.c_void_p), out.size)
        return out

    def gradient(self, n, boxlen, ijk):
        """gradient (imaginary parts) and vfac_corr at modes ijk (m x 3) of an n-grid"""
        ijk = np.ascontiguousarray(ijk, dtype=np.int64).reshape(-1, 3)
        m = ijk.shape[0]
        g = np.zeros((m, 3))
        v = np.zeros(m)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        call("mfx_plt_gradient", self._h, int(n), float(boxlen), p(ijk), m, p(g), p(v))
        return g, v


# ------------------------------------------------------------------ LPT driver (src/ic_generator.cc:90-971)
GENERATORS = {"MUSIC1": 0, "NGENIC": 1, "PANPHASIA_THREEFRY": 2, "PANPHASIA": 3}
LATTICES = {"sc": 0, "bcc": 1, "fcc": 2, "rsc": 3}  # [setup] ParticleLoad (ic_generator.cc:131-139)


class ICGenerator:
    """Device-resident ic_generator::run (FP64 particles) for the sc, bcc, fcc
    and rsc particle loads."""

    def __init__(self, ctx, GridRes, BoxLength, zstart, LPTorder=3, precision="double", generator="MUSIC1",
                 seed=12345, DoFixing=False, DoInversion=False, H0=70.0, Omega_m=0.3, Omega_b=0.045, n_s=0.96,
                 sigma_8=0.81, Tcmb=2.7255, ZeroRadiation=False, fused=True, PLT=False, plt_ngrid=64,
                 autotune=True, assemble_each=False, field_output=False, own_particles=False, ParticleLoad="sc",
                 DoBaryons=False, generic_out_eulerian=False, descriptor=None, cubesize=32, amplitude_table=None):
        self.ctx = ctx
        self.n = int(GridRes)
        self.lattice = str(ParticleLoad)
        self.overload = 1 << LATTICES[self.lattice]  # particles per cell
        self.boxlen, self.zstart, self.lpt_order = float(BoxLength), float(zstart), int(LPTorder)
        self.cosmo = dict(H0=H0, Omega_m=Omega_m, Omega_b=Omega_b)
        self.generator, self.plt = generator, bool(PLT)
        self.field_output = bool(field_output)
        self.baryons = bool(DoBaryons)
        # [output] generic_out_eulerian: semiclassical density and velocity fields
        self.eulerian = bool(generic_out_eulerian)
        self._host = [None] * 13
        p = _lib.IcgenParams()
        p.n, p.boxlen, p.zstart, p.lpt_order = self.n, float(BoxLength), float(zstart), int(LPTorder)
        p.precision, p.generator, p.seed = _prec(precision), GENERATORS[generator], int(seed)
        p.fixing, p.inversion = int(DoFixing), int(DoInversion)
        p.H0, p.Omega_m, p.Omega_b, p.n_s, p.sigma_8, p.Tcmb = H0, Omega_m, Omega_b, n_s, sigma_8, Tcmb
        p.zero_radiation, p.use_fused_lpt = int(ZeroRadiation), int(fused)
        p.use_plt, p.plt_ngrid = int(PLT), int(plt_ngrid)
        p.autotune = int(autotune)
        p.assemble_each = int(assemble_each)
        p.field_output, p.own_particles = int(field_output), int(own_particles)
        p.lattice = LATTICES[self.lattice]
        p.baryons = int(self.baryons)
        p.eulerian = int(self.eulerian)
        # [random] descriptor of the PANPHASIA plugin (kept alive with the generator)
        self._descriptor = None if descriptor is None else str(descriptor).encode()
        p.descriptor = self._descriptor
        # [random] cubesize of the MUSIC1 plugin (random_music.cc:120)
        p.cubesize = int(cubesize)
        # A(kfac sqrt(q)) for q = 0 .. 3(n/2)^2 of any transfer function; None:
        # Eisenstein-Hu from the cosmology parameters (read at creation only)
        if amplitude_table is not None:
            amp = np.ascontiguousarray(amplitude_table, dtype=np.float64)
            p.amplitude_table = amp.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            p.amplitude_ntable = amp.size
        self._h = ctypes.c_void_p()
        call("mfx_icgen_create", ctx.handle, ctypes.byref(p), ctypes.byref(self._h))

    def __del__(self):
        try:
            if self._h:
                call("mfx_icgen_destroy", self._h)
                self._h = ctypes.c_void_p()
            for b in self._host:
                if b is not None:
                    call("mfx_host_free", ctypes.c_void_p(b[0]))
            self._host = [None] * 13
        except Exception:
            pass

    def _shape(self, w):
        n = self.n
        return (n, n, n) if (w == 6 or self.overload == 1) else (self.overload, n, n, n)

    def _pinned(self, w):
        """pinned host buffer of output array w (n^3 doubles per sublattice),
        as a numpy view"""
        if self._host[w] is None:
            cnt = int(np.prod(self._shape(w)))
            ptr = ctypes.c_void_p()
            call("mfx_host_alloc", cnt * 8, ctypes.byref(ptr))
            arr = np.ctypeslib.as_array((ctypes.c_double * cnt).from_address(ptr.value))
            self._host[w] = (ptr.value, arr.reshape(self._shape(w)))
        return self._host[w][1]

    def download_async(self, what=range(6)):
        """start copying output arrays `what` (0..5, 6 = density of the field
        output, 7..12 the baryons') of the last realisation into pinned host
        buffers; the next run overlaps these copies until it overwrites the
        arrays"""
        mask = 0
        dst = (ctypes.c_void_p * 13)()
        for w in what:
            self._pinned(w)
            dst[w] = self._host[w][0]
            mask |= 1 << w
        call("mfx_icgen_download_async", self._h, mask, dst)

    def download_wait(self):
        """wait for download_async; returns the pinned host arrays (views
        that the next download overwrites)"""
        call("mfx_icgen_download_wait", self._h)
        return [h[1] if h is not None else None for h in self._host]

    def write_generic(self, fname, fields=None):
        """generic HDF5 output (src/plugins/output_generic.cc): DM_dx..DM_vz
        and DM_delta from the downloaded arrays (field output only:
        displacements without the lattice, lunit = vunit = 1); with
        generic_out_eulerian the semiclassical DM_delta and DM_vx..DM_vz"""
        from .io import GenericOutput
        if not self.field_output:
            raise ValueError("generic output needs field_output=True")
        arrs = fields if fields is not None else self.download_wait()
        names = ["DM_dx", "DM_dy", "DM_dz", "DM_vx", "DM_vy", "DM_vz", "DM_delta"]
        astart = 1.0 / (1.0 + self.zstart)
        Om, Ob = self.cosmo["Omega_m"], self.cosmo["Omega_b"]
        with GenericOutput(fname, self.boxlen, astart, Ob, Om, self._omega_de()) as out:
            for w, name in enumerate(names):
                if arrs[w] is not None and not (self.eulerian and w < 3):
                    out.write_field(name, np.ascontiguousarray(arrs[w]))

    def write_gadget(self, fname, double_prec=True, long_ids=False, arrays=None):
        """Gadget-HDF5 particle output (src/plugins/output_gadget_hdf5.cc):
        positions in Mpc/h (lunit = BoxLength), velocities with vunit =
        BoxLength / sqrt(a), from the downloaded arrays 0..5"""
        from .io import GadgetOutput
        if self.field_output:
            raise ValueError("particle output needs field_output=False")
        arrs = arrays if arrays is not None else self.download_wait()
        a = 1.0 / (1.0 + self.zstart)
        L = self.boxlen
        h = self.cosmo["H0"] / 100.0
        Om, Ob = self.cosmo["Omega_m"], self.cosmo["Omega_b"]
        with GadgetOutput(fname, self.zstart, L, Om, self._omega_de(), Ob, h,
                          self.lpt_order, self.n, double_prec=double_prec, long_ids=long_ids, rng=self.generator,
                          plt="yes" if self.plt else "no", particle_load=self.lattice,
                          do_baryons=self.baryons) as out:
            # ic_generator.cc:158-168: with baryons the dark matter carries Omega_m - Omega_b
            out.write_particles(arrs[0:3], arrs[3:6], self.n, Om - Ob if self.baryons else Om, L, L / np.sqrt(a))
            if self.baryons:
                out.write_particles(arrs[7:10], arrs[10:13], self.n, Ob, L, L / np.sqrt(a), species=0)

    def _omega_de(self):
        c = CosmologyCalculator(H0=self.cosmo["H0"], Omega_m=self.cosmo["Omega_m"], Omega_b=self.cosmo["Omega_b"])
        return c.omega_de()

    def run(self):
        call("mfx_icgen_run", self._h)

    def set_seed(self, seed):
        """stage the noise seed of the next realisation (host part of the fill)"""
        call("mfx_icgen_set_seed", self._h, int(seed))

    def capture(self):
        """run once and capture the device sequence of a realisation as a HIP graph"""
        call("mfx_icgen_capture", self._h)

    def launch(self):
        """replay the captured realisation for the staged seed (asynchronous)"""
        call("mfx_icgen_launch", self._h)

    def hbar(self):
        """hbar of the last realisation's semiclassical fields (Eulerian output)"""
        h = ctypes.c_double()
        call("mfx_icgen_hbar", self._h, ctypes.byref(h))
        return h.value

    def scalars(self):
        d, v = ctypes.c_double(), ctypes.c_double()
        call("mfx_icgen_scalars", self._h, ctypes.byref(d), ctypes.byref(v))
        return d.value, v.value

    def particles_dev(self, what):
        p = ctypes.c_void_p()
        call("mfx_icgen_particles", self._h, what, ctypes.byref(p))
        return p

    def density(self):
        """host copy of the field output's 1st-order density (n,n,n)"""
        n = self.n
        return self.ctx.d2h(self.particles_dev(6), n ** 3 * 8, np.float64).reshape(n, n, n)

    def particles(self, what, species="dm"):
        """host copy of positions (0..2) or velocities (3..5) of the dark
        matter or (species "baryon", DoBaryons) the baryons: shape (n,n,n)
        for the sc load, (sublattice, n, n, n) for bcc, fcc and rsc"""
        if species == "baryon":
            what = what + 7
        shp = self._shape(what)
        return self.ctx.d2h(self.particles_dev(what), int(np.prod(shp)) * 8, np.float64).reshape(shp)

    def grid_handle(self, which):
        p = ctypes.c_void_p()
        call("mfx_icgen_grid", self._h, which, ctypes.byref(p))
        return p

    def grid_data(self, which):
        """download grid `which` (0 wnoise, 1 phi, 2 phi2, 3 phi3, 4..6 A3, 7 tmp) as its raw padded array"""
        h = self.grid_handle(which)
        nt, sp, pr = ctypes.c_size_t(), ctypes.c_int(), ctypes.c_int()
        call("mfx_grid_info", h, ctypes.byref(nt), ctypes.byref(sp), ctypes.byref(pr))
        dt = np.float64 if pr.value == 8 else np.float32
        a = np.empty(nt.value, dtype=dt)
        call("mfx_grid_download", h, a.ctypes.data_as(ctypes.c_void_p), a.nbytes)
        return a.reshape(self.n, self.n, self.n + 2)


class Batch:
    """many realisations (any mix of grid sizes) as one captured HIP graph:
    launch() replays every generator's realisation for its staged seed"""

    def __init__(self, generators, nstreams=8):
        self.gens = list(generators)
        arr = (ctypes.c_void_p * len(self.gens))(*[g._h.value for g in self.gens])
        self._h = ctypes.c_void_p()
        call("mfx_batch_create", arr, len(self.gens), int(nstreams), ctypes.byref(self._h))

    def launch(self):
        call("mfx_batch_launch", self._h)

    def synchronize(self):
        call("mfx_batch_synchronize", self._h)

    def close(self):
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            call("mfx_batch_destroy", h)
        self.gens = []  # the generators may now be destroyed

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
