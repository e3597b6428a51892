"""Python view of the host output boundary (include/mfx_io.h, libmfx_io.so):
the generic HDF5 field writer (src/plugins/output_generic.cc) and the
Gadget-HDF5 particle writer (src/plugins/output_gadget_hdf5.cc)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
IO_LIB_PATH = os.path.join(HERE, "libmfx_io.so")
_io = None


class MfxIoError(RuntimeError):
    pass


class GadgetParams(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("zstart", "boxlength", "omega_m", "omega_de", "omega_b", "h")] + \
               [(k, ctypes.c_int) for k in ("lpt_order", "grid_res", "double_prec", "long_ids", "gadget2_compat",
                                            "do_fixing", "do_inversion", "do_baryons", "do_baryons_vrel")] + \
               [(k, ctypes.c_char_p) for k in ("particle_load", "transfer", "param_set", "rng", "precision",
                                               "convolutions", "plt", "git_rev", "git_tag", "git_branch")]


def load_io():
    global _io
    if _io is not None:
        return _io
    if not os.path.exists(IO_LIB_PATH):
        raise MfxIoError(f"{IO_LIB_PATH} not found: build it with `python monofonic_amd/build.py`")
    L = ctypes.CDLL(IO_LIB_PATH)
    vp, sz, i32, dbl = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.c_double
    pp = ctypes.POINTER(ctypes.c_void_p)
    L.mfx_io_last_error.restype = ctypes.c_char_p
    sig = {
        "mfx_out_generic_create": [ctypes.c_char_p, dbl, dbl, dbl, dbl, dbl, pp],
        "mfx_out_write_field": [vp, ctypes.c_char_p, vp, i32, sz, sz, sz, sz],
        "mfx_out_gadget_create": [ctypes.c_char_p, ctypes.POINTER(GadgetParams), pp],
        "mfx_out_gadget_write_particles": [vp, i32, ctypes.POINTER(vp), ctypes.POINTER(vp), sz, dbl, dbl, dbl],
        "mfx_out_close": [vp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _io = L
    return L


def _call(name, *args):
    L = load_io()
    if getattr(L, name)(*args) != 0:
        raise MfxIoError(f"{name}: {L.mfx_io_last_error().decode()}")


class _Writer:
    def __init__(self):
        self._h = ctypes.c_void_p()

    def close(self):
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            _call("mfx_out_close", h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GenericOutput(_Writer):
    """generic_output_plugin: Header attributes and one dataset per field"""

    def __init__(self, fname, boxsize, astart, Omega_b, Omega_m, Omega_L):
        super().__init__()
        _call("mfx_out_generic_create", os.fsencode(fname), float(boxsize), float(astart), float(Omega_b),
              float(Omega_m), float(Omega_L), ctypes.byref(self._h))

    def write_field(self, name, a):
        """a: (n0, n1, n2) or padded (n0, n1, pitch) array with the field in
        its first n2 entries per row; n2 = a.shape[2] unless n2 is given"""
        return self.write_field_padded(name, a, a.shape[2])

    def write_field_padded(self, name, a, n2):
        a = np.ascontiguousarray(a)
        if a.dtype not in (np.float32, np.float64) or a.ndim != 3:
            raise ValueError("field must be a 3D float32/float64 array")
        _call("mfx_out_write_field", self._h, name.encode(), a.ctypes.data_as(ctypes.c_void_p), a.itemsize,
              a.shape[0], a.shape[1], int(n2), a.shape[2])


class GadgetOutput(_Writer):
    """gadget_hdf5 (float) / gadget_hdf5_double particle files"""

    def __init__(self, fname, zstart, boxlength, Omega_m, Omega_DE, Omega_b, h, lpt_order, grid_res,
                 double_prec=False, long_ids=False, gadget2_compat=False, rng="MUSIC1", plt="no",
                 precision="DOUBLE", transfer="eisenstein", particle_load="sc", do_baryons=False):
        super().__init__()
        self.overload = {"sc": 1, "bcc": 2, "fcc": 4, "rsc": 8}[particle_load]
        p = GadgetParams()
        p.zstart, p.boxlength, p.omega_m, p.omega_de, p.omega_b, p.h = zstart, boxlength, Omega_m, Omega_DE, \
            Omega_b, h
        p.lpt_order, p.grid_res = int(lpt_order), int(grid_res)
        p.do_baryons = int(do_baryons)
        p.double_prec, p.long_ids, p.gadget2_compat = int(double_prec), int(long_ids), int(gadget2_compat)
        p.particle_load, p.transfer, p.param_set = particle_load.encode(), transfer.encode(), b"none"
        p.rng, p.precision, p.convolutions, p.plt = rng.encode(), precision.encode(), b"ORSZAG", plt.encode()
        self._keep = p
        _call("mfx_out_gadget_create", os.fsencode(fname), ctypes.byref(p), ctypes.byref(self._h))

    def write_particles(self, pos, vel, n, omega_species, lscale, vscale, species=1):
        """pos, vel: three FP64 arrays of n^3 particles per sublattice each
        (host), sublattice-major"""
        arrs = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in list(pos) + list(vel)]
        for x in arrs:
            if x.size != self.overload * n ** 3:
                raise ValueError("each component array must hold n^3 values per sublattice")
        P = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in arrs[:3]])
        V = (ctypes.c_void_p * 3)(*[x.ctypes.data for x in arrs[3:]])
        _call("mfx_out_gadget_write_particles", self._h, int(species), P, V, int(n), float(omega_species),
              float(lscale), float(vscale))
