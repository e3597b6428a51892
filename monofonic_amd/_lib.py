"""Loader for the gfx950 library libmfx.so (C ABI declared in include/mfx.h).

There is no CPU fallback: if the shared library is missing or no HIP device is
present, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmfx.so")
_lib = None

MFX_FLOAT, MFX_DOUBLE = 4, 8
MFX_KSPACE, MFX_RSPACE = 0, 1
MFX_OP_ASSIGN, MFX_OP_ADD, MFX_OP_SUB, MFX_OP_MULADD = 0, 1, 2, 3

# every symbol declared in include/mfx.h
EXPORTS = [
    "mfx_last_error", "mfx_version", "mfx_ctx_create", "mfx_ctx_destroy", "mfx_ctx_synchronize",
    "mfx_event_record", "mfx_event_elapsed", "mfx_profile_enable", "mfx_profile_reset", "mfx_profile_count",
    "mfx_profile_get", "mfx_device_alloc", "mfx_device_free", "mfx_memcpy_d2h", "mfx_memcpy_h2d",
    "mfx_autotune", "mfx_autotune_report", "mfx_set_column_tile",
    "mfx_grid_create", "mfx_grid_destroy", "mfx_grid_info", "mfx_grid_data", "mfx_grid_upload",
    "mfx_grid_download", "mfx_grid_zero", "mfx_grid_copy_from", "mfx_grid_fft_forward", "mfx_grid_fft_backward",
    "mfx_grid_scale", "mfx_grid_inverse_laplacian", "mfx_grid_zero_dc", "mfx_grid_stats",
    "mfx_grid_shift_field", "mfx_grid_fourier_interpolate_copy",
    "mfx_grid_power_spectrum", "mfx_noise_normalise", "mfx_phi1", "mfx_conv_create", "mfx_conv_destroy",
    "mfx_conv2", "mfx_conv3", "mfx_conv_gradients", "mfx_conv_gradient_and_hessian", "mfx_conv_hessians",
    "mfx_conv_hessians3", "mfx_conv_sum_of_hessians", "mfx_conv_difference_of_hessians",
    "mfx_conv_multiply_field", "mfx_lpt_sources", "mfx_conv_set_lpt_path", "mfx_conv_lpt_path", "mfx_plt_create", "mfx_plt_destroy", "mfx_plt_tables",
    "mfx_plt_dmatrix", "mfx_plt_gradient", "mfx_assemble_kspace_plt", "mfx_assemble_kspace", "mfx_set_particles",
    "mfx_rng_music1_fill", "mfx_rng_music1_fill_noalloc", "mfx_rng_ngenic_fill", "mfx_rng_threefry_fill", "mfx_rng_panphasia_fill",
    "mfx_panphasia_info", "mfx_panphasia_tables", "mfx_rng_mt19937_raw",
    "mfx_rng_ranlxd1_raw", "mfx_rng_threefry_uniforms", "mfx_cosmo_create", "mfx_cosmo_destroy",
    "mfx_cosmo_query", "mfx_cosmo_amplitude_table", "mfx_icgen_create", "mfx_icgen_destroy", "mfx_icgen_run",
    "mfx_icgen_set_seed", "mfx_icgen_capture", "mfx_icgen_launch",
    "mfx_icgen_particles", "mfx_icgen_grid", "mfx_icgen_scalars", "mfx_icgen_hbar", "mfx_icgen_download_async",
    "mfx_icgen_download_wait", "mfx_host_alloc", "mfx_host_free", "mfx_batch_create", "mfx_batch_launch",
    "mfx_batch_synchronize", "mfx_batch_destroy",
]


class MfxError(RuntimeError):
    pass


class Mult(ctypes.Structure):
    """mfx_mult: sum_t coef_t * prod_d f_d^pow_t[d]"""
    _fields_ = [("nterms", ctypes.c_int), ("coef_re", ctypes.c_double * 2), ("coef_im", ctypes.c_double * 2),
                ("pow", (ctypes.c_int * 3) * 2), ("nyq", ctypes.c_int * 2)]


class IcgenParams(ctypes.Structure):
    _fields_ = [("n", ctypes.c_size_t), ("boxlen", ctypes.c_double), ("zstart", ctypes.c_double),
                ("lpt_order", ctypes.c_int), ("precision", ctypes.c_int), ("generator", ctypes.c_int),
                ("seed", ctypes.c_long), ("fixing", ctypes.c_int), ("inversion", ctypes.c_int),
                ("H0", ctypes.c_double), ("Omega_m", ctypes.c_double), ("Omega_b", ctypes.c_double),
                ("n_s", ctypes.c_double), ("sigma_8", ctypes.c_double), ("Tcmb", ctypes.c_double),
                ("zero_radiation", ctypes.c_int), ("use_fused_lpt", ctypes.c_int), ("use_plt", ctypes.c_int),
                ("plt_ngrid", ctypes.c_int), ("autotune", ctypes.c_int),
                ("assemble_each", ctypes.c_int), ("field_output", ctypes.c_int), ("own_particles", ctypes.c_int),
                ("lattice", ctypes.c_int), ("baryons", ctypes.c_int),
                ("eulerian", ctypes.c_int), ("descriptor", ctypes.c_char_p), ("cubesize", ctypes.c_int),
                ("amplitude_table", ctypes.POINTER(ctypes.c_double)), ("amplitude_ntable", ctypes.c_size_t)]


def load():
    """Load libmfx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MfxError(f"{LIB_PATH} not found: build it with `python monofonic_amd/build.py`")
    L = ctypes.CDLL(LIB_PATH)
    L.mfx_last_error.restype = ctypes.c_char_p
    vp = ctypes.c_void_p
    sz = ctypes.c_size_t
    dbl = ctypes.c_double
    i32 = ctypes.c_int
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "mfx_ctx_create": [i32, pp],
        "mfx_ctx_destroy": [vp], "mfx_ctx_synchronize": [vp],
        "mfx_event_record": [vp, i32], "mfx_event_elapsed": [vp, i32, i32, ctypes.POINTER(ctypes.c_float)],
        "mfx_profile_enable": [vp, i32], "mfx_profile_reset": [vp],
        "mfx_profile_count": [vp, ctypes.POINTER(i32)],
        "mfx_profile_get": [vp, i32, ctypes.c_char_p, sz, ctypes.POINTER(i32), ctypes.POINTER(dbl),
                            ctypes.POINTER(dbl)],
        "mfx_device_alloc": [vp, sz, pp], "mfx_device_free": [vp, vp],
        "mfx_memcpy_d2h": [vp, vp, vp, sz], "mfx_memcpy_h2d": [vp, vp, vp, sz],
        "mfx_autotune": [vp, sz, i32], "mfx_autotune_report": [vp, ctypes.c_char_p, sz],
        "mfx_set_column_tile": [vp, i32, i32, i32, i32],
        "mfx_grid_create": [vp, sz, sz, sz, dbl, dbl, dbl, i32, i32, pp],
        "mfx_grid_destroy": [vp],
        "mfx_grid_info": [vp, ctypes.POINTER(sz), ctypes.POINTER(i32), ctypes.POINTER(i32)],
        "mfx_grid_data": [vp, pp],
        "mfx_grid_upload": [vp, vp, sz], "mfx_grid_download": [vp, vp, sz],
        "mfx_grid_zero": [vp], "mfx_grid_copy_from": [vp, vp],
        "mfx_grid_fft_forward": [vp, i32], "mfx_grid_fft_backward": [vp, i32],
        "mfx_grid_scale": [vp, dbl], "mfx_grid_inverse_laplacian": [vp], "mfx_grid_zero_dc": [vp],
        "mfx_grid_stats": [vp, ctypes.POINTER(dbl), ctypes.POINTER(dbl)],
        "mfx_grid_shift_field": [vp, dbl, dbl, dbl, i32], "mfx_grid_fourier_interpolate_copy": [vp, vp],
        "mfx_grid_power_spectrum": [vp, vp, vp, vp, vp, i32, ctypes.POINTER(i32)],
        "mfx_noise_normalise": [vp, dbl, i32, i32],
        "mfx_phi1": [vp, vp, vp, sz],
        "mfx_conv_create": [vp, sz, dbl, i32, pp], "mfx_conv_destroy": [vp],
        "mfx_conv2": [vp, vp, vp, vp, vp, vp, i32, dbl],
        "mfx_conv3": [vp, vp, vp, vp, vp, vp, vp, vp, i32, dbl],
        "mfx_conv_gradients": [vp, vp, i32, vp, i32, vp, i32, dbl],
        "mfx_conv_gradient_and_hessian": [vp, vp, i32, vp, vp, vp, i32, dbl],
        "mfx_conv_hessians": [vp, vp, vp, vp, vp, vp, i32, dbl],
        "mfx_conv_hessians3": [vp, vp, vp, vp, vp, vp, vp, vp, i32, dbl],
        "mfx_conv_sum_of_hessians": [vp, vp, vp, vp, vp, vp, vp, i32, dbl],
        "mfx_conv_difference_of_hessians": [vp, vp, vp, vp, vp, vp, vp, i32, dbl],
        "mfx_conv_multiply_field": [vp, vp, vp, vp, i32, dbl],
        "mfx_lpt_sources": [vp, vp, vp, vp, vp, vp, vp, i32],
        "mfx_conv_set_lpt_path": [vp, i32], "mfx_conv_lpt_path": [vp, ctypes.POINTER(i32)],
        "mfx_assemble_kspace": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, dbl, dbl],
        "mfx_assemble_kspace_plt": [vp, vp, vp, vp, vp, vp, vp, i32, i32, i32, dbl, dbl, vp],
        "mfx_plt_create": [vp, i32, pp], "mfx_plt_destroy": [vp],
        "mfx_plt_tables": [vp, vp, sz], "mfx_plt_dmatrix": [vp, vp, sz],
        "mfx_plt_gradient": [vp, sz, dbl, vp, i32, vp, vp],
        "mfx_set_particles": [vp, vp, i32, dbl, i32],
        "mfx_rng_music1_fill": [vp, ctypes.c_long, i32],
        "mfx_rng_music1_fill_noalloc": [vp, ctypes.c_long, i32],
        "mfx_rng_ngenic_fill": [vp, ctypes.c_long],
        "mfx_rng_threefry_fill": [vp, ctypes.c_ulonglong],
        "mfx_rng_panphasia_fill": [vp, ctypes.c_char_p],
        "mfx_panphasia_info": [ctypes.c_char_p, i32, ctypes.POINTER(ctypes.c_longlong)],
        "mfx_panphasia_tables": [vp, vp, vp],
        "mfx_rng_mt19937_raw": [vp, vp, i32, i32, vp],
        "mfx_rng_ranlxd1_raw": [vp, vp, i32, i32, vp],
        "mfx_rng_threefry_uniforms": [vp, vp, i32, ctypes.c_ulonglong, i32, vp],
        "mfx_cosmo_create": [dbl, dbl, dbl, dbl, dbl, dbl, i32, pp],
        "mfx_cosmo_destroy": [vp],
        "mfx_cosmo_query": [vp, i32, dbl, ctypes.POINTER(dbl)],
        "mfx_cosmo_amplitude_table": [vp, sz, dbl, vp, sz],
        "mfx_icgen_create": [vp, vp, pp], "mfx_icgen_destroy": [vp], "mfx_icgen_run": [vp],
        "mfx_icgen_set_seed": [vp, ctypes.c_long], "mfx_icgen_capture": [vp], "mfx_icgen_launch": [vp],
        "mfx_icgen_particles": [vp, i32, pp], "mfx_icgen_grid": [vp, i32, pp],
        "mfx_icgen_download_async": [vp, ctypes.c_uint, ctypes.POINTER(vp)], "mfx_icgen_download_wait": [vp],
        "mfx_host_alloc": [sz, pp], "mfx_host_free": [vp],
        "mfx_batch_create": [ctypes.POINTER(vp), i32, i32, pp], "mfx_batch_launch": [vp],
        "mfx_batch_synchronize": [vp], "mfx_batch_destroy": [vp],
        "mfx_icgen_scalars": [vp, ctypes.POINTER(dbl), ctypes.POINTER(dbl)],
        "mfx_icgen_hbar": [vp, ctypes.POINTER(dbl)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    _lib = L
    return L


def check(status):
    if status != 0:
        raise MfxError(f"mfx error {status}: {load().mfx_last_error().decode()}")
    return status


def call(name, *args):
    return check(getattr(load(), name)(*args))
