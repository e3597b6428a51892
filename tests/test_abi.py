"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/mfx.h declares; the Python mirror binds exactly that set."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "mfx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(mfx_\w+)\s*\(", txt, flags=re.M)))


def test_header_parses():
    fns = header_functions()
    assert "mfx_grid_fft_forward" in fns and "mfx_conv2" in fns and len(fns) > 50


def test_library_exports_every_declared_symbol():
    from monofonic_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_python_binding_covers_header():
    from monofonic_amd import _lib
    assert sorted(_lib.EXPORTS) == header_functions()


def test_exports_are_c_symbols():
    from monofonic_amd import _lib
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    syms = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    for f in header_functions():
        assert f in syms, f


def test_library_fails_loudly_without_device():
    """without a HIP device the context cannot be created and the call reports an error"""
    import monofonic_amd as M
    lib = M.load()
    h = ctypes.c_void_p()
    st = lib.mfx_ctx_create(12345, ctypes.byref(h))
    assert st != 0
    assert lib.mfx_last_error()
