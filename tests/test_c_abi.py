"""The C ABI from C: tests/c/abi_roundtrip.c includes include/mfx.h as C99
and links libmfx.so.  It compiles and links on the CPU; on a GPU it runs a
16^3 forward/backward transform and checks the plane-wave mode and the
round trip."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build(tmp_path):
    exe = str(tmp_path / "abi_roundtrip")
    lib = os.path.join(ROOT, "monofonic_amd")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "abi_roundtrip.c"), "-o", exe, "-L", lib, "-lmfx", f"-Wl,-rpath,{lib}",
           "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_c_program_compiles_and_links(tmp_path):
    build(tmp_path)


@pytest.mark.gpu
def test_c_program_round_trip(tmp_path):
    exe = build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"


def build_shim(tmp_path):
    """tests/c/grid_shim.cpp: the Grid_FFT shim of the integration notes as C++17"""
    exe = str(tmp_path / "grid_shim")
    lib = os.path.join(ROOT, "monofonic_amd")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "grid_shim.cpp"), "-o", exe, "-L", lib, "-lmfx", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_grid_shim_compiles_and_links(tmp_path):
    build_shim(tmp_path)


def test_integration_notes_show_the_compiled_shim():
    """the Grid_FFT and OrszagConvolver blocks of the integration notes are the compiled ones"""
    src = open(os.path.join(ROOT, "tests", "c", "grid_shim.cpp")).read()
    notes = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    for block in ("SHIM", "CONV"):
        code = src.split(f"// ---- {block} BEGIN\n")[1].split(f"// ---- {block} END")[0].strip()
        assert code in notes, block


@pytest.mark.gpu
def test_grid_shim_runs(tmp_path):
    """plane wave through the host mirror, transforms, stats, power spectrum,
    shift_field and FourierInterpolateCopyTo through the C++ shim; the
    convolver shim on a closed-form Hessian product with assign_to and
    multiply_add_to"""
    exe = build_shim(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip() == "ok"
