"""C ABI of liblagrbf.so without a GPU: the library loads, exports every symbol
include/*.h declares (and the Python mirror binds), and its host-only parts
(options channel, option validation, error reporting) behave as the reference's
(PETSc error codes, messages kept for lagrbf_last_error).  No compute call."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))
LIB = os.path.join(ROOT, "lagrbf_amd", "liblagrbf.so")

# symbols the reference's Python layer binds (register_cfunc.py:58-195)
REFERENCE_SYMBOLS = {
    "lagrbf_create", "lagrbf_destroy", "lagrbf_fill_swarm", "lagrbf_interpolation", "lagrbf_output",
    "lagrbf_avect_analytical", "lagrbf_advect_interpolate", "lagrbf_advection_function",
    "lagrbf_advection_function_batch", "lagrbf_reset_coefficients", "lagrbf_interpolate_driver",
}


def header_symbols():
    syms = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        syms |= set(re.findall(r"\b(lagrbf_\w+)\s*\(", text))
    return syms


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("liblagrbf.so not built (python -m lagrbf_amd.build)")
    return ctypes.CDLL(LIB)


def test_header_declares_reference_abi():
    syms = header_symbols()
    assert REFERENCE_SYMBOLS <= syms, REFERENCE_SYMBOLS - syms


def test_library_exports_every_declared_symbol(lib):
    for s in sorted(header_symbols()):
        assert hasattr(lib, s), s
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = header_symbols() - exported
    assert not missing, missing


def test_python_mirror_binds_the_header():
    from lagrbf_amd import _lib
    assert set(_lib.EXPORTED_SYMBOLS) == header_symbols()


def last_error(lib):
    buf = ctypes.create_string_buffer(1024)
    assert lib.lagrbf_last_error(buf, len(buf)) == 0
    return buf.value.decode()


def create(lib):
    lib.lagrbf_create.restype = ctypes.c_void_p
    return lib.lagrbf_create()


def test_missing_required_options_reported(lib):
    assert lib.lagrbf_clear_options() == 0
    assert create(lib) is None
    assert "error code 62" in last_error(lib) and "pu_mesh_dim" in last_error(lib)
    assert lib.lagrbf_set_option(b"pu_mesh_dim", b"4") == 0
    assert create(lib) is None
    msg = last_error(lib)
    assert "pu_mesh_n" in msg or "dimension" in msg.lower()
    lib.lagrbf_clear_options()


def test_option_validation_before_device(lib):
    lib.lagrbf_clear_options()
    for k, v in (("pu_mesh_dim", "3"), ("pu_mesh_n", "4,4,4"), ("kernel_type", "9")):
        assert lib.lagrbf_set_option(k.encode(), v.encode()) == 0
    assert create(lib) is None
    assert "error code" in last_error(lib)
    lib.lagrbf_clear_options()


def test_options_file_channel(lib, tmp_path):
    f = tmp_path / "opts.txt"
    f.write_text("# PU grid\n-pu_mesh_dim 2\n-pu_mesh_n 5,5\n-kernel_type 2   # Wendland C2\n")
    lib.lagrbf_clear_options()
    assert lib.lagrbf_options_insert_file(str(f).encode()) == 0
    assert lib.lagrbf_options_insert_file(str(tmp_path / "absent.txt").encode()) != 0
    lib.lagrbf_clear_options()


def test_null_context_rejected(lib):
    lib.lagrbf_interpolation.argtypes = [ctypes.c_void_p]
    assert lib.lagrbf_interpolation(None) == 85          # PETSC_ERR_ARG_NULL
    lib.lagrbf_advect_steps.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double]
    assert lib.lagrbf_advect_steps(None, 1, 1e-3) == 85
    v = ctypes.c_double(0)
    lib.lagrbf_get_info.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
    assert lib.lagrbf_get_info(None, b"removed", ctypes.byref(v)) != 0
