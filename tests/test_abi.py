"""The C-ABI libraries load and export every symbol their headers declare (no GPU
calls: this runs on machines without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rflu_\w+)\s*\(", text)))


def test_rocflu_amd_exports_header_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "rocflu_amd", "librocflu_amd.so"))
    names = declared("rocflu_amd.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_header():
    from rocflu_amd import _native
    assert set(declared("rocflu_amd.h")) == set(_native.SIGNATURES)


def test_mesh_library_exports_header_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "rocflu_amd", "librflu_mesh.so"))
    missing = [n for n in declared("rflu_mesh.h") if not hasattr(lib, n)]
    assert not missing, missing


def test_device_count_without_gpu_is_safe():
    from rocflu_amd import _native
    n = _native.lib().rflu_gpu_device_count()
    assert n >= 0


def test_create_region_rejects_bad_descriptor():
    from rocflu_amd import _native
    lib = _native.lib()
    d = _native.RegionDesc()
    d.n_cells = -1
    h = ctypes.c_int32(0)
    rc = lib.rflu_gpu_create_region(ctypes.byref(d), ctypes.byref(h))
    assert rc == _native.RFLU_ERR_ARG
    assert lib.rflu_gpu_destroy_region(12345) == _native.RFLU_ERR_HANDLE
