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
    for header in ("rflu_mesh.h", "rflu_input.h"):
        names = declared(header)
        assert len(names) >= 3, header
        missing = [n for n in names if not hasattr(lib, n)]
        assert not missing, (header, missing)


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


def test_create_region_checks_patch_arguments():
    """The descriptor checks of the total-condition inflow, farfield and injection
    patches run before any device call, so they report here without a GPU."""
    import pytest
    from rocflu_amd import _native, mesh as rmesh
    from rocflu_amd.solver import GpuRegion, Numerics
    S = rmesh.BC_SLIPWALL
    rg = Numerics().gas_prime()[1]

    def code(bc, vals, opt=0, opt2=0, order=1, creconst=0):
        m = rmesh.box(rmesh.HEX, 3, 3, 3, bc_type=(bc,) + (S,) * 5,
                      bc_opt=(opt, 0, 0, 0, 0, 0), stencils=order == 2)
        num = Numerics(order=order, reconst=0, patch_vals=[vals] + [[0.0] * 5] * 5,
                       patch_opt2=[opt2, 0, 0, 0, 0, 0], creconst=creconst)
        with pytest.raises(_native.RfluError) as e:
            GpuRegion(m, num)
        return e.value.code

    ARG = _native.RFLU_ERR_ARG
    # farfield: vortex correction, non-positive pressure or temperature
    assert code(rmesh.BC_FARFIELD, [0.5, 0, 0, 1.0, 1.0 / rg], opt2=1) == ARG
    assert code(rmesh.BC_FARFIELD, [0.5, 0, 0, 0.0, 1.0 / rg]) == ARG
    # total-condition inflow: supersonic option without a Mach number, bad switches
    assert code(rmesh.BC_INFLOW_TOTANG, [2.0, 1.0, 0, 0, 0.0], opt=0) == ARG
    assert code(rmesh.BC_INFLOW_TOTANG, [2.0, 1.0, 0, 0, 0.0], opt=1, opt2=2) == ARG
    assert code(rmesh.BC_INFLOW_TOTANG, [2.0, 1.0, 0, 0, 0.0], opt=3) == ARG
    # injection: a burning patch needs a temperature; no constrained gradients
    assert code(rmesh.BC_INJECTION, [1.0, 0.0, 0, 0, 0]) == ARG
    assert code(rmesh.BC_INJECTION, [0.0, 0.0, 0, 0, 0], order=2, creconst=1) == ARG
    # valid descriptors get past the checks and stop only at the missing device
    if _native.lib().rflu_gpu_device_count() == 0:
        assert code(rmesh.BC_FARFIELD, [0.5, 0, 0, 1.0, 1.0 / rg]) == _native.RFLU_ERR_HIP
        assert code(rmesh.BC_INFLOW_TOTANG, [2.0, 1.0, 0, 0, 0.0], opt=1, opt2=1) == \
            _native.RFLU_ERR_HIP
        assert code(rmesh.BC_INJECTION, [0.0, 0.0, 0, 0, 0], order=2) == _native.RFLU_ERR_HIP
