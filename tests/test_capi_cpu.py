"""CPU-side checks of the C ABI library: it loads, exports every symbol the
header declares, and fails cleanly (status code, message) without a GPU."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from gpu_cross_amd import capi
from gpu_cross_amd.manager import ProfileDriver, _round_half_away, _round_half_away_vec, make_grid


def test_library_exports_every_declared_symbol():
    L = capi.load()
    names = capi.declared_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(names) <= exported


def test_status_codes_without_device_or_with_bad_args():
    L = capi.load()
    assert L.gx_version().startswith(b"gx")
    h = ctypes.c_void_p()
    # an unserved profile (PSUEDO_VOIGT, Input.h:13) is rejected before any device work
    assert L.gx_create(0, 3, 0, ctypes.byref(h)) == capi.GX_INVALID_ARG
    assert b"profile" in L.gx_last_error()
    assert L.gx_create(0, capi.VOIGT, 0, None) == capi.GX_INVALID_ARG
    # NULL contexts
    assert L.gx_execute(None, 1, 1, 0) == capi.GX_INVALID_ARG
    assert L.gx_transfer_results(None, None, None, 1) == capi.GX_INVALID_ARG
    assert L.gx_ready(None) < 0
    assert L.gx_get_ntrans(None) == 0
    assert L.gx_destroy(None) == capi.GX_OK


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU path")
def test_create_reports_no_device():
    L = capi.load()
    h = ctypes.c_void_p()
    assert L.gx_create(0, capi.VOIGT, 0, ctypes.byref(h)) == capi.GX_NO_DEVICE
    assert h.value is None


def test_round_half_away_matches_c_round():
    libm = ctypes.CDLL("libm.so.6")
    libm.round.restype = ctypes.c_double
    libm.round.argtypes = [ctypes.c_double]
    vals = np.array([0.5, 1.5, 2.5, -0.5, -2.5, 0.49999999999999994, 1e7 + 0.5, 3.4999999, -7.5000001])
    for v in vals:
        assert _round_half_away(float(v)) == libm.round(float(v))
    np.testing.assert_array_equal(_round_half_away_vec(vals), [libm.round(float(v)) for v in vals])


def test_grid_matches_reference_formula():
    f, d = make_grid(0.0, 12000.0, 1_200_000)
    assert d == 12000.0 / 1_200_000
    assert f[777] == 0.0 + 777.0 * d
    assert f[-1] == 12000.0 - d or abs(f[-1] - (12000.0 - d)) < 1e-9


@pytest.mark.skipif(not os.path.exists("/root/reference/BaseManager.h"), reason="needs the reference checkout")
def test_cpp_manager_compiles_against_the_reference_basemanager(tmp_path, root):
    """mi355x_manager.h built against the reference's own BaseManager.h (the route a
    maintainer takes), every virtual overridden with its exact signature, linked to libgx.so."""
    src = tmp_path / "use_manager.cpp"
    src.write_text(
        '#include "mi355x_manager.h"\n'
        "BaseManager* make(ProfileType p) { return new Mi355xManager(p); }\n"
        "void drive(BaseManager* m, double* f, double* I, double* e, double* nu, double* a, int* g,\n"
        "           double* gam, double* n) {\n"
        "  m->InitializeVectors(16); m->TransferFreq(f, I, 16);\n"
        "  m->InitializeConstants(0.0, 1000.0, 1.0, 0.01, 17.03, 1.0, 500.0);\n"
        "  m->TransferVectors(4, e, nu, a, g, gam, n); m->ExecuteCrossSection(16, 4, 0);\n"
        "  m->TransferResults(f, I, 16); (void)m->ReadyForWork(); (void)m->GetNtrans(); m->Cleanup();\n"
        "  delete static_cast<Mi355xManager*>(m);   // the reference's ~BaseManager is protected\n"
        "}\n"
        "int main() { return make(VOIGT) == nullptr; }\n")
    exe = tmp_path / "use_manager"
    cmd = ["g++", "-std=c++11", "-Wall", "-Werror=overloaded-virtual", '-DGX_BASE_MANAGER_HEADER="BaseManager.h"',
           "-I/root/reference", "-I" + os.path.join(root, "gpu_cross_amd", "csrc"), "-I" + os.path.join(root, "include"),
           str(src), "/root/reference/BaseManager.cpp", "-L" + os.path.join(root, "gpu_cross_amd"), "-lgx",
           "-Wl,-rpath-link,/opt/rocm/lib",
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
