"""The C++ Mi355xManager (gpu_cross_amd/csrc/mi355x_manager.h) driven by the
reference's driver flow (tests/cpp/drive_manager.cpp) on the exact inputs the
reference's own CPU manager received for each golden fixture."""
import glob
import json
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests.golden.caseio import read_grid, write_case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "drive_manager")
GOLDEN = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def rel_err(out, ref, floor=1e-6):
    mask = ref >= floor * ref.max()
    return float(np.max(np.abs(out[mask] - ref[mask]) / ref[mask]))


def _cases():
    """(fixture, precision, tolerance): FP64 on every fixture, mixed precision on the Voigt ones
    (the mixed path serves Voigt only)."""
    out = []
    for p in GOLDEN:
        name = os.path.basename(p)[:-4]
        out.append(pytest.param(p, 0, 1e-10, id=f"{name}-fp64"))
        if name.startswith("voigt"):
            out.append(pytest.param(p, 1, 1e-5, id=f"{name}-mixed"))
    return out


@pytest.mark.parametrize("path,precision,tol", _cases())
def test_cpp_manager_on_reference_inputs(path, precision, tol):
    assert os.path.exists(DRIVER), "tests/cpp/drive_manager is not built (make -C tests/cpp)"
    z = np.load(path)
    meta = json.loads(bytes(z["meta"]).decode())
    lines = {k: z[k] for k in ("elow", "nu", "aif", "gns")}
    if meta["profile"] == 2:
        lines["gamma"], lines["nexp"] = z["gamma"], z["nexp"]
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_case(fin, meta, lines)
        print(f"cpp manager {meta['name']} precision={precision}", flush=True)
        subprocess.run([DRIVER, fin, fout, str(precision)], check=True, timeout=120)
        out = read_grid(fout)
    assert rel_err(out, z["intens"]) <= tol
