"""Binary case file shared by oracle/ref_harness.cpp (the reference's own CPU
manager) and tests/cpp/drive_manager.cpp (the C++ Mi355xManager):

  int32  profile, nthreads, npoints, max_points, capacity
  double start_nu, end_nu, temperature, partition, mean_mass, pressure, ref_temp
  int64  nlines
  double energies[n], nu[n], aif[n]; int32 gns[n]; double gamma[n], nexp[n] (Voigt only)
"""
import struct

import numpy as np

VOIGT = 2


def write_case(path, case: dict, lines: dict) -> None:
    prof = case["profile"]
    n = lines["nu"].size
    with open(path, "wb") as f:
        f.write(struct.pack("<5i", prof, case["nthreads"], case["npoints"], case["max_points"],
                            min(case["capacity"], n)))
        f.write(struct.pack("<7d", case["start"], case["end"], case["T"], case["Q"], case["mass"], case["P"],
                            case["ref_temp"]))
        f.write(struct.pack("<q", n))
        for k in ("elow", "nu", "aif"):
            f.write(np.ascontiguousarray(lines[k], "<f8").tobytes())
        f.write(np.ascontiguousarray(lines["gns"], "<i4").tobytes())
        if prof == VOIGT:
            f.write(np.ascontiguousarray(lines["gamma"], "<f8").tobytes())
            f.write(np.ascontiguousarray(lines["nexp"], "<f8").tobytes())


def read_grid(path) -> np.ndarray:
    return np.fromfile(path, dtype="<f8")
