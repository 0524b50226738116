"""ctypes binding of the C oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker the HIP library is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.

``run_reference_driver`` restates the chunk loop of the reference's profile
drivers (VoigtProfile.cpp:60-118, DopplerProfile.cpp:47-99, with
BaseProfile.cpp:49-62) on top of ``orc_cross_section``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

GAUSSIAN, DOPPLER, VOIGT = 0, 1, 2
CONV_OPENMP, CONV_GPU = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class OrcConsts(ctypes.Structure):
    _fields_ = [("profile", ctypes.c_int), ("convention", ctypes.c_int), ("temperature", ctypes.c_double),
                ("partition", ctypes.c_double), ("dfreq", ctypes.c_double), ("mean_mass", ctypes.c_double),
                ("pressure", ctypes.c_double), ("ref_temp", ctypes.c_double), ("cutoff", ctypes.c_double),
                ("doppler_widths", ctypes.c_double), ("half_width", ctypes.c_double)]


def build() -> str:
    """Compile liboracle.so (gcc; no reference code involved)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_humlic.restype = ctypes.c_double
        L.orc_humlic.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_humlic_vec.argtypes = [ctypes.c_size_t, _dp, _dp, _dp]
        L.orc_erf_vec.argtypes = [ctypes.c_size_t, _dp, _dp]
        L.orc_erfcx_vec.argtypes = [ctypes.c_size_t, _dp, _dp]
        L.orc_voigt916_vec.argtypes = [ctypes.c_size_t, _dp, _dp, ctypes.c_double, _dp]
        L.orc_gh3.restype = ctypes.c_double
        L.orc_gh3.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_line_intensity_plain.restype = ctypes.c_double
        L.orc_line_intensity_plain.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double, ctypes.c_int]
        L.orc_partition.restype = ctypes.c_double
        L.orc_partition.argtypes = [ctypes.c_size_t, _dp, _ip, ctypes.c_double]
        for name in ("orc_beta",):
            getattr(L, name).restype = ctypes.c_double
            getattr(L, name).argtypes = [ctypes.c_double]
        L.orc_cmcoef.restype = ctypes.c_double
        L.orc_cmcoef.argtypes = [ctypes.c_int]
        L.orc_dpwcoef.restype = ctypes.c_double
        L.orc_dpwcoef.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int]
        L.orc_reference_cutoff.restype = ctypes.c_double
        L.orc_reference_cutoff.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_line_intensity.restype = ctypes.c_double
        L.orc_line_intensity.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_int]
        L.orc_gamma_t.restype = ctypes.c_double
        L.orc_gamma_t.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_double, ctypes.c_double]
        L.orc_line_cutoff.restype = ctypes.c_double
        L.orc_line_cutoff.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_double]
        L.orc_window.restype = ctypes.c_long
        L.orc_window.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_cross_section.restype = ctypes.c_int
        L.orc_cross_section.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_size_t, _dp, _dp, _dp, _ip, _dp, _dp,
                                        _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_chunk_range.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, _ip, _ip]
        L.orc_count_evals.restype = ctypes.c_longlong
        L.orc_count_evals.argtypes = [ctypes.POINTER(OrcConsts), ctypes.c_size_t, _dp, _dp, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int]
        _lib = L
    return _lib


def _d(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def consts(profile: int, temperature: float, partition: float, dfreq: float, mean_mass: float,
           pressure: float = 1.0, ref_temp: float = 500.0, cutoff: Optional[float] = None,
           convention: int = CONV_OPENMP, doppler_widths: float = 0.0, half_width: float = 0.0) -> OrcConsts:
    """Constants of one point.  Default cutoff: the reference rule of the profile (Gaussian:
    10 half-widths, cross_kernal.cu:578).  A Voigt line list without gamma takes the CUDA
    manager's fixed window NOGAMMA_CUTOFF (pass it as cutoff)."""
    if cutoff is None:
        cutoff = 10.0 * half_width if profile == GAUSSIAN else lib().orc_reference_cutoff(profile, pressure)
    return OrcConsts(profile, convention, temperature, partition, dfreq, mean_mass, pressure, ref_temp, cutoff,
                     doppler_widths, half_width)


# window of the Voigt sum without per-line gamma: 500 gamma_L, gamma_L = 0.05 (cross_kernal.cu:870-889)
NOGAMMA_CUTOFF = 500.0 * 0.05


def voigt916(x, y, a: float = 1.0) -> np.ndarray:
    x = np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), np.broadcast(x, y).shape)).ravel()
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float64), x.shape)).ravel()
    out = np.empty_like(x)
    lib().orc_voigt916_vec(x.size, _d(x), _d(y), a, _d(out))
    return out


def gh3(x: float, y: float) -> float:
    """The 3-point Gauss-Hermite expression of the Voigt sum without per-line gamma
    (cross_kernal.cu:183-186), at one (x, y)."""
    return lib().orc_gh3(float(x), float(y), float(x) * float(x) + float(y) * float(y))


def erfcx(x) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.float64)).ravel()
    out = np.empty_like(x)
    lib().orc_erfcx_vec(x.size, _d(x), _d(out))
    return out


def humlic(x, y) -> np.ndarray:
    x = np.ascontiguousarray(np.broadcast_to(np.asarray(x, np.float64), np.broadcast(x, y).shape)).ravel()
    y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float64), x.shape)).ravel()
    out = np.empty_like(x)
    lib().orc_humlic_vec(x.size, _d(x), _d(y), _d(out))
    return out


def erf(x) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.float64)).ravel()
    out = np.empty_like(x)
    lib().orc_erf_vec(x.size, _d(x), _d(out))
    return out


def partition(state_e: np.ndarray, state_g: np.ndarray, temperature: float) -> float:
    return lib().orc_partition(state_e.size, _d(np.ascontiguousarray(state_e, np.float64)),
                               _i(np.ascontiguousarray(state_g, np.int32)), temperature)


def make_grid(start: float, end: float, npoints: int):
    """BaseProfile.cpp:10,19 -- dfreq = (end-start)/N, f_i = start + i*dfreq."""
    dfreq = (end - start) / float(npoints)
    return start + np.arange(npoints, dtype=np.float64) * dfreq, dfreq


def cross_section(c: OrcConsts, elow, nu, aif, gns, gamma, nexp, freq, points: int, start_idx: int,
                  nthreads: int = 1, out: Optional[np.ndarray] = None) -> np.ndarray:
    """One ExecuteCrossSection chunk of the CPU manager, accumulated into ``out``."""
    if out is None:
        out = np.zeros(freq.size, np.float64)
    rc = lib().orc_cross_section(ctypes.byref(c), nu.size, _d(elow), _d(nu), _d(aif), _i(gns), _d(gamma), _d(nexp),
                                 _d(freq), freq.size, points, start_idx, nthreads, _d(out))
    if rc != 0:
        raise RuntimeError(f"orc_cross_section failed ({rc})")
    return out


def windows(c: OrcConsts, nu: np.ndarray, freq: np.ndarray, points: int, start_idx: int):
    L = lib()
    ib = np.empty(nu.size, np.int32)
    ie = np.empty(nu.size, np.int32)
    a, b = ctypes.c_int(), ctypes.c_int()
    for k, v in enumerate(nu):
        L.orc_window(ctypes.byref(c), float(v), L.orc_line_cutoff(ctypes.byref(c), float(v)), _d(freq), freq.size,
                     points, start_idx, ctypes.byref(a), ctypes.byref(b))
        ib[k], ie[k] = a.value, b.value
    return ib, ie


def count_evals(c: OrcConsts, nu: np.ndarray, freq: np.ndarray, points: int, start_idx: int) -> int:
    return int(lib().orc_count_evals(ctypes.byref(c), nu.size, _d(nu), _d(freq), freq.size, points, start_idx))


def run_reference_driver(c: OrcConsts, elow, nu, aif, gns, gamma, nexp, start_nu: float, end_nu: float,
                         npoints: int, capacity: int, max_points: int = 80000, nthreads: int = 1) -> np.ndarray:
    """The reference's whole-grid flow over lines in file order (chunking included)."""
    L = lib()
    freq, dfreq = make_grid(start_nu, end_nu, npoints)
    assert abs(dfreq - c.dfreq) == 0.0
    hw = L.orc_reference_cutoff(c.profile, c.pressure) if c.cutoff <= 0 else c.cutoff
    out = np.zeros(npoints, np.float64)
    voigt = c.profile == VOIGT
    chunk = []
    mn, mx = 99999.0, 0.0
    ib, ie = ctypes.c_int(), ctypes.c_int()

    def perform(idx):
        idx = np.asarray(idx, dtype=np.int64)
        L.orc_chunk_range(float(nu[idx].min()), float(nu[idx].max()), hw, start_nu, dfreq, npoints,
                          ctypes.byref(ib), ctypes.byref(ie))
        g = lambda a: None if a is None else np.ascontiguousarray(a[idx])
        cross_section(c, g(elow), g(nu), g(aif), g(gns), g(gamma) if voigt else None, g(nexp) if voigt else None,
                      freq, ie.value - ib.value, ib.value, nthreads, out)

    for i in range(nu.size):
        v = float(nu[i])
        if v < start_nu - hw:
            continue
        if v > end_nu + hw:
            break
        chunk.append(i)
        mn, mx = min(mn, v), max(mx, v)
        L.orc_chunk_range(mn, mx, hw, start_nu, dfreq, npoints, ctypes.byref(ib), ctypes.byref(ie))
        span = voigt and (ie.value - ib.value) > max_points
        if len(chunk) >= capacity or span:
            perform(chunk)
            chunk, mn, mx = [], 99999.0, 0.0
    if chunk:
        perform(chunk)
    return out


# --- ExoMol reader restatement (ingest_oracle.cpp) --------------------------------------

_ingest_lib = None


def ingest_lib():
    global _ingest_lib
    if _ingest_lib is None:
        path = os.path.join(HERE, "liboracle_ingest.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, "liboracle_ingest.so"], check=True)
        L = ctypes.CDLL(path)
        L.orc_states.restype = ctypes.c_long
        L.orc_ingest.restype = ctypes.c_long
        L.orc_hitran.restype = ctypes.c_long
        _ingest_lib = L
    return _ingest_lib


def read_states(text: bytes) -> dict:
    """.states text read as initialize_states reads it (exomol_functions.cpp:8-48)."""
    L = ingest_lib()
    n = L.orc_states(text, ctypes.c_long(len(text)), ctypes.c_long(0), None, None, None, None)
    out = dict(id=np.zeros(n, np.int32), energy=np.zeros(n), gtot=np.zeros(n, np.int32), J=np.zeros(n, np.int32))
    L.orc_states(text, ctypes.c_long(len(text)), ctypes.c_long(n), _i(out["id"]), _d(out["energy"]),
                 _i(out["gtot"]), _i(out["J"]))
    return out


def read_trans(states: dict, broad: Optional[bytes], pressure: float, lo: float, hi: float, files,
               use_broad: bool = True):
    """The reference's read loop over .trans texts (ReadNextState + GetGammaN + the filter of
    VoigtProfile.cpp:56-68).  Returns ([E'', nu, A, g, gamma, n] arrays, records read)."""
    L = ingest_lib()
    e = np.ascontiguousarray(states["energy"], np.float64)
    g = np.ascontiguousarray(states["gtot"], np.int32)
    J = np.ascontiguousarray(states["J"], np.int32)
    files = [bytes(f) for f in files]
    cap = sum(len(f) for f in files) // 5 + 16
    arrs = [np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(cap, np.int32), np.zeros(cap), np.zeros(cap)]
    fptrs = (ctypes.c_char_p * len(files))(*files)
    lens = (ctypes.c_long * len(files))(*[len(f) for f in files])
    rec = ctypes.c_long(0)
    b = broad or b""
    n = L.orc_ingest(_d(e), _i(g), _i(J), ctypes.c_long(e.size), b, ctypes.c_long(len(b)),
                     ctypes.c_int(1 if use_broad else 0), ctypes.c_double(pressure), ctypes.c_double(lo),
                     ctypes.c_double(hi), fptrs, lens, ctypes.c_int(len(files)), ctypes.c_long(cap),
                     _d(arrs[0]), _d(arrs[1]), _d(arrs[2]), _i(arrs[3]), _d(arrs[4]), _d(arrs[5]), ctypes.byref(rec))
    if n < 0:
        raise ValueError(f"oracle reader: state id out of range ({n})")
    return [a[:n] for a in arrs], rec.value


def read_hitran(text: bytes, pressure: float, mixture: float, ref_partition: float, lo: float, hi: float):
    """HITRANStateReader's read of one .par text with the Voigt driver's filter
    (HITRANStateReader.cpp:3-250, VoigtProfile.cpp:61-75).  Returns ([E'', nu, A, g, gamma, n],
    records read, molecule id, isotope)."""
    L = ingest_lib()
    cap = bytes(text).count(b"\n") + len(text) // 1023 + 2      # fgets pieces at most
    arrs = [np.zeros(cap), np.zeros(cap), np.zeros(cap), np.zeros(cap, np.int32), np.zeros(cap), np.zeros(cap)]
    rec, mol, iso = ctypes.c_long(0), ctypes.c_int(0), ctypes.c_int(0)
    n = L.orc_hitran(bytes(text), ctypes.c_long(len(text)), ctypes.c_double(pressure), ctypes.c_double(mixture),
                     ctypes.c_double(ref_partition), ctypes.c_double(lo), ctypes.c_double(hi), ctypes.c_long(cap),
                     _d(arrs[0]), _d(arrs[1]), _d(arrs[2]), _i(arrs[3]), _d(arrs[4]), _d(arrs[5]), ctypes.byref(rec),
                     ctypes.byref(mol), ctypes.byref(iso))
    if n < 0:
        raise OSError("oracle HITRAN reader: fmemopen failed")
    if n > cap:
        raise ValueError("oracle HITRAN reader: more lines than the capacity")
    return [a[:n] for a in arrs], rec.value, mol.value, iso.value
