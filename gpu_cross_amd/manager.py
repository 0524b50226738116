"""Python mirror of the reference's compute-backend interface.

``Mi355xManager`` has the method names, argument meaning and call order of
``BaseManager`` (BaseManager.h:22-30) and forwards each call to the C ABI of
libgx.so.  ``ProfileDriver`` restates what the reference's profile drivers do
around the manager -- grid construction (BaseProfile.cpp:7-21), the range
filter and chunk flush (VoigtProfile.cpp:60-118, DopplerProfile.cpp:47-99)
and the chunk bin range (BaseProfile.cpp:49-62) -- so a whole cross-section
runs through the library the way the reference's main() runs it.
Errors raise ``GxError`` (the reference prints and exit()s instead,
cuda_utils.cu:8-17).
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional

import numpy as np

from . import capi
from .capi import DOPPLER, VOIGT, check, dptr, iptr

__all__ = ["Mi355xManager", "ProfileDriver", "make_grid", "DOPPLER", "VOIGT"]


def make_grid(start: float, end: float, npoints: int):
    """BaseProfile.cpp:10,19: dfreq = (end-start)/N and f_i = start + i*dfreq."""
    dfreq = (end - start) / float(npoints)
    return start + np.arange(npoints, dtype=np.float64) * dfreq, dfreq


class Mi355xManager:
    """Single-MI355X manager with the BaseManager lifecycle."""

    def __init__(self, profile: int, device: int = 0, memory: int = 0, convention: int = capi.CONV_OPENMP,
                 precision: int = capi.PREC_FP64, cutoff: float = 0.0, doppler_widths: float = 0.0):
        self._L = capi.load()
        self.profile = profile
        self.convention = convention
        self.precision = precision
        self.cutoff = cutoff
        self.doppler_widths = doppler_widths
        h = ctypes.c_void_p()
        check(self._L.gx_create(device, profile, memory, ctypes.byref(h)))
        self._h = h
        self.N_trans = 0

    # --- BaseManager.h:22-30 ---
    def InitializeVectors(self, Npoints: int) -> None:
        n = ctypes.c_size_t()
        check(self._L.gx_initialize_vectors(self._h, Npoints, ctypes.byref(n)))
        self.N_trans = n.value

    def InitializeConstants(self, half_width: float, temperature: float, partition: float, dfreq: float,
                            meanmass: float, pressure: float = 10.0, ref_temp: float = 296.0) -> None:
        c = capi.Consts(half_width, temperature, partition, dfreq, meanmass, pressure, ref_temp, self.cutoff,
                        self.doppler_widths, self.convention, self.precision)
        check(self._L.gx_initialize_constants(self._h, ctypes.byref(c)))

    def TransferVectors(self, Nener: int, h_energies, h_nu, h_aif, h_gns, h_gamma=None, h_n=None) -> None:
        arrs = [np.ascontiguousarray(a, np.float64) for a in (h_energies, h_nu, h_aif)]
        g = np.ascontiguousarray(h_gns, np.int32)
        gam = None if h_gamma is None else np.ascontiguousarray(h_gamma, np.float64)
        nx = None if h_n is None else np.ascontiguousarray(h_n, np.float64)
        for a in arrs + [g] + [x for x in (gam, nx) if x is not None]:
            if a.shape[0] < Nener:
                raise ValueError("line arrays shorter than Nener")
        check(self._L.gx_transfer_vectors(self._h, Nener, dptr(arrs[0]), dptr(arrs[1]), dptr(arrs[2]), iptr(g),
                                          dptr(gam), dptr(nx)))

    def TransferFreq(self, h_freq, h_intens, N: int) -> None:
        f = np.ascontiguousarray(h_freq, np.float64)
        it = None if h_intens is None else np.ascontiguousarray(h_intens, np.float64)
        check(self._L.gx_transfer_freq(self._h, dptr(f), dptr(it), N))

    def ExecuteCrossSection(self, N: int, N_ener: int, start_idx: int) -> None:
        check(self._L.gx_execute(self._h, N, N_ener, start_idx))

    def TransferResults(self, h_freq: Optional[np.ndarray], h_intens: np.ndarray, N: int) -> None:
        assert h_intens.dtype == np.float64 and h_intens.flags.c_contiguous and h_intens.size >= N
        check(self._L.gx_transfer_results(self._h, dptr(h_freq), dptr(h_intens), N))

    def Cleanup(self) -> None:
        check(self._L.gx_cleanup(self._h))

    def ReadyForWork(self) -> bool:
        r = self._L.gx_ready(self._h)
        if r < 0:
            check(-r)
        return r == 1

    def GetNtrans(self) -> int:
        return int(self._L.gx_get_ntrans(self._h))

    # --- extras ---
    def synchronize(self) -> None:
        check(self._L.gx_synchronize(self._h))

    def last_timing(self) -> capi.Timing:
        t = capi.Timing()
        check(self._L.gx_last_timing(self._h, ctypes.byref(t)))
        return t

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.gx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class ProfileDriver:
    """The reference's profile driver flow (grid, filter, chunking) over in-memory lines."""

    def __init__(self, profile: int, start_nu: float, end_nu: float, npoints: int, temperature: float,
                 partition: float, mean_mass: float, pressure: float = 1.0, ref_temp: Optional[float] = None,
                 max_points: int = 80000, capacity: Optional[int] = None, **manager_kw):
        self.profile = profile
        self.start_nu, self.end_nu, self.npoints = start_nu, end_nu, npoints
        self.freq, self.dfreq = make_grid(start_nu, end_nu, npoints)
        self.intens = np.zeros(npoints, np.float64)
        self.pressure = pressure
        self.max_points = max_points
        # VoigtProfile.cpp:14 (ExoMol T_ref = 500), DopplerProfile.cpp:25 (296)
        ref_temp = ref_temp if ref_temp is not None else (500.0 if profile == VOIGT else 296.0)
        self.manager = Mi355xManager(profile, **manager_kw)
        self.manager.InitializeVectors(npoints)
        self.manager.TransferFreq(self.freq, self.intens, npoints)
        self.capacity = min(capacity or self.manager.GetNtrans(), self.manager.GetNtrans())
        self.manager.InitializeConstants(0.0, temperature, partition, self.dfreq, mean_mass, pressure, ref_temp)
        cut = self.manager.cutoff
        self.hw = cut if cut > 0 else (100.0 if profile == DOPPLER else min(25.0 * pressure, 100.0))

    def _range(self, lo: float, hi: float):
        ib = int(max(_round_half_away((lo - self.hw - self.start_nu) / self.dfreq), 0.0))
        ie = int(min(_round_half_away((hi + self.hw - self.start_nu) / self.dfreq), float(self.npoints)))
        return ib, ie

    def run(self, elow, nu, aif, gns, gamma=None, nexp=None) -> np.ndarray:
        """Filter and chunk the lines in the given order, execute every chunk, return the grid."""
        voigt = self.profile == VOIGT
        nu = np.asarray(nu, np.float64)
        lo_lim, hi_lim = self.start_nu - self.hw, self.end_nu + self.hw
        keep = nu >= lo_lim
        over = np.nonzero(nu > hi_lim)[0]
        stop = over[0] if over.size else nu.size          # `break` on the first line above the range
        idx = np.nonzero(keep[:stop])[0]
        chunks = []
        start = 0
        while start < idx.size:
            end = min(start + self.capacity, idx.size)
            if voigt and self.max_points:
                # flush when the chunk's bin span exceeds MAX-POINTS (VoigtProfile.cpp:84-90)
                run_min = np.minimum.accumulate(nu[idx[start:end]])
                run_max = np.maximum.accumulate(nu[idx[start:end]])
                lo_b = np.maximum(_round_half_away_vec((run_min - self.hw - self.start_nu) / self.dfreq), 0.0)
                hi_b = np.minimum(_round_half_away_vec((run_max + self.hw - self.start_nu) / self.dfreq),
                                  self.npoints)
                over_span = np.nonzero(hi_b - lo_b > self.max_points)[0]
                if over_span.size:
                    end = start + int(over_span[0]) + 1
            chunks.append(idx[start:end])
            start = end
        g = lambda a, c: None if a is None else np.ascontiguousarray(np.asarray(a)[c])
        for c in chunks:
            self.manager.TransferVectors(c.size, g(elow, c), g(nu, c), g(aif, c), g(gns, c).astype(np.int32),
                                         g(gamma, c) if voigt else None, g(nexp, c) if voigt else None)
            ib, ie = self._range(float(nu[c].min()), float(nu[c].max()))
            self.manager.ExecuteCrossSection(ie - ib, c.size, ib)
        out = np.zeros(self.npoints, np.float64)
        tmp = np.zeros(self.npoints, np.float64)
        self.manager.TransferResults(None, tmp, self.npoints)
        out += tmp                                         # HybridManager.cpp:66-73
        return out


def _round_half_away(v: float) -> float:
    """C round(): halves go away from zero (exact, no v + 0.5 rounding)."""
    a = abs(v)
    f = math.floor(a)
    return math.copysign(f + (1.0 if a - f >= 0.5 else 0.0), v)


def _round_half_away_vec(v: np.ndarray) -> np.ndarray:
    a = np.abs(v)
    f = np.floor(a)
    return np.copysign(f + (a - f >= 0.5), v)
