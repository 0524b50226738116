"""Python binding of the request server (gx_server_* in include/gx.h)."""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import capi
from .capi import check, dptr

__all__ = ["Server"]


class Server:
    """Continuous-batching cross-section server over one ResidentLines list."""

    def __init__(self, lines, mean_mass: float, ref_temp: float = 500.0, convention: int = capi.CONV_OPENMP,
                 max_inflight: int = 4, max_batch: int = 16, hbm_budget: int = 0):
        self._L = capi.load()
        self.lines = lines
        opts = capi.ServerOpts(mean_mass, ref_temp, convention, max_inflight, max_batch, hbm_budget)
        h = ctypes.c_void_p()
        check(self._L.gx_server_create(lines._h, ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self._npoints = {}

    def submit(self, T: float, P: float, Q: float, start_nu: float, end_nu: float, npoints: int,
               profile: int = capi.VOIGT, precision: int = capi.PREC_MIXED, cutoff: float = 25.0,
               doppler_widths: float = 0.0) -> int:
        r = capi.Request(T, P, Q, profile, precision, start_nu, end_nu, npoints, cutoff, doppler_widths)
        rid = ctypes.c_uint64()
        check(self._L.gx_server_submit(self._h, ctypes.byref(r), ctypes.byref(rid)))
        self._npoints[rid.value] = npoints
        return rid.value

    def poll(self, rid: int) -> int:
        st = ctypes.c_int()
        check(self._L.gx_server_poll(self._h, rid, ctypes.byref(st)))
        return st.value

    def wait(self, rid: int, timeout: float = -1.0) -> int:
        st = ctypes.c_int()
        check(self._L.gx_server_wait(self._h, rid, timeout, ctypes.byref(st)))
        return st.value

    def result(self, rid: int, out: Optional[np.ndarray] = None):
        n = self._npoints[rid]
        if out is None:
            out = np.empty(n, np.float64)
        t = capi.Timing()
        check(self._L.gx_server_result(self._h, rid, dptr(out), n, ctypes.byref(t)))
        del self._npoints[rid]
        return out, t

    def cancel(self, rid: int) -> None:
        check(self._L.gx_server_cancel(self._h, rid))

    def stats(self) -> capi.ServerSummary:
        s = capi.ServerSummary()
        check(self._L.gx_server_stats(self._h, ctypes.byref(s)))
        return s

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.gx_server_destroy(self._h)
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
