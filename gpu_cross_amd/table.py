"""Python binding of the batched (T,P) table API (gx_lines_* / gx_table_* in include/gx.h)."""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import capi
from .capi import check, dptr, iptr

__all__ = ["ResidentLines", "Table"]


class ResidentLines:
    """A line list uploaded once to one GPU and kept sorted by nu there."""

    def __init__(self, elow, nu, aif, gns, gamma_ref=None, nexp=None, device: int = 0):
        self._L = capi.load()
        a = [np.ascontiguousarray(x, np.float64) for x in (elow, nu, aif)]
        g = np.ascontiguousarray(gns, np.int32)
        gr = None if gamma_ref is None else np.ascontiguousarray(gamma_ref, np.float64)
        nx = None if nexp is None else np.ascontiguousarray(nexp, np.float64)
        h = ctypes.c_void_p()
        check(self._L.gx_lines_create(device, a[1].size, dptr(a[0]), dptr(a[1]), dptr(a[2]), iptr(g), dptr(gr),
                                      dptr(nx), ctypes.byref(h)))
        self._h = h

    @classmethod
    def from_linelist(cls, ll, device: int = 0) -> "ResidentLines":
        return cls(ll.elow, ll.nu, ll.aif, ll.gns, ll.gamma_ref, ll.nexp, device)

    def __len__(self) -> int:
        return int(self._L.gx_lines_count(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.gx_lines_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Table:
    """Cross-sections of many (T, P, Q) points on one grid, computed together."""

    def __init__(self, lines: ResidentLines, start_nu: float, end_nu: float, npoints: int,
                 points: Sequence[tuple], mean_mass: float, ref_temp: float = 500.0, cutoff: float = 25.0,
                 profile: int = capi.VOIGT, precision: int = capi.PREC_MIXED, convention: int = capi.CONV_OPENMP,
                 doppler_widths: float = 0.0, use_graph: bool = True, autotune: bool = True):
        self._L = capi.load()
        self.lines = lines
        self.npoints = npoints
        self.points = list(points)
        self.start_nu, self.end_nu = float(start_nu), float(end_nu)
        self.profile, self.precision = profile, precision
        self.mean_mass, self.ref_temp, self.cutoff = mean_mass, ref_temp, cutoff
        grid = capi.Grid(start_nu, end_nu, npoints)
        pts = (capi.Point * len(self.points))(*[capi.Point(float(T), float(P), float(Q)) for T, P, Q in self.points])
        opts = capi.TableOpts(profile, precision, convention, mean_mass, ref_temp, cutoff, doppler_widths,
                              int(use_graph), int(autotune))
        h = ctypes.c_void_p()
        check(self._L.gx_table_create(lines._h, ctypes.byref(grid), pts, len(self.points), ctypes.byref(opts),
                                      ctypes.byref(h)))
        self._h = h

    def run(self) -> None:
        check(self._L.gx_table_run(self._h))

    def wait(self) -> capi.Timing:
        t = capi.Timing()
        check(self._L.gx_table_wait(self._h, ctypes.byref(t)))
        return t

    def read(self, point: int, out: Optional[np.ndarray] = None) -> np.ndarray:
        if out is None:
            out = np.empty(self.npoints, np.float64)
        check(self._L.gx_table_read(self._h, point, dptr(out), self.npoints))
        return out

    def save(self, path) -> None:
        """Write every grid of the last run to an opacity-table file (gpu_cross_amd.opacity)."""
        from .opacity import write_table

        grids = np.empty((len(self.points), self.npoints), np.float64)
        for k in range(len(self.points)):
            self.read(k, grids[k])
        write_table(path, self.points, grids, self.start_nu, self.end_nu, self.profile, self.precision,
                    self.mean_mass, self.ref_temp, self.cutoff)

    def shape(self):
        r, s = ctypes.c_int(), ctypes.c_int()
        check(self._L.gx_table_shape(self._h, ctypes.byref(r), ctypes.byref(s)))
        return r.value, s.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.gx_table_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
