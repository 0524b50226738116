"""TIPS-2011 partition sums for HITRAN line lists (gx_tips_* in include/gx_tips.h).

Replaces the reference's Fortran ``bd_tips_2011_`` (HITRAN_files/BD_TIPS_2011_v1p0.for),
reading its tables from the caller's copy of that source file::

    tips = Tips("/path/to/gpu_cross/HITRAN_files/BD_TIPS_2011_v1p0.for")
    gi, q296 = tips.q(2, 1, 296.0)        # CO2, isotopologue 626
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import capi
from .capi import check

__all__ = ["Tips", "lagrange"]


class Tips:
    """Tables of one TIPS-2011 source file; ``q(mol, iso, T)`` is bd_tips_2011_(MOL, T, ISO)."""

    def __init__(self, path: str = None, text: bytes = None):
        self._L = capi.load()
        h = ctypes.c_void_p()
        if text is not None:
            buf = ctypes.create_string_buffer(bytes(text), len(text))
            check(self._L.gx_tips_load_text(ctypes.cast(buf, ctypes.c_void_p), len(text), ctypes.byref(h)))
        else:
            check(self._L.gx_tips_load(str(path).encode(), ctypes.byref(h)))
        self._h = h

    def q(self, mol: int, iso: int, temperature: float):
        """(gi, Q(T)); raises GxError outside 70-3000 K or for an unknown molecule/isotopologue."""
        gi, q = ctypes.c_double(), ctypes.c_double()
        check(self._L.gx_tips_q(self._h, mol, iso, temperature, ctypes.byref(gi), ctypes.byref(q)))
        return gi.value, q.value

    def isotopes(self, mol: int) -> int:
        return int(self._L.gx_tips_isotopes(self._h, mol))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._L.gx_tips_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lagrange(a, b, x: float) -> float:
    """The routine's AtoB interpolation on any increasing table (a, b) at x."""
    L = capi.load()
    a = np.ascontiguousarray(a, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    out = ctypes.c_double()
    check(L.gx_tips_lagrange(capi.dptr(a), capi.dptr(b), a.size, x, ctypes.byref(out)))
    return out.value
