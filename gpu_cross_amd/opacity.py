"""Output writers (include/gx_output.h): the reference's per-bin text output and the
opacity-table file of the batched (T,P) tables.

    text = format_xsec_text(freq, intens)        # "%12.6f %13.8E\\n" per bin (BaseProfile.cpp:65-74)
    write_xsec_text("out.txt", freq, intens)
    tab.save("table.gxopac")                     # gpu_cross_amd.table.Table
    t = read_table("table.gxopac")               # header, points, grids (memory-mapped)
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Sequence

import numpy as np

from . import capi

HEADER_BYTES = 128
MAGIC = b"GXOPAC01"
# the header fields at their byte offsets (the layout include/gx_output.h documents)
_HEADER = np.dtype({"names": ["magic", "header_bytes", "version", "npoints", "npts", "start_nu", "end_nu", "dfreq",
                              "profile", "precision", "mean_mass", "ref_temp", "cutoff", "data_offset"],
                    "formats": ["S8", "<u4", "<u4", "<u8", "<u8", "<f8", "<f8", "<f8", "<i4", "<i4", "<f8", "<f8",
                                "<f8", "<u8"],
                    "offsets": [0, 8, 12, 16, 24, 32, 40, 48, 56, 60, 64, 72, 80, 88],
                    "itemsize": HEADER_BYTES})


def format_xsec_text(freq: np.ndarray, intens: np.ndarray) -> bytes:
    """The reference's output lines, "%12.6f %13.8E\\n" per bin (BaseProfile.cpp:69-71)."""
    f = np.ascontiguousarray(freq, np.float64)
    x = np.ascontiguousarray(intens, np.float64)
    L = capi.load()
    need = ctypes.c_size_t(0)
    capi.check(L.gx_xsec_text(capi.dptr(f), capi.dptr(x), f.size, None, 0, ctypes.byref(need)))
    buf = np.empty(need.value + 1, np.uint8)
    capi.check(L.gx_xsec_text(capi.dptr(f), capi.dptr(x), f.size, buf.ctypes.data, buf.size, ctypes.byref(need)))
    return buf[:need.value].tobytes()


def write_xsec_text(path, freq: np.ndarray, intens: np.ndarray) -> None:
    f = np.ascontiguousarray(freq, np.float64)
    x = np.ascontiguousarray(intens, np.float64)
    capi.check(capi.load().gx_xsec_text_write(os.fsencode(path), capi.dptr(f), capi.dptr(x), f.size))


@dataclasses.dataclass
class OpacityTable:
    """A read opacity-table file."""

    npoints: int
    start_nu: float
    end_nu: float
    dfreq: float
    profile: int
    precision: int
    mean_mass: float
    ref_temp: float
    cutoff: float
    points: np.ndarray          # (npts, 3): T, P, Q
    grids: np.ndarray           # (npts, npoints), memory-mapped

    @property
    def freq(self) -> np.ndarray:
        return self.start_nu + np.arange(self.npoints, dtype=np.float64) * self.dfreq


def write_table(path, points: Sequence[tuple], grids: np.ndarray, start_nu: float, end_nu: float,
                profile: int = capi.VOIGT, precision: int = capi.PREC_MIXED, mean_mass: float = 0.0,
                ref_temp: float = 500.0, cutoff: float = 25.0) -> None:
    g = np.ascontiguousarray(grids, np.float64)
    npts, m = g.shape
    if len(points) != npts:
        raise ValueError("one (T, P, Q) per grid")
    info = capi.OpacityInfo(m, npts, float(start_nu), float(end_nu), (float(end_nu) - float(start_nu)) / m,
                            int(profile), int(precision), float(mean_mass), float(ref_temp), float(cutoff), 0)
    pts = (capi.Point * npts)(*[capi.Point(float(T), float(P), float(Q)) for T, P, Q in points])
    capi.check(capi.load().gx_opacity_write(os.fsencode(path), ctypes.byref(info), pts, capi.dptr(g)))


def read_table(path, mmap: bool = True) -> OpacityTable:
    """Read a table file with numpy alone (header fields at their documented offsets)."""
    h = np.fromfile(path, dtype=_HEADER, count=1)[0]
    if bytes(h["magic"]) != MAGIC or int(h["header_bytes"]) != HEADER_BYTES or int(h["version"]) != 1:
        raise ValueError(f"{path}: not a GXOPAC01 file")
    npts, m = int(h["npts"]), int(h["npoints"])
    pts = np.fromfile(path, dtype="<f8", count=3 * npts, offset=HEADER_BYTES).reshape(npts, 3)
    off = int(h["data_offset"])
    if mmap:
        grids = np.memmap(path, dtype="<f8", mode="r", offset=off, shape=(npts, m))
    else:
        grids = np.fromfile(path, dtype="<f8", count=npts * m, offset=off).reshape(npts, m)
    return OpacityTable(m, float(h["start_nu"]), float(h["end_nu"]), float(h["dfreq"]), int(h["profile"]),
                        int(h["precision"]), float(h["mean_mass"]), float(h["ref_temp"]), float(h["cutoff"]), pts,
                        grids)


def read_table_c(path):
    """The same file through the library's reader: (info, points (npts, 3), grids)."""
    L = capi.load()
    info = capi.OpacityInfo()
    capi.check(L.gx_opacity_read_info(os.fsencode(path), ctypes.byref(info)))
    pts = (capi.Point * info.npts)()
    grids = np.empty((info.npts, info.npoints), np.float64)
    capi.check(L.gx_opacity_read(os.fsencode(path), pts, info.npts, capi.dptr(grids), grids.size))
    p = np.array([[q.temperature, q.pressure, q.partition] for q in pts]).reshape(info.npts, 3)
    return info, p, grids
