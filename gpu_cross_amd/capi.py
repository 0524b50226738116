"""ctypes binding of libgx.so's C ABI (include/gx.h).

The library is built in-tree (``gpu_cross_amd/libgx.so``) by
``make -C gpu_cross_amd/csrc``; ``load()`` never falls back to anything else:
a missing library raises ``LibraryMissing``.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
# GX_LIB selects another build of the same library (the sanitizer build, libgx_asan.so)
LIB_PATH = os.environ.get("GX_LIB") or os.path.join(HERE, "libgx.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "gx.h")
HEADERS = [HEADER] + [os.path.join(os.path.dirname(HERE), "include", h)
                      for h in ("gx_ingest.h", "gx_output.h", "gx_hitran.h", "gx_tips.h")]

GX_OK, GX_INVALID_ARG, GX_OUT_OF_MEMORY, GX_NO_DEVICE, GX_DEVICE_ERROR, GX_BUSY, GX_NOT_FOUND = range(7)
GAUSSIAN, DOPPLER, VOIGT = 0, 1, 2
CONV_OPENMP, CONV_GPU = 0, 1
PREC_FP64, PREC_MIXED = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class LibraryMissing(RuntimeError):
    pass


class GxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gx status {status}: {msg}")
        self.status = status


class Consts(ctypes.Structure):
    _fields_ = [("half_width", ctypes.c_double), ("temperature", ctypes.c_double), ("partition", ctypes.c_double),
                ("dfreq", ctypes.c_double), ("mean_mass", ctypes.c_double), ("pressure", ctypes.c_double),
                ("ref_temp", ctypes.c_double), ("cutoff", ctypes.c_double), ("doppler_widths", ctypes.c_double),
                ("convention", ctypes.c_int), ("precision", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("intensity_ms", ctypes.c_float), ("accumulate_ms", ctypes.c_float), ("total_ms", ctypes.c_float),
                ("evals", ctypes.c_longlong), ("lines", ctypes.c_longlong), ("region1_ms", ctypes.c_float),
                ("core_ms", ctypes.c_float), ("core_evals", ctypes.c_longlong), ("far_ms", ctypes.c_float),
                ("far_evals", ctypes.c_longlong), ("interp_pairs", ctypes.c_longlong)]


class HitranOpts(ctypes.Structure):
    _fields_ = [("pressure", ctypes.c_double), ("mixture", ctypes.c_double), ("ref_partition", ctypes.c_double),
                ("lo", ctypes.c_double), ("hi", ctypes.c_double), ("nthreads", ctypes.c_int)]


class Grid(ctypes.Structure):
    _fields_ = [("start_nu", ctypes.c_double), ("end_nu", ctypes.c_double), ("npoints", ctypes.c_int)]


class Point(ctypes.Structure):
    _fields_ = [("temperature", ctypes.c_double), ("pressure", ctypes.c_double), ("partition", ctypes.c_double)]


class TableOpts(ctypes.Structure):
    _fields_ = [("profile", ctypes.c_int), ("precision", ctypes.c_int), ("convention", ctypes.c_int),
                ("mean_mass", ctypes.c_double), ("ref_temp", ctypes.c_double), ("cutoff", ctypes.c_double),
                ("doppler_widths", ctypes.c_double), ("use_graph", ctypes.c_int), ("autotune", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("temperature", ctypes.c_double), ("pressure", ctypes.c_double), ("partition", ctypes.c_double),
                ("profile", ctypes.c_int), ("precision", ctypes.c_int), ("start_nu", ctypes.c_double),
                ("end_nu", ctypes.c_double), ("npoints", ctypes.c_int), ("cutoff", ctypes.c_double),
                ("doppler_widths", ctypes.c_double)]


class ServerOpts(ctypes.Structure):
    _fields_ = [("mean_mass", ctypes.c_double), ("ref_temp", ctypes.c_double), ("convention", ctypes.c_int),
                ("max_inflight", ctypes.c_int), ("max_batch", ctypes.c_int), ("hbm_budget", ctypes.c_size_t)]


class ServerSummary(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_longlong), ("completed", ctypes.c_longlong), ("cancelled", ctypes.c_longlong),
                ("failed", ctypes.c_longlong), ("batches", ctypes.c_longlong), ("batched_requests", ctypes.c_longlong),
                ("evals", ctypes.c_longlong), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("busy_s", ctypes.c_double),
                ("peak_bytes", ctypes.c_size_t), ("shared_requests", ctypes.c_longlong)]


class OpacityInfo(ctypes.Structure):
    _fields_ = [("npoints", ctypes.c_uint64), ("npts", ctypes.c_uint64), ("start_nu", ctypes.c_double),
                ("end_nu", ctypes.c_double), ("dfreq", ctypes.c_double), ("profile", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("mean_mass", ctypes.c_double), ("ref_temp", ctypes.c_double),
                ("cutoff", ctypes.c_double), ("data_offset", ctypes.c_uint64)]


# scratch-array limits of gx_execute_noalloc (include/gx.h)
NOALLOC_MAX_LINES = 1 << 20
NOALLOC_MAX_BINS = 1 << 24

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_FAILED, REQ_CANCELLED = range(5)


def build() -> str:
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc")], check=True)
    return LIB_PATH


def declared_symbols() -> list:
    """Every function name declared in the headers under include/ (gx.h, gx_ingest.h)."""
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(gx_\w+)\s*\(", text, flags=re.M))
    return sorted(names)


_lib = None


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(f"{LIB_PATH} is not built; run `make -C gpu_cross_amd/csrc`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    _sig(L, "gx_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(P)])
    _sig(L, "gx_initialize_vectors", ctypes.c_int, [P, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)])
    _sig(L, "gx_initialize_constants", ctypes.c_int, [P, ctypes.POINTER(Consts)])
    _sig(L, "gx_transfer_freq", ctypes.c_int, [P, _dp, _dp, ctypes.c_int])
    _sig(L, "gx_transfer_vectors", ctypes.c_int, [P, ctypes.c_size_t, _dp, _dp, _dp, _ip, _dp, _dp])
    _sig(L, "gx_execute", ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int])
    _sig(L, "gx_execute_noalloc", ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_int])
    _sig(L, "gx_ready", ctypes.c_int, [P])
    _sig(L, "gx_transfer_results", ctypes.c_int, [P, _dp, _dp, ctypes.c_int])
    _sig(L, "gx_cleanup", ctypes.c_int, [P])
    _sig(L, "gx_destroy", ctypes.c_int, [P])
    _sig(L, "gx_get_ntrans", ctypes.c_size_t, [P])
    _sig(L, "gx_synchronize", ctypes.c_int, [P])
    _sig(L, "gx_last_timing", ctypes.c_int, [P, ctypes.POINTER(Timing)])
    _sig(L, "gx_lines_create", ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, _dp, _dp, _dp, _ip, _dp, _dp,
                                              ctypes.POINTER(P)])
    _sig(L, "gx_lines_destroy", ctypes.c_int, [P])
    _sig(L, "gx_lines_count", ctypes.c_size_t, [P])
    _sig(L, "gx_table_create", ctypes.c_int, [P, ctypes.POINTER(Grid), ctypes.POINTER(Point), ctypes.c_int,
                                              ctypes.POINTER(TableOpts), ctypes.POINTER(P)])
    _sig(L, "gx_table_run", ctypes.c_int, [P])
    _sig(L, "gx_table_wait", ctypes.c_int, [P, ctypes.POINTER(Timing)])
    _sig(L, "gx_table_read", ctypes.c_int, [P, ctypes.c_int, _dp, ctypes.c_int])
    _sig(L, "gx_table_shape", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)])
    _sig(L, "gx_table_destroy", ctypes.c_int, [P])
    _sig(L, "gx_server_create", ctypes.c_int, [P, ctypes.POINTER(ServerOpts), ctypes.POINTER(P)])
    _sig(L, "gx_server_submit", ctypes.c_int, [P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)])
    _sig(L, "gx_server_poll", ctypes.c_int, [P, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int)])
    _sig(L, "gx_server_wait", ctypes.c_int, [P, ctypes.c_uint64, ctypes.c_double, ctypes.POINTER(ctypes.c_int)])
    _sig(L, "gx_server_result", ctypes.c_int, [P, ctypes.c_uint64, _dp, ctypes.c_int, ctypes.POINTER(Timing)])
    _sig(L, "gx_server_cancel", ctypes.c_int, [P, ctypes.c_uint64])
    _sig(L, "gx_server_stats", ctypes.c_int, [P, ctypes.POINTER(ServerSummary)])
    _sig(L, "gx_server_destroy", ctypes.c_int, [P])
    sz, szp = ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)
    _sig(L, "gx_partition", ctypes.c_int, [_dp, _ip, sz, ctypes.c_double, _dp])
    _sig(L, "gx_states_count", ctypes.c_int, [P, sz, szp])
    _sig(L, "gx_states_parse", ctypes.c_int, [P, sz, sz, ctypes.c_int, _ip, _dp, _ip, _ip])
    _sig(L, "gx_ingest_create", ctypes.c_int, [ctypes.c_int, sz, _dp, _ip, _ip, ctypes.POINTER(P)])
    _sig(L, "gx_ingest_destroy", None, [P])
    _sig(L, "gx_ingest_set_broadening", ctypes.c_int, [P, P, sz, ctypes.c_double])
    _sig(L, "gx_ingest_set_range", ctypes.c_int, [P, ctypes.c_double, ctypes.c_double])
    _sig(L, "gx_ingest_begin_file", ctypes.c_int, [P])
    _sig(L, "gx_ingest_feed", ctypes.c_int, [P, P, sz, ctypes.c_int])
    _sig(L, "gx_ingest_count", ctypes.c_int, [P, szp, szp])
    _sig(L, "gx_ingest_download", ctypes.c_int, [P, sz, _dp, _dp, _dp, _ip, _dp, _dp])
    _sig(L, "gx_ingest_clear", ctypes.c_int, [P])
    _sig(L, "gx_ingest_last_feed", ctypes.c_int, [P, ctypes.POINTER(ctypes.c_float), szp, szp])
    _sig(L, "gx_format_trans", ctypes.c_int, [sz, _ip, _ip, _dp, P, sz, szp])
    _sig(L, "gx_hitran_read", ctypes.c_int, [P, sz, ctypes.POINTER(HitranOpts), sz, _dp, _dp, _dp, _ip, _dp, _dp,
                                             szp, szp, _ip, _ip])
    _sig(L, "gx_format_hitran", ctypes.c_int, [sz, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                               _dp, P, sz, szp])
    _sig(L, "gx_xsec_text", ctypes.c_int, [_dp, _dp, ctypes.c_int, P, sz, szp])
    _sig(L, "gx_xsec_text_write", ctypes.c_int, [ctypes.c_char_p, _dp, _dp, ctypes.c_int])
    _sig(L, "gx_opacity_write", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(OpacityInfo), ctypes.POINTER(Point),
                                               _dp])
    _sig(L, "gx_opacity_read_info", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(OpacityInfo)])
    _sig(L, "gx_opacity_read", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Point), sz, _dp, sz])
    _sig(L, "gx_tips_load", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(P)])
    _sig(L, "gx_tips_load_text", ctypes.c_int, [P, sz, ctypes.POINTER(P)])
    _sig(L, "gx_tips_free", None, [P])
    _sig(L, "gx_tips_q", ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp])
    _sig(L, "gx_tips_isotopes", ctypes.c_int, [P, ctypes.c_int])
    _sig(L, "gx_tips_lagrange", ctypes.c_int, [_dp, _dp, ctypes.c_int, ctypes.c_double, _dp])
    _sig(L, "gx_last_error", ctypes.c_char_p, [])
    _sig(L, "gx_version", ctypes.c_char_p, [])
    _lib = L
    return L


def check(status: int) -> None:
    if status != GX_OK:
        raise GxError(status, load().gx_last_error().decode(errors="replace"))


def dptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def iptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_ip)
