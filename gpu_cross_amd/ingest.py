"""ExoMol reader on the GPU: the reference's ExomolStateReader (ExomolStateReader.cpp:88-178)
and the read loop of VoigtProfile::ExecuteCrossSection (VoigtProfile.cpp:56-107), over the
C ABI of include/gx_ingest.h.

    states = read_states("mol.states")
    with ExomolReader(states, partition=-1.0, pressure=P, broadener="mol.broad") as rd:
        Q = rd.ComputePartition(T)
        rd.set_range(start - C, end + C)
        lines = rd.read_trans(["a.trans", "b.trans.bz2"])

The lines are what the reference hands BaseManager::TransferVectors, in read order: E'', nu,
A, g and (Voigt) gamma, n.  .bz2 files are decompressed here (Python's bz2, all streams) and
read like text; the reference's own bz2 path reads fixed 37-byte records
(read_compress_trans.c:35-89), which is the same thing for ExoMol's "%12d %12d %10.4E" lines.
"""
from __future__ import annotations

import bz2
import ctypes
import dataclasses
import os
from typing import Iterable, Optional, Union

import numpy as np

from . import capi

_BLOCK = 256 << 20


@dataclasses.dataclass
class States:
    """One molecule's .states records (positional: state id k is entry k - 1)."""

    id: np.ndarray        # int32, the file's own id column minus one
    energy: np.ndarray    # float64
    gtot: np.ndarray      # int32
    J: np.ndarray         # int32

    @property
    def n(self) -> int:
        return int(self.energy.shape[0])


def _text(b) -> np.ndarray:
    """Zero-copy uint8 view of bytes / bytearray / memoryview / uint8 array."""
    if isinstance(b, np.ndarray):
        return np.ascontiguousarray(b, np.uint8).reshape(-1)
    return np.frombuffer(b, np.uint8)


def _read_bytes(src: Union[str, bytes, os.PathLike]) -> bytes:
    if isinstance(src, (bytes, bytearray, memoryview)):
        return bytes(src)
    p = os.fspath(src)
    opener = bz2.open if p.endswith(".bz2") else open
    with opener(p, "rb") as f:
        return f.read()


def parse_states(text, nthreads: int = 0) -> States:
    """.states text -> States (exomol_functions.cpp:8-48, read with the same C calls)."""
    L = capi.load()
    t = _text(text)
    n = ctypes.c_size_t(0)
    capi.check(L.gx_states_count(t.ctypes.data, t.size, ctypes.byref(n)))
    st = States(np.empty(n.value, np.int32), np.empty(n.value, np.float64), np.empty(n.value, np.int32),
                np.empty(n.value, np.int32))
    capi.check(L.gx_states_parse(t.ctypes.data, t.size, n.value, nthreads, capi.iptr(st.id), capi.dptr(st.energy),
                                 capi.iptr(st.gtot), capi.iptr(st.J)))
    return st


def read_states(src: Union[str, bytes, os.PathLike], nthreads: int = 0) -> States:
    return parse_states(_read_bytes(src), nthreads)


def partition(energy: np.ndarray, gtot: np.ndarray, temperature: float) -> float:
    """Q(T) summed in state order (compute_partition, exomol_functions.cpp:50-73)."""
    e = np.ascontiguousarray(energy, np.float64)
    g = np.ascontiguousarray(gtot, np.int32)
    q = ctypes.c_double(0.0)
    capi.check(capi.load().gx_partition(capi.dptr(e), capi.iptr(g), e.size, float(temperature), ctypes.byref(q)))
    return q.value


@dataclasses.dataclass
class Lines:
    """Kept lines in read order, in the types TransferVectors takes (BaseManager.h:24)."""

    elow: np.ndarray
    nu: np.ndarray
    aif: np.ndarray
    gns: np.ndarray
    gamma: np.ndarray
    nexp: np.ndarray
    records: int          # records read (ReadNextState calls that returned true)

    @property
    def n(self) -> int:
        return int(self.nu.shape[0])


class ExomolReader:
    """GPU ExoMol reader over one molecule's states (ExomolStateReader.cpp:88-99)."""

    def __init__(self, states: Union[States, str, bytes, os.PathLike], partition: float = -1.0,
                 pressure: float = 1.0, broadener: Union[None, str, bytes, os.PathLike] = None, device: int = 0):
        self.states = states if isinstance(states, States) else read_states(states)
        self.partition = float(partition)
        self.pressure = float(pressure)
        L = capi.load()
        h = ctypes.c_void_p()
        st = self.states
        capi.check(L.gx_ingest_create(device, st.n, capi.dptr(st.energy), capi.iptr(st.gtot), capi.iptr(st.J),
                                      ctypes.byref(h)))
        self._h = h
        self.use_broadeners = False
        if broadener is not None:
            text = None
            if isinstance(broadener, (bytes, bytearray)):
                text = bytes(broadener)
            elif os.path.exists(os.fspath(broadener)):
                text = _read_bytes(broadener)
            # a missing file leaves gamma and n unassigned (ExomolStateReader.cpp:41-45)
            if text is not None:
                self.set_broadening(text)

    def set_broadening(self, text) -> None:
        """.broad text (InitializeBroadener, ExomolStateReader.cpp:39-87)."""
        t = _text(text)
        capi.check(capi.load().gx_ingest_set_broadening(self._h, t.ctypes.data if t.size else None, t.size,
                                                         self.pressure))
        self.use_broadeners = True

    def ComputePartition(self, temperature: float) -> float:
        """ExomolStateReader.cpp:172-178: computed once when the given partition is negative."""
        if self.partition < 0.0:
            self.partition = partition(self.states.energy, self.states.gtot, temperature)
        return self.partition

    def set_range(self, lo: float, hi: float) -> None:
        """Keep nu >= lo, stop each file at its first nu > hi (VoigtProfile.cpp:65-68)."""
        capi.check(capi.load().gx_ingest_set_range(self._h, float(lo), float(hi)))

    def begin_file(self) -> None:
        capi.check(capi.load().gx_ingest_begin_file(self._h))

    def feed(self, text, eof: bool) -> None:
        """The next piece of the current file (bytes or a uint8 array)."""
        t = _text(text)
        capi.check(capi.load().gx_ingest_feed(self._h, t.ctypes.data if t.size else None, t.size, 1 if eof else 0))

    def last_feed(self):
        """(kernel ms, text bytes, records read) of the last feed."""
        ms, nb, nr = ctypes.c_float(0), ctypes.c_size_t(0), ctypes.c_size_t(0)
        capi.check(capi.load().gx_ingest_last_feed(self._h, ctypes.byref(ms), ctypes.byref(nb), ctypes.byref(nr)))
        return ms.value, nb.value, nr.value

    def read_file(self, src: Union[str, os.PathLike, bytes], block: int = _BLOCK):
        """One .trans file (path, .bz2 path, bytes or uint8 array), fed in blocks.
        Returns (device ms, text bytes, records read) summed over the feeds."""
        self.begin_file()
        tot = [0.0, 0, 0]

        def feed(piece, eof):
            self.feed(piece, eof)
            ms, nb, nr = self.last_feed()
            tot[0] += ms
            tot[1] += nb
            tot[2] += nr

        if isinstance(src, (bytes, bytearray, memoryview, np.ndarray)):
            data = _text(src)
            for a in range(0, max(data.size, 1), block):
                feed(data[a:a + block], a + block >= data.size)
            return tuple(tot)
        p = os.fspath(src)
        opener = bz2.open if p.endswith(".bz2") else open
        with opener(p, "rb") as f:
            chunk = f.read(block)
            while True:
                nxt = f.read(block)
                feed(chunk, not nxt)
                if not nxt:
                    break
                chunk = nxt
        return tuple(tot)

    def read_trans(self, files: Iterable[Union[str, os.PathLike, bytes]], block: int = _BLOCK) -> Lines:
        for f in files:
            self.read_file(f, block)
        return self.lines()

    def count(self):
        n, r = ctypes.c_size_t(0), ctypes.c_size_t(0)
        capi.check(capi.load().gx_ingest_count(self._h, ctypes.byref(n), ctypes.byref(r)))
        return n.value, r.value

    def lines(self) -> Lines:
        n, r = self.count()
        out = Lines(np.empty(n), np.empty(n), np.empty(n), np.empty(n, np.int32), np.empty(n), np.empty(n), r)
        capi.check(capi.load().gx_ingest_download(self._h, n, capi.dptr(out.elow), capi.dptr(out.nu),
                                                  capi.dptr(out.aif), capi.iptr(out.gns), capi.dptr(out.gamma),
                                                  capi.dptr(out.nexp)))
        return out

    def clear(self) -> None:
        capi.check(capi.load().gx_ingest_clear(self._h))

    def close(self) -> None:
        if self._h:
            capi.load().gx_ingest_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def format_trans(id_f: np.ndarray, id_i: np.ndarray, aif: np.ndarray) -> np.ndarray:
    """ExoMol .trans text (uint8 array), one "%12d %12d %10.4E" line per transition."""
    f = np.ascontiguousarray(id_f, np.int32)
    i = np.ascontiguousarray(id_i, np.int32)
    a = np.ascontiguousarray(aif, np.float64)
    cap = 48 * f.size + 1
    buf = np.empty(cap, np.uint8)
    wrote = ctypes.c_size_t(0)
    capi.check(capi.load().gx_format_trans(f.size, capi.iptr(f), capi.iptr(i), capi.dptr(a), buf.ctypes.data, cap,
                                           ctypes.byref(wrote)))
    return buf[:wrote.value]


def read_hitran(src: Union[str, bytes, os.PathLike, np.ndarray], pressure: float = 1.0, mixture: float = 1.0,
                ref_partition: float = 0.0, lo: float = -np.inf, hi: float = np.inf, nthreads: int = 0,
                tips=None):
    """HITRAN .par text read as HITRANStateReader and the Voigt driver read it
    (HITRANStateReader.cpp:3-250, VoigtProfile.cpp:61-75): (Lines, molecule id, isotope).
    ref_partition = Q(296 K), used for records whose A field is 0; when it is not given and
    ``tips`` (a gpu_cross_amd.tips.Tips) is, it is TIPS-2011 Q(296 K) of the first record's
    molecule and isotopologue, as HITRANStateReader::OpenFile takes it (HITRANStateReader.cpp:81-95)."""
    t = _text(src if isinstance(src, np.ndarray) else _read_bytes(src))
    L = capi.load()
    if ref_partition <= 0 and tips is not None and t.size:
        nl = np.nonzero(t == 10)[0]
        first = np.ascontiguousarray(t[: (nl[0] + 1) if nl.size else t.size])
        _, mol0, iso0 = read_hitran(first, pressure, mixture, 1.0)
        ref_partition = tips.q(mol0, iso0, 296.0)[1]
    o = capi.HitranOpts(float(pressure), float(mixture), float(ref_partition), float(lo), float(hi), int(nthreads))
    n, rec = ctypes.c_size_t(0), ctypes.c_size_t(0)
    mol, iso = ctypes.c_int(0), ctypes.c_int(0)
    # 161 bytes per HITRAN record; a text with shorter records reports its count and is read again
    cap = t.size // 161 + 16
    while True:
        out = [np.empty(cap), np.empty(cap), np.empty(cap), np.empty(cap, np.int32), np.empty(cap), np.empty(cap)]
        rc = L.gx_hitran_read(t.ctypes.data, t.size, ctypes.byref(o), cap, capi.dptr(out[0]), capi.dptr(out[1]),
                              capi.dptr(out[2]), capi.iptr(out[3]), capi.dptr(out[4]), capi.dptr(out[5]),
                              ctypes.byref(n), ctypes.byref(rec), ctypes.byref(mol), ctypes.byref(iso))
        if rc == capi.GX_INVALID_ARG and n.value > cap:
            cap = n.value
            continue
        capi.check(rc)
        break
    m = n.value
    return Lines(*[a[:m] for a in out], rec.value), mol.value, iso.value


def format_hitran(nu, intens, aif, gamma_air, gamma_self, elow, nexp, gup, glow, mol_id: int = 11,
                  iso: int = 1) -> np.ndarray:
    """HITRAN .par text (uint8 array), one 160-character record per line."""
    arrs = [np.ascontiguousarray(a, np.float64) for a in (nu, intens, aif, gamma_air, gamma_self, elow, nexp, gup, glow)]
    n = arrs[0].size
    buf = np.empty(161 * n + 1, np.uint8)
    wrote = ctypes.c_size_t(0)
    capi.check(capi.load().gx_format_hitran(n, mol_id, iso, *[capi.dptr(a) for a in arrs], buf.ctypes.data, buf.size,
                                            ctypes.byref(wrote)))
    return buf[:wrote.value]
