"""ExoMol reader on the CPU: the decimal parser of the ingest kernels, the oracle's
restatement of the reference's reading path, and the product's host-side readers.

The known-answer case below is worked out by hand from the reference's source
(exomol_functions.cpp:8-73, ExomolStateReader.cpp:7-167, VoigtProfile.cpp:56-107); every
expected number is written out, so the oracle is pinned by it, not by itself.
"""
import ctypes
import math
import os
import subprocess

import numpy as np
import pytest

from gpu_cross_amd import capi, ingest
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# --- the known-answer case -------------------------------------------------------------
# states: id, E, g_tot, J.  The last text line is 1023 + 11 characters long: fgets(buf, 1024)
# reads it as two records (exomol_functions.cpp:15,28), the second being state 10.
LONG_LINE = "9 7.5 2 4" + " " * (1023 - 9) + "10 8.5 5 6\n"
STATES = ("           1        0.000000      1      0\n"
          "           2      100.000000      3      1\n"
          "           3      250.500000      3      2\n"
          "           4     1000.000000      1     12\n"
          "           5     1200.250000      6      1\n"
          "           6     5000.000000      3     11\n"
          "           7     6000.125000      1      2\n"
          "           8    20000.000000      3      3\n" + LONG_LINE).encode()
EXP_STATES = dict(id=list(range(10)),
                  energy=[0.0, 100.0, 250.5, 1000.0, 1200.25, 5000.0, 6000.125, 20000.0, 7.5, 8.5],
                  gtot=[1, 3, 3, 1, 6, 3, 1, 3, 2, 5],
                  J=[0, 1, 2, 12, 1, 11, 2, 3, 4, 6])

# keys (code + Ji + Jf): a00-1, a112, a1112 (twice: Ji=1,Jf=12 then Ji=11,Jf=2, the later
# wins), a121, a101; the blank line is skipped
BROAD = (b"a0 0.0700 0.500 0\n"
         b"a1 0.0800 0.600 1 2\n"
         b"a1 0.0900 0.700 1 12\n"
         b"a1 0.1000 0.800 11 2\n"
         b"a1 0.0500 0.400 2 1\n"
         b"\n"
         b"a1 0.0600 0.450 0 1\n")
PRESSURE = 2.0
LO, HI = 200.0, 6000.0
TRANS = [
    # file 1: r0 below LO (skipped, but its key a101 sets gamma/n), r6 above HI (read, then
    # break), r7 never read
    (b"           2            1 1.0000E-01\n"
     b"           5            2 2.5000E+00\n"
     b"           4            2 3.0000E-02\n"
     b"           7            3 1.2345E-05\n"
     b"           3            1 4.0000E+00\n"
     b"           6            5 5.0000E-07\n"
     b"           8            1 1.0000E+00\n"
     b"           2            1 7.0000E-01\n"),
    # file 2: the stream reads 2.0 of "2.0D-05" and fails at "D" on the next record
    (b"  4 2 1.0E+00\n"
     b"  3 1 2.0D-05\n"
     b"  2 1 1.0E+00\n"),
    # file 3: records are token triples, not lines; a lone "6" at the end is not a record
    (b"5 2\n1.5E+00 7 3 2.5E-01\n6\n"),
]
# (E'', nu, A, g, gamma, n) of the kept lines, in read order
EXP_LINES = [
    (100.0, 1100.25, 2.5, 6, 0.12, 0.45),        # key a111 misses (and a111a01-1): a101 carried
    (100.0, 900.0, 0.03, 1, 0.2, 0.8),           # key a1112: the Ji=11, Jf=2 line
    (250.5, 5749.625, 1.2345e-05, 1, 0.2, 0.8),  # a122 misses
    (0.0, 250.5, 4.0, 3, 0.2, 0.8),              # a102 misses; a00-1 exists but the key is a102a00-1
    (1200.25, 3799.75, 5e-07, 3, 0.2, 0.8),      # a1111 misses
    (100.0, 900.0, 1.0, 1, 0.2, 0.8),            # file 2
    (0.0, 250.5, 2.0, 3, 0.2, 0.8),
    (100.0, 1100.25, 1.5, 6, 0.2, 0.8),          # file 3
    (250.5, 5749.625, 0.25, 1, 0.2, 0.8),
]
EXP_RECORDS = 7 + 2 + 2


def oracle_states(text: bytes):
    return po.read_states(text)


def oracle_ingest(states, broad, pressure, lo, hi, files, use_broad=True):
    """The oracle's read loop; returns (E'', nu, A, g, gamma, n) arrays and the records read."""
    return po.read_trans(states, broad, pressure, lo, hi, files, use_broad)


def test_oracle_states_known_answer():
    st = oracle_states(STATES)
    for k, v in EXP_STATES.items():
        assert list(st[k]) == v, k


def test_oracle_read_loop_known_answer():
    (el, nu, a, g, gam, n), rec = oracle_ingest(EXP_STATES, BROAD, PRESSURE, LO, HI, TRANS)
    assert rec == EXP_RECORDS
    got = list(zip(el, nu, a, g, gam, n))
    assert len(got) == len(EXP_LINES)
    for row, exp in zip(got, EXP_LINES):
        assert row[0] == exp[0] and row[1] == exp[1] and row[2] == exp[2] and row[3] == exp[3]
        assert row[4] == exp[4] * 1.0 and row[5] == exp[5], (row, exp)


def test_oracle_without_broadener_leaves_gamma_unassigned():
    (_, _, _, _, gam, n), _ = oracle_ingest(EXP_STATES, None, PRESSURE, LO, HI, TRANS, use_broad=False)
    assert np.all(gam == 0.0) and np.all(n == 0.0)


def test_product_states_parser_known_answer():
    st = ingest.parse_states(STATES)
    assert list(st.id) == EXP_STATES["id"] and list(st.energy) == EXP_STATES["energy"]
    assert list(st.gtot) == EXP_STATES["gtot"] and list(st.J) == EXP_STATES["J"]


def test_product_states_parser_matches_oracle_on_odd_text():
    # base-0 ints (0x.., leading-zero octal), signs, tabs, CR, blank lines, missing fields,
    # lines of exactly 1022 / 1023 / 1024 characters, no final newline
    rng = np.random.default_rng(7)
    lines = []
    for k in range(3000):
        kind = k % 9
        if kind == 0:
            lines.append(f"{k + 1:12d} {rng.uniform(0, 4e4):12.6f} {int(rng.integers(1, 99)):6d} {k % 50:7d}  extra q n")
        elif kind == 1:
            lines.append(f"0x{k + 1:x}\t{rng.uniform(0, 100):.3e}\t0{int(rng.integers(1, 7))}\t-{k % 5}\r")
        elif kind == 2:
            lines.append("")
        elif kind == 3:
            lines.append(f"  {k + 1}  {rng.uniform(0, 10):.17g}")
        elif kind == 4:
            lines.append(f"{k + 1} 1.5 2 3" + " " * (1022 - len(f"{k + 1} 1.5 2 3")))
        elif kind == 5:
            lines.append(f"{k + 1} 2.5 3 4" + "x" * (1023 - len(f"{k + 1} 2.5 3 4")))
        elif kind == 6:
            lines.append(f"{k + 1} 3.5 4 5" + " " * (1024 - len(f"{k + 1} 3.5 4 5")) + "77 1.25 9 8")
        elif kind == 7:
            lines.append(f"+{k + 1} -{rng.uniform(0, 5):.6f} +08 012")
        else:
            lines.append(f"{k + 1} nan inf 5")
    text = ("\n".join(lines)).encode()
    exp = oracle_states(text)
    for threads in (1, 4):
        st = ingest.parse_states(text, nthreads=threads)
        assert st.n == exp["id"].size
        assert np.array_equal(st.id, exp["id"]) and np.array_equal(st.gtot, exp["gtot"])
        assert np.array_equal(st.J, exp["J"])
        assert np.array_equal(st.energy.view(np.uint64), exp["energy"].view(np.uint64))


def test_partition_closed_form_and_oracle():
    beta = 6.6260693e-27 * 2.99792458e10 / (1.380658e-16 * 300.0)
    q = ingest.partition(np.array([0.0, 100.0]), np.array([1, 3]), 300.0)
    assert q == 1.0 + 3 * math.exp(-beta * 100.0)
    rng = np.random.default_rng(1)
    e = np.sort(rng.uniform(0, 2e4, 20000))
    g = rng.integers(1, 60, 20000).astype(np.int32)
    for T in (296.0, 1000.0, 3000.0):
        assert ingest.partition(e, g, T) == po.partition(e, g, T)


def test_format_trans_lines():
    t = bytes(ingest.format_trans(np.array([1, 123456789]), np.array([7, 2]), np.array([1.5e-3, 3.25e-40])))
    assert t == b"           1            7 1.5000E-03\n   123456789            2 3.2500E-40\n"


def test_decimal_parser_against_stream_and_strtod():
    exe = os.path.join(ROOT, "tests", "cpp", "decimal_check")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), "decimal_check"], check=True)
    r = subprocess.run([exe, "300000", "11"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    # the ExoMol columns are always converted on the device
    exo = [l for l in r.stdout.splitlines() if l.startswith("exomol")][0].split()
    assert exo[exo.index("back") + 1] == "0"


# --- HITRAN .par reader ---------------------------------------------------------------
# One 160-character record laid out column by column as HITRAN's format fixes it (molecule
# I2, isotope I1, nu F12.6, S E10.3, A E10.3, gamma_air F5.4, gamma_self F5.3, E'' F10.4,
# n F4.2, delta F8.6, four 15-character quanta, error 6, references 12, flag 1, g' F7.1,
# g'' F7.1), written out by hand so that the expected values below do not come from the
# code under test.
def _hitran_record(mol="12", iso="1", nu="  1234.567890", s=" 1.234E-20", a=" 5.678E+01", gair=".0712",
                   gself="0.345", elow="  123.4567", n="0.75", delta="-.001234", gup="   27.0", glow="   25.0"):
    rec = (mol + iso + nu[-12:] + s + a + gair + gself + elow + n + delta + " " * 60 + " " * 6 + " " * 12 + " "
           + gup + glow)
    assert len(rec) == 160
    return rec + "\n"


HITRAN_PLANCK, HITRAN_VELLGT, HITRAN_BOLTZ = 6.6260693e-27, 2.99792458e10, 1.380658e-16


def test_oracle_hitran_known_answer():
    """Fields at their columns, gamma = (gamma_air x + gamma_self (1 - x)) P, g' stored as int,
    molecule and isotope from the first record (HITRANStateReader.cpp:3-58, 207-247)."""
    text = (_hitran_record() + _hitran_record(nu="   250.000000", gup="   10.9")).encode()
    (e, nu, a, g, gam, n), rec, mol, iso = po.read_hitran(text, 2.0, 0.7, 100.0, 0.0, 1e9)
    assert (mol, iso, rec) == (12, 1, 2)
    assert list(nu) == [1234.56789, 250.0]
    assert list(e) == [123.4567, 123.4567] and list(a) == [56.78, 56.78] and list(n) == [0.75, 0.75]
    assert list(g) == [27, 10]                     # 10.9 truncated by the int store
    assert gam[0] == (0.0712 * 0.7 + 0.345 * (1.0 - 0.7)) * 2.0


def test_oracle_hitran_einstein_a_from_intensity():
    """A record with A == 0 gets A from its intensity and Q(296 K), with g = 1: fed back into
    the path's intensity formula at T = 296 K and Q = Q(296 K), it returns the HITRAN
    intensity times the path's ISQRTPI factor (the two formulas are inverses)."""
    text = _hitran_record(a=" 0.000E+00", s=" 3.250E-21", nu="  2100.250000", elow=" 1500.0000").encode()
    (e, nu, a, g, gam, n), rec, mol, iso = po.read_hitran(text, 1.0, 1.0, 286.09, 0.0, 1e9)
    assert g[0] == 1
    beta = HITRAN_PLANCK * HITRAN_VELLGT / (HITRAN_BOLTZ * 296.0)
    cmcoef = 1.0 / (8.0 * math.pi * HITRAN_VELLGT)
    S = cmcoef * a[0] * g[0] * math.exp(-beta * e[0]) * (1.0 - math.exp(-beta * nu[0])) / (nu[0] * nu[0] * 286.09)
    assert abs(S / 3.25e-21 - 1.0) < 1e-14


def test_oracle_hitran_short_record_reads_the_previous_buffer():
    """The reader's buffer is kept for the whole file: a record too short for the fields it
    reads takes the missing bytes from the record before (its fgets left them there); the
    driver's filter skips below lo and stops at the first line above hi."""
    long = _hitran_record()
    short = long[:50] + "\n"                       # ends inside gamma_self: E'' onwards are stale
    text = (long + short + _hitran_record(nu="  3000.000000") + _hitran_record(nu="  1300.000000")).encode()
    (e, nu, a, g, gam, n), rec, mol, iso = po.read_hitran(text, 1.0, 1.0, 100.0, 1000.0, 2000.0)
    assert rec == 3                                 # the record above hi ends the read
    assert list(nu) == [1234.56789, 1234.56789]
    # the short record (51 characters with its line feed): gamma_air and gamma_self (columns
    # 35-44) are its own; E'' (columns 45-54) holds "  123", its line feed and terminator, so
    # atof reads 123; n (55-58) and g' (146-152) are the previous record's bytes
    assert gam[1] == 0.0712
    assert e[1] == 123.0 and n[1] == 0.75 and g[1] == 27


@pytest.fixture(scope="module")
def hitran_text():
    rng = np.random.default_rng(5)
    n = 4000
    nu = np.sort(rng.uniform(10.0, 9000.0, n))
    A = 10 ** rng.uniform(-3, 2, n)
    A[::9] = 0.0
    args = (nu, 10 ** rng.uniform(-30, -19, n), A, rng.uniform(0.01, 0.12, n), rng.uniform(0.05, 0.5, n),
            rng.uniform(0, 8000, n), rng.uniform(0.3, 0.8, n), rng.integers(1, 90, n).astype(float),
            rng.integers(1, 90, n).astype(float))
    text = bytes(ingest.format_hitran(*args, mol_id=2, iso=1))
    # odd records: a short one, one longer than fgets' 1023 characters, no final newline
    lines = text.split(b"\n")
    lines[100] = lines[100][:70]
    lines[200] = lines[200] + b" " * 1200
    return b"\n".join(lines[:-1])


@pytest.mark.parametrize("nthreads,lo,hi", [(1, -1e9, 1e9), (3, 500.0, 7000.0), (0, 2500.0, 2600.0)])
def test_product_hitran_reader_matches_oracle(hitran_text, nthreads, lo, hi):
    L, mol, iso = ingest.read_hitran(hitran_text, pressure=3.0, mixture=0.6, ref_partition=286.09, lo=lo, hi=hi,
                                     nthreads=nthreads)
    ref, rec, m2, i2 = po.read_hitran(hitran_text, 3.0, 0.6, 286.09, lo, hi)
    assert (L.records, mol, iso) == (rec, m2, i2)
    # a blank piece (the tail of the overlong line) reads nu = 0 and A = 0, and the
    # reference's A from the intensity is then 0/0: NaN on both sides
    for got, exp in zip([L.elow, L.nu, L.aif, L.gns, L.gamma, L.nexp], ref):
        assert np.array_equal(got, exp, equal_nan=True)


def test_product_hitran_reader_empty_and_format():
    L, mol, iso = ingest.read_hitran(b"")
    assert L.n == 0 and L.records == 0 and (mol, iso) == (0, 0)
    rec = bytes(ingest.format_hitran([1234.56789], [1.234e-20], [56.78], [0.0712], [0.345], [123.4567], [0.75],
                                     [27.0], [25.0], mol_id=12, iso=1))
    exp = _hitran_record(delta="0.000000")
    assert rec.decode() == exp


def test_product_hitran_reader_short_records():
    """Records far shorter than 160 characters (more records than the reader's first guess
    of the count) read as the oracle reads them."""
    text = b"".join(b"02%13.6f%10.3E\n" % (100.0 + 0.5 * i, 1e-21 * (i % 7)) for i in range(5000))
    L, mol, iso = ingest.read_hitran(text, ref_partition=300.0, lo=-1e9, hi=1e9)
    ref, rec, m2, i2 = po.read_hitran(text, 1.0, 1.0, 300.0, -1e9, 1e9)
    assert L.n == 5000 and (L.records, mol, iso) == (rec, m2, i2)
    for got, exp in zip([L.elow, L.nu, L.aif, L.gns, L.gamma, L.nexp], ref):
        assert np.array_equal(got, exp, equal_nan=True)

