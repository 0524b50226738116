"""TIPS-2011 partition sums (gx_tips_*): the interpolation against what Lagrange
interpolation fixes (polynomial reproduction, node values), the fixed-form reader on a
synthetic source with every construct it handles, and -- where the reference checkout
exists -- its own tables (node values, monotonic Q(T), published Q(296 K))."""
import os

import numpy as np
import pytest

from gpu_cross_amd import capi, ingest
from gpu_cross_amd.tips import Tips, lagrange

REF_TIPS = "/root/reference/HITRAN_files/BD_TIPS_2011_v1p0.for"
TDAT = [60.0 + 25.0 * k for k in range(119)]


def test_lagrange_reproduces_cubics_inside_and_quadratics_at_the_ends():
    rng = np.random.default_rng(3)
    a = np.cumsum(rng.uniform(5.0, 40.0, 60)) + 50.0
    cubic = lambda x: 3.0 - 0.2 * x + 1e-3 * x ** 2 + 2e-6 * x ** 3
    quad = lambda x: 7.0 + 0.3 * x - 2e-4 * x ** 2
    b3, b2 = cubic(a), quad(a)
    for x in rng.uniform(a[2] + 1e-9, a[-2], 200):          # 4-point intervals
        assert lagrange(a, b3, x) == pytest.approx(cubic(x), rel=1e-11)
    for x in list(rng.uniform(a[0] - 10.0, a[2], 50)) + list(rng.uniform(a[-2] + 1e-9, a[-1], 50)):
        assert lagrange(a, b2, x) == pytest.approx(quad(x), rel=1e-11)   # 3-point at both ends
    for k in range(1, a.size):                                # node values come back exactly
        assert lagrange(a, b3, a[k]) == b3[k]
    with pytest.raises(capi.GxError):
        lagrange(a, b3, a[-1] + 1.0)


def _fortran_source():
    """A two-molecule TIPS source in the reference's fixed form: comments, tab lines,
    continuation lines, repeat counts, a D-exponent literal and a molecule with no table."""
    temps = ",".join(f"{t:.0f}." for t in TDAT)
    chunks = [temps[i:i + 60] for i in range(0, len(temps), 60)]
    lines = ["      Subroutine BD_TIPS_2011(", "     I MOL,\t\t! HITRAN molecule number", "     O QT)",
             "C  a comment line", "      PARAMETER (NT=119)", "      data Tdat/" + chunks[0][:40]]
    lines += ["     +" + c for c in [chunks[0][40:]] + chunks[1:]]
    lines[-1] += "/"
    lines += ["      IF(MOL.EQ.7) THEN", "      CALL QT_XY(Temp,ISO,gi,QT)", "      go to 100", "      ENDIF",
              "      IF(MOL.EQ.34) THEN", "c...not applicable", "      gi=0", "      QT = 0.", "      go to 100",
              "      ENDIF", " 100  RETURN", "      END", "c     *****************",
              "      Subroutine QT_XY   (", "     I T, \t! temperature in K", "     O QT)",
              "      dimension xgj( 2), QofT( 2,119),Q(NT)", "      data xgj/ 1.,", "     + 6./"]
    rows = {1: [0.1 * (1.0 + k) ** 1.5 for k in range(119)], 2: [2.5 + 0.01 * k for k in range(119)]}
    for iso, vals in rows.items():
        toks = [f"{v:.5E}" for v in vals]
        if iso == 2:
            toks = ["2*" + toks[0]] + toks[2:-1] + [toks[-1].replace("E", "D")]
        body = ", ".join(toks)
        parts = [body[i:i + 60] for i in range(0, len(body), 60)]    # tokens may split across lines
        lines.append(f"      data (QofT( {iso},J),J=1,119)/")
        lines += ["     + " + p for p in parts]
        lines[-1] += "/"
    lines += ["\tdo I=1,NT", "\t  Q(I)=QofT(iso,I)", "\tenddo", "   99 return", "      end"]
    assert max(len(x) for x in lines) <= 72
    return "\n".join(lines) + "\n", rows


def test_reader_on_a_synthetic_source():
    text, rows = _fortran_source()
    t = Tips(text=text.encode())
    assert t.isotopes(7) == 2 and t.isotopes(34) == 0 and t.isotopes(1) == 0
    assert t.q(34, 1, 500.0) == (0.0, 0.0)                    # molecule answered with Q = 0
    for iso in (1, 2):
        vals = [float(np.float32(float(f"{v:.5E}"))) for v in rows[iso]]    # default REAL literals
        if iso == 2:
            vals[1] = vals[0]                                                # 2*v repeat count
            vals[-1] = float(f"{rows[2][-1]:.5E}")                           # D literal: double
        for k in range(1, 118):
            gi, q = t.q(7, iso, TDAT[k])
            assert q == vals[k], (iso, k)
            assert gi == (1.0 if iso == 1 else 6.0)
        want = lagrange(np.array(TDAT), np.array(vals), 1234.5)
        assert t.q(7, iso, 1234.5)[1] == want
    with pytest.raises(capi.GxError):
        t.q(7, 3, 296.0)                                        # no third isotopologue
    with pytest.raises(capi.GxError):
        t.q(5, 1, 296.0)                                        # no such molecule in this source
    with pytest.raises(capi.GxError):
        t.q(7, 1, 3000.5)                                       # outside 70-3000 K: Q = -1
    with pytest.raises(capi.GxError):
        Tips(text=b"      Subroutine QT_XY(\n      end\n")      # no tables at all


@pytest.mark.skipif(not os.path.exists(REF_TIPS), reason="needs the reference checkout's TIPS source")
def test_reference_tables():
    t = Tips(REF_TIPS)
    counts = {m: t.isotopes(m) for m in range(1, 52)}
    assert counts[1] == 6 and counts[2] == 11 and counts[3] == 18 and counts[6] == 4
    assert sum(counts.values()) == 141 and counts[34] == 0 and counts[39] == 0
    assert t.q(34, 1, 296.0) == (0.0, 0.0) and t.q(39, 1, 296.0) == (0.0, 0.0)
    # H2O 161 at the 285 K node: the DATA literal 0.16503E+03 as a single-precision constant
    assert t.q(1, 1, 285.0)[1] == float(np.float32(165.03))
    # Q(T) rises with T for every tabulated isotopologue but one: the CF4 table (molecule 42)
    # jumps eightfold between its 585 K and 610 K nodes (0.67764E+06, 0.55684E+07), so its
    # 4-point interpolation dips between 560 and 585 K; the reader keeps the table as given
    T = np.linspace(70.0, 3000.0, 300)
    for m, n in counts.items():
        for iso in range(1, n + 1):
            q = np.array([t.q(m, iso, x)[1] for x in T])
            if q.max() > 0 and m != 42:
                assert np.all(np.diff(q) > 0), (m, iso)
    assert t.q(42, 1, 585.0)[1] == float(np.float32(0.67764e6))
    assert t.q(42, 1, 610.0)[1] == float(np.float32(0.55684e7))
    # published HITRAN Q(296 K) of the main isotopologues (TIPS versions differ by < 0.5 %)
    published = {(1, 1): 174.58, (2, 1): 286.09, (3, 1): 3483.71, (4, 1): 4984.9, (5, 1): 107.42,
                 (6, 1): 590.48, (7, 1): 215.73, (8, 1): 1142.13, (11, 1): 1725.22}
    for (m, iso), q in published.items():
        assert t.q(m, iso, 296.0)[1] == pytest.approx(q, rel=5e-3), (m, iso)


@pytest.mark.skipif(not os.path.exists(REF_TIPS), reason="needs the reference checkout's TIPS source")
def test_hitran_reader_takes_q296_from_tips():
    """A record with A = 0 gets A from its intensity with Q(296 K); with a TIPS table and no
    explicit Q(296 K), the reader takes TIPS Q(296 K) of the first record's molecule."""
    t = Tips(REF_TIPS)
    q296 = t.q(2, 1, 296.0)[1]
    n = 5
    rng = np.random.default_rng(1)
    text = ingest.format_hitran(np.linspace(2000.0, 2001.0, n), 10 ** rng.uniform(-22, -19, n), np.zeros(n),
                                np.full(n, 0.07), np.full(n, 0.09), rng.uniform(0, 3000, n), np.full(n, 0.75),
                                np.full(n, 3.0), np.full(n, 1.0), mol_id=2, iso=1)
    a, mol, iso = ingest.read_hitran(text, tips=t)
    b, _, _ = ingest.read_hitran(text, ref_partition=q296)
    assert (mol, iso) == (2, 1)
    assert np.array_equal(a.aif, b.aif) and np.all(a.aif > 0)
