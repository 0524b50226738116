"""Output writers on the CPU: the reference's "%12.6f %13.8E" text (BaseProfile.cpp:65-74)
against Python's own formatting, and the opacity-table file written by the library and read
back by the library and by numpy alone."""
import numpy as np
import pytest

from gpu_cross_amd import capi, opacity


def test_text_lines_match_printf_format():
    rng = np.random.default_rng(3)
    f = np.concatenate([[0.0, 0.0000005, 0.0000015, 9999.9999995, 123456.123456789, -1.5],
                        rng.uniform(0, 30000, 5000)])
    x = np.concatenate([[0.0, 1e-300, 5e-324, 1.7976931348623157e308, -2.5e-10, 1.23456785e-20],
                        10.0 ** rng.uniform(-320, 300, 5000)])
    text = opacity.format_xsec_text(f, x)
    exp = "".join("%12.6f %13.8E\n" % (a, b) for a, b in zip(f, x)).encode()
    assert text == exp


def test_text_file(tmp_path):
    f = np.linspace(0, 10, 11)
    x = np.arange(11) * 1e-20
    p = tmp_path / "xsec.txt"
    opacity.write_xsec_text(p, f, x)
    assert p.read_bytes() == opacity.format_xsec_text(f, x)


def test_opacity_file_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    pts = [(300.0 * 10 ** (k / 3), 10.0 ** (k - 2), 1000.0 + k) for k in range(5)]
    grids = rng.uniform(0, 1e-18, (5, 1234))
    p = tmp_path / "t.gxopac"
    opacity.write_table(p, pts, grids, 0.0, 12000.0, capi.VOIGT, capi.PREC_MIXED, 17.03, 500.0, 25.0)
    t = opacity.read_table(p)
    assert t.npoints == 1234 and t.start_nu == 0.0 and t.end_nu == 12000.0 and t.dfreq == 12000.0 / 1234
    assert t.profile == capi.VOIGT and t.precision == capi.PREC_MIXED and t.mean_mass == 17.03
    assert np.array_equal(t.points, np.array(pts)) and np.array_equal(np.asarray(t.grids), grids)
    assert t.freq[5] == 0.0 + 5 * (12000.0 / 1234)
    info, p2, g2 = opacity.read_table_c(p)
    assert info.data_offset % 4096 == 0 and info.npts == 5
    assert np.array_equal(p2, np.array(pts)) and np.array_equal(g2, grids)


def test_opacity_file_rejects_other_files(tmp_path):
    p = tmp_path / "bad"
    p.write_bytes(b"x" * 256)
    with pytest.raises(ValueError):
        opacity.read_table(p)
    with pytest.raises(capi.GxError):
        opacity.read_table_c(p)
