"""GPU ExoMol reader (include/gx_ingest.h) against the oracle's restatement of the
reference's reading path, bit for bit, and against the hand-worked known-answer case.
Needs an MI355X."""
import numpy as np
import pytest

from gpu_cross_amd import capi, ingest
from tests.test_ingest_cpu import (BROAD, EXP_LINES, EXP_RECORDS, EXP_STATES, HI, LO, PRESSURE, STATES, TRANS,
                             oracle_ingest, oracle_states)

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind == "f":
        return a.shape == b.shape and np.array_equal(a.view(np.uint64), np.asarray(b, np.float64).view(np.uint64))
    return np.array_equal(a, b)


def _check(lines, exp, records):
    (el, nu, a, g, gam, n), rec = exp
    assert lines.records == rec == records if records is not None else lines.records == rec
    assert lines.n == nu.size
    for got, want, name in ((lines.elow, el, "elow"), (lines.nu, nu, "nu"), (lines.aif, a, "aif"),
                            (lines.gns, g, "gns"), (lines.gamma, gam, "gamma"), (lines.nexp, n, "nexp")):
        assert _same(got, want), name


def _reader(states_text=STATES, broad=BROAD, pressure=PRESSURE, lo=LO, hi=HI):
    st = ingest.parse_states(states_text)
    rd = ingest.ExomolReader(st, pressure=pressure, broadener=broad)
    rd.set_range(lo, hi)
    return st, rd


@pytest.mark.parametrize("block", [1 << 20, 7, 1])
def test_known_answer(block):
    print(f"ingest known answer, block {block}", flush=True)
    _, rd = _reader()
    with rd:
        got = rd.read_trans(TRANS, block=block)
    assert got.records == EXP_RECORDS
    assert got.n == len(EXP_LINES)
    for k, exp in enumerate(EXP_LINES):
        row = (got.elow[k], got.nu[k], got.aif[k], got.gns[k], got.gamma[k], got.nexp[k])
        assert row == exp, (k, row, exp)


def _synthetic(seed, nstates, nlines, noise):
    """States text and .trans text with a mix of number forms, spacing and line endings."""
    rng = np.random.default_rng(seed)
    e = np.sort(rng.uniform(0.0, 20000.0, nstates))
    g = np.where(rng.integers(0, 2, nstates) == 0, 1, 3) * (2 * rng.integers(0, 40, nstates) + 1)
    J = rng.integers(0, 40, nstates)
    st = "".join(f"{k + 1:12d} {e[k]:12.6f} {g[k]:6d} {J[k]:7d}\n" for k in range(nstates)).encode()
    f = rng.integers(1, nstates + 1, nlines).astype(np.int32)
    i = rng.integers(1, nstates + 1, nlines).astype(np.int32)
    a = 10.0 ** rng.uniform(-45, 4, nlines)
    text = bytes(ingest.format_trans(f, i, a))
    if noise:
        # the same records with other spellings, spacing and line ends
        out = []
        forms = ["{:.17g}", "{:.6e}", "{:.3f}", "+{:.5E}", "{:.25e}", "{:.4e}\t", "{:.12g}"]
        for k in range(nlines):
            form = forms[k % len(forms)]
            sep = [" ", "\t", "  "][k % 3]
            end = ["\n", "\r\n", " \n"][k % 3]
            out.append(f"{f[k]}{sep}{i[k]}{sep}{form.format(a[k])}{end}")
        extra = ["1 2 1e-310\n", "3 4 4.9406564584124654e-324\n", "5 6 1.7976931348623157e308\n",
                 "7 8 0.000000000000000000000000000000123456789012345678901\n", "9 10 9007199254740993\n",
                 "11 12 2.2250738585072011e-308\n", "13 14 .5\n", "15 16 5.\n", "17 18 -0\n"]
        text = "".join(out[: nlines // 2] + extra + out[nlines // 2:]).encode()
    broad = "".join(f"a1 {rng.uniform(0.02, 0.12):.4f} {rng.uniform(0.3, 0.8):.3f} {ji} {jf}\n"
                    for ji in range(0, 40, 3) for jf in range(0, 40, 2)).encode()
    broad += b"a0 0.0700 0.500 5\n"
    return st, text, broad


@pytest.mark.parametrize("noise", [False, True])
def test_random_files_match_oracle(noise):
    st_text, text, broad = _synthetic(3 if noise else 4, 5000, 200_000, noise)
    print(f"ingest random noise={noise}: {len(text)} bytes", flush=True)
    states = oracle_states(st_text)
    for lo, hi in [(-1e30, 1e30), (1000.0, 15000.0)]:
        # cut the file in two so the second part goes through the break / filter of the first
        files = [text[: len(text) // 2], text[len(text) // 2:]]
        exp = oracle_ingest(states, broad, 0.7, lo, hi, files)
        st = ingest.parse_states(st_text)
        with ingest.ExomolReader(st, pressure=0.7, broadener=broad) as rd:
            rd.set_range(lo, hi)
            got = rd.read_trans(files, block=1 << 20)
        print(f"  lo={lo} hi={hi}: kept {got.n}, records {got.records}", flush=True)
        _check(got, exp, None)


def test_pieces_do_not_matter():
    st_text, text, broad = _synthetic(5, 3000, 20_000, True)
    states = oracle_states(st_text)
    exp = oracle_ingest(states, broad, 1.0, -1e30, 1e30, [text])
    st = ingest.parse_states(st_text)
    for block in (997, 37, 4096 + 3):
        with ingest.ExomolReader(st, pressure=1.0, broadener=broad) as rd:
            got = rd.read_trans([text], block=block)
        _check(got, exp, None)


def test_large_file_crosses_device_passes():
    # 1e7 records = 370 MB: more than one 256 MiB device pass
    rng = np.random.default_rng(9)
    nst = 100_000
    e = np.sort(rng.uniform(0.0, 20000.0, nst))
    st = ingest.States(np.arange(nst, dtype=np.int32), e, np.full(nst, 3, np.int32),
                       rng.integers(0, 30, nst).astype(np.int32))
    n = 10_000_000
    f = rng.integers(1, nst + 1, n).astype(np.int32)
    i = rng.integers(1, nst + 1, n).astype(np.int32)
    a = 10.0 ** rng.uniform(-40, 2, n)
    text = ingest.format_trans(f, i, a)
    print(f"ingest large: {text.size} bytes", flush=True)
    with ingest.ExomolReader(st, pressure=1.0) as rd:
        rd.set_range(0.0, 1e30)
        got = rd.read_trans([text])
    nu = e[f - 1] - e[i - 1]
    keep = nu >= 0.0
    assert got.records == n and got.n == int(keep.sum())
    assert _same(got.nu, nu[keep]) and _same(got.elow, e[i - 1][keep])
    # the %10.4E text of A read back: what the C library reads from the same text
    exp_a = np.array([float(x) for x in np.char.mod("%10.4E", a[keep][:200000])])
    assert _same(got.aif[:200000], exp_a)


def test_state_id_out_of_range_is_an_error():
    st = ingest.parse_states(STATES)
    with ingest.ExomolReader(st) as rd:
        with pytest.raises(capi.GxError):
            rd.read_trans([b"1 2 1.0\n11 1 1.0\n"])


def test_tokens_that_run_together_are_refused():
    st = ingest.parse_states(STATES)
    with ingest.ExomolReader(st) as rd:
        with pytest.raises(capi.GxError):
            rd.read_trans([b"1 2 1.0\n2 3-1 1.0\n4 1 1.0\n"])


def test_without_broadener_gamma_is_never_assigned():
    st = ingest.parse_states(STATES)
    with ingest.ExomolReader(st, pressure=PRESSURE, broadener="/nonexistent/file.broad") as rd:
        rd.set_range(LO, HI)
        got = rd.read_trans(TRANS)
    assert got.n == len(EXP_LINES) and np.all(got.gamma == 0.0) and np.all(got.nexp == 0.0)
