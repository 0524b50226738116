"""Batched (T,P) tables on the GPU: every grid of a table against the FP64
oracle of the same point, graph replays against the eager run, FP64 tables
against the oracle at the parity bar."""
import numpy as np
import pytest

from gpu_cross_amd import capi
from gpu_cross_amd.manager import make_grid
from gpu_cross_amd.synth import make_lines
from gpu_cross_amd.table import ResidentLines, Table
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def rel_err(out, ref, floor=1e-6):
    mask = ref >= floor * ref.max()
    return float(np.max(np.abs(out[mask] - ref[mask]) / ref[mask]))


@pytest.fixture(scope="module")
def setup():
    ll = make_lines(21, 20000, 20000.0, 300_000, 12000.0, 17.03)
    sel = np.nonzero((ll.nu > 1950.0) & (ll.nu < 2250.0))[0]
    sub = ll.subset(sel)
    lines = ResidentLines.from_linelist(sub)
    temps = [300.0, 1200.0, 3000.0]
    press = [1e-3, 0.3, 30.0]
    pts = [(T, P, po.partition(ll.state_e, ll.state_g, T)) for T in temps for P in press]
    return ll, sub, lines, pts


def oracle_grid(sub, T, P, Q, freq, d, cutoff):
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=cutoff)
    return po.cross_section(c, sub.elow, sub.nu, sub.aif, sub.gns, sub.gamma_at(P), sub.nexp, freq, freq.size, 0, 8)


def test_mixed_table_vs_oracle_and_graph_replays(setup):
    ll, sub, lines, pts = setup
    M = 30000
    freq, d = make_grid(2000.0, 2200.0, M)
    print(f"table: {len(pts)} points, {sub.n} lines, {M} bins", flush=True)
    eager = Table(lines, 2000.0, 2200.0, M, pts, 17.03, cutoff=25.0, use_graph=False, autotune=True)
    eager.run()
    t = eager.wait()
    assert t.evals > 0
    # the split of the accumulation into region I and the W4 core pass
    assert 0 < t.core_evals < t.evals
    assert t.region1_ms > 0 and t.core_ms > 0
    assert t.region1_ms + t.core_ms <= t.accumulate_ms * 1.01 + 0.05
    ref_grids = [eager.read(p) for p in range(len(pts))]
    for p, (T, P, Q) in enumerate(pts):
        err = rel_err(ref_grids[p], oracle_grid(sub, T, P, Q, freq, d, 25.0))
        print(f"  point T={T} P={P}: max rel err {err:.2e}", flush=True)
        assert err <= 1e-5
    # same launch shape (the autotuner's cached choice) so the summation grouping is identical
    graph = Table(lines, 2000.0, 2200.0, M, pts, 17.03, cutoff=25.0, use_graph=True, autotune=True)
    assert graph.shape() == eager.shape()
    for rep in range(3):
        graph.run()
        graph.wait()
        for p in range(len(pts)):
            assert np.array_equal(graph.read(p), ref_grids[p]), (rep, p)
        eager.run()                     # eager work between the replays
        eager.wait()
    graph.close()
    eager.close()


@pytest.mark.parametrize("R,autotune", [("1", False), ("2", True), ("8", False)])
def test_mixed_table_fixed_block_size(setup, monkeypatch, R, autotune):
    """GX_MIXED_R fixes a table's block size (64 R bins) for its ranges, far field and
    kernels alike; every grid against the oracle, with the far field engaged."""
    monkeypatch.setenv("GX_MIXED_R", R)
    ll, sub, lines, pts = setup
    M = 30000
    freq, d = make_grid(2000.0, 2200.0, M)
    tab = Table(lines, 2000.0, 2200.0, M, pts, 17.03, cutoff=25.0, use_graph=False, autotune=autotune)
    assert tab.shape()[0] == int(R)
    tab.run()
    t = tab.wait()
    print(f"table R={R}: far field {t.far_evals / t.evals:.1%} of pairs", flush=True)
    assert t.far_evals > 0
    for p, (T, P, Q) in enumerate(pts):
        err = rel_err(tab.read(p), oracle_grid(sub, T, P, Q, freq, d, 25.0))
        assert err <= 1e-5, (p, err)
    tab.close()


def test_fp64_table_matches_oracle(setup):
    ll, sub, lines, pts = setup
    M = 20000
    freq, d = make_grid(2050.0, 2150.0, M)
    tab = Table(lines, 2050.0, 2150.0, M, pts[:4], 17.03, cutoff=0.0, precision=capi.PREC_FP64, use_graph=True)
    tab.run()
    tab.wait()
    for p, (T, P, Q) in enumerate(pts[:4]):
        ref = oracle_grid(sub, T, P, Q, freq, d, min(25.0 * P, 100.0))     # reference cutoff rule per point
        assert rel_err(tab.read(p), ref) <= 1e-10
    tab.close()


def test_table_grid_equals_single_point_manager(setup):
    """A table grid equals the manager path's grid for that point alone (same kernels, same order)."""
    from gpu_cross_amd.manager import Mi355xManager
    ll, sub, lines, pts = setup
    M = 12000
    freq, d = make_grid(2080.0, 2140.0, M)
    T, P, Q = pts[4]
    tab = Table(lines, 2080.0, 2140.0, M, [pts[4]], 17.03, cutoff=25.0, use_graph=False, autotune=False)
    tab.run()
    tab.wait()
    srt = sub.sorted_by_nu()
    with Mi355xManager(capi.VOIGT, precision=capi.PREC_MIXED, cutoff=25.0) as m:
        m.InitializeVectors(M)
        m.TransferFreq(freq, np.zeros(M), M)
        m.InitializeConstants(0.0, T, Q, d, 17.03, P, 500.0)
        m.TransferVectors(srt.n, srt.elow, srt.nu, srt.aif, srt.gns, srt.gamma_at(P), srt.nexp)
        m.ExecuteCrossSection(M, srt.n, 0)
        out = np.zeros(M)
        m.TransferResults(None, out, M)
    np.testing.assert_allclose(tab.read(0), out, rtol=1e-6, atol=0)
    tab.close()


def test_table_saved_to_an_opacity_file(setup, tmp_path):
    from gpu_cross_amd import opacity
    ll, sub, lines, pts = setup
    M = 8000
    tab = Table(lines, 2050.0, 2150.0, M, pts[:3], 17.03, cutoff=25.0, use_graph=False, autotune=False)
    tab.run()
    tab.wait()
    path = tmp_path / "table.gxopac"
    tab.save(path)
    t = opacity.read_table(path)
    assert t.npoints == M and t.start_nu == 2050.0 and t.end_nu == 2150.0
    assert np.array_equal(t.points, np.array(pts[:3]))
    for k in range(3):
        assert np.array_equal(np.asarray(t.grids[k]), tab.read(k))
    freq, _ = make_grid(2050.0, 2150.0, M)
    assert np.array_equal(t.freq, freq)


def test_mixed_table_at_the_48_point_set_vs_oracle():
    """The 48-point opacity-table workload's own (T,P) set -- 8 temperatures log-spaced
    300-3000 K x 6 pressures log-spaced 1e-4..100 atm, 25 cm^-1 cutoff, 0.01 cm^-1 bins --
    as ONE graph-captured mixed-precision table on a 200 cm^-1 sub-range: every grid within
    1e-5 of the FP64 oracle, the far field engaged (real and complex-centred moments)."""
    from gpu_cross_amd.synth import WORKLOADS
    wl = WORKLOADS["table_48"]
    ll = make_lines(41, 40000, 20000.0, 600_000, 12000.0, 17.03)
    sel = np.nonzero((ll.nu > 5970.0) & (ll.nu < 6230.0))[0]
    sub = ll.subset(sel)
    lines = ResidentLines.from_linelist(sub)
    pts = [(T, P, po.partition(ll.state_e, ll.state_g, T)) for T in wl["temperatures"] for P in wl["pressures"]]
    assert len(pts) == 48 and pts[0][1] == 1e-4 and abs(pts[5][1] - 100.0) < 1e-9
    M = 20000
    freq, d = make_grid(6000.0, 6200.0, M)
    assert d == 0.01
    print(f"48-point table: {sub.n} lines, {M} bins", flush=True)
    tab = Table(lines, 6000.0, 6200.0, M, pts, 17.03, cutoff=25.0, use_graph=True, autotune=True)
    tab.run()
    t = tab.wait()
    print(f"  far field {t.far_evals / t.evals:.1%} of pairs, core pass {t.core_evals / t.evals:.2%}", flush=True)
    assert t.far_evals > 0 and t.core_evals > 0
    worst = 0.0
    for p, (T, P, Q) in enumerate(pts):
        err = rel_err(tab.read(p), oracle_grid(sub, T, P, Q, freq, d, 25.0))
        worst = max(worst, err)
        assert err <= 1e-5, (T, P, err)
    print(f"  worst grid: max rel err {worst:.2e}", flush=True)
    tab.close()
    lines.close()
