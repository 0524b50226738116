"""Parity of the HIP library (through its C ABI) with the oracle and with the
reference's own golden vectors.  Needs an MI355X.

Bar: FP64 path within 1e-10 relative per bin for bins >= 1e-6 of the grid
maximum; the mixed-precision path within 1e-5 of the FP64 oracle.
"""
import glob
import json
import os

import numpy as np
import pytest

from gpu_cross_amd import capi
from gpu_cross_amd.manager import Mi355xManager, ProfileDriver, make_grid
from gpu_cross_amd.synth import make_lines
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
FP64_TOL = 1e-10
MIXED_TOL = 1e-5


def rel_err(out, ref, floor=1e-6):
    mask = ref >= floor * ref.max()
    assert mask.sum() > 0
    return float(np.max(np.abs(out[mask] - ref[mask]) / ref[mask]))


def run_manager(profile, L, freq, d, T, Q, mass, P, points=None, start_idx=0, gamma=None, ref_temp=500.0, **kw):
    points = freq.size if points is None else points
    with Mi355xManager(profile, **kw) as m:
        m.InitializeVectors(freq.size)
        m.TransferFreq(freq, np.zeros_like(freq), freq.size)
        m.InitializeConstants(0.0, T, Q, d, mass, P, ref_temp)
        m.TransferVectors(L.n, L.elow, L.nu, L.aif, L.gns, gamma, L.nexp if gamma is not None else None)
        m.ExecuteCrossSection(points, L.n, start_idx)
        out = np.zeros_like(freq)
        m.TransferResults(None, out, freq.size)
        t = m.last_timing()
    return out, t


@pytest.fixture(scope="module")
def voigt_lines():
    ll = make_lines(2, 20000, 20000.0, 400_000, 12000.0, 17.03)
    return ll


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_golden_through_driver(path):
    z = np.load(path)
    meta = json.loads(bytes(z["meta"]).decode())
    print(f"golden {meta['name']}", flush=True)
    voigt = meta["profile"] == capi.VOIGT
    drv = ProfileDriver(meta["profile"], meta["start"], meta["end"], meta["npoints"], meta["T"], meta["Q"],
                        meta["mass"], meta["P"], meta["ref_temp"], max_points=meta["max_points"],
                        capacity=meta["capacity"])
    out = drv.run(z["elow"], z["nu"], z["aif"], z["gns"], z["gamma"] if voigt else None,
                  z["nexp"] if voigt else None)
    assert rel_err(out, z["intens"]) <= FP64_TOL


@pytest.mark.parametrize("T,P,conv", [(1000.0, 1.0, 0), (296.0, 0.1, 0), (2500.0, 3.0, 0), (700.0, 0.5, 1)])
def test_voigt_fp64_vs_oracle(voigt_lines, T, P, conv):
    ll = voigt_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 2900.0) & (ll.nu < 3200.0))[0]
    L = ll.subset(sel)                                # unsorted order: the library sorts on device
    freq, d = make_grid(3000.0, 3100.0, 40000)
    gam = L.gamma_at(P)
    print(f"voigt fp64 T={T} P={P} conv={conv} lines={L.n}", flush=True)
    out, t = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, convention=conv)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, convention=conv)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    assert rel_err(out, ref) <= FP64_TOL
    assert t.evals == po.count_evals(c, L.nu, freq, freq.size, 0)


def test_voigt_subrange_and_accumulation(voigt_lines):
    """ExecuteCrossSection adds into bins [start_idx, start_idx+points) of a persistent grid."""
    ll = voigt_lines
    T, P = 1200.0, 1.0
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 4970.0) & (ll.nu < 5130.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(5000.0, 5100.0, 20000)
    gam = L.gamma_at(P)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0)
    with Mi355xManager(capi.VOIGT) as m:
        m.InitializeVectors(freq.size)
        m.TransferFreq(freq, np.zeros_like(freq), freq.size)
        m.InitializeConstants(0.0, T, Q, d, 17.03, P, 500.0)
        half = L.n // 2
        first = L.subset(np.arange(half))
        second = L.subset(np.arange(half, L.n))
        m.TransferVectors(first.n, first.elow, first.nu, first.aif, first.gns, first.gamma_at(P), first.nexp)
        m.ExecuteCrossSection(7000, first.n, 3000)
        m.TransferVectors(second.n, second.elow, second.nu, second.aif, second.gns, second.gamma_at(P), second.nexp)
        m.ExecuteCrossSection(12000, second.n, 8000)
        out = np.zeros_like(freq)
        m.TransferResults(None, out, freq.size)
    ref = po.cross_section(c, first.elow, first.nu, first.aif, first.gns, first.gamma_at(P), first.nexp, freq, 7000,
                           3000)
    ref = po.cross_section(c, second.elow, second.nu, second.aif, second.gns, second.gamma_at(P), second.nexp, freq,
                           12000, 8000, out=ref)
    assert np.all(out[:3000] == 0.0)
    assert rel_err(out, ref) <= FP64_TOL


@pytest.mark.parametrize("T,widths", [(296.0, 0.0), (1500.0, 0.0), (296.0, 25.0), (2000.0, 25.0)])
def test_doppler_fp64_vs_oracle(T, widths):
    ll = make_lines(1, 20000, 15000.0, 200_000, 10000.0, 30.01, voigt=False)
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 1900.0) & (ll.nu < 2200.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(2000.0, 2100.0, 20000)
    print(f"doppler T={T} widths={widths} lines={L.n}", flush=True)
    out, t = run_manager(capi.DOPPLER, L, freq, d, T, Q, 30.01, 1.0, ref_temp=296.0, doppler_widths=widths)
    c = po.consts(po.DOPPLER, T, Q, d, 30.01, 1.0, 296.0)       # reference window: 100 cm^-1
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, None, None, freq, freq.size, 0, 4)
    assert rel_err(out, ref) <= FP64_TOL


@pytest.mark.parametrize("T,widths", [(296.0, 0.0), (1500.0, 0.0), (296.0, 25.0), (2500.0, 25.0)])
def test_doppler_gpu_convention_vs_oracle(T, widths):
    """The CUDA manager's Doppler conventions (cross_kernal.cu:645-647, 667, 686-687): the 1/e
    width, a line at nu = 0 adds nothing (S = 0), and bin j takes a line iff
    -C <= nu - f_j <= C.  Lines at nu = 0 and within the window of a grid starting at 0."""
    ll = make_lines(1, 20000, 15000.0, 200_000, 10000.0, 30.01, voigt=False)
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero(ll.nu < 220.0)[0]
    L = ll.subset(sel)
    freq, d = make_grid(0.0, 100.0, 10000)
    # two lines at nu = 0 and one exactly C above a node (the window's inclusive edge)
    extra = np.array([0.0, 0.0, freq[4000] + 100.0])
    L.nu = np.concatenate([L.nu, extra])
    L.elow = np.concatenate([L.elow, [100.0, 2000.0, 500.0]])
    L.aif = np.concatenate([L.aif, [1.0, 10.0, 5.0]])
    L.gns = np.concatenate([L.gns, np.array([3, 5, 7], np.int32)])
    print(f"doppler gpu convention T={T} widths={widths} lines={L.n}", flush=True)
    out, t = run_manager(capi.DOPPLER, L, freq, d, T, Q, 30.01, 1.0, ref_temp=296.0, doppler_widths=widths,
                         convention=capi.CONV_GPU)
    c = po.consts(po.DOPPLER, T, Q, d, 30.01, 1.0, 296.0, convention=po.CONV_GPU, doppler_widths=widths)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, None, None, freq, freq.size, 0, 4)
    assert np.isfinite(out).all()
    assert rel_err(out, ref) <= FP64_TOL
    # lines with S = 0 add no evaluations on the device; every other window as the oracle's
    nz = L.nu != 0.0
    assert t.evals == po.count_evals(c, L.nu[nz], freq, freq.size, 0)


@pytest.mark.parametrize("T,hw,cutoff", [(296.0, 0.05, 0.0), (1500.0, 0.2, 0.0), (800.0, 0.01, 0.0),
                                         (800.0, 0.3, 1.5)])
def test_gaussian_vs_oracle(T, hw, cutoff):
    """The fixed-half-width Gaussian (cross_kernal.cu:471-625): window 10 HWHM (or a given
    cutoff), erf-binned, FP64, against the oracle at the parity bar; windows counted as the
    oracle's per-bin rule counts them."""
    ll = make_lines(1, 20000, 15000.0, 200_000, 10000.0, 30.01, voigt=False)
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 1990.0) & (ll.nu < 2110.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(2000.0, 2100.0, 20000)
    print(f"gaussian T={T} hw={hw} cutoff={cutoff} lines={L.n}", flush=True)
    with Mi355xManager(capi.GAUSSIAN, cutoff=cutoff) as m:
        m.InitializeVectors(freq.size)
        m.TransferFreq(freq, np.zeros_like(freq), freq.size)
        m.InitializeConstants(hw, T, Q, d, 30.01, 1.0, 296.0)
        m.TransferVectors(L.n, L.elow, L.nu, L.aif, L.gns)
        m.ExecuteCrossSection(freq.size, L.n, 0)
        out = np.zeros_like(freq)
        m.TransferResults(None, out, freq.size)
        t = m.last_timing()
    c = po.consts(po.GAUSSIAN, T, Q, d, 30.01, 1.0, 296.0, half_width=hw, cutoff=cutoff if cutoff > 0 else None)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, None, None, freq, freq.size, 0, 4)
    assert rel_err(out, ref) <= FP64_TOL
    assert t.evals == po.count_evals(c, L.nu, freq, freq.size, 0)


@pytest.mark.parametrize("mass", [17.03, 17.03 / 6.0221412927e23], ids=["amu", "grams"])
@pytest.mark.parametrize("T", [300.0, 2000.0])
def test_voigt_without_gamma_vs_oracle(voigt_lines, T, mass):
    """A Voigt chunk uploaded without gamma and n takes the CUDA manager's sum for it
    (cross_kernal.cu:860-922): gamma_L = 0.05, window 25 cm^-1, Alg. 916 near the centre and the
    3-point form beyond.  The mean mass in amu (the reference drivers' unit: its Doppler width
    there lacks Avogadro's number, so every bin takes the 3-point form) and in grams (physical
    width: Alg. 916 within ~10 Doppler widths).  FP64 against the oracle, mixed constants too."""
    ll = voigt_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 2950.0) & (ll.nu < 3150.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(3000.0, 3100.0, 20000)
    print(f"voigt without gamma T={T} mass={mass:.3e} lines={L.n}", flush=True)
    outs = []
    for prec in (capi.PREC_FP64, capi.PREC_MIXED):    # no mixed form for this sum: FP64 either way
        with Mi355xManager(capi.VOIGT, precision=prec) as m:
            m.InitializeVectors(freq.size)
            m.TransferFreq(freq, np.zeros_like(freq), freq.size)
            m.InitializeConstants(0.0, T, Q, d, mass, 1.0, 500.0)
            m.TransferVectors(L.n, L.elow, L.nu, L.aif, L.gns)
            m.ExecuteCrossSection(freq.size, L.n, 0)
            out = np.zeros_like(freq)
            m.TransferResults(None, out, freq.size)
            t = m.last_timing()
        outs.append(out)
    assert np.array_equal(outs[0], outs[1])
    c = po.consts(po.VOIGT, T, Q, d, mass, 1.0, 500.0, cutoff=po.NOGAMMA_CUTOFF)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, None, None, freq, freq.size, 0, 4)
    assert rel_err(outs[0], ref) <= FP64_TOL
    cw = po.consts(po.VOIGT, T, Q, d, mass, 1.0, 500.0, cutoff=po.NOGAMMA_CUTOFF, convention=po.CONV_GPU)
    assert t.evals == po.count_evals(cw, L.nu, freq, freq.size, 0)


@pytest.mark.parametrize("T,P", [(1500.0, 0.1), (300.0, 1e-3), (3000.0, 10.0), (1000.0, 1.0)])
def test_voigt_mixed_vs_fp64_oracle(voigt_lines, T, P):
    ll = voigt_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 5950.0) & (ll.nu < 6250.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(6000.0, 6200.0, 40000)
    gam = L.gamma_at(P)
    print(f"voigt mixed T={T} P={P} lines={L.n}", flush=True)
    out, _ = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                         cutoff=25.0)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=25.0)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(out, ref)
    print(f"  mixed max rel err {err:.3e}", flush=True)
    assert err <= MIXED_TOL


@pytest.mark.parametrize("T,P", [(300.0, 1e-4), (1500.0, 0.1), (3000.0, 10.0), (1000.0, 100.0), (300.0, 100.0),
                                 (3000.0, 100.0)])
def test_voigt_mixed_far_field(voigt_lines, monkeypatch, T, P):
    """Region I from far-field moments (the default) against the FP64 oracle, against the
    same run with the far field off (GX_FAR=0), and engaged: a share of the pairs came
    from moments.  At 10 and 100 atm every Lorentz width is >= 0.05 cm^-1 and the moments
    are complex-centred."""
    ll = voigt_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 5900.0) & (ll.nu < 6300.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(6000.0, 6200.0, 40000)
    gam = L.gamma_at(P)
    print(f"voigt mixed far field T={T} P={P} lines={L.n}", flush=True)
    on, t_on = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                           cutoff=25.0)
    monkeypatch.setenv("GX_FAR", "0")
    off, t_off = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                             cutoff=25.0)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=25.0)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    e_on, e_off = rel_err(on, ref), rel_err(off, ref)
    share = t_on.far_evals / max(t_on.evals, 1)
    print(f"  far field: {share:.1%} of pairs, {t_on.far_ms:.2f} ms; max rel err on {e_on:.3e} off {e_off:.3e}",
          flush=True)
    assert e_on <= MIXED_TOL and e_off <= MIXED_TOL
    assert t_off.far_evals == 0 and t_on.far_evals > 0
    assert t_on.far_evals + t_on.core_evals < t_on.evals


@pytest.fixture(scope="module")
def dense_lines(voigt_lines):
    """300k lines of the Voigt list with their centres mapped affinely into 6000-6100 cm^-1
    (~3000 per cm^-1, 64 lines within ~4 bins of 0.005 cm^-1, as dense as the 1e8-line
    workloads), so that leaves are narrow enough to have near ranges."""
    ll = voigt_lines
    L = ll.subset(np.nonzero(ll.nu > 1000.0)[0][:300_000])
    lo, hi = L.nu.min(), L.nu.max()
    L.nu = 6000.0 + (L.nu - lo) * (100.0 / (hi - lo))
    return L


@pytest.mark.parametrize("R,T,P", [("1", 300.0, 1e-4), ("1", 3000.0, 0.025), ("2", 1000.0, 0.4), ("1", 1500.0, 0.1),
                                   ("4", 2000.0, 1e-3)])
def test_voigt_mixed_near_leaves(voigt_lines, dense_lines, monkeypatch, R, T, P):
    """Near leaves (the default; GX_NEAR=0 switches them off): leaves whose lines' cores reach
    a block are summed from their moments at its other bins and their near-range bins go to the
    core pass.  Both forms against the FP64 oracle; with near leaves more pairs come from
    moments and more from the core pass, and the two forms agree far inside the tolerance."""
    monkeypatch.setenv("GX_MIXED_R", R)
    ll, L = voigt_lines, dense_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    freq, d = make_grid(6040.0, 6060.0, 4000)
    gam = L.gamma_at(P)
    cut = 3.0
    print(f"voigt mixed near leaves R={R} T={T} P={P} lines={L.n}", flush=True)
    on, t_on = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                           cutoff=cut)
    monkeypatch.setenv("GX_NEAR", "0")
    off, t_off = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                             cutoff=cut)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=cut)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    e_on, e_off = rel_err(on, ref), rel_err(off, ref)
    direct_on = t_on.evals - t_on.far_evals - t_on.core_evals
    direct_off = t_off.evals - t_off.far_evals - t_off.core_evals
    print(f"  region-I pairs one by one: {direct_on:.3e} with near leaves, {direct_off:.3e} without; "
          f"max rel err {e_on:.3e} / {e_off:.3e}", flush=True)
    assert e_on <= MIXED_TOL and e_off <= MIXED_TOL
    assert t_on.evals == t_off.evals
    assert t_on.far_evals > t_off.far_evals and t_on.core_evals > t_off.core_evals
    assert direct_on < direct_off
    assert rel_err(on, off) <= 2e-6


@pytest.mark.parametrize("T,P,width", [(3000.0, 100.0, 0.01), (300.0, 6.3, 0.01), (1500.0, 100.0, 0.005),
                                     (3000.0, 30.0, 0.0025)])
def test_voigt_mixed_interpolated_lines(voigt_lines, dense_lines, monkeypatch, T, P, width):
    """Interpolated lines (the default with 64-bin blocks; GX_INTERP=0 switches them off): a line
    whose window covers a block, whose core misses it and whose Lorentz width is at least 1.5
    half-blocks is evaluated at 16 Chebyshev-Lobatto nodes of the block and its sum interpolated
    to the 64 bins.  Both forms against the FP64 oracle; the interpolated pairs are counted, and
    the two forms agree far inside the tolerance."""
    monkeypatch.setenv("GX_MIXED_R", "1")
    ll, L = voigt_lines, dense_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    n = int(round(20.0 / width))
    freq, d = make_grid(6040.0, 6060.0, n)
    gam = L.gamma_at(P)
    cut = 5.0
    print(f"voigt mixed interpolated lines T={T} P={P} dnu={width} lines={L.n}", flush=True)
    on, t_on = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                           cutoff=cut)
    monkeypatch.setenv("GX_INTERP", "0")
    off, t_off = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                             cutoff=cut)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=cut)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    e_on, e_off = rel_err(on, ref), rel_err(off, ref)
    direct = t_on.evals - t_on.far_evals - t_on.core_evals
    print(f"  interpolated pairs {t_on.interp_pairs:.3e} of {direct:.3e} region-I pairs; "
          f"max rel err {e_on:.3e} / {e_off:.3e}", flush=True)
    assert e_on <= MIXED_TOL and e_off <= MIXED_TOL
    assert t_on.evals == t_off.evals and t_on.far_evals == t_off.far_evals
    assert t_off.interp_pairs == 0 and 0 < t_on.interp_pairs <= direct
    assert rel_err(on, off) <= 1e-6


@pytest.mark.parametrize("R,T,P", [("1", 3000.0, 1e-3), ("1", 1000.0, 100.0), ("2", 300.0, 10.0)])
def test_voigt_mixed_far_field_small_blocks(voigt_lines, monkeypatch, R, T, P):
    """The far field with blocks of 64 and 128 bins (real and complex centres) against the
    FP64 oracle, engaged."""
    monkeypatch.setenv("GX_MIXED_R", R)
    ll = voigt_lines
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 5900.0) & (ll.nu < 6300.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(6000.0, 6200.0, 40000)
    gam = L.gamma_at(P)
    print(f"voigt mixed far field R={R} T={T} P={P} lines={L.n}", flush=True)
    out, t = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                         cutoff=25.0)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=25.0)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(out, ref)
    print(f"  far field: {t.far_evals / max(t.evals, 1):.1%} of pairs; max rel err {err:.3e}", flush=True)
    assert err <= MIXED_TOL
    assert t.far_evals > 0


@pytest.mark.parametrize("T,cutoff", [(300.0, 0.3), (300.0, 1.2), (3000.0, 1.2), (3000.0, 0.3), (1000.0, 3.0)])
def test_voigt_mixed_short_windows(voigt_lines, T, cutoff):
    # Windows of 120-1200 bins against 256-bin blocks: both window ends in one block, one end
    # per block, and (T = 3000 K, P = 1e-3: W4 cores of ~300 bins) a core and a window end in
    # the same block, or a core that fills the whole window (cutoff 0.3).
    ll = voigt_lines
    P = 1e-3
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 5990.0) & (ll.nu < 6210.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(6000.0, 6200.0, 40000)
    gam = L.gamma_at(P)
    print(f"voigt mixed short windows T={T} cutoff={cutoff} lines={L.n}", flush=True)
    out, _ = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                         cutoff=cutoff)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=cutoff)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(out, ref)
    print(f"  mixed max rel err {err:.3e}", flush=True)
    assert err <= MIXED_TOL


@pytest.mark.parametrize("width,kind", [(1.0, "core row in LDS, grid values from HBM"), (0.5, "FP64 fallback")])
def test_voigt_mixed_fine_grids(voigt_lines, width, kind):
    # 4000 bins over `width` cm^-1 at 10000 cm^-1, T = 3000 K: the W4 cores span ~2900 (5700) bins.
    # At 1.0 cm^-1 the core pass keeps its bin row in LDS but reads the grid from HBM; at 0.5 cm^-1
    # the row no longer fits and the point runs on the FP64 path, held to the FP64 bar.
    ll = voigt_lines
    T, P = 3000.0, 1e-3
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 9970.0) & (ll.nu < 10030.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(10000.0, 10000.0 + width, 4000)
    gam = L.gamma_at(P)
    print(f"voigt mixed fine grid width={width} ({kind}) lines={L.n}", flush=True)
    out, t = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED,
                         cutoff=25.0)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=25.0)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(out, ref)
    print(f"  max rel err {err:.3e}, core evals {t.core_evals}", flush=True)
    if kind == "FP64 fallback":
        assert err <= FP64_TOL and t.core_evals == 0
    else:
        assert err <= MIXED_TOL and t.core_evals > 0


def test_fp64_deterministic_and_linear(voigt_lines):
    ll = voigt_lines
    T, P = 900.0, 1.0
    Q = po.partition(ll.state_e, ll.state_g, T)
    freq, d = make_grid(0.0, 12000.0, 120000)
    gam = ll.gamma_at(P)
    a, _ = run_manager(capi.VOIGT, ll, freq, d, T, Q, 17.03, P, gamma=gam)
    b, _ = run_manager(capi.VOIGT, ll, freq, d, T, Q, 17.03, P, gamma=gam)
    assert np.array_equal(a, b)                                 # fixed summation order
    ll2 = ll.subset(slice(None))
    ll2.aif = ll.aif * 2.0
    c2, _ = run_manager(capi.VOIGT, ll2, freq, d, T, Q, 17.03, P, gamma=gam)
    assert np.array_equal(c2, 2.0 * a)                          # linear in S, bit for bit


def test_edge_cases():
    freq, d = make_grid(100.0, 110.0, 1000)
    L = make_lines(4, 1000, 20000.0, 50, 12000.0, 17.03)
    with Mi355xManager(capi.VOIGT) as m:
        m.InitializeVectors(freq.size)
        m.TransferFreq(freq, np.full(freq.size, 1.5), freq.size)      # non-zero initial grid is kept
        m.InitializeConstants(0.0, 500.0, 100.0, d, 17.03, 1.0, 500.0)
        m.TransferVectors(0, L.elow[:0], L.nu[:0], L.aif[:0], L.gns[:0], L.gamma_ref[:0], L.nexp[:0])
        m.ExecuteCrossSection(freq.size, 0, 0)                         # empty chunk
        m.TransferVectors(L.n, L.elow, L.nu, L.aif, L.gns, L.gamma_ref, L.nexp)
        m.ExecuteCrossSection(0, L.n, 500)                             # empty bin range
        out = np.zeros_like(freq)
        m.TransferResults(None, out, freq.size)
        assert np.all(out == 1.5)
        with pytest.raises(capi.GxError):
            m.ExecuteCrossSection(10, L.n, 995)                        # past the grid end
        with pytest.raises(capi.GxError):
            m.ExecuteCrossSection(10, L.n + 1, 0)                      # more lines than uploaded
        assert m.ReadyForWork()


def test_voigt_mixed_tiny_nu_lines():
    """Lines at nu ~ 1e-5..30 cm^-1 have 1/alpha up to ~1e10: their far wings must not
    overflow the FP32 region-I evaluation (they dominate the bins near nu = 0)."""
    rng = np.random.default_rng(17)
    n = 3000
    nu = 10 ** rng.uniform(-5, np.log10(30.0), n)
    L = make_lines(5, 2000, 20000.0, n, 12000.0, 17.03)
    L.nu = np.ascontiguousarray(nu)
    L.elow = np.ascontiguousarray(rng.uniform(0, 3000, n))
    T, P, Q = 1500.0, 0.1, 1234.0
    freq, d = make_grid(0.0, 30.0, 6000)
    gam = L.gamma_at(P)
    out, _ = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED, cutoff=25.0)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=25.0)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(out, ref)
    print(f"  tiny-nu mixed max rel err {err:.3e}", flush=True)
    assert err <= MIXED_TOL


@pytest.mark.parametrize("R", ["1", "2", "4", "8"])
def test_voigt_mixed_bins_per_lane_and_repeatability(voigt_lines, monkeypatch, R):
    """Every launch shape of the mixed kernel (1, 2, 4 and 8 bins per lane: blocks of 64 R
    bins) against the oracle, and the same inputs twice give the same grid bit for bit
    (fixed summation order)."""
    monkeypatch.setenv("GX_MIXED_R", R)
    ll = voigt_lines
    T, P = 800.0, 0.01
    Q = po.partition(ll.state_e, ll.state_g, T)
    sel = np.nonzero((ll.nu > 4450.0) & (ll.nu < 4750.0))[0]
    L = ll.subset(sel)
    freq, d = make_grid(4500.0, 4700.0, 40000)
    gam = L.gamma_at(P)
    print(f"voigt mixed R={R} lines={L.n}", flush=True)
    a, t = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED, cutoff=25.0)
    b, _ = run_manager(capi.VOIGT, L, freq, d, T, Q, 17.03, P, gamma=gam, precision=capi.PREC_MIXED, cutoff=25.0)
    assert np.array_equal(a, b)
    c = po.consts(po.VOIGT, T, Q, d, 17.03, P, 500.0, cutoff=25.0)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, gam, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(a, ref)
    print(f"  max rel err {err:.3e}, core evals {t.core_evals} of {t.evals}", flush=True)
    assert err <= MIXED_TOL
    assert 0 < t.core_evals < t.evals


@pytest.mark.parametrize("T,P,x", [(296.0, 1.0, 0.79), (1200.0, 0.05, 1.0)])
def test_hitran_par_through_manager(T, P, x):
    """HITRAN .par text read by the library's reader (gamma = (gamma_air x + gamma_self
    (1 - x)) P, A from the intensity for A = 0 records, the driver's range filter) and then
    through the FP64 manager, against the oracle's cross-section of the lines read; the
    mixed path against the same at its own tolerance.  VoigtProfile.cpp:56-107 is the
    reference's loop for this input."""
    from gpu_cross_amd import ingest
    rng = np.random.default_rng(11)
    n = 30000
    nu = np.sort(rng.uniform(1950.0, 2250.0, n))
    A = 10 ** rng.uniform(-3, 2, n)
    A[::7] = 0.0
    text = bytes(ingest.format_hitran(nu, 10 ** rng.uniform(-24, -19, n), A, rng.uniform(0.02, 0.1, n),
                                      rng.uniform(0.05, 0.4, n), rng.uniform(0, 5000, n), rng.uniform(0.4, 0.8, n),
                                      rng.integers(1, 60, n).astype(float), rng.integers(1, 60, n).astype(float),
                                      mol_id=5, iso=1))
    start, end, cutoff = 2000.0, 2200.0, 25.0
    L, mol, iso = ingest.read_hitran(text, pressure=P, mixture=x, ref_partition=107.4, lo=start - cutoff,
                                     hi=end + cutoff)
    assert (mol, iso) == (5, 1) and L.n > 20000 and np.all(L.aif > 0)
    freq, d = make_grid(start, end, 20000)
    Q = 107.4 * (T / 296.0)                       # Q(T) from the caller (TIPS is not part of the library)
    print(f"hitran lines={L.n} T={T} P={P} x={x}", flush=True)
    out, t = run_manager(capi.VOIGT, L, freq, d, T, Q, 28.0, P, gamma=L.gamma, ref_temp=296.0, cutoff=cutoff)
    c = po.consts(po.VOIGT, T, Q, d, 28.0, P, 296.0, cutoff=cutoff)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, L.gamma, L.nexp, freq, freq.size, 0, 4)
    err = rel_err(out, ref)
    mixed, _ = run_manager(capi.VOIGT, L, freq, d, T, Q, 28.0, P, gamma=L.gamma, ref_temp=296.0, cutoff=cutoff,
                           precision=capi.PREC_MIXED)
    err_m = rel_err(mixed, ref)
    print(f"  fp64 max rel err {err:.3e}, mixed {err_m:.3e}", flush=True)
    assert err <= FP64_TOL and err_m <= MIXED_TOL
    assert t.evals == po.count_evals(c, L.nu, freq, freq.size, 0)
