"""Pins of the C oracle (CPU only).

Each oracle function is checked against something other than its own formula:
the golden vectors of the reference's own CPU manager, closed forms,
invariants, scipy's Faddeeva function, or brute force on tiny inputs.
"""
import glob
import json
import math
import os

import numpy as np
import pytest

from oracle import pyoracle as po

GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def load_golden(path):
    z = np.load(path)
    meta = json.loads(bytes(z["meta"]).decode())
    voigt = meta["profile"] == po.VOIGT
    lines = dict(elow=z["elow"], nu=z["nu"], aif=z["aif"], gns=z["gns"],
                 gamma=z["gamma"] if voigt else None, nexp=z["nexp"] if voigt else None)
    return meta, lines, z["intens"]


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_reference_golden(path):
    meta, L, ref = load_golden(path)
    freq, d = po.make_grid(meta["start"], meta["end"], meta["npoints"])
    c = po.consts(meta["profile"], meta["T"], meta["Q"], d, meta["mass"], meta["P"], meta["ref_temp"])
    out = po.run_reference_driver(c, L["elow"], L["nu"], L["aif"], L["gns"], L["gamma"], L["nexp"], meta["start"],
                                  meta["end"], meta["npoints"], meta["capacity"], meta["max_points"],
                                  meta["nthreads"])
    assert np.count_nonzero(ref) > 0
    # single-chunk runs are bit-identical; chunked runs differ only in the thread-sum order across chunks
    np.testing.assert_allclose(out, ref, rtol=1e-14, atol=0)
    if "chunked" not in meta["name"]:
        assert np.array_equal(out, ref)


def test_golden_manifest_covers_edge_cases():
    names = {os.path.basename(p)[:-4] for p in GOLDEN}
    assert {"voigt_halfbin_ties", "voigt_T2000_P2_t4_unsorted", "voigt_T296_P0.1_t3_chunked",
            "doppler_T1500_t2_chunked"} <= names


# ---------------------------------------------------------------- partition
def test_partition_two_level_closed_form():
    T = 296.0
    beta = 6.6260693e-27 * 2.99792458e10 / (1.380658e-16 * T)
    e = np.array([0.0, 1234.5])
    g = np.array([3, 7], np.int32)
    assert po.partition(e, g, T) == pytest.approx(3 + 7 * math.exp(-beta * 1234.5), rel=1e-15)


def test_partition_brute_force_fsum():
    rng = np.random.default_rng(5)
    e = np.sort(rng.uniform(0, 15000, 5000))
    g = rng.integers(1, 200, 5000).astype(np.int32)
    for T in (100.0, 1000.0, 3000.0):
        b = 1.4387769 / T   # c2/T to 1e-7 (the reference's h, c, k give c2 = 1.438769...)
        exact = math.fsum(float(gi) * math.exp(-b * ei) for gi, ei in zip(g, e))
        assert po.partition(e, g, T) == pytest.approx(exact, rel=1e-5)


def test_beta_is_second_radiation_constant_over_T():
    # c2 = hc/k with the reference's constants (defines.h) = 1.43877 cm K
    assert po.lib().orc_beta(1.0) == pytest.approx(1.4387700, rel=1e-5)


# ---------------------------------------------------------------- intensity
def test_intensity_closed_form_and_guards():
    L = po.lib()
    c = po.consts(po.VOIGT, 1000.0, 500.0, 0.01, 17.03, 1.0, 500.0)
    E, nu, A, g = 3000.0, 2500.0, 0.37, 9
    b = L.orc_beta(1000.0)
    cm = 1.0 / (8.0 * 3.14159265359 * 2.99792458e10)
    want = cm * A * g * 0.564189584 * math.exp(-b * E) * -math.expm1(-b * nu) / (nu * nu * 500.0)
    got = L.orc_line_intensity(c, E, nu, A, g)
    assert got == pytest.approx(want, rel=1e-14)
    # Doppler drops the 1/sqrt(pi) factor
    cd = po.consts(po.DOPPLER, 1000.0, 500.0, 0.01, 17.03)
    assert L.orc_line_intensity(cd, E, nu, A, g) == pytest.approx(got / 0.564189584, rel=1e-15)
    # nu == 0 is replaced by 1e-6 (OpenMPManager.cpp:198), g == 0 gives 0
    assert L.orc_line_intensity(c, E, 0.0, A, g) == L.orc_line_intensity(c, E, 1e-6, A, g)
    assert L.orc_line_intensity(c, E, nu, A, 0) == 0.0
    # gpu convention: g == -1 means S = A (cross_kernal.cu:777); Doppler nu == 0 gives 0
    cg = po.consts(po.VOIGT, 1000.0, 500.0, 0.01, 17.03, convention=po.CONV_GPU)
    assert L.orc_line_intensity(cg, E, nu, A, -1) == A
    cgd = po.consts(po.DOPPLER, 1000.0, 500.0, 0.01, 17.03, convention=po.CONV_GPU)
    assert L.orc_line_intensity(cgd, E, 0.0, A, g) == 0.0
    # pi literal vs acos(-1) differ at 1e-12
    assert L.orc_cmcoef(0) / L.orc_cmcoef(1) == pytest.approx(math.pi / 3.14159265359, rel=1e-15)


def test_intensity_linear_in_A_and_boltzmann_ratio():
    L = po.lib()
    c = po.consts(po.VOIGT, 700.0, 321.0, 0.01, 17.03)
    s1 = L.orc_line_intensity(c, 1000.0, 1500.0, 1.0, 5)
    s2 = L.orc_line_intensity(c, 2000.0, 1500.0, 1.0, 5)
    assert s1 / s2 == pytest.approx(math.exp(L.orc_beta(700.0) * 1000.0), rel=1e-13)
    assert L.orc_line_intensity(c, 1000.0, 1500.0, 2.0, 5) == 2 * s1


def test_gamma_scaling():
    L = po.lib()
    c = po.consts(po.VOIGT, 500.0, 1.0, 0.01, 17.03, ref_temp=500.0)
    assert L.orc_gamma_t(c, 0.07, 0.66) == 0.07
    c2 = po.consts(po.VOIGT, 1000.0, 1.0, 0.01, 17.03, ref_temp=500.0)
    assert L.orc_gamma_t(c2, 0.07, 1.0) == pytest.approx(0.035, rel=1e-15)


def test_doppler_width_conventions():
    L = po.lib()
    for T, m in ((296.0, 30.01), (1500.0, 17.03)):
        a = L.orc_dpwcoef(T, m, 0)
        b = L.orc_dpwcoef(T, m, 1)
        assert a / b == pytest.approx(math.sqrt(0.69314718056), rel=1e-15)
        # HWHM/nu = sqrt(2 ln2 k T / (m u)) / c
        k, u, c = 1.380658e-16, 1.0 / 6.0221412927e23, 2.99792458e10
        assert a == pytest.approx(math.sqrt(2 * math.log(2) * k * T / (m * u)) / c, rel=1e-10)


# ---------------------------------------------------------------- W4 Voigt
def test_humlic_against_scipy_wofz():
    from scipy.special import wofz
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.uniform(0, 30, 4000), rng.uniform(0, 6, 4000)])
    y = np.concatenate([10 ** rng.uniform(-3, 2, 4000), 10 ** rng.uniform(-4, 0.5, 4000)])
    h = po.humlic(x, y)
    w = wofz(x + 1j * y).real
    # Humlicek (1982) states ~1e-4 relative accuracy for W4
    assert np.max(np.abs(h - w) / w) < 2e-4


def test_humlic_symmetry_and_limits():
    x = np.linspace(0, 20, 401)
    for y in (1e-3, 0.1, 1.0, 7.0):
        np.testing.assert_array_equal(po.humlic(x, y), po.humlic(-x, y))
    # region I equals its closed real form c*y*(0.5+x^2+y^2)/((0.5+y^2-x^2)^2+4x^2y^2)
    xx, yy = np.meshgrid(np.linspace(15, 300, 50), np.linspace(0.01, 50, 50))
    real_form = 0.5641896 * yy * (0.5 + xx ** 2 + yy ** 2) / ((0.5 + yy ** 2 - xx ** 2) ** 2 + 4 * xx ** 2 * yy ** 2)
    np.testing.assert_allclose(po.humlic(xx.ravel(), yy.ravel()), real_form.ravel(), rtol=1e-13)
    # Lorentz asymptote y/(sqrt(pi)(x^2+y^2)) far out
    big = po.humlic(np.array([1e4]), np.array([0.3]))[0]
    assert big == pytest.approx(0.3 / (math.sqrt(math.pi) * (1e8 + 0.09)), rel=1e-7)
    # y -> 0 is the Gaussian exp(-x^2) in the core
    g = po.humlic(np.array([0.5, 1.0, 2.0]), np.array([1e-9] * 3))
    np.testing.assert_allclose(g, np.exp(-np.array([0.25, 1.0, 4.0])), rtol=2e-4)


def test_humlic_region_boundaries_continuous():
    # across S = 15 and S = 5.5 the regions agree to W4's accuracy
    for s in (15.0, 5.5):
        for y in (0.05, 1.0, 3.0):
            x0 = s - y
            lo, hi = po.humlic(np.array([x0 - 1e-9, x0 + 1e-9]), np.array([y, y]))
            assert abs(lo - hi) / hi < 3e-4


# ---------------------------------------------------------------- windows
def test_window_openmp_brute_force_and_ties():
    freq, d = po.make_grid(0.0, 128.0, 16384)       # d = 1/128 exactly
    c = po.consts(po.VOIGT, 700.0, 1.0, d, 17.03, 1.0)  # C = 25 = 3200 bins
    # nu = (k + 1/2) d puts both window edges on exact half-bin ties: round() goes away from
    # zero (9800.5 -> 9801, 9801.5 -> 9802; banker's rounding would give 9800 and 9802)
    cases = {100: (0, 3301), 777: (0, 3978), 5000: (1801, 8201), 13000: (9801, 16201), 13001: (9802, 16202),
             15000: (11801, 16384)}
    for k, want in cases.items():
        ib, ie = po.windows(c, np.array([(k + 0.5) * d]), freq, 16384, 0)
        assert (ib[0], ie[0]) == want, k
    # a chunk clamps the window to [start_idx, start_idx + points)
    ib, ie = po.windows(c, np.array([(5000 + 0.5) * d]), freq, 1000, 2000)
    assert (ib[0], ie[0]) == (2000, 3000)
    # a line at a bin centre covers exactly 2C/d bins when not clamped
    ib, ie = po.windows(c, np.array([64.0]), freq, 16384, 0)
    assert ie[0] - ib[0] == 6400


def test_window_gpu_convention_brute_force():
    freq, d = po.make_grid(10.0, 20.0, 1000)
    c = po.consts(po.VOIGT, 700.0, 1.0, d, 17.03, 0.1, convention=po.CONV_GPU)   # C = 2.5
    rng = np.random.default_rng(3)
    nus = rng.uniform(5, 25, 200)
    ib, ie = po.windows(c, nus, freq, 400, 300)
    for nu, a, b in zip(nus, ib, ie):
        member = [j for j in range(300, 700) if -2.5 <= nu - freq[j] <= 2.5]
        if member:
            assert (a, b) == (member[0], member[-1] + 1)
        else:
            assert b - a == 0


# ---------------------------------------------------------------- accumulation invariants
def _one_line(profile, nu, T=1000.0, P=1.0, gamma=0.08, npts=200000, start=900.0, end=1100.0, widths=0.0):
    freq, d = po.make_grid(start, end, npts)
    c = po.consts(profile, T, 100.0, d, 17.03, P, 500.0, doppler_widths=widths)
    L = dict(elow=np.array([500.0]), nu=np.array([nu]), aif=np.array([1.0]), gns=np.array([3], np.int32),
             gamma=np.array([gamma]), nexp=np.array([0.6]))
    out = po.cross_section(c, L["elow"], L["nu"], L["aif"], L["gns"], L["gamma"] if profile == 2 else None,
                           L["nexp"] if profile == 2 else None, freq, npts, 0)
    S = po.lib().orc_line_intensity(c, 500.0, nu, 1.0, 3)
    return out, S, d, c


def test_voigt_area_invariant():
    out, S, d, c = _one_line(po.VOIGT, 1000.0)
    gT = po.lib().orc_gamma_t(c, 0.08, 0.6)
    C = c.cutoff
    # sum I d = S * sqrt(pi) * (1 - 2 gamma_T/(pi C)) to W4 accuracy (S already holds 1/sqrt(pi))
    want = S * math.sqrt(math.pi) * (1 - 2 * gT / (math.pi * C))
    assert out.sum() * d == pytest.approx(want, rel=5e-4)


def test_doppler_telescoping_sum():
    freq, d = po.make_grid(900.0, 1100.0, 20000)
    nu = float(freq[10000])               # line on a bin centre
    out, S, _, _ = _one_line(po.DOPPLER, nu, npts=20000)
    assert out.sum() == pytest.approx(2 * S, rel=1e-12)


def test_doppler_width_cutoff_changes_nothing():
    freq, d = po.make_grid(900.0, 1100.0, 20000)
    a, _, _, _ = _one_line(po.DOPPLER, 1000.00123, T=296.0, npts=20000)
    b, _, _, _ = _one_line(po.DOPPLER, 1000.00123, T=296.0, npts=20000, widths=25.0)
    assert np.array_equal(a, b)


def test_linearity_superposition_threads_translation():
    rng = np.random.default_rng(11)
    n = 300
    freq, d = po.make_grid(512.0, 528.0, 8192)          # d = 1/512: every grid node is exact
    c = po.consts(po.VOIGT, 800.0, 222.0, d, 17.03, 0.2, 500.0)
    L = dict(elow=rng.uniform(0, 8000, n), nu=np.sort(rng.uniform(507, 533, n)), aif=10 ** rng.uniform(-4, 1, n),
             gns=rng.integers(1, 40, n).astype(np.int32), gamma=rng.uniform(0.01, 0.024, n),
             nexp=rng.uniform(0.5, 0.8, n))
    run = lambda sel=slice(None), A=None, nthreads=1, grid=freq: po.cross_section(
        c, L["elow"][sel], L["nu"][sel], (L["aif"] if A is None else A)[sel], L["gns"][sel],
        L["gamma"][sel], L["nexp"][sel], grid, 8192, 0, nthreads)
    full = run()
    assert np.array_equal(run(A=2 * L["aif"]), 2 * full)                      # linear in S
    parts = run(slice(0, 150)) + run(slice(150, None))
    np.testing.assert_allclose(parts, full, rtol=1e-13)                       # superposition
    np.testing.assert_allclose(run(nthreads=7), full, rtol=1e-13)             # thread split
    # translation: a grid that starts 40 bins later sees the same bins shifted by 40
    shifted_grid, d2 = po.make_grid(512.0 + 40 * d, 528.0 + 40 * d, 8192)
    assert d2 == d
    shifted = run(grid=shifted_grid)
    np.testing.assert_allclose(shifted[:8000], full[40:8040], rtol=1e-15)


# ------------------------------------------ Gaussian and no-gamma Voigt (CUDA-manager sums)
def _plain_line(profile, nu, S_target, freq, d, T=1000.0, mass=17.03, **kw):
    """Grid of one line of intensity S_target (A chosen so that the plain intensity equals it)."""
    c = po.consts(profile, T, 500.0, d, mass, 1.0, 500.0, **kw)
    s1 = po.lib().orc_line_intensity_plain(c, 1000.0, nu, 1.0, 1)
    aif = np.array([S_target / s1])
    out = po.cross_section(c, np.array([1000.0]), np.array([nu]), aif, np.array([1], np.int32), None, None, freq,
                           freq.size, 0, 1)
    return out, c


def test_plain_intensity_closed_form():
    T, Q = 1500.0, 3210.0
    c = po.consts(po.GAUSSIAN, T, Q, 0.01, 17.03, 1.0, 500.0, half_width=0.1)
    beta = 6.6260693e-27 * 2.99792458e10 / (1.380658e-16 * T)
    E, nu, A, g = 2345.6, 1234.5, 3.2, 7
    want = A * g * math.exp(-beta * E) * (-math.expm1(-beta * nu)) / (8 * math.pi * 2.99792458e10 * nu * nu * Q)
    assert po.lib().orc_line_intensity_plain(c, E, nu, A, g) == pytest.approx(want, rel=1e-13)
    assert po.lib().orc_line_intensity_plain(c, E, 0.0, A, g) == 0.0      # nu = 0 adds nothing


def test_erfcx_against_scipy():
    from scipy import special
    y = np.concatenate([np.linspace(0.0, 60.0, 6001), 10 ** np.linspace(-8, 3, 400)])
    got, want = po.erfcx(y), special.erfcx(y)
    assert np.max(np.abs(got - want) / want) <= 2e-15


def test_voigt916_against_scipy_wofz():
    """Alg. 916 truncated to four terms (a = 1) against the Faddeeva function: within the
    3e-4 the truncation leaves, exact at x = 0 (erfcx) and y = 0 (Gaussian)."""
    from scipy import special
    X, Y = np.meshgrid(np.linspace(0.001, 9.9, 240), 10 ** np.linspace(-3, 1, 160))
    m = X ** 2 + Y ** 2 < 100
    v = po.voigt916(X[m], Y[m])
    w = special.wofz(X[m] + 1j * Y[m]).real
    rel = np.abs(v - w) / w
    assert rel.max() <= 3e-4 and np.median(rel) <= 1e-4
    y = np.linspace(0.01, 9.0, 50)
    assert np.allclose(po.voigt916(np.zeros_like(y), y), special.wofz(1j * y).real, rtol=1e-14, atol=0)
    x = np.linspace(0.01, 9.0, 50)
    assert np.allclose(po.voigt916(x, np.zeros_like(x)), np.exp(-x * x), rtol=1e-15, atol=0)


def test_gaussian_telescoping_sum_and_half_width():
    """The binned Gaussian telescopes to 2S over a window of +-10 HWHM, and its value falls to
    one half one HWHM from the centre (erf difference of a narrow bin ~ bin x Gaussian)."""
    hw = 0.05
    freq, d = po.make_grid(990.0, 1010.0, 200_000)             # 1e-4 cm^-1 bins: HW / 500
    nu = 1000.0 + 0.37 * d
    out, c = _plain_line(po.GAUSSIAN, nu, 2.5e-20, freq, d, half_width=hw)
    assert out.sum() == pytest.approx(2 * 2.5e-20, rel=1e-12)
    prof = lambda f: np.interp(f, freq, out)
    assert prof(nu + hw) / prof(nu) == pytest.approx(0.5, rel=1e-5)   # O((bin/HW)^2) from binning
    assert prof(nu - hw) / prof(nu) == pytest.approx(0.5, rel=1e-5)
    # per-bin window |nu - f| <= 10 HW, inclusive at both ends
    nz = np.nonzero(out)[0]
    inside = np.nonzero(np.abs(nu - freq) <= 10 * hw)[0]
    assert nz.min() >= inside.min() and nz.max() <= inside.max()
    assert po.count_evals(c, np.array([nu]), freq, freq.size, 0) == inside.size


def test_gh3_is_a_three_node_gauss_hermite_sum():
    """The 3-point expression is a sum over the Gauss-Hermite nodes 0 and +-sqrt(3/2) of
    Lorentzians: w0 y / (pi (x^2 + y^2)) + w1 (1/((x - a)^2 + y^2) + 1/((x + a)^2 + y^2)),
    a^2 = 3/2, w1 = sqrt(pi)/6 (the rule's outer weight), with the reference's centre weight
    1.1181635900 (the rule's 2 sqrt(pi)/3 = 1.1816359006 with two digits swapped) and no
    y/pi on the outer pair, both as the reference writes them.  Summing the nodes one by one
    checks the oracle's combined rational form (the pair's common denominator
    (x^2 + y^2 + a^2)^2 - 4 a^2 x^2)."""
    a2 = 1.5
    w0, w1 = 1.1181635900, 0.2954089751                 # the literals (w1 = sqrt(pi)/6 to 10 digits)
    assert abs(w1 - math.sqrt(math.pi) / 6.0) < 1e-10 and abs(2.0 * math.sqrt(math.pi) / 3.0 - 1.1816359006) < 1e-10
    rng = np.random.default_rng(11)
    for x, y in zip(rng.uniform(-30.0, 30.0, 400), 10 ** rng.uniform(-3.0, 1.5, 400)):
        if x * x + y * y < 100.0:
            continue                                  # the sum uses the 3-point form from 100 up
        want = (w0 * y / (3.14159265359 * (x * x + y * y)) +
                w1 * (1.0 / ((x - math.sqrt(a2)) ** 2 + y * y) + 1.0 / ((x + math.sqrt(a2)) ** 2 + y * y)))
        assert po.gh3(x, y) == pytest.approx(want, rel=1e-12)


def test_nogamma_voigt_lorentz_limit_with_mass_in_amu():
    """With the mean mass in amu the no-gamma sum's Doppler width (no Avogadro factor) is
    ~1e-16 cm^-1, every bin takes the 3-point form, and the profile reduces (to O(gG/gamma_L)) to
    S 1.1181635900 (SQRTLN2PI/SQRTLN2)/PI gamma_L/(dnu^2 + gamma_L^2), gamma_L = 0.05."""
    freq, d = po.make_grid(3000.0, 3100.0, 20_000)
    nu = 3050.0 + 0.3 * d
    S = 1.7e-21
    out, c = _plain_line(po.VOIGT, nu, S, freq, d, cutoff=po.NOGAMMA_CUTOFF)
    g = 0.05
    dnu = nu - freq
    want = S * 1.1181635900 * (0.46971863935 / 0.83255461115) / 3.14159265359 * g / (dnu ** 2 + g ** 2)
    win = np.abs(dnu) <= 25.0
    np.testing.assert_allclose(out[win], want[win], rtol=1e-11, atol=0)   # x, y ~ 1e16: squares ~ 1e32
    assert not out[~win].any()


def test_nogamma_voigt_alg916_centre_with_mass_in_grams():
    """With the molecular mass in grams the Doppler width is physical: near the centre
    (x^2 + y^2 < 100) the sum is S SQRTLN2PI/gG y V916(|x|, y), within 3e-4 of
    S SQRTLN2PI/gG y Re w(x + iy), and symmetric about the centre."""
    from scipy import special
    freq, d = po.make_grid(2999.0, 3001.0, 4000)
    nu = 3000.0
    S = 1e-20
    mass_g = 17.03 / 6.0221412927e23
    out, c = _plain_line(po.VOIGT, nu, S, freq, d, T=1000.0, mass=mass_g, cutoff=po.NOGAMMA_CUTOFF)
    gG = nu * math.sqrt(2.0 * 0.69314718056 * 1.380658e-16 * 1000.0 / mass_g) / 2.99792458e10
    y = 0.83255461115 * 0.05 / gG
    x = 0.83255461115 * (nu - freq) / gG
    core = x ** 2 + y ** 2 < 100
    assert core.sum() > 100 and (~core).sum() > 100
    w = special.wofz(np.abs(x[core]) + 1j * y).real
    want = S * 0.46971863935 / gG * y * w
    assert np.max(np.abs(out[core] - want) / want) <= 3e-4
    mid = 2000                                                      # freq[2000] == nu
    assert freq[mid] == nu
    np.testing.assert_allclose(out[mid - 500:mid][::-1], out[mid + 1:mid + 501], rtol=1e-12)
