"""gx_execute_noalloc (the FP64 path with its intermediate buffers in the
library's fixed device array) against gx_execute, through the C ABI.

Both run the same kernels on the same uploaded lines, so the grids must be
equal bit for bit; the second call of each pair has fewer lines and fewer bins
than the first, so anything left in the shared array by the first call would
show.  The oracle holds both to 1e-10 relative per bin (bins >= 1e-6 of the
maximum).  Needs an MI355X.
"""
import ctypes

import numpy as np
import pytest

from gpu_cross_amd import capi
from gpu_cross_amd.manager import make_grid
from gpu_cross_amd.synth import make_lines
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
MASS, T, P, TREF = 17.03, 1200.0, 0.5, 500.0


def rel_err(out, ref, floor=1e-6):
    mask = ref >= floor * ref.max()
    assert mask.sum() > 0
    return float(np.max(np.abs(out[mask] - ref[mask]) / ref[mask]))


class Ctx:
    """A bare context driven through the C ABI."""

    def __init__(self, profile, npoints):
        self.L = capi.load()
        self.profile = profile
        h = ctypes.c_void_p()
        capi.check(self.L.gx_create(0, profile, 0, ctypes.byref(h)))
        self.h = h
        n = ctypes.c_size_t()
        capi.check(self.L.gx_initialize_vectors(self.h, npoints, ctypes.byref(n)))

    def constants(self, Q, dfreq, precision=capi.PREC_FP64):
        c = capi.Consts(0.0, T, Q, dfreq, MASS, P, TREF, 0.0, 0.0, capi.CONV_OPENMP, precision)
        capi.check(self.L.gx_initialize_constants(self.h, ctypes.byref(c)))

    def lines(self, L, gamma):
        voigt = self.profile == capi.VOIGT
        capi.check(self.L.gx_transfer_vectors(self.h, L.n, capi.dptr(L.elow), capi.dptr(L.nu), capi.dptr(L.aif),
                                              capi.iptr(L.gns), capi.dptr(gamma if voigt else None),
                                              capi.dptr(L.nexp if voigt else None)))

    def reset_grid(self, freq):
        z = np.zeros_like(freq)
        capi.check(self.L.gx_transfer_freq(self.h, capi.dptr(freq), capi.dptr(z), freq.size))

    def result(self, n):
        out = np.zeros(n)
        capi.check(self.L.gx_transfer_results(self.h, None, capi.dptr(out), n))
        return out

    def close(self):
        self.L.gx_destroy(self.h)


@pytest.fixture(scope="module")
def lines():
    return make_lines(7, 20000, 20000.0, 300_000, 12000.0, MASS)


def run_pair(ctx, L, gam, freq, calls, noalloc):
    """Run each (n_lines, points, start_idx) call on a fresh grid; return the grids."""
    outs = []
    for n, points, start in calls:
        ctx.reset_grid(freq)
        ctx.lines(L.subset(np.arange(n)), gam[:n])
        fn = ctx.L.gx_execute_noalloc if noalloc else ctx.L.gx_execute
        capi.check(fn(ctx.h, points, n, start))
        outs.append(ctx.result(freq.size))
    return outs


@pytest.mark.parametrize("profile", [capi.VOIGT, capi.DOPPLER], ids=["voigt", "doppler"])
def test_noalloc_equals_execute_on_two_calls(lines, profile):
    sel = np.nonzero((lines.nu > 2950.0) & (lines.nu < 3160.0))[0]
    L = lines.subset(sel)                       # unsorted: the upload sorts on the device
    gam = L.gamma_at(P)
    freq, d = make_grid(3000.0, 3100.0, 20000)
    Q = po.partition(lines.state_e, lines.state_g, T)
    # second call: fewer lines, fewer bins, another start
    calls = [(L.n, freq.size, 0), (L.n // 3, 7000, 5000)]
    print(f"noalloc profile={profile} lines={L.n} bins={freq.size} calls={calls}", flush=True)
    a = Ctx(profile, freq.size)
    b = Ctx(profile, freq.size)
    try:
        a.constants(Q, d)
        b.constants(Q, d)
        want = run_pair(a, L, gam, freq, calls, noalloc=False)
        got = run_pair(b, L, gam, freq, calls, noalloc=True)
    finally:
        a.close()
        b.close()
    c = po.consts(profile, T, Q, d, MASS, P, TREF)
    for (n, points, start), w, g in zip(calls, want, got):
        assert np.array_equal(w, g), f"noalloc differs from execute (n={n}, points={points}, start={start})"
        Ls = L.subset(np.arange(n))
        Ls = Ls.subset(np.argsort(Ls.nu, kind="stable"))
        gs = Ls.gamma_at(P) if profile == capi.VOIGT else None
        ref = po.cross_section(c, Ls.elow, Ls.nu, Ls.aif, Ls.gns, gs, Ls.nexp if gs is not None else None, freq,
                               points, start, 4)
        assert rel_err(g, ref) <= FP64_TOL
        outside = np.ones(freq.size, bool)
        outside[start:start + points] = False
        assert not g[outside].any()


def test_noalloc_refuses_what_it_cannot_hold(lines):
    freq, d = make_grid(3000.0, 3100.0, 4096)
    Q = po.partition(lines.state_e, lines.state_g, T)
    ctx = Ctx(capi.VOIGT, freq.size)
    try:
        ctx.reset_grid(freq)
        # mixed-precision constants: the scratch array holds the FP64 path only
        ctx.constants(Q, d, precision=capi.PREC_MIXED)
        L = lines.subset(np.arange(1000))
        ctx.lines(L, L.gamma_at(P))
        assert ctx.L.gx_execute_noalloc(ctx.h, freq.size, L.n, 0) == capi.GX_INVALID_ARG
        assert b"FP64" in ctx.L.gx_last_error()
        # one line more than the array holds (sorted centres: no sort staging)
        ctx.constants(Q, d)
        n = capi.NOALLOC_MAX_LINES + 1
        nu = np.linspace(2990.0, 3110.0, n)
        one = np.ones(n)
        g = np.ones(n, np.int32)
        capi.check(ctx.L.gx_transfer_vectors(ctx.h, n, capi.dptr(one), capi.dptr(nu), capi.dptr(one),
                                             capi.iptr(g), capi.dptr(one * 0.05), capi.dptr(one * 0.5)))
        assert ctx.L.gx_execute_noalloc(ctx.h, freq.size, n, 0) == capi.GX_INVALID_ARG
        assert b"GX_NOALLOC_MAX_LINES" in ctx.L.gx_last_error()
        # at the limit it runs, and equals gx_execute
        m = capi.NOALLOC_MAX_LINES
        capi.check(ctx.L.gx_execute_noalloc(ctx.h, freq.size, m, 0))
        got = ctx.result(freq.size)
        ctx.reset_grid(freq)
        capi.check(ctx.L.gx_execute(ctx.h, freq.size, m, 0))
        assert np.array_equal(got, ctx.result(freq.size))
    finally:
        ctx.close()
    # more bins than the array's block slices cover
    M = capi.NOALLOC_MAX_BINS + 256
    big = Ctx(capi.VOIGT, M)
    try:
        fb, db = make_grid(3000.0, 3100.0, M)
        big.reset_grid(fb)
        big.constants(Q, db)
        L = lines.subset(np.arange(10))
        big.lines(L, L.gamma_at(P))
        assert big.L.gx_execute_noalloc(big.h, M, L.n, 0) == capi.GX_INVALID_ARG
        assert b"GX_NOALLOC_MAX_BINS" in big.L.gx_last_error()
    finally:
        big.close()
