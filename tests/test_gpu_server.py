"""The request server on the GPU: every result equals the same request run
alone (bit for bit), matches the oracle, and cancellation / admission
failures are reported per request."""
import numpy as np
import pytest

from gpu_cross_amd import capi
from gpu_cross_amd.manager import make_grid
from gpu_cross_amd.server import Server
from gpu_cross_amd.synth import make_lines
from gpu_cross_amd.table import ResidentLines
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world():
    ll = make_lines(31, 20000, 20000.0, 400_000, 12000.0, 17.03)
    return ll, ResidentLines.from_linelist(ll)


def requests(ll):
    Q = lambda T: po.partition(ll.state_e, ll.state_g, T)
    grids = [(1000.0, 1040.0, 8000), (2500.0, 2520.0, 2000), (6000.0, 6100.0, 20000)]
    reqs = []
    rng = np.random.default_rng(3)
    for k in range(14):
        g = grids[k % 3]
        T = float(rng.uniform(200, 3000))
        P = float(10 ** rng.uniform(-3, 1))
        if k % 5 == 4:
            reqs.append(dict(T=T, P=P, Q=Q(T), grid=g, profile=capi.DOPPLER, precision=capi.PREC_FP64, cutoff=0.0,
                             doppler_widths=25.0))
        elif k % 5 == 3:
            reqs.append(dict(T=T, P=P, Q=Q(T), grid=g, profile=capi.VOIGT, precision=capi.PREC_FP64, cutoff=25.0,
                             doppler_widths=0.0))
        else:
            reqs.append(dict(T=T, P=P, Q=Q(T), grid=g, profile=capi.VOIGT, precision=capi.PREC_MIXED, cutoff=25.0,
                             doppler_widths=0.0))
    return reqs


def submit(srv, r):
    return srv.submit(r["T"], r["P"], r["Q"], *r["grid"], profile=r["profile"], precision=r["precision"],
                      cutoff=r["cutoff"], doppler_widths=r["doppler_widths"])


def test_alone_equals_mix_and_oracle(world):
    ll, lines = world
    reqs = requests(ll)
    print(f"server: {len(reqs)} requests over {len(lines)} lines", flush=True)
    with Server(lines, 17.03, max_inflight=3, max_batch=8) as srv:
        ids = [submit(srv, r) for r in reqs]                    # all arrive together: co-batched, overlapping
        mix = []
        for rid in ids:
            assert srv.wait(rid, 120.0) == capi.REQ_DONE
            mix.append(srv.result(rid)[0])
        st = srv.stats()
        assert st.completed == len(reqs) and st.failed == 0
        assert st.batches < len(reqs)                           # some requests shared a launch
        assert st.p99_ms >= st.p50_ms > 0
    with Server(lines, 17.03, max_inflight=1, max_batch=1) as srv:
        for r, m in zip(reqs, mix):
            rid = submit(srv, r)
            assert srv.wait(rid, 120.0) == capi.REQ_DONE
            alone = srv.result(rid)[0]
            assert np.array_equal(alone, m)
    # spot checks against the oracle
    for r, m in list(zip(reqs, mix))[:5]:
        start, end, npts = r["grid"]
        freq, d = make_grid(start, end, npts)
        voigt = r["profile"] == capi.VOIGT
        c = po.consts(r["profile"], r["T"], r["Q"], d, 17.03, r["P"], 500.0 if voigt else 500.0,
                      cutoff=r["cutoff"] if voigt else 100.0, doppler_widths=r["doppler_widths"])
        sel = np.nonzero((ll.nu > start - 120) & (ll.nu < end + 120))[0]
        L = ll.subset(sel)
        ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, L.gamma_at(r["P"]) if voigt else None,
                               L.nexp if voigt else None, freq, npts, 0, 8)
        mask = ref >= 1e-6 * ref.max()
        err = float(np.max(np.abs(m[mask] - ref[mask]) / ref[mask]))
        tol = 1e-5 if r["precision"] == capi.PREC_MIXED else 1e-10
        print(f"  request profile={r['profile']} prec={r['precision']}: max rel err {err:.2e}", flush=True)
        assert err <= tol


def test_cancel_and_budget_rejection(world):
    ll, lines = world
    Q = po.partition(ll.state_e, ll.state_g, 1000.0)
    with Server(lines, 17.03, max_inflight=1, max_batch=1) as srv:
        ids = [srv.submit(1000.0, 1.0, Q, 100.0, 11000.0, 200_000) for _ in range(4)]
        srv.cancel(ids[-1])
        states = [srv.wait(rid, 300.0) for rid in ids]
        assert states[:3] == [capi.REQ_DONE] * 3
        assert states[3] in (capi.REQ_CANCELLED, capi.REQ_DONE)
    with Server(lines, 17.03, hbm_budget=1 << 20) as srv:       # 1 MiB: nothing fits
        rid = srv.submit(1000.0, 1.0, Q, 100.0, 11000.0, 200_000)
        assert srv.wait(rid, 60.0) == capi.REQ_FAILED
        with pytest.raises(capi.GxError):
            srv.result(rid)


def _occupy(srv, ll, profile=capi.VOIGT, grid=(100.0, 11000.0, 400_000)):
    """A larger request launched first, so the one slot is busy while the requests that follow
    queue (and batch); returns once it runs -- the scheduler otherwise may launch the first of
    the others alone, or batch them with it."""
    import time
    Q = po.partition(ll.state_e, ll.state_g, 1000.0)
    prec = capi.PREC_MIXED if profile == capi.VOIGT else capi.PREC_FP64
    rid = srv.submit(1000.0, 1.0, Q, *grid, profile=profile, precision=prec, cutoff=0.0)
    t0 = time.time()
    while srv.poll(rid) == capi.REQ_QUEUED:
        assert time.time() - t0 < 60.0
        time.sleep(1e-4)
    return rid


def test_doppler_batch_of_two_temperatures_near_the_top(world):
    """Per-line Doppler windows (k alpha + dfreq) widen with temperature: a batch of a cold
    and a hot request on one grid near the top of the list takes the hot one's line slice,
    so each equals its own run alone bit for bit, and the oracle within 1e-10."""
    ll, _ = world
    rng = np.random.default_rng(11)
    n = 40_000                                   # 40 lines per cm^-1 in [11000, 12000]
    nu = np.sort(rng.uniform(11000.0, 12000.0, n))
    elow = rng.uniform(0.0, 8000.0, n)
    aif = 10 ** rng.uniform(-6, 2, n)
    gns = rng.integers(1, 40, n).astype(np.int32)
    lines = ResidentLines(elow, nu, aif, gns)
    grid = (11500.0, 11950.0, 45_000)
    reqs = [dict(T=T, P=1.0, Q=po.partition(ll.state_e, ll.state_g, T), grid=grid, profile=capi.DOPPLER,
                 precision=capi.PREC_FP64, cutoff=0.0, doppler_widths=3.0) for T in (300.0, 3000.0)]
    try:
        with Server(lines, 17.03, max_inflight=1, max_batch=4) as srv:
            first = _occupy(srv, ll, capi.DOPPLER, (11000.0, 12000.0, 1_000_000))
            ids = [submit(srv, r) for r in reqs]
            mix = []
            for rid in ids:
                assert srv.wait(rid, 120.0) == capi.REQ_DONE
                mix.append(srv.result(rid)[0])
            assert srv.wait(first, 120.0) == capi.REQ_DONE
            st = srv.stats()
            print(f"doppler batch: {st.batches} batches for {st.completed} requests", flush=True)
            assert st.batches == 2 and st.shared_requests == 2     # both Doppler requests in one launch
        with Server(lines, 17.03, max_inflight=1, max_batch=1) as srv:
            for r, m in zip(reqs, mix):
                rid = submit(srv, r)
                assert srv.wait(rid, 120.0) == capi.REQ_DONE
                assert np.array_equal(srv.result(rid)[0], m)
    finally:
        lines.close()
    start, end, npts = grid
    freq, d = make_grid(start, end, npts)
    for r, m in zip(reqs, mix):
        c = po.consts(capi.DOPPLER, r["T"], r["Q"], d, 17.03, 1.0, 500.0, cutoff=100.0, doppler_widths=3.0)
        ref = po.cross_section(c, elow, nu, aif, gns, None, None, freq, npts, 0, 8)
        mask = ref >= 1e-6 * ref.max()
        err = float(np.max(np.abs(m[mask] - ref[mask]) / ref[mask]))
        print(f"  T={r['T']}: max rel err {err:.2e}", flush=True)
        assert err <= 1e-10
        assert np.all(m[ref > 1e-9 * ref.max()] > 0)   # edge bins reached by the hot lines included


def test_tight_budget_lets_idle_slots_release(world):
    """With a budget for one large request and one small one, alternating large and small
    requests all finish: idle slots give back their grown buffers before admission waits."""
    ll, lines = world
    Q = po.partition(ll.state_e, ll.state_g, 1200.0)
    large = (500.0, 9500.0, 900_000)
    small = (3000.0, 3010.0, 1000)
    sub = lambda srv, g: srv.submit(1200.0, 0.5, Q, *g, profile=capi.VOIGT, precision=capi.PREC_FP64, cutoff=25.0)
    with Server(lines, 17.03, max_inflight=1, max_batch=1) as srv:
        rid = sub(srv, large)
        assert srv.wait(rid, 120.0) == capi.REQ_DONE
        need = srv.stats().peak_bytes
    budget = need + need // 2
    print(f"budget {budget} B for a large request of {need} B", flush=True)
    with Server(lines, 17.03, max_inflight=2, max_batch=1, hbm_budget=budget) as srv:
        ids = [sub(srv, large if k % 2 == 0 else small) for k in range(6)]
        states = [srv.wait(rid, 120.0) for rid in ids]
        st = srv.stats()
        assert states == [capi.REQ_DONE] * 6, states
        assert st.failed == 0
        assert st.peak_bytes <= budget, (st.peak_bytes, budget)


def test_failed_and_cancelled_requests_are_released(world):
    ll, lines = world
    Q = po.partition(ll.state_e, ll.state_g, 1000.0)
    with Server(lines, 17.03, hbm_budget=1 << 20) as srv:        # 1 MiB: nothing fits
        rid = srv.submit(1000.0, 1.0, Q, 100.0, 11000.0, 200_000)
        assert srv.wait(rid, 60.0) == capi.REQ_FAILED
        with pytest.raises(capi.GxError) as ei:
            srv.result(rid)
        assert ei.value.status == capi.GX_DEVICE_ERROR
        with pytest.raises(capi.GxError) as ei:           # released: the id is gone
            srv.poll(rid)
        assert ei.value.status == capi.GX_NOT_FOUND


def _fp64_requests(ll):
    """FP64 requests on six different grids: Doppler (reference 100 cm^-1 window, and per-line
    windows) and FP64 Voigt, coarse grids whose blocks each take thousands of lines."""
    Q = lambda T: po.partition(ll.state_e, ll.state_g, T)
    specs = [(capi.DOPPLER, 300.0, 1.0, (1000.0, 1100.0, 1000), 0.0, 0.0),
             (capi.DOPPLER, 2500.0, 1.0, (4000.0, 4006.0, 3000), 0.0, 25.0),
             (capi.VOIGT, 900.0, 0.3, (2000.0, 2050.0, 500), 10.0, 0.0),
             (capi.DOPPLER, 1200.0, 1.0, (7000.0, 7400.0, 40000), 0.0, 25.0),
             (capi.VOIGT, 2200.0, 3.0, (5000.0, 5002.0, 2000), 25.0, 0.0),
             (capi.DOPPLER, 700.0, 1.0, (9000.0, 9500.0, 500), 0.0, 0.0)]
    return [dict(T=T, P=P, Q=Q(T), grid=g, profile=prof, precision=capi.PREC_FP64, cutoff=cut, doppler_widths=dw)
            for prof, T, P, g, cut, dw in specs]


def _oracle_check(ll, r, m, tol):
    start, end, npts = r["grid"]
    freq, d = make_grid(start, end, npts)
    voigt = r["profile"] == capi.VOIGT
    cut = r["cutoff"] if voigt else 100.0
    c = po.consts(r["profile"], r["T"], r["Q"], d, 17.03, r["P"], 500.0, cutoff=cut,
                  doppler_widths=r["doppler_widths"])
    sel = np.nonzero((ll.nu > start - cut - 1.0) & (ll.nu < end + cut + 1.0))[0]
    L = ll.subset(sel)
    ref = po.cross_section(c, L.elow, L.nu, L.aif, L.gns, L.gamma_at(r["P"]) if voigt else None,
                           L.nexp if voigt else None, freq, npts, 0, 8)
    mask = ref >= 1e-6 * ref.max()
    err = float(np.max(np.abs(m[mask] - ref[mask]) / ref[mask]))
    assert err <= tol, (r["profile"], r["grid"], err)
    return err


def test_fp64_requests_of_different_grids_share_a_launch(world):
    ll, lines = world
    reqs = _fp64_requests(ll)
    with Server(lines, 17.03, max_inflight=1, max_batch=8) as srv:
        first = _occupy(srv, ll)
        ids = [submit(srv, r) for r in reqs]
        mix = []
        for rid in ids:
            assert srv.wait(rid, 120.0) == capi.REQ_DONE
            out, t = srv.result(rid)
            mix.append((out, t.evals))
        assert srv.wait(first, 120.0) == capi.REQ_DONE
        st = srv.stats()
    print(f"fp64 grids: {st.batches} batches for {st.completed} requests", flush=True)
    assert st.batches == 2 and st.shared_requests == len(reqs)     # the six in one launch
    with Server(lines, 17.03, max_inflight=1, max_batch=1) as srv:
        for r, (m, ev) in zip(reqs, mix):
            rid = submit(srv, r)
            assert srv.wait(rid, 120.0) == capi.REQ_DONE
            alone, t = srv.result(rid)
            assert np.array_equal(alone, m), r["grid"]
            assert t.evals == ev                        # each request counts its own windows
    for r, (m, _) in zip(reqs, mix):
        err = _oracle_check(ll, r, m, 1e-10)
        print(f"  profile={r['profile']} grid={r['grid']}: max rel err {err:.2e}", flush=True)


def test_poisson_arrivals_alone_equals_mix(world):
    """Seeded Poisson arrivals while batches run on three slots: every request's grid equals
    its grid alone, bit for bit, and FP64 requests share launches."""
    import time
    ll, lines = world
    rng = np.random.default_rng(17)
    base = requests(ll) + _fp64_requests(ll)
    order = rng.permutation(len(base))
    gaps = rng.exponential(0.004, len(base))
    with Server(lines, 17.03, max_inflight=3, max_batch=8) as srv:
        ids = []
        for k, gap in zip(order, gaps):
            time.sleep(gap)
            ids.append((k, submit(srv, base[k])))
        mix = {}
        for k, rid in ids:
            assert srv.wait(rid, 120.0) == capi.REQ_DONE
            mix[k] = srv.result(rid)[0]
        st = srv.stats()
    print(f"poisson: {st.batches} batches for {st.completed} requests, p50 {st.p50_ms:.1f} ms "
          f"p99 {st.p99_ms:.1f} ms", flush=True)
    assert st.completed == len(base) and st.failed == 0
    with Server(lines, 17.03, max_inflight=1, max_batch=1) as srv:
        for k, r in enumerate(base):
            rid = submit(srv, r)
            assert srv.wait(rid, 120.0) == capi.REQ_DONE
            assert np.array_equal(srv.result(rid)[0], mix[k]), (k, r["grid"], r["profile"], r["precision"])
