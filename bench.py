"""Benchmark of the cross-section hot path on one MI355X.

Default workload (table_48): a step is one 48-point (T, P) Voigt opacity table
over 1e8 resident synthetic ExoMol-format lines (8 T log-spaced 300-3000 K x 6 P
log-spaced 1e-4-100 atm, grid 0-12000 cm^-1 at 0.01 cm^-1, 25 cm^-1 cutoff),
computed as one hipGraph replay of the mixed-precision form (FP32 profiles with
compensated sums, FP64 intensities/offsets).  The same table also runs once
through the plain FP64 per-point path for the speed ratio and the precision
report (max/RMS relative error of every mixed grid against its FP64 grid, bins
>= 1e-6 of the maximum; the FP64 path itself is held to 1e-10 of the oracle by
the parity tests).  Other workloads: voigt_1e8 / voigt_1e7 (one cross-section
per step), serving_3e8 (request server latency), doppler_1e5.

Prints one JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_EVAL = 20          # region-I W4 closed form + offset + scaling + compensated add
# HBM bytes per accumulation launch, from rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
# same command scaled by a plain-copy calibration (profiles/k_accum_voigt_mixed.table_48.hbm.md)
# From the rocprofv3 summaries under profiles/, per workload and block size (64 R bins; the
# table's autotuner picks R): HBM bytes per launch of the region-I kernel (2 FETCH_SIZE +
# WRITE_SIZE, copy-calibrated), and the step's VALU issue use, sum over every kernel of one
# 48-point table of SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x clock x step time).
PROFILED_TRAFFIC = {("table_48", 1): 1.725e11, "ingest_1e8": None}
PROFILED_STEP_VALU = {("table_48", 1): {"value": 0.874, "step_ms": 381.7, "valu_insts": 1.970e11, "clock_ghz": 2.31}}
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12
# ExoMol reader: algorithmic bytes per .trans record = the text read (37 B of
# "%12d %12d %10.4E\n") + the line written (E'', nu, A, gamma, n f64 + g i32 = 44 B)
INGEST_BYTES_PER_RECORD = 37 + 44


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(ll, wl, Q, _gam=None, target_s=15.0):
    """CPU baseline on a contiguous bin sub-range of the same grid and line list, at P = 1 atm
    (where the reference's own window rule, 25 P cm^-1, is the workload's 25 cm^-1 cutoff), all
    host threads up to 16.  The sample is sized with the oracle (the C restatement of the
    reference's CPU manager); where the reference's own OpenMPManager has been built
    (oracle/_ref/ref_harness) that binary is timed on the same sample (kind "reference"),
    otherwise the oracle (kind "port")."""
    import subprocess
    import tempfile

    from oracle import pyoracle as po
    from gpu_cross_amd.manager import make_grid
    from tests.golden.caseio import write_case

    start, end, M = wl["grid"]
    freq, d = make_grid(start, end, M)
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        pass
    affinity = cores
    # the host's CPU share for this GPU: OMP_NUM_THREADS when the environment sets it, else every
    # core this process may run on
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = max(1, min(cores, int(os.environ["OMP_NUM_THREADS"])))
    P = 1.0
    gam = ll.gamma_at(P)
    c = po.consts(po.VOIGT, wl["T"], Q, d, ll.mass, P, 500.0, cutoff=wl["cutoff"])
    mid = M // 2
    bins = 200
    while True:
        j0 = mid - bins // 2
        lo_nu, hi_nu = freq[j0] - wl["cutoff"] - 2 * d, freq[j0 + bins - 1] + wl["cutoff"] + 2 * d
        a, b = np.searchsorted(ll.nu, [lo_nu, hi_nu])
        sl = slice(a, b)
        args = (ll.elow[sl], ll.nu[sl], ll.aif[sl], ll.gns[sl], np.ascontiguousarray(gam[sl]), ll.nexp[sl])
        n_ev = po.count_evals(c, ll.nu[sl], freq, bins, j0)
        t0 = time.perf_counter()
        po.cross_section(c, *args, freq, bins, j0, cores)
        dt = time.perf_counter() - t0
        if dt > target_s / 4 or bins >= M // 4:
            break
        bins = int(min(M // 4, bins * max(2.0, min(8.0, target_s / max(dt, 1e-3)))))
    port_rate = n_ev / dt
    sample = (f"bins [{j0},{j0 + bins}) of the same grid ({bins} bins) with every line touching them "
              f"({b - a} lines, {n_ev:.3e} evals), P = 1 atm")
    res = {"value": port_rate, "unit": "line-bin pairs/s", "cores": cores, "kind": "port", "sample": sample,
           "port_seconds": dt, "host_cpu": cpu_model(), "affinity_cores": affinity}
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if os.path.exists(harness) and bins * 8 >= 64 * 1024:
        with tempfile.TemporaryDirectory() as tmp:
            case = dict(profile=po.VOIGT, nthreads=cores, npoints=bins, max_points=bins + 1, capacity=b - a,
                        start=float(freq[j0]), end=float(freq[j0]) + bins * d, T=wl["T"], Q=Q, mass=ll.mass, P=P,
                        ref_temp=500.0)
            lines = dict(elow=ll.elow[sl], nu=ll.nu[sl], aif=ll.aif[sl], gns=ll.gns[sl], gamma=ll.gamma_ref[sl],
                         nexp=ll.nexp[sl])
            write_case(os.path.join(tmp, "case.bin"), case, lines)
            r = subprocess.run([harness, os.path.join(tmp, "case.bin"), os.path.join(tmp, "out.bin")],
                               capture_output=True, text=True, timeout=600)
            secs = None
            for line in r.stderr.splitlines():
                if line.startswith("manager_seconds"):
                    secs = float(line.split()[1])
            if r.returncode == 0 and secs:
                res.update(value=n_ev / secs, kind="reference", reference_seconds=secs,
                           sample=sample + "; the reference's own OpenMPManager (oracle/_ref, built from its "
                                           "sources), timed around its BaseManager calls")
    return res


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def calibrate_copy(mib: int) -> None:
    """A plain device copy of known size, for scaling FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(5):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": n, "copies": 5}), flush=True)


def bench_table(args, wl):
    """48-point (T,P) opacity table over 1e8 resident lines: one graph replay per step."""
    from gpu_cross_amd import capi
    from gpu_cross_amd.synth import workload
    from gpu_cross_amd.ingest import partition
    from gpu_cross_amd.table import ResidentLines, Table

    t0 = time.time()
    ll = workload("table_48").sorted_by_nu()
    log(f"generated {ll.n} lines in {time.time() - t0:.1f} s")
    start, end, M = wl["grid"]
    pts = [(T, P, partition(ll.state_e, ll.state_g, T)) for T in wl["temperatures"] for P in wl["pressures"]]
    t1 = time.time()
    lines = ResidentLines.from_linelist(ll)
    log(f"uploaded {len(lines)} lines in {time.time() - t1:.1f} s")
    t1 = time.time()
    tab = Table(lines, start, end, M, pts, ll.mass, 500.0, wl["cutoff"], use_graph=True, autotune=True)
    log(f"planned + autotuned table in {time.time() - t1:.1f} s, shape (bins/lane, segments) = {tab.shape()}")
    for _ in range(max(args.warmup, 1)):          # first run: eager timed run + graph capture
        tab.run()
        t = tab.wait()
    eager_int, eager_acc = t.intensity_ms, t.accumulate_ms
    r1_ms, core_ms, eager_far_ms = t.region1_ms, t.core_ms, t.far_ms
    log(f"eager run: prep {eager_int:.1f} ms, accumulate {eager_acc:.1f} ms "
        f"(far field {eager_far_ms:.1f} ms for {t.far_evals / max(t.evals, 1):.1%} of the pairs, region I "
        f"{r1_ms:.1f} ms, W4 core pass {core_ms:.1f} ms)")
    walls = []
    for _ in range(args.steps):
        t1 = time.perf_counter()
        tab.run()
        t = tab.wait()
        walls.append(time.perf_counter() - t1)
        log(f"table step: {walls[-1] * 1e3:.1f} ms wall, graph {t.total_ms:.1f} ms, evals {t.evals:.4e}")
    evals = t.evals
    core_evals = t.core_evals
    far_evals, far_ms = t.far_evals, eager_far_ms
    r1_pairs = evals - core_evals - far_evals     # region-I pairs the region-I kernel summed
    interp = t.interp_pairs                       # of them: interpolated from 16 nodes per 64 bins
    r1_evals = r1_pairs - interp + interp // 4    # evaluations it made
    log(f"region-I kernel: {r1_pairs:.4e} pairs, {interp:.4e} of them interpolated, {r1_evals:.4e} evaluations")
    wall = float(np.sum(walls))
    value = evals * args.steps / wall
    ms_step = wall / args.steps * 1e3
    res = {
        "metric": "Voigt line-bin evals/s + ms per (T,P) xsec at 1e-5 rel err; serving p50/p99",
        "value": value, "unit": "line-bin pairs/s",
        "value_definition": "line-bin pairs (line i, bin j) with ib_i <= j < ie_i from the FP64 windows, "
                            "summed over the 48 grids, per second of table wall time; far_field.share of them "
                            "are summed from multipole moments, roofline.interpolated_pairs interpolated from "
                            "16 node evaluations per line and 64-bin block, the rest evaluated one by one "
                            "(direct_evals_per_s: evaluations made, node evaluations included)",
        "direct_evals_per_s": (r1_evals + core_evals) * args.steps / wall,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"table_48: {ll.n} lines, 8 T log-spaced 300-3000 K x 6 P log-spaced 1e-4-100 atm, "
                               f"{M} bins {start}-{end} cm^-1, cutoff {wl['cutoff']} cm^-1, one hipGraph replay "
                               f"of the whole table per step", "lines": ll.n, "bins": M, "points": len(pts),
                   "evals_per_step": evals},
        "ms_per_xsec": ms_step / len(pts),
        "roofline": {"bound": "valu", "achieved": FLOP_PER_EVAL * r1_evals / (r1_ms * 1e-3) / 1e12,
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": FLOP_PER_EVAL * r1_evals / (r1_ms * 1e-3) / PEAK_FP32,
                     "traffic": PROFILED_TRAFFIC.get(("table_48", tab.shape()[0])),
                     "traffic_source": "rocprofv3 2*FETCH_SIZE+WRITE_SIZE per launch (copy-calibrated), profiles/",
                     "kernel": "k_accum_voigt_mixed", "kernel_ms": r1_ms, "evals": r1_evals,
                     "pairs": r1_pairs, "interpolated_pairs": interp, "flop_per_eval": FLOP_PER_EVAL},
        "core_pass": {"kernel": "k_core_mixed (W4 regions II-IV; region I of near ranges)", "ms": core_ms, "evals": core_evals,
                      "evals_per_s": core_evals / (core_ms * 1e-3) if core_ms > 0 else None},
        "far_field": {"kernels": "k_far_kind, k_far_geom, k_far_leaf, k_far_up, k_far_plan, k_edge_plan, k_far_eval, "
                                 "k_near_eval, k_far_edge_parts, k_far_edges (FP32 leaf sums and records, FP64 tree shifts and series; near leaves summed "
                                 "per bin)",
                      "ms": far_ms, "evals": far_evals, "share": far_evals / max(evals, 1)},
        "intensity_kernels_ms": eager_int,
        "step_valu_issue": dict(PROFILED_STEP_VALU.get(("table_48", tab.shape()[0]), {}),
                                source="rocprofv3 --pmc SQ_INSTS_VALU GRBM_GUI_ACTIVE over one 48-point table "
                                       "(profiles/), at the block size this run picked")
        if ("table_48", tab.shape()[0]) in PROFILED_STEP_VALU else None,
    }
    mixed_grids = [tab.read(p) for p in range(len(pts))]
    tab.close()
    if not args.no_fp64:
        # the same mixed table with the far field off: every pair evaluated one by one (256-bin
        # blocks, the launch shape that is fastest without the far field)
        os.environ["GX_FAR"], os.environ["GX_MIXED_R"] = "0", "4"
        try:
            tnf = Table(lines, start, end, M, pts, ll.mass, 500.0, wl["cutoff"], use_graph=True, autotune=False)
            tnf.run()
            tnf.wait()
            ws = []
            for _ in range(2):
                t1 = time.perf_counter()
                tnf.run()
                tnf.wait()
                ws.append(time.perf_counter() - t1)
            tnf.close()
        finally:
            del os.environ["GX_FAR"], os.environ["GX_MIXED_R"]
        nofar_ms = min(ws) * 1e3
        log(f"mixed table without the far field: {nofar_ms:.1f} ms")
        res["no_far_field"] = {"ms_per_table": nofar_ms, "pairs_per_s": evals / (nofar_ms * 1e-3),
                               "form": "GX_FAR=0, 256-bin blocks: every pair evaluated (region I + core pass)"}
        f64 = Table(lines, start, end, M, pts, ll.mass, 500.0, wl["cutoff"], precision=capi.PREC_FP64,
                    use_graph=False, autotune=False)
        t1 = time.perf_counter()
        f64.run()
        tf = f64.wait()
        fwall = time.perf_counter() - t1
        log(f"fp64 per-point table: {fwall * 1e3:.1f} ms")
        worst, rms_acc, nb = 0.0, 0.0, 0
        for p in range(len(pts)):
            ref = f64.read(p)
            mask = ref >= 1e-6 * ref.max()
            rel = np.abs(mixed_grids[p][mask] - ref[mask]) / ref[mask]
            worst = max(worst, float(rel.max()))
            rms_acc += float(np.sum(rel ** 2))
            nb += int(mask.sum())
        f64.close()
        res["fp64_path"] = {"evals_per_s": tf.evals / fwall, "ms_per_table": fwall * 1e3,
                            "frac_fp64_peak": FLOP_PER_EVAL * tf.evals / fwall / PEAK_FP64,
                            "form": "48 separate per-point launches (intensity, ranges, FP64 accumulation)"}
        res["speedup_vs_fp64"] = value / res["fp64_path"]["evals_per_s"]
        res["speedup_vs_fp64_direct"] = fwall * 1e3 / nofar_ms     # both evaluate every pair
        res["precision"] = {"max_rel_err": worst, "rms_rel_err": (rms_acc / max(nb, 1)) ** 0.5, "bins_checked": nb,
                            "against": "FP64 GPU path (oracle-parity 1e-10), all 48 grids"}
    lines.close()
    if not args.no_cpu:
        log("cpu baseline ...")
        T, _, Q = pts[len(pts) // 2]
        res["cpu_baseline"] = cpu_baseline(ll, dict(wl, T=T), Q, None)
        res["cpu_baseline"]["sample"] += f", T={T:.0f} K"
    print(json.dumps(res), flush=True)


def serving_requests(rng, n, ll, Qfun):
    """Requests of the serving workload: T ~ U[200,3000] K, P log-U[1e-5,100] atm, Doppler or
    Voigt 50/50, window width log-U[10,12000] cm^-1 placed uniformly in [0,12000], resolution
    log-U[0.001,0.1] cm^-1, at most 1e7 bins."""
    from gpu_cross_amd import capi
    out = []
    for _ in range(n):
        T = float(rng.uniform(200.0, 3000.0))
        P = float(10 ** rng.uniform(-5.0, 2.0))
        width = float(10 ** rng.uniform(1.0, np.log10(12000.0)))
        start = float(rng.uniform(0.0, 12000.0 - width))
        res = float(10 ** rng.uniform(-3.0, -1.0))
        npts = int(min(max(round(width / res), 1), 10_000_000))
        voigt = rng.uniform() < 0.5
        out.append(dict(T=T, P=P, Q=Qfun(T), start=start, end=start + width, npoints=npts,
                        profile=capi.VOIGT if voigt else capi.DOPPLER,
                        precision=capi.PREC_MIXED if voigt else capi.PREC_FP64,
                        cutoff=25.0 if voigt else 0.0, doppler_widths=0.0 if voigt else 25.0))
    return out


def bench_serving(args, wl):
    """Continuous batching over a resident 3e8-line list with seeded Poisson arrivals."""
    from gpu_cross_amd import capi
    from gpu_cross_amd.server import Server
    from gpu_cross_amd.synth import workload
    from gpu_cross_amd.ingest import partition
    from gpu_cross_amd.table import ResidentLines

    t0 = time.time()
    ll = workload("serving_3e8")
    log(f"generated {ll.n} lines in {time.time() - t0:.1f} s")
    t0 = time.time()
    lines = ResidentLines.from_linelist(ll)           # sorted on the device
    log(f"uploaded + sorted on device in {time.time() - t0:.1f} s")
    qcache = {}
    Qfun = lambda T: qcache.setdefault(T, partition(ll.state_e, ll.state_g, T))
    submit = lambda srv, r: srv.submit(r["T"], r["P"], r["Q"], r["start"], r["end"], r["npoints"],
                                       profile=r["profile"], precision=r["precision"], cutoff=r["cutoff"],
                                       doppler_widths=r["doppler_widths"])
    rng = np.random.default_rng(55)
    # capacity: the same request set as the open-loop levels, submitted as one closed-loop burst
    # (request costs span orders of magnitude, so a smaller or different set misjudges it)
    reqs = serving_requests(rng, args.requests, ll, Qfun)
    with Server(lines, ll.mass, max_inflight=4, max_batch=16) as srv:
        t1 = time.perf_counter()
        ids = [submit(srv, r) for r in reqs]
        for rid in ids:
            srv.wait(rid)
            srv.result(rid)
        cap_rps = len(reqs) / (time.perf_counter() - t1)
    log(f"closed-loop capacity ~{cap_rps:.2f} requests/s")
    levels = []
    keep = []
    # the 90 % level twice, with its own arrival seeds, so its p99 is not one draw
    for load, seed in ((0.5, 1), (0.9, 2), (0.9, 3)):
        gaps = np.random.default_rng(seed).exponential(1.0 / (load * cap_rps), len(reqs))
        with Server(lines, ll.mass, max_inflight=4, max_batch=16) as srv:
            t1 = time.perf_counter()
            ids = []
            for r, gap in zip(reqs, gaps):
                time.sleep(gap)
                ids.append(submit(srv, r))
            results = []
            for rid in ids:
                srv.wait(rid)
                results.append(srv.result(rid)[0])
            wall = time.perf_counter() - t1
            st = srv.stats()
        levels.append({"offered_load": load, "arrival_seed": seed, "arrival_rate_rps": load * cap_rps,
                       "requests": len(reqs),
                       "p50_ms": st.p50_ms, "p99_ms": st.p99_ms, "mean_ms": st.mean_ms, "max_ms": st.max_ms,
                       "throughput_rps": len(reqs) / wall, "evals_per_s": st.evals / max(st.busy_s, 1e-9),
                       "batches": st.batches, "batched_requests": st.batched_requests,
                       "shared_launch_share": st.shared_requests / max(st.batched_requests, 1),
                       "failed": st.failed, "peak_hbm_bytes": st.peak_bytes})
        log(f"load {load}: p50 {st.p50_ms:.1f} ms p99 {st.p99_ms:.1f} ms, {len(reqs) / wall:.2f} req/s, "
            f"{st.batches} batches, {st.shared_requests} requests in shared launches")
        keep = list(zip(reqs, results))
    # alone == in the mix, on a sample of the last level
    same = 0
    sample = keep[:: max(1, len(keep) // 6)][:6]
    with Server(lines, ll.mass, max_inflight=1, max_batch=1) as srv:
        for r, mixed in sample:
            rid = submit(srv, r)
            srv.wait(rid)
            same += int(np.array_equal(srv.result(rid)[0], mixed))
    lines.close()
    top = max(levels[1:], key=lambda lv: lv["p99_ms"])
    res = {"metric": "serving p50/p99 latency (ms) at 90% offered load; throughput", "value": top["p99_ms"],
           "unit": "ms (p99)", "n_gpus": 1, "steps": args.requests, "warmup": 0,
           "ms_per_step": top["mean_ms"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"serving_3e8: {ll.n} resident lines, Poisson arrivals, Doppler/Voigt 50/50, "
                                  "T U[200,3000] K, P log-U[1e-5,100] atm, width log-U[10,12000] cm^-1, "
                                  "resolution log-U[1e-3,0.1] cm^-1 (<= 1e7 bins), Voigt cutoff 25 cm^-1 mixed, "
                                  "Doppler 25 HWHM FP64", "closed_loop_capacity_rps": cap_rps},
           "value_definition": "the larger p99 of two 90 % offered-load runs (arrival seeds 2 and 3) over the "
                               "same request set; offered rate = 0.9 x the closed-loop capacity of that set",
           "levels": levels, "alone_equals_mix": f"{same}/{len(sample)} sampled requests bit-identical"}
    print(json.dumps(res), flush=True)


def cpu_baseline_doppler(ll, T, Q, freq, d):
    """The CPU baseline of the Doppler workload: the oracle (C restatement of the reference's
    OpenMPManager) on the whole cross-section with the reference's 100 cm^-1 window, up to 16
    threads; its grid is also the checker of the GPU result."""
    from oracle import pyoracle as po

    t2 = time.perf_counter()
    c = po.consts(po.DOPPLER, T, Q, d, ll.mass, 1.0, 296.0)
    cores = min(len(os.sched_getaffinity(0)), 16)
    ref = po.cross_section(c, ll.elow, ll.nu, ll.aif, ll.gns, None, None, freq, freq.size, 0, cores)
    return ref, time.perf_counter() - t2, cores


def bench_doppler(args, wl):
    """Doppler cross-section of 1e5 lines on a 1e6-bin grid (25-HWHM per-line windows), FP64."""
    from gpu_cross_amd import capi
    from gpu_cross_amd.ingest import partition
    from gpu_cross_amd.manager import Mi355xManager, make_grid
    from gpu_cross_amd.synth import workload

    ll = workload("doppler_1e5").sorted_by_nu()
    T = wl["T"]
    Q = partition(ll.state_e, ll.state_g, T)
    start, end, M = wl["grid"]
    freq, d = make_grid(start, end, M)
    m = Mi355xManager(capi.DOPPLER, doppler_widths=wl["doppler_widths"])
    m.InitializeVectors(M)
    m.TransferFreq(freq, np.zeros(M), M)
    m.InitializeConstants(0.0, T, Q, d, ll.mass, 1.0, 296.0)
    m.TransferVectors(ll.n, ll.elow, ll.nu, ll.aif, ll.gns)
    for _ in range(max(args.warmup, 1)):
        m.ExecuteCrossSection(M, ll.n, 0)
    m.synchronize()
    m.TransferFreq(freq, np.zeros(M), M)
    t1 = time.perf_counter()
    for _ in range(args.steps):
        m.ExecuteCrossSection(M, ll.n, 0)
    m.synchronize()
    wall = time.perf_counter() - t1
    t = m.last_timing()
    out = np.zeros(M)
    m.TransferResults(None, out, M)
    m.close()
    out /= args.steps
    ref, cpu_s, cores = cpu_baseline_doppler(ll, T, Q, freq, d)
    mask = ref >= 1e-6 * ref.max()
    rel = np.abs(out[mask] - ref[mask]) / ref[mask]
    res = {"metric": "Doppler cross-sections/s (FP64)", "value": args.steps / wall, "unit": "xsec/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"doppler_1e5: {ll.n} lines, {M} bins {start}-{end} cm^-1, T={T} K, windows "
                                  f"{wl['doppler_widths']} HWHM + 1 bin (the reference's 100 cm^-1 window gives "
                                  "the same grid)", "evals_per_step": t.evals},
           "precision": {"max_rel_err_vs_oracle_100cm": float(rel.max()), "bins_checked": int(mask.sum())},
           "cpu_baseline": {"value": 1.0 / cpu_s, "unit": "xsec/s", "cores": cores, "kind": "port",
                            "sample": "the whole cross-section with the reference's 100 cm^-1 window"}}
    print(json.dumps(res), flush=True)


def cpu_baseline_ingest(arrays, broad, wl, trans, nrec, target_s=10.0):
    """The CPU baseline of the ExoMol reader: the oracle's restatement of the reference's
    read loop (one std::istream over the text, ExomolStateReader.cpp:145, one thread as the
    reference) on the file's first records, about target_s seconds of work."""
    from oracle import pyoracle as po

    start, end, _ = wl["grid"]
    lo, hi = start - wl["cutoff"], end + wl["cutoff"]
    st = dict(energy=arrays["energy"], gtot=arrays["gtot"], J=arrays["J"])
    probe = 200_000
    t0 = time.perf_counter()
    po.read_trans(st, broad, wl["pressure"], lo, hi, [bytes(trans[:37 * probe])])
    per = (time.perf_counter() - t0) / probe
    n = int(min(nrec, max(probe, target_s / per)))
    t0 = time.perf_counter()
    _, rec = po.read_trans(st, broad, wl["pressure"], lo, hi, [bytes(trans[:37 * n])])
    dt = time.perf_counter() - t0
    return {"value": rec / dt, "unit": "records/s", "cores": 1, "kind": "port",
            "sample": f"the first {rec} records of the same .trans text ({dt:.1f} s), the oracle's istream "
                      "read loop with the same states, broadener and range"}


def bench_ingest(args, wl):
    """ExoMol reader: 1e8 .trans records (3.7 GB of text) over 1e6 states with a broadener,
    read for a 0-12000 cm^-1 grid with a 25 cm^-1 cutoff; a step reads the whole file."""
    from gpu_cross_amd import ingest
    from gpu_cross_amd.synth import make_exomol

    t0 = time.time()
    st_text, trans, broad, arrays = make_exomol(wl["seed"], wl["n_states"], wl["e_max"], wl["n_lines"],
                                                wl["de_max"])
    log(f"generated {wl['n_states']} states and {wl['n_lines']} records ({trans.size / 1e9:.2f} GB) "
        f"in {time.time() - t0:.1f} s")
    t0 = time.perf_counter()
    states = ingest.parse_states(st_text)
    states_s = time.perf_counter() - t0
    start, end, _ = wl["grid"]
    walls, kms = [], []
    with ingest.ExomolReader(states, pressure=wl["pressure"], broadener=broad) as rd:
        rd.set_range(start - wl["cutoff"], end + wl["cutoff"])
        for _ in range(max(args.warmup, 1)):
            rd.clear()
            rd.read_file(trans)
        for _ in range(args.steps):
            rd.clear()
            t1 = time.perf_counter()
            ms, nbytes, nrec = rd.read_file(trans)
            walls.append(time.perf_counter() - t1)
            kms.append(ms / 1e3)
            log(f"read step: {walls[-1] * 1e3:.1f} ms wall (text from host memory), device {ms:.1f} ms, "
                f"{nrec} records")
        nkept, nread = rd.count()
    dev_s = float(np.mean(kms))
    wall_s = float(np.mean(walls))
    achieved = INGEST_BYTES_PER_RECORD * nread / dev_s / 1e9
    res = {"metric": "ExoMol .trans records read per second (GPU reader, text in HBM)", "value": nread / dev_s,
           "unit": "records/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": dev_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "u8", "data": "synthetic",
           "config": {"workload": f"ingest_1e8: {wl['n_lines']} records ({trans.size} bytes of .trans text), "
                                  f"{wl['n_states']} states, broadener, range {start - wl['cutoff']}-"
                                  f"{end + wl['cutoff']} cm^-1", "records": nread, "lines_kept": nkept},
           "end_to_end_records_per_s": nread / wall_s,
           "states_parse_s": states_s,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                        "frac": achieved / (PEAK_HBM / 1e9), "traffic": PROFILED_TRAFFIC.get("ingest_1e8"),
                        "kernel": "ExoMol reader passes (k_tok_count, k_tok_emit, k_records, k_keep_count, "
                                  "k_keep_emit)", "bytes_per_record": INGEST_BYTES_PER_RECORD}}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline_ingest(arrays, broad, wl, trans, nread)
    print(json.dumps(res), flush=True)


def bench_hitran(args, wl):
    """HITRAN .par reader: 1e7 160-character records (1.6 GB of text), read for a 0-12000
    cm^-1 grid with a 25 cm^-1 cutoff at 1 atm; a step reads the whole text (host threads,
    the text in host memory)."""
    from gpu_cross_amd import ingest

    rng = np.random.default_rng(wl["seed"])
    n = wl["n_lines"]
    t0 = time.time()
    nu = np.sort(rng.uniform(0.0, 12000.0, n))
    aif = 10 ** rng.uniform(-3, 2, n)
    aif[::50] = 0.0                                      # some records carry only the intensity
    text = ingest.format_hitran(nu, 10 ** rng.uniform(-30, -19, n), aif, rng.uniform(0.01, 0.12, n),
                                rng.uniform(0.05, 0.5, n), rng.uniform(0, 8000, n), rng.uniform(0.3, 0.8, n),
                                rng.integers(1, 90, n).astype(float), rng.integers(1, 90, n).astype(float),
                                mol_id=2, iso=1)
    log(f"generated {n} records ({text.size / 1e9:.2f} GB) in {time.time() - t0:.1f} s")
    start, end, _ = wl["grid"]
    lo, hi = start - wl["cutoff"], end + wl["cutoff"]
    walls = []
    for i in range(max(args.warmup, 1) + args.steps):
        t1 = time.perf_counter()
        lines, mol, iso = ingest.read_hitran(text, pressure=1.0, mixture=0.8, ref_partition=286.09, lo=lo, hi=hi)
        if i >= max(args.warmup, 1):
            walls.append(time.perf_counter() - t1)
            log(f"read step: {walls[-1] * 1e3:.1f} ms, {lines.records} records, {lines.n} kept")
    wall = float(np.mean(walls))
    threads = os.cpu_count()
    res = {"metric": "HITRAN .par records read per second (host reader, text in host memory)",
           "value": lines.records / wall, "unit": "records/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u8", "data": "synthetic",
           "config": {"workload": f"hitran_1e7: {n} records ({text.size} bytes of .par text), range {lo}-{hi} cm^-1",
                      "records": lines.records, "lines_kept": lines.n, "threads": threads},
           "bytes_per_s": text.size / wall}
    if not args.no_cpu:
        from oracle import pyoracle as po

        probe = min(n, 2_000_000)
        t0 = time.perf_counter()
        _, rec, _, _ = po.read_hitran(bytes(text[:161 * probe]), 1.0, 0.8, 286.09, lo, hi)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": rec / dt, "unit": "records/s", "cores": 1, "kind": "port",
                               "sample": f"the first {rec} records of the same text ({dt:.1f} s), the oracle's "
                                         "fgets/atof read loop, one thread as the reference"}
    print(json.dumps(res), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="table_48",
                    choices=["table_48", "voigt_1e8", "voigt_1e7", "serving_3e8", "doppler_1e5", "ingest_1e8",
                             "hitran_1e7"])
    ap.add_argument("--requests", type=int, default=120, help="serving: requests per load level")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB 5 times (HBM counter calibration)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)

    from gpu_cross_amd import capi
    from gpu_cross_amd.manager import Mi355xManager, make_grid
    from gpu_cross_amd.ingest import partition
    from gpu_cross_amd.synth import WORKLOADS, workload

    if args.workload == "hitran_1e7":
        return bench_hitran(args, dict(seed=11, n_lines=10_000_000, grid=(0.0, 12000.0, 1_200_000), cutoff=25.0))
    wl = WORKLOADS[args.workload]
    if args.workload == "table_48":
        return bench_table(args, wl)
    if args.workload == "serving_3e8":
        return bench_serving(args, wl)
    if args.workload == "doppler_1e5":
        return bench_doppler(args, wl)
    if args.workload == "ingest_1e8":
        return bench_ingest(args, wl)
    t0 = time.time()
    ll = workload(args.workload).sorted_by_nu()
    log(f"generated {ll.n} lines in {time.time() - t0:.1f} s")
    T, P = wl["T"], wl["P"]
    Q = partition(ll.state_e, ll.state_g, T)
    start, end, M = wl["grid"]
    freq, d = make_grid(start, end, M)
    gam = ll.gamma_at(P)

    # the workload is one launch over the whole list: GetNtrans() (2^26 by default, the chunk
    # the reference's driver would cut the list into) raised to hold it
    os.environ["GX_MAX_NTRANS"] = str(max(1 << 26, 1 << int(np.ceil(np.log2(ll.n)))))

    def run(precision, steps, warmup):
        m = Mi355xManager(capi.VOIGT, precision=precision, cutoff=wl["cutoff"])
        m.InitializeVectors(M)
        m.TransferFreq(freq, np.zeros(M), M)
        m.InitializeConstants(0.0, T, Q, d, ll.mass, P, 500.0)
        t1 = time.time()
        m.TransferVectors(ll.n, ll.elow, ll.nu, ll.aif, ll.gns, gam, ll.nexp)
        log(f"uploaded lines in {time.time() - t1:.1f} s (precision={precision})")
        for _ in range(warmup):
            m.ExecuteCrossSection(M, ll.n, 0)
        m.synchronize()
        acc_ms, int_ms, r1_ms, core_ms, far_ms = [], [], [], [], []
        m.TransferFreq(freq, np.zeros(M), M)
        t1 = time.perf_counter()
        for _ in range(steps):
            m.ExecuteCrossSection(M, ll.n, 0)
            t = m.last_timing()                      # events on the manager's stream
            acc_ms.append(t.accumulate_ms)
            int_ms.append(t.intensity_ms)
            r1_ms.append(t.region1_ms)
            core_ms.append(t.core_ms)
            far_ms.append(t.far_ms)
            log(f"step: accumulate {t.accumulate_ms:.2f} ms (far field {t.far_ms:.2f} ms for "
                f"{t.far_evals / max(t.evals, 1):.1%} of the pairs, region I {t.region1_ms:.2f} ms, core "
                f"{t.core_ms:.2f} ms), intensity {t.intensity_ms:.3f} ms, evals {t.evals:.4e}")
        m.synchronize()
        wall = time.perf_counter() - t1
        out = np.zeros(M)
        m.TransferResults(None, out, M)
        evals = t.evals
        m.close()
        return dict(wall=wall, acc_ms=float(np.mean(acc_ms)), int_ms=float(np.mean(int_ms)), evals=evals,
                    r1_ms=float(np.mean(r1_ms)), core_ms=float(np.mean(core_ms)), core_evals=t.core_evals,
                    far_ms=float(np.mean(far_ms)), far_evals=t.far_evals, interp=t.interp_pairs, out=out / steps)

    mixed = run(capi.PREC_MIXED, args.steps, args.warmup)
    evals = mixed["evals"]
    r1_pairs = evals - mixed["core_evals"] - mixed["far_evals"]   # region-I pairs of the region-I kernel
    r1_evals = r1_pairs - mixed["interp"] + mixed["interp"] // 4  # evaluations it made (16 nodes per 64 bins)
    ms_step = mixed["wall"] / args.steps * 1e3
    value = evals * args.steps / mixed["wall"]
    res = {
        "metric": "Voigt line-bin evals/s (mixed FP32+Kahan, FP64 intensity; 1e-5 rel err target)",
        "value": value, "unit": "line-bin pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {ll.n} lines, {M} bins {start}-{end} cm^-1, T={T} K, P={P} atm, "
                               f"cutoff {wl['cutoff']} cm^-1, one (T,P) cross-section per step",
                   "lines": ll.n, "bins": M, "evals_per_step": evals},
        "ms_per_xsec": ms_step,
        "roofline": {"bound": "valu",
                     "achieved": FLOP_PER_EVAL * r1_evals / (mixed["r1_ms"] * 1e-3) / 1e12,
                     "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": FLOP_PER_EVAL * r1_evals / (mixed["r1_ms"] * 1e-3) / PEAK_FP32,
                     "traffic": None, "kernel": "k_accum_voigt_mixed", "kernel_ms": mixed["r1_ms"],
                     "evals": r1_evals, "pairs": r1_pairs, "interpolated_pairs": mixed["interp"],
                     "flop_per_eval": FLOP_PER_EVAL},
        "core_pass": {"kernel": "k_core_mixed (W4 regions II-IV)", "ms": mixed["core_ms"],
                      "evals": mixed["core_evals"]},
        "far_field": {"ms": mixed["far_ms"], "evals": mixed["far_evals"], "share": mixed["far_evals"] / max(evals, 1)},
        "intensity_kernel": {"ms": mixed["int_ms"], "bytes": ll.n * 68,
                             "GB_per_s": ll.n * 68 / (mixed["int_ms"] * 1e-3) / 1e9},
    }
    if not args.no_fp64:
        f64 = run(capi.PREC_FP64, 1, 0)
        ref = f64["out"]
        mask = ref >= 1e-6 * ref.max()
        rel = np.abs(mixed["out"][mask] - ref[mask]) / ref[mask]
        res["fp64_path"] = {"evals_per_s": f64["evals"] / (f64["wall"]), "accumulate_ms": f64["acc_ms"],
                            "frac_fp64_peak": FLOP_PER_EVAL * f64["evals"] / (f64["acc_ms"] * 1e-3) / PEAK_FP64}
        res["speedup_vs_fp64"] = value / res["fp64_path"]["evals_per_s"]
        res["precision"] = {"max_rel_err": float(rel.max()), "rms_rel_err": float(np.sqrt(np.mean(rel ** 2))),
                            "bins_checked": int(mask.sum()), "against": "FP64 GPU path (oracle-parity 1e-10)"}
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(ll, wl, Q, gam)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
