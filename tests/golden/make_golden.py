"""Write the golden fixtures under tests/golden/ from the reference's own CPU manager.

The reference's OpenMPManager is compiled unmodified from its sources by
``make -C oracle ref`` (only possible where the reference checkout exists) and
driven by oracle/ref_harness.cpp through the BaseManager lifecycle.  Each
fixture stores the exact inputs it was given and the FP64 grid it returned;
the reference code itself never enters the repository.

Run:  python tests/golden/make_golden.py   (needs the reference checkout)
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gpu_cross_amd.synth import make_lines  # noqa: E402
from oracle import pyoracle  # noqa: E402
from tests.golden.caseio import read_grid, write_case  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def run_reference(case: dict, lines: dict) -> np.ndarray:
    case["capacity"] = min(case["capacity"], lines["nu"].size)   # keeps the reference's host buffers small
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        write_case(fin, case, lines)
        subprocess.run([HARNESS, fin, fout], check=True, stdout=subprocess.DEVNULL)
        return read_grid(fout)


def select(ll, lo, hi, order="sorted"):
    idx = np.nonzero((ll.nu >= lo) & (ll.nu <= hi))[0]
    if order == "sorted":
        idx = idx[np.argsort(ll.nu[idx], kind="stable")]
    return idx


def tie_lines(case):
    """Lines whose window edges fall exactly on half bins (dfreq = 1/128, all exact)."""
    d = (case["end"] - case["start"]) / case["npoints"]
    c = min(25.0 * case["P"], 100.0)
    ks = np.arange(40, 16000, 397)
    nu = case["start"] + (ks + 0.5) * d             # (nu - C - start)/d = k + 0.5 - C/d exactly
    nu = np.concatenate([nu, case["start"] + c + (ks[:10] + 0.5) * d])
    m = nu.size
    rng = np.random.Generator(np.random.PCG64(99))
    return dict(elow=rng.uniform(0, 5000, m), nu=nu, aif=10 ** rng.uniform(-3, 1, m),
                gns=rng.integers(1, 30, m).astype(np.int32), gamma=rng.uniform(0.05, 0.12, m),
                nexp=rng.uniform(0.5, 0.8, m))


CASES = [
    # name, line source, selection window, reference deck parameters
    dict(name="voigt_T1000_P1_t1", src=("voigt", 2, 120_000), sel=(975.0, 1125.0, "sorted"),
         profile=2, T=1000.0, P=1.0, ref_temp=500.0, mass=17.03, start=1000.0, end=1100.0, npoints=20000,
         nthreads=1, capacity=1 << 30, max_points=80000),
    dict(name="voigt_T296_P0.1_t3_chunked", src=("voigt", 7, 120_000), sel=(2497.5, 2602.5, "sorted"),
         profile=2, T=296.0, P=0.1, ref_temp=500.0, mass=17.03, start=2500.0, end=2600.0, npoints=40000,
         nthreads=3, capacity=300, max_points=8000),
    dict(name="voigt_T2000_P2_t4_unsorted", src=("voigt", 11, 120_000), sel=(4450.0, 4650.0, "unsorted"),
         profile=2, T=2000.0, P=2.0, ref_temp=500.0, mass=17.03, start=4500.0, end=4600.0, npoints=16384,
         nthreads=4, capacity=1 << 30, max_points=80000),
    dict(name="doppler_T296_t1", src=("doppler", 1, 100_000), sel=(900.0, 1200.0, "sorted"),
         profile=1, T=296.0, P=1.0, ref_temp=296.0, mass=30.01, start=1000.0, end=1100.0, npoints=10000,
         nthreads=1, capacity=1 << 30, max_points=80000),
    dict(name="doppler_T1500_t2_chunked", src=("doppler", 13, 100_000), sel=(100.0, 400.0, "sorted"),
         profile=1, T=1500.0, P=1.0, ref_temp=296.0, mass=30.01, start=200.0, end=300.0, npoints=12000,
         nthreads=2, capacity=700, max_points=80000),
    # the 1e7-line Voigt workload's own grid (0-12000 cm^-1 at 0.01 cm^-1, 1.2e6 bins), T = 1000 K,
    # P = 1 atm, its state count and energy ranges; a thinned line list around 3000-3100 cm^-1
    dict(name="voigt_T1000_P1_grid1e7", src=("voigt", 2, 400_000, 200_000), sel=(2990.0, 3110.0, "sorted"),
         profile=2, T=1000.0, P=1.0, ref_temp=500.0, mass=17.03, start=0.0, end=12000.0, npoints=1_200_000,
         nthreads=4, capacity=1 << 30, max_points=80000),
    dict(name="voigt_halfbin_ties", src=("ties",), sel=None,
         profile=2, T=700.0, P=1.0, ref_temp=500.0, mass=17.03, start=0.0, end=128.0, npoints=16384,
         nthreads=2, capacity=1 << 30, max_points=80000),
]


def build_case(case):
    if case["src"][0] == "ties":
        lines = tie_lines(case)
        case["Q"] = 1234.5
    else:
        kind, seed, n = case["src"][:3]
        n_states = case["src"][3] if len(case["src"]) > 3 else 20_000
        voigt = kind == "voigt"
        ll = make_lines(seed, n_states, 20000.0 if voigt else 15000.0, n, 12000.0 if voigt else 10000.0,
                        case["mass"], voigt=voigt)
        idx = select(ll, *case["sel"])
        lines = dict(elow=ll.elow[idx], nu=ll.nu[idx], aif=ll.aif[idx], gns=ll.gns[idx])
        if voigt:
            lines["gamma"] = ll.gamma_at(case["P"])[idx]
            lines["nexp"] = ll.nexp[idx]
        case["Q"] = pyoracle.partition(ll.state_e, ll.state_g, case["T"])
    lines = {k: np.ascontiguousarray(v) for k, v in lines.items()}
    return lines


def main(only=None):
    """Write every fixture (or only the named ones) and the manifest of all fixtures."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", "liboracle.so"], check=True)
    manifest = []
    for case in CASES:
        case = dict(case)
        if only and case["name"] not in only:
            z = np.load(os.path.join(HERE, case["name"] + ".npz"))
            out, nl = z["intens"], int(z["nu"].size)
            manifest.append(dict(name=case["name"], nlines=nl, npoints=case["npoints"], max=float(out.max()),
                                 nonzero=int(np.count_nonzero(out))))
            continue
        lines = build_case(case)
        out = run_reference(case, lines)
        meta = {k: v for k, v in case.items() if k not in ("src", "sel")}
        np.savez_compressed(os.path.join(HERE, case["name"] + ".npz"), intens=out,
                            meta=np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8), **lines)
        manifest.append(dict(name=case["name"], nlines=int(lines["nu"].size), npoints=case["npoints"],
                             max=float(out.max()), nonzero=int(np.count_nonzero(out))))
        print(manifest[-1])
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main(set(sys.argv[1:]) or None)
