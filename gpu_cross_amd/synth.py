"""Seeded synthetic ExoMol-format line lists.

The reference reads ExoMol ``.states`` (id, E, g_tot, J) and ``.trans``
(id_f, id_i, A) files (exomol_functions.cpp:8-48, ExomolStateReader.cpp:130-167)
and derives per transition nu = E_f - E_i, E'' = E_i and g = g_tot of the
upper state (ExomolStateReader.cpp:156-160).  No data files ship with it, so
every workload here is generated from a seed:

* states: E ~ U[0, E_max] sorted, J ~ U{0..40}, g_ns in {1, 3}, g_tot = (2J+1) g_ns;
* transitions: state pairs (u, l) drawn uniformly and kept when
  0 < E_u - E_l <= dE_max (rejection sampling), A log-uniform in [1e-6, 1e2] s^-1;
* Voigt broadening per line: gamma_ref ~ U[0.05, 0.12] cm^-1/atm and
  n ~ U[0.5, 0.8]; the reference's readers hand the manager gamma already
  multiplied by P (ExomolStateReader.cpp:17,26), see ``LineList.gamma_at``.

Arrays are structure-of-arrays in the dtypes the reference's BaseManager takes
(BaseManager.h:24): float64 energies/nu/A/gamma/n and int32 g.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

__all__ = ["LineList", "make_states", "make_lines", "WORKLOADS", "workload"]


@dataclasses.dataclass
class LineList:
    """One synthetic molecule: its states and its (unsorted) transitions."""

    mass: float                 # amu
    state_e: np.ndarray         # float64 [n_states], sorted
    state_g: np.ndarray         # int32   [n_states]
    elow: np.ndarray            # float64 [n]  E''
    nu: np.ndarray              # float64 [n]  line centre
    aif: np.ndarray             # float64 [n]  Einstein A
    gns: np.ndarray             # int32   [n]  upper-state g_tot
    gamma_ref: Optional[np.ndarray] = None   # float64 [n] cm^-1/atm
    nexp: Optional[np.ndarray] = None        # float64 [n]

    @property
    def n(self) -> int:
        return int(self.nu.shape[0])

    def gamma_at(self, pressure: float) -> np.ndarray:
        """gamma * P, as the reference's ExoMol reader delivers it."""
        return self.gamma_ref * pressure

    def sorted_by_nu(self) -> "LineList":
        order = np.argsort(self.nu, kind="stable")
        take = lambda a: None if a is None else np.ascontiguousarray(a[order])
        return LineList(self.mass, self.state_e, self.state_g, take(self.elow), take(self.nu),
                        take(self.aif), take(self.gns), take(self.gamma_ref), take(self.nexp))

    def subset(self, idx) -> "LineList":
        take = lambda a: None if a is None else np.ascontiguousarray(a[idx])
        return LineList(self.mass, self.state_e, self.state_g, take(self.elow), take(self.nu),
                        take(self.aif), take(self.gns), take(self.gamma_ref), take(self.nexp))


def make_states(rng: np.random.Generator, n_states: int, e_max: float):
    e = np.sort(rng.uniform(0.0, e_max, n_states))
    j = rng.integers(0, 41, n_states)
    gns = np.where(rng.integers(0, 2, n_states) == 0, 1, 3)
    g_tot = ((2 * j + 1) * gns).astype(np.int32)
    return e, g_tot


def make_lines(seed: int, n_states: int, e_max: float, n_lines: int, de_max: float, mass: float,
               voigt: bool = True) -> LineList:
    """Generate ``n_lines`` transitions between ``n_states`` random states."""
    rng = np.random.Generator(np.random.PCG64(seed))
    state_e, state_g = make_states(rng, n_states, e_max)
    up = np.empty(n_lines, dtype=np.int64)
    lo = np.empty(n_lines, dtype=np.int64)
    filled = 0
    while filled < n_lines:
        want = n_lines - filled
        batch = int(min(max(2.6 * want + 1024, 4096), 1 << 25))
        u = rng.integers(0, n_states, batch)
        l = rng.integers(0, n_states, batch)
        d = state_e[u] - state_e[l]
        keep = np.nonzero((d > 0.0) & (d <= de_max))[0][:want]
        k = keep.shape[0]
        up[filled:filled + k] = u[keep]
        lo[filled:filled + k] = l[keep]
        filled += k
    aif = 10.0 ** rng.uniform(-6.0, 2.0, n_lines)
    nu = state_e[up] - state_e[lo]
    elow = state_e[lo]
    gns = state_g[up]
    del up, lo
    gamma_ref = nexp = None
    if voigt:
        gamma_ref = rng.uniform(0.05, 0.12, n_lines)
        nexp = rng.uniform(0.5, 0.8, n_lines)
    return LineList(mass, state_e, state_g, np.ascontiguousarray(elow), np.ascontiguousarray(nu),
                    np.ascontiguousarray(aif), np.ascontiguousarray(gns, dtype=np.int32), gamma_ref, nexp)


def make_exomol(seed: int, n_states: int, e_max: float, n_lines: int, de_max: float):
    """Seeded ExoMol-format texts (the reference's input files, ExomolStateReader.cpp:130-167):

    * .states: "%12d %12.6f %6d %7d" = id, E, g_tot, J, energies sorted;
    * .trans:  "%12d %12d %10.4E" = upper id, lower id, A, ordered by nu as ExoMol's
      wavenumber-range files are (the reader's `break` assumes it, VoigtProfile.cpp:67);
    * .broad:  "a1 gamma n J'' J'" for |J' - J''| <= 1 and "a0 gamma n J''" per J''.

    Returns (states text bytes, trans text as a uint8 array, broad text bytes, arrays) with
    arrays = dict(energy, gtot, J, id_f, id_i, aif)."""
    from .ingest import format_trans

    rng = np.random.Generator(np.random.PCG64(seed))
    e = np.sort(rng.uniform(0.0, e_max, n_states))
    j = rng.integers(0, 41, n_states)
    gtot = ((2 * j + 1) * np.where(rng.integers(0, 2, n_states) == 0, 1, 3)).astype(np.int32)
    up = np.empty(n_lines, dtype=np.int64)
    lo = np.empty(n_lines, dtype=np.int64)
    filled = 0
    while filled < n_lines:
        want = n_lines - filled
        batch = int(min(max(2.6 * want + 1024, 4096), 1 << 25))
        u = rng.integers(0, n_states, batch)
        l = rng.integers(0, n_states, batch)
        d = e[u] - e[l]
        keep = np.nonzero((d > 0.0) & (d <= de_max))[0][:want]
        k = keep.shape[0]
        up[filled:filled + k] = u[keep]
        lo[filled:filled + k] = l[keep]
        filled += k
    order = np.argsort(e[up] - e[lo], kind="stable")
    id_f = (up[order] + 1).astype(np.int32)
    id_i = (lo[order] + 1).astype(np.int32)
    del up, lo, order
    aif = 10.0 ** rng.uniform(-6.0, 2.0, n_lines)
    states = "".join(f"{k + 1:12d} {e[k]:12.6f} {gtot[k]:6d} {j[k]:7d}\n" for k in range(n_states)).encode()
    trans = format_trans(id_f, id_i, aif)
    broad = []
    for ji in range(41):
        broad.append(f"a0 {rng.uniform(0.05, 0.12):.4f} {rng.uniform(0.5, 0.8):.3f} {ji}\n")
        for jf in (ji - 1, ji, ji + 1):
            if 0 <= jf <= 40:
                broad.append(f"a1 {rng.uniform(0.05, 0.12):.4f} {rng.uniform(0.5, 0.8):.3f} {ji} {jf}\n")
    arrays = dict(energy=e, gtot=gtot, J=j.astype(np.int32), id_f=id_f, id_i=id_i, aif=aif)
    return states, trans, "".join(broad).encode(), arrays


# Workload shapes (one seed each).  Grid = (start, end, bins); the reference's
# grid is f_i = start + i*(end-start)/bins (BaseProfile.cpp:10,19).
WORKLOADS = {
    "doppler_1e5": dict(seed=1, mass=30.01, n_states=20_000, e_max=15000.0, n_lines=100_000, de_max=10000.0,
                        voigt=False, profile="doppler", T=296.0, P=1.0, grid=(0.0, 10000.0, 1_000_000),
                        doppler_widths=25.0),
    "voigt_1e7": dict(seed=2, mass=17.03, n_states=200_000, e_max=20000.0, n_lines=10_000_000, de_max=12000.0,
                      voigt=True, profile="voigt", T=1000.0, P=1.0, grid=(0.0, 12000.0, 1_200_000), cutoff=25.0),
    "voigt_1e8": dict(seed=3, mass=17.03, n_states=1_000_000, e_max=20000.0, n_lines=100_000_000,
                      de_max=12000.0, voigt=True, profile="voigt", T=1500.0, P=0.1,
                      grid=(0.0, 12000.0, 2_400_000), cutoff=25.0),
    "table_48": dict(seed=3, mass=17.03, n_states=1_000_000, e_max=20000.0, n_lines=100_000_000,
                     de_max=12000.0, voigt=True, profile="voigt",
                     temperatures=[300.0 * 10 ** (k / 7.0) for k in range(8)],
                     pressures=[1e-4 * 10 ** (6.0 * k / 5.0) for k in range(6)],
                     grid=(0.0, 12000.0, 1_200_000), cutoff=25.0),
    "serving_3e8": dict(seed=5, mass=17.03, n_states=3_000_000, e_max=20000.0, n_lines=300_000_000,
                        de_max=12000.0, voigt=True),
    # ExoMol reader: .trans text of 1e8 records over 1e6 states, read for a 0-12000 cm^-1 grid
    "ingest_1e8": dict(seed=7, mass=17.03, n_states=1_000_000, e_max=20000.0, n_lines=100_000_000,
                       de_max=12000.0, voigt=True, pressure=1.0, grid=(0.0, 12000.0, 1_200_000), cutoff=25.0),
}


def workload(name: str, n_lines: Optional[int] = None, n_states: Optional[int] = None) -> LineList:
    """Line list of a named workload, optionally scaled down (same distributions)."""
    w = WORKLOADS[name]
    return make_lines(w["seed"], n_states or w["n_states"], w["e_max"], n_lines or w["n_lines"], w["de_max"],
                      w["mass"], voigt=w["voigt"])
