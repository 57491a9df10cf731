"""Seeded synthetic proteomes, L1 tables and MS2 spectra for the benchmark workloads.

Workloads (WORKLOADS[i]; numpy PCG64 `default_rng`; proteins seed 100+i, spectra 200+i):
  0 plain-2k-proteins-10k-spectra: 2,000 proteins, 0 missed cleavages, length 7-30, fixed C+57.021464, reversed
        decoys (~1e5 rows); 10,000 spectra: 70% from a random target peptidoform (charge 2/3/4 at 60/30/10%),
        40% of b/y 1+ ions kept, +-10 ppm fragment jitter, lognormal intensities, 100 uniform noise peaks in
        150-2000, top-150 peaks kept, precursor jitter U(-5,5) ppm; 30% noise spectra with a random precursor.
        Search: precursor +-10 ppm, fragment +-0.02 Th (Da units), top-k 50.
  1 proteome-plain-20k-proteins-100k-spectra: 20,000 proteins, 1 missed cleavage, variable Met-ox <= 2 (~3M
        rows); 100,000 spectra as workload 0 plus b/y 2+ ions for precursor charge >= 3 (max_fragment_charge 2).
  2 uneven-16-runs-500k-spectra: the table of workload 1; 500,000 spectra from 16 runs, peak counts U(20,500),
        charges 2-6, per-run calibration offset U(-5,5) ppm.
  3 open-500da-50k-spectra: the table of workload 1; 50,000 spectra as workload 1, 40% with a precursor-only
        mass shift U(-150,+400) Da; open window +-500 Da, top-k 100.
  4 phospho-prefilter-20k-proteins-200k-spectra: 20,000 proteins, variable STY phospho <= 3 and Met-ox <= 2
        (~36M rows); 200,000 spectra as workload 2; index top-200.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ..gpu.host import PROTON, SearchParams, Spectra
from . import build as B
from ._core import H2O_F64, NTERM_BIT, RESIDUE_F64

AA = "ARNDCQEGHILKMFPSTWYV"
AA_PCT = np.array([7.0, 5.6, 3.6, 4.7, 2.3, 4.7, 7.1, 6.6, 2.6, 4.3, 10.0, 5.7, 2.1, 3.7, 6.3, 8.3, 5.4, 1.2, 2.7,
                   6.0])
PHOSPHO = 79.966331
OX = 15.9949


@dataclass
class Workload:
    name: str
    n_proteins: int
    missed: int
    var_mods: dict
    max_var: int
    var_caps: dict | None
    n_spectra: int
    max_fragment_charge: int
    charges: tuple = (2, 3, 4)
    charge_p: tuple = (0.6, 0.3, 0.1)
    n_runs: int = 1
    peak_range: tuple | None = None        # workloads 2/4: total peaks per spectrum U(lo, hi)
    calib_ppm: float = 0.0
    shift_frac: float = 0.0
    shift_range: tuple = (-150.0, 400.0)
    topk: int = 50
    open_da: tuple | None = None
    key: int = 0
    table_of: int = 0
    extra: dict = field(default_factory=dict)
    geometry: dict = field(default_factory=dict)   # index / K1 geometry tuned for this scale (chunk_size, ...)
    leg: str = "plain"                             # benchmark leg: plain | uneven | open | prefilter

    def search_params(self) -> SearchParams:
        p = SearchParams(precursor_ppm=(-10.0, 10.0), fragment_ppm=(-0.02, 0.02), fragment_unit="da",
                         max_fragment_charge=self.max_fragment_charge, topk=self.topk,
                         precursor_charge=(min(self.charges), max(self.charges)))
        for k, v in self.geometry.items():
            setattr(p, k, v)
        return p

    def index_params(self) -> dict:
        return {"enzyme": {"cleave_at": "KR", "restrict": "P", "c_terminal": True, "missed_cleavages": self.missed,
                           "min_len": 7, "max_len": 30},
                "static_mods": {"C": 57.021464}, "variable_mods": dict(self.var_mods),
                "max_variable_mods": self.max_var, "variable_mod_caps": self.var_caps}


WORKLOADS = {
    0: Workload("plain-2k-proteins-10k-spectra", 2000, 0, {}, 0, None, 10_000, 1, key=0, table_of=0),
    1: Workload("proteome-plain-20k-proteins-100k-spectra", 20_000, 1, {"M": OX}, 2, None, 100_000, 2, key=1, table_of=1),
    2: Workload("uneven-16-runs-500k-spectra", 20_000, 1, {"M": OX}, 2, None, 500_000, 2, charges=(2, 3, 4, 5, 6),
                charge_p=(0.4, 0.3, 0.15, 0.1, 0.05), n_runs=16, peak_range=(20, 500), calib_ppm=5.0, key=2,
                table_of=1, leg="uneven"),
    3: Workload("open-500da-50k-spectra", 20_000, 1, {"M": OX}, 2, None, 50_000, 2, shift_frac=0.4, topk=100,
                open_da=(-500.0, 500.0), key=3, table_of=1, leg="open"),
    4: Workload("phospho-prefilter-20k-proteins-200k-spectra", 20_000, 1, {"S": PHOSPHO, "T": PHOSPHO, "Y": PHOSPHO, "M": OX}, 5,
                {"S": 3, "T": 3, "Y": 3, "M": 2}, 200_000, 2, charges=(2, 3, 4, 5, 6),
                charge_p=(0.4, 0.3, 0.15, 0.1, 0.05), n_runs=16, peak_range=(20, 500), calib_ppm=5.0, topk=200,
                key=4, table_of=4, geometry={"chunk_size": 1024, "bin_ppm": 320.0, "tile": 1024},
                leg="prefilter"),
}


def proteome(n: int, seed: int) -> list[str]:
    """n random proteins: lognormal length (mean 450, sigma 0.6) clipped to [50, 5000], human-like AA mix."""
    rng = np.random.default_rng(seed)
    sigma = 0.6
    mu = math.log(450.0) - sigma * sigma / 2
    lens = np.clip(np.rint(rng.lognormal(mu, sigma, n)), 50, 5000).astype(np.int64)
    letters = np.frombuffer(AA.encode(), np.uint8)[rng.choice(20, size=int(lens.sum()), p=AA_PCT / AA_PCT.sum())]
    raw = letters.tobytes().decode()
    off = np.concatenate([[0], np.cumsum(lens)])
    return [raw[off[i]:off[i + 1]] for i in range(n)]


def build_index(workload: int, n_proteins: int | None = None, **kw):
    """L1 table of workload `workload` (n_proteins overrides the size for small tests)."""
    w = WORKLOADS[WORKLOADS[workload].table_of]
    prots = proteome(n_proteins or w.n_proteins, 100 + w.key)
    return B.build_tables(prots, w.index_params(), **kw)


def _ladders(l1, rows: np.ndarray, static: dict, var: dict):
    """f64 neutral b and y ladders (k = 0..L-2) and f64 peptidoform masses for `rows` (padded 2D arrays)."""
    L = l1.length[rows].astype(np.int64)
    Lm = int(L.max()) if len(L) else 1
    k = np.arange(Lm)
    idx = l1.seq_off[rows][:, None] + np.minimum(k[None, :], (L - 1)[:, None])
    res = l1.residues[idx].astype(np.int64) - 65
    st = np.zeros(26)
    vr = np.zeros(26)
    for r, m in static.items():
        st[ord(r) - 65] = m
    nterm = 0.0
    for r, m in var.items():
        if r == "^":
            nterm = m
        else:
            vr[ord(r) - 65] = m
    bits = ((l1.mask[rows][:, None] >> k[None, :].astype(np.uint64)) & np.uint64(1)).astype(bool)
    rm = RESIDUE_F64[res] + st[res] + np.where(bits, vr[res], 0.0)
    rm[k[None, :] >= L[:, None]] = 0.0
    nt = np.where((l1.mask[rows] & NTERM_BIT) != 0, nterm, 0.0)
    cum = np.cumsum(rm, axis=1) + nt[:, None]
    M = H2O_F64 + cum[:, -1]
    b = cum[:, :Lm - 1] if Lm > 1 else np.zeros((len(rows), 0))
    y = M[:, None] - b
    valid = k[None, :Lm - 1] < (L - 1)[:, None]
    return b, y, valid, M


def spectra(l1, labels: dict, workload: int, n: int | None = None, seed: int | None = None,
            index_params: dict | None = None):
    """Synthetic spectra of workload `workload` against table `l1`. Returns (Spectra, truth_row, meta)."""
    w = WORKLOADS[workload]
    n = n or w.n_spectra
    rng = np.random.default_rng(200 + workload if seed is None else seed)
    ip = index_params or WORKLOADS[w.table_of].index_params()
    z = rng.choice(np.array(w.charges), size=n, p=np.array(w.charge_p))
    sig = rng.random(n) >= 0.3
    targets = np.nonzero(~labels["is_decoy"])[0]
    truth = np.full(n, -1, np.int64)
    truth[sig] = rng.choice(targets, size=int(sig.sum()))
    run = (np.arange(n) * w.n_runs) // n
    calib = rng.uniform(-w.calib_ppm, w.calib_ppm, w.n_runs) if w.calib_ppm else np.zeros(w.n_runs)
    si = np.nonzero(sig)[0]
    b, y, valid, M = _ladders(l1, truth[si], ip["static_mods"], ip["variable_mods"])
    nk = b.shape[1]
    # ions (spectrum, m/z): b and y at fragment charge 1, plus 2+ for precursor charge >= 3 when allowed
    ion_s, ion_mz = [], []
    for zf in (1, 2):
        if zf > w.max_fragment_charge:
            break
        ok = valid & (z[si][:, None] > zf) if zf > 1 else valid
        for lad in (b, y):
            keep = ok & (rng.random((len(si), nk)) < 0.4)
            r_, c_ = np.nonzero(keep)
            ion_s.append(si[r_])
            ion_mz.append((lad[r_, c_] + zf * float(PROTON)) / zf)
    ion_s = np.concatenate(ion_s) if ion_s else np.zeros(0, np.int64)
    ion_mz = np.concatenate(ion_mz) if ion_mz else np.zeros(0)
    ion_mz = ion_mz * (1.0 + rng.uniform(-10e-6, 10e-6, len(ion_mz)))
    n_sig_ions = np.bincount(ion_s, minlength=n)
    if w.peak_range is None:
        target_peaks = None
        n_noise = np.full(n, 100, np.int64)
    else:
        target_peaks = rng.integers(w.peak_range[0], w.peak_range[1] + 1, n)
        n_noise = np.maximum(target_peaks - n_sig_ions, 0)
        n_noise[~sig] = target_peaks[~sig]
    noise_s = np.repeat(np.arange(n), n_noise)
    noise_mz = rng.uniform(150.0, 2000.0, len(noise_s))
    s_all = np.concatenate([ion_s, noise_s])
    mz_all = np.concatenate([ion_mz, noise_mz])
    it_all = rng.lognormal(0.0, 1.0, len(s_all))
    cal = 1.0 + calib[run[s_all]] * 1e-6
    mz_all = mz_all * cal
    # per spectrum: keep the top peaks by intensity (150, or the drawn peak count), then m/z order
    cap = np.full(n, 150, np.int64) if target_peaks is None else target_peaks
    o = np.lexsort((-it_all, s_all))
    s_o = s_all[o]
    first = np.searchsorted(s_o, s_o, side="left")
    rank = np.arange(len(s_o)) - first
    keep = o[rank < cap[s_o]]
    s_k, mz_k, it_k = s_all[keep], mz_all[keep], it_all[keep]
    o2 = np.lexsort((mz_k, s_k))
    s_k, mz_k, it_k = s_k[o2], mz_k[o2], it_k[o2]
    off = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(s_k, minlength=n), out=off[1:])
    # precursor m/z from the f64 peptidoform mass, jitter U(-5,5) ppm; noise spectra get a random mass
    mass = np.empty(n)
    mass[si] = M
    mass[~sig] = rng.uniform(600.0, 4000.0, int((~sig).sum()))
    shift = np.zeros(n)
    if w.shift_frac:
        sh = sig & (rng.random(n) < w.shift_frac)
        shift[sh] = rng.uniform(w.shift_range[0], w.shift_range[1], int(sh.sum()))
    pmz = (mass + shift + z * float(PROTON)) / z * (1.0 + rng.uniform(-5e-6, 5e-6, n)) * (1.0 + calib[run] * 1e-6)
    sp = Spectra(f"{w.name}", np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32),
                 np.linspace(0.0, 120.0, n), pmz.astype(np.float64), z.astype(np.int8),
                 mz_k.astype(np.float32), it_k.astype(np.float32), off)
    return sp, truth, {"run": run, "shift": shift}
