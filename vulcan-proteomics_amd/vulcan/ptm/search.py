"""Search modes on the L1 table: closed, mass-offset and open (wide precursor window), with shifted +
unshifted ions in K2. Call surface of the reference's vulcan/ptm/search.py (ModeParams, make_mode_hypotheses,
search_modes; MODE_CLOSED / MODE_OFFSET / MODE_OPEN).

Hypotheses (ptm/search.py:48-77): the closed hypotheses of a spectrum, then one per (mass offset, charge) at
isotope 0 (window M - offset +- precursor ppm), then one open hypothesis per charge (calc mass in
[M - open_da[1], M - open_da[0]]). The reference lists all closed hypotheses of the run first; here they are
grouped by spectrum with the same order inside a spectrum, which is the only order the top-K tie rule and the
dedupe (lowest hypothesis wins) look at. K1 runs the same fused kernel with the open threshold for open
hypotheses and a larger LDS tile; K2 scores candidates of offset / open hypotheses with shift = M - calc mass
(vh_candidate_shift + vh_k2 shifted path).
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import pyarrow as pa

from ..gpu import engine as E
from ..gpu.host import NEUTRON, SearchParams, make_hypotheses

MODE_CLOSED, MODE_OFFSET, MODE_OPEN = 0, 1, 2


@dataclass
class ModeParams:
    """ptm/search.py:38-45. spectra_per_batch / append_cap are not needed: the fused K1 keeps candidate lists in
    LDS per spectrum, so there is no append buffer to overflow."""
    offsets: tuple = ()                 # mass-offset mode: Da (delta = observed - peptide)
    open_da: tuple | None = None        # open mode: (lo, hi) for delta = observed - peptide
    open_tile: int = 4096               # K1 LDS histogram slots per wave when open hypotheses are present
    k1_threshold_open: int = 1
    spectra_per_batch: int = 1000
    append_cap: int = 600_000_000


def make_mode_hypotheses(sp, valid, idx_mass32, p: SearchParams, mp: ModeParams):
    """Host restatement (numpy, f32) of the mode hypotheses, grouped by spectrum (see module docstring)."""
    hyp = make_hypotheses(sp, valid, idx_mass32, p)
    n0 = len(hyp["spec"])
    hyp["mode"] = np.zeros(n0, np.uint8)
    hyp["off"] = np.zeros(n0, np.float32)
    parts = [hyp]
    base = hyp["iso"] == 0
    s, z = hyp["spec"][base], hyp["z"][base]
    M0, pm0, mfz = hyp["M"][base], hyp["pm"][base], hyp["maxfz"][base]
    plo, phi = np.float32(p.precursor_ppm[0]), np.float32(p.precursor_ppm[1])
    for o in mp.offsets:
        Mo = (M0 - np.float32(o)).astype(np.float32)
        lo = (Mo + (Mo * plo) / np.float32(1e6)).astype(np.float32)
        hi = (Mo + (Mo * phi) / np.float32(1e6)).astype(np.float32)
        parts.append({"spec": s, "z": z, "iso": np.zeros_like(z), "pm": pm0, "M": M0,
                      "lo": np.searchsorted(idx_mass32, lo, side="left").astype(np.int64),
                      "hi": np.searchsorted(idx_mass32, hi, side="right").astype(np.int64), "maxfz": mfz,
                      "mode": np.full(len(s), MODE_OFFSET, np.uint8), "off": np.full(len(s), o, np.float32)})
    if mp.open_da is not None:
        lo = (M0 - np.float32(mp.open_da[1])).astype(np.float32)
        hi = (M0 - np.float32(mp.open_da[0])).astype(np.float32)
        parts.append({"spec": s, "z": z, "iso": np.zeros_like(z), "pm": pm0, "M": M0,
                      "lo": np.searchsorted(idx_mass32, lo, side="left").astype(np.int64),
                      "hi": np.searchsorted(idx_mass32, hi, side="right").astype(np.int64), "maxfz": mfz,
                      "mode": np.full(len(s), MODE_OPEN, np.uint8), "off": np.full(len(s), np.nan, np.float32)})
    out = {k: np.concatenate([q[k] for q in parts]) for k in parts[0]}
    o = np.argsort(out["spec"], kind="stable")      # group by spectrum, keeping closed, offset, open order
    return {k: v[o] for k, v in out.items()}


def mode_search_params(p: SearchParams, mp: ModeParams) -> SearchParams:
    q = SearchParams(**p.to_dict())
    if mp.open_da is not None:
        q.tile = mp.open_tile
    return q


def search_modes(sp, gi: E.GpuIndex, l1, p: SearchParams, mp: ModeParams, report: int | None = None):
    """Closed + mass-offset + open search of one run (ptm/search.py:178-215). Returns (PSM table, stats); the
    table carries the reference's extra columns mode, hyp_offset and mass_shift."""
    st = E.Stats()
    t0 = time.perf_counter()
    report = p.report_psms if report is None else report
    q = mode_search_params(p, mp)
    raw = E.upload_spectra(sp, gi.dev, st)
    ds = E.DeviceSearch(gi, q, len(sp), report=1, offsets=tuple(mp.offsets), open_da=mp.open_da,
                        k1_threshold_open=mp.k1_threshold_open)
    with E.Timer(st, "search_device"):
        ds.run(raw)
    with E.Timer(st, "features"):
        tbl, hh, dbg = E.device_table(ds, sp, gi, l1, q, report)
        hyp = dbg["hyp"]
        rows = tbl.column("pf_row").to_numpy()
        iso = tbl.column("isotope_error").to_numpy().astype(np.float32)
        pm = tbl.column("expmass").to_numpy()
        ms = (pm - (iso * NEUTRON).astype(np.float32) - l1.mass32[rows]).astype(np.float32)
        tbl = tbl.append_column("mode", pa.array(hyp["mode"][hh], pa.uint8()))
        tbl = tbl.append_column("hyp_offset", pa.array(hyp["off"][hh], pa.float32()))
        tbl = tbl.append_column("mass_shift", pa.array(ms, pa.float32()))
    st.wall_s["total"] = time.perf_counter() - t0
    st.counts.update(n_spectra=len(sp), n_hypotheses=len(hyp["spec"]), n_rescored=dbg["n_rescored"],
                     n_psms=tbl.num_rows)
    return tbl, st
