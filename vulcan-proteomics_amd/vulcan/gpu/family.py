"""Family top-2 retention on a multi-replicate table (the reference's vulcan/gpu/family.py surface:
FamilyIndex, search_family, family_table, reconstruct, SubIndex; family.py:31-243).

A table from vulcan.index.build.build_multi_decoy_tables holds targets T, entrapment families E_s and, per decoy
set j, the decoys of both. One search keeps, per spectrum, the two best candidates of every retention group
g = ((f * J) + j) * 2 + is_decoy (f = forward family: 0 for T, 1 + i for the i-th entrapment seed), from which
each replicate's competition (decoy set j, entrapment seed s) is rebuilt exactly as a separate search of its
rows would give it:

  * shortlist (vh_family_select, on the bucket-join K1's per-row window counts): top-K per spectrum and K1
    family (forward family x side), closure over every decoy set, one entry per (spectrum, row) with its
    canonical hypothesis. Spectra whose shortlist does not fit one wave's LDS go through the wide path
    (vh_family_wide_stats / vh_family_wide / vh_family_wide_copy: device radix sorts, the same rule);
  * exact rescoring (vh_k2) of the shortlist;
  * top-2 table (vh_family_top2): per decoy context j the admissible candidates (shortlisted, or the set-j
    partner of a shortlisted row) with at least min_matched_peaks matches, the best two per (spectrum, group)
    by (hyperscore desc, row asc).

Every step is a HIP kernel; torch holds the buffers. The family codes come from the label table (the
reference's sanctioned use of labels for retention; the rule is symmetric in the target / decoy roles).
Whole-table features (scored_candidates, poisson) cannot be rebuilt per replicate and are not emitted.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np
import pyarrow as pa
import torch

from . import _lib as K
from . import engine as E

ROW_BITS = E.ROW_BITS


def family_codes(labels: dict):
    """Per-row codes from the label columns: forward family (0 = T, 1 + i = entrapment family of the i-th
    smallest entrapment seed), decoy side (-1 forward, j = decoy set j) and the sorted entrapment seeds."""
    fam = np.asarray(labels["family"])
    seed = np.asarray(labels["entrap_seed"])
    entrap = (fam % 2) == 1                       # 1 = E, 3 = dec(E)
    seeds = np.unique(seed[entrap]).tolist()
    fwd = np.zeros(len(fam), np.int64)
    fwd[entrap] = 1 + np.searchsorted(np.asarray(seeds, seed.dtype), seed[entrap])
    return fwd, np.asarray(labels["decoy_set"], np.int64), seeds


class FamilyIndex:
    """Device codes of a multi-replicate table: fwd_fam and side (int8 per row), partners (J x n_rows, the pair
    partner in every decoy set, the row itself where none) and the K1 family fwd_fam * (J + 1) + side + 1."""

    def __init__(self, fwd_fam, side, partners, n_fwd_fam: int, dev, seeds=None):
        fwd_fam, side = np.asarray(fwd_fam, np.int64), np.asarray(side, np.int64)
        self.J, self.F = len(partners), int(n_fwd_fam)
        if not (1 <= self.J <= 8) or self.F * (self.J + 1) > 256 or self.F * self.J * 2 > 256:
            raise ValueError("1..8 decoy sets and at most 256 K1 families / retention groups")
        if len(fwd_fam) and (fwd_fam.min() < 0 or fwd_fam.max() >= self.F or side.min() < -1 or side.max() >= self.J):
            raise ValueError("family codes outside [0, F) x [-1, J)")
        self.dev = dev
        self.seeds = seeds or {}
        self.fwd_fam = torch.as_tensor(fwd_fam.astype(np.int8), device=dev)
        self.side = torch.as_tensor(side.astype(np.int8), device=dev)
        self.kfam = torch.as_tensor((fwd_fam * (self.J + 1) + side + 1).astype(np.uint8), device=dev)
        self.partners = torch.as_tensor(np.stack([np.asarray(q, np.int64) for q in partners]), device=dev)
        self.searches = {}

    @classmethod
    def from_l1(cls, l1, gi, labels: dict | None = None):
        from ..index.build import read_labels
        lab = labels if labels is not None else read_labels(l1.path)
        if not np.array_equal(np.asarray(lab["pf_id"]), l1.pf_id):
            raise ValueError("labels / peptidoforms row order differ")
        fwd, side, seeds = family_codes(lab)
        return cls(fwd, side, l1.partners, 1 + len(seeds), gi.dev,
                   {"entrap_seeds": seeds, "decoy_seeds": l1.meta.get("params", {}).get("decoy_seeds")})

    def group(self, f: int, j: int, decoy: bool) -> int:
        return (f * self.J + j) * 2 + int(decoy)

    def replicate_groups(self, j: int, s_index: int) -> list:
        """The four groups of replicate (decoy set j, entrapment family 1 + s_index)."""
        return [self.group(f, j, d) for f in (0, 1 + s_index) for d in (False, True)]


class Shortlist:
    """Device shortlist: per spectrum its entries in row order at [off[s], off[s + 1]) -- row, canonical
    hypothesis, K1 count (-1: partner only) and the shortlisted flag; `n_wide` spectra took the wide path."""

    def __init__(self, off, row, hyp, cnt, top, n_wide: int):
        self.off, self.row, self.hyp, self.cnt, self.top, self.n_wide = off, row, hyp, cnt, top, n_wide

    def __len__(self):
        return self.row.numel()


def _wide_path(ds, fi: FamilyIndex, n_wide: int, wide, out_off):
    """Shortlists of the spectra vh_family_select listed as too large for one wave: returns the packed entries
    and their per-spectrum placement (written into out_off before the scan)."""
    L, dev = K.lib(), ds.gi.dev
    wspec = wide[1:1 + n_wide]
    j, hy = ds.join, ds.hyp
    w_off = torch.empty(n_wide + 1, dtype=torch.int64, device=dev)
    rmin = torch.empty(n_wide, dtype=torch.int64, device=dev)
    gstats = torch.empty(2, dtype=torch.int64, device=dev)
    ws = torch.empty(int(L.vh_family_wide_stats_workspace_bytes(n_wide)), dtype=torch.uint8, device=dev)
    K.check(L.vh_family_wide_stats(C.byref(ds.vp), n_wide, K.ptr(wspec), K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]),
                                   K.ptr(hy["hi"]), K.ptr(j["h_slot"]), j["slot_cap"], K.ptr(j["counts"]), ds.gi.n,
                                   K.ptr(w_off), K.ptr(rmin), K.ptr(gstats), K.ptr(ws), ws.numel(), K.stream()),
            "family_wide_stats")
    w_host = w_off.cpu().numpy()
    cmax, span = (int(x) for x in gstats.cpu().numpy())
    total = int(w_host[-1])
    uniq = torch.empty(max(total * (fi.J + 1), 1), dtype=torch.int64, device=dev)
    u_off = torch.empty(n_wide, dtype=torch.int64, device=dev)
    u_cnt = torch.empty(n_wide, dtype=torch.int64, device=dev)
    ws = torch.empty(int(L.vh_family_wide_workspace_bytes(total, fi.J)), dtype=torch.uint8, device=dev)
    K.check(L.vh_family_wide(C.byref(ds.vp), n_wide, K.ptr(wspec), K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]),
                             K.ptr(hy["hi"]), K.ptr(j["h_slot"]), j["slot_cap"], K.ptr(j["counts"]),
                             K.ptr(fi.kfam), fi.F * (fi.J + 1), ds.gi.n, K.ptr(fi.partners), fi.J, K.ptr(w_off),
                             w_host.ctypes.data, K.ptr(rmin), cmax, span, K.ptr(uniq), K.ptr(u_off), K.ptr(u_cnt),
                             K.ptr(out_off), K.ptr(ws), ws.numel(), K.stream()), "family_wide")
    return wspec, uniq, u_off, u_cnt


def select_family_join(ds: "E.DeviceSearch", fi: FamilyIndex, n: int) -> Shortlist:
    """The family shortlist of the n spectra whose front and join K1 counts `ds` has run (vh_family_select, the
    wide path for spectra beyond one wave's LDS)."""
    L, dev = K.lib(), ds.gi.dev
    j, hy = ds.join, ds.hyp
    out_off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    wide = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(int(L.vh_family_select_workspace_bytes(n)), dtype=torch.uint8, device=dev)
    head = (C.byref(ds.vp), n, K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]), K.ptr(j["h_slot"]),
            j["slot_cap"], K.ptr(j["counts"]), K.ptr(fi.kfam), ds.gi.n, K.ptr(fi.partners), fi.J)

    def call(mode, outs=(None, None, None, None)):
        K.check(L.vh_family_select(*head, mode, K.ptr(out_off), *[K.ptr(t) for t in outs], K.ptr(wide),
                                   K.ptr(status), K.ptr(ws), ws.numel(), K.stream()), f"family_select ({mode})")
    call(0)
    n_wide = int(wide[0].item())
    wd = _wide_path(ds, fi, n_wide, wide, out_off) if n_wide else None
    call(2)
    tot = int(out_off[n].item())
    outs = (torch.empty(max(tot, 1), dtype=torch.int32, device=dev), torch.empty(max(tot, 1), dtype=torch.int32,
            device=dev), torch.empty(max(tot, 1), dtype=torch.int32, device=dev),
            torch.empty(max(tot, 1), dtype=torch.uint8, device=dev))
    call(1, outs)
    if wd is not None:
        wspec, uniq, u_off, u_cnt = wd
        K.check(L.vh_family_wide_copy(n_wide, K.ptr(wspec), K.ptr(hy["hyp_off"]), K.ptr(u_off), K.ptr(u_cnt),
                                      K.ptr(uniq), K.ptr(out_off), *[K.ptr(t) for t in outs], K.stream()),
                "family_wide_copy")
    if int(status.item()):
        raise RuntimeError("family_select: a spectrum has more than 64 hypotheses or a window beyond the join K1's "
                           "slot capacity")
    row, hyp, cnt, top = (t[:tot] for t in outs)
    return Shortlist(out_off, row, hyp, cnt, top, n_wide)


def _device_search(gi, fi: FamilyIndex, p, n: int) -> "E.DeviceSearch":
    """Preallocated step buffers, reused by searches with the same index, batch size and parameters."""
    key = (id(gi), n, json.dumps(p.to_dict(), sort_keys=True, default=list))
    ds = fi.searches.get(key)
    if ds is None:
        ds = E.DeviceSearch(gi, p, n, k1_impl="join")
        fi.searches = {key: ds}
    return ds


def search_family(sp: E.Spectra, gi: E.GpuIndex, l1, fi: FamilyIndex, p: E.SearchParams):
    """Closed search with family top-2 retention. Returns (table, stats); one table row per (spectrum, retention
    group, grank in {1, 2})."""
    st = E.Stats()
    n = len(sp)
    with E.Timer(st, "upload"):
        raw = E.upload_spectra(sp, gi.dev, st)
    ds = _device_search(gi, fi, p, n)
    mark = E._marker(None)
    ds.status.zero_()
    with E.Timer(st, "k0"):
        ds.front(raw, n, mark)             # K0, hypotheses, processing order
    with E.Timer(st, "k1"):
        ds.join_front(n, mark)             # join K1: pairs + per-row window counts
        if int(ds.status.item()) & 8:      # this batch outgrew the pair / slot capacities: refit, count again
            ds.join["fitted"] = False
            ds.status.zero_()
            ds.join_front(n, mark)
    with E.Timer(st, "topk"):
        sl = select_family_join(ds, fi, n)
    with E.Timer(st, "k2"):
        k2 = E._k2_out(len(sl), gi.dev)
        if len(sl):
            E._k2_call(gi, ds.vp, n, None, sl.off, None, 0, sl.row, sl.hyp, ds.hyp["maxfz"], ds.peaks, k2)
    with E.Timer(st, "features"):
        tbl = family_table(sp, l1, fi, ds, sl, k2, p)
    st.counts.update(n_spectra=n, n_hypotheses=int(ds.hyp["hyp_off"][n].item()), n_rescored=len(sl),
                     n_wide=sl.n_wide, n_rows=tbl.num_rows)
    return tbl, st


def family_table(sp, l1, fi: FamilyIndex, ds, sl: Shortlist, k2: dict, p) -> pa.Table:
    """The top-2 table (vh_family_top2, gathered on the device) with the reference's columns
    (family.py:190-203)."""
    from ._abi import FAMILY_COLS, VhFamilyCols
    L, dev = K.lib(), fi.dev
    n = len(sp)
    out_off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    ws = torch.empty(int(L.vh_family_top2_workspace_bytes(n)), dtype=torch.uint8, device=dev)
    hy = ds.hyp
    head = (C.byref(ds.vp), n, K.ptr(sl.off), K.ptr(sl.row), K.ptr(sl.hyp), K.ptr(sl.top), K.ptr(k2["hyperscore"]),
            K.ptr(k2["matched_b"]), K.ptr(k2["matched_y"]), K.ptr(hy["z"]), K.ptr(hy["iso"]), K.ptr(fi.fwd_fam),
            K.ptr(fi.side), ds.gi.n, K.ptr(fi.partners), fi.J, fi.F)
    for mode, what in ((0, "counts"), (2, "offsets")):
        K.check(L.vh_family_top2(*head, mode, K.ptr(out_off), None, K.ptr(ws), ws.numel(), K.stream()),
                f"family_top2 ({what})")
    m = int(out_off[n].item())
    types = {"idx": torch.int64, "group": torch.int16, "rank": torch.int8, "row": torch.int32, "hyp": torch.int32,
             "hs": torch.float64, "mb": torch.int16, "my": torch.int16, "z": torch.int8, "iso": torch.int8}
    cols = {k: torch.empty(max(m, 1), dtype=types[k], device=dev) for k in FAMILY_COLS}
    oc = VhFamilyCols(**{k: K.ptr(v) for k, v in cols.items()})
    K.check(L.vh_family_top2(*head, 1, K.ptr(out_off), C.byref(oc), K.ptr(ws), ws.numel(), K.stream()),
            "family_top2")
    h = E._to_host({**{k: v[:m] for k, v in cols.items() if k != "idx"}, "off": out_off})
    spec = np.repeat(np.arange(n, dtype=np.int64), np.diff(h["off"]))
    rows = h["row"].astype(np.int64)
    return pa.table({
        "run": E.string_column(sp.run, m), "spec_idx": spec, "scan": sp.scan[spec],
        "fgroup": h["group"], "grank": h["rank"], "pf_row": rows, "pf_id": l1.pf_id[rows], "hyperscore": h["hs"],
        "charge": h["z"], "isotope_error": h["iso"], "matched_b": h["mb"], "matched_y": h["my"]})


def reconstruct(tbl: pa.Table, groups, n: int = 2) -> dict:
    """One replicate's competition from a family table: per spectrum the best n rows over `groups` by
    (hyperscore desc, row asc), as numpy arrays in (spectrum, rank) order."""
    col = {k: tbl.column(k).to_numpy() for k in ("fgroup", "spec_idx", "hyperscore", "pf_row", "pf_id")}
    pick = np.isin(col["fgroup"], np.asarray(list(groups)))
    s, hs, row, pid = (col[k][pick] for k in ("spec_idx", "hyperscore", "pf_row", "pf_id"))
    order = np.lexsort((row, -hs, s))
    s, hs, row, pid = s[order], hs[order], row[order], pid[order]
    rank = np.arange(len(s)) - np.searchsorted(s, s, side="left")     # position inside the spectrum's run
    keep = rank < n
    return {"spec_idx": s[keep], "rank": rank[keep] + 1, "hyperscore": hs[keep], "pf_id": pid[keep],
            "pf_row": row[keep]}


class SubIndex:
    """One replicate's rows as a table of their own (order kept, so mass order and row tie-breaks are kept),
    with the decoy set j partner renumbered into it: the index of a separate search of that replicate."""

    def __init__(self, l1, keep: np.ndarray, j: int):
        self.rows = np.flatnonzero(np.asarray(keep, bool))
        new_of = np.full(len(l1), -1, np.int64)
        new_of[self.rows] = np.arange(len(self.rows))
        pj = new_of[np.asarray(l1.partners[j])[self.rows]]
        if len(pj) and pj.min() < 0:
            raise ValueError("a partner lies outside the replicate's rows")
        self.path, self.meta = getattr(l1, "path", None), l1.meta
        for name in ("pf_id", "mass32", "seq_off", "length", "mask", "n_var", "missed"):
            setattr(self, name, np.asarray(getattr(l1, name))[self.rows])
        self.residues = l1.residues
        self.partner = pj
        self.partners = [pj]

    def __len__(self):
        return len(self.rows)
