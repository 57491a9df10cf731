"""Family top-2 retention over a multi-replicate table (reference vulcan/gpu/family.py:90-203).

A table built by vulcan.index.build.build_multi_decoy_tables holds targets T, entrapment families E_s and, for
every decoy set j, the decoys of all of them. One search over it keeps, per spectrum, the top-2 candidates of
every retention group g = ((f * J) + j) * 2 + is_decoy (f: forward family, 0 = T, 1 + i = E of the i-th
entrapment seed), from which the competition of each replicate (decoy set j, entrapment seed s) is rebuilt
exactly as a separate search over that replicate's rows would give it:

  * shortlist: top-K per spectrum per K1 family (forward family x side, side = forward or decoy set j) by
    (K1 count desc, row desc; equal keys from two hypotheses: lower hypothesis first) -- a separate search
    ranks the same rows of its four families the same way, since nothing else competes inside a family;
  * closure over every decoy set (each shortlisted row adds its set-j partner for every j), deduplicated by
    (spectrum, row): lowest source hypothesis, highest count, shortlisted if any source was;
  * canonical hypothesis per (spectrum, row): the lowest-index hypothesis of the spectrum whose window holds the
    row, else the lowest source hypothesis (it depends on the spectrum and the row mass only, not on the table);
  * exact rescoring (K2) of the survivors, then per decoy context j the admissible candidates (shortlisted, or
    partner of a shortlisted row in set j) with at least min_matched_peaks matches, top-2 per (spectrum, group)
    by (hyperscore desc, row asc).

K0, the hypotheses, the K1 counts and K2 run on the HIP kernels of vulcan.gpu.engine (the hypotheses stay on the
device); the selection steps are device sorts and scans over the appended K1 list. The family codes come from
the label table (the reference's sanctioned use of labels for retention, symmetric in the target / decoy roles).
Features that depend on the whole table (scored_candidates, poisson) cannot be rebuilt per replicate and are
not emitted.
"""
from __future__ import annotations

import numpy as np
import pyarrow as pa
import torch

from . import engine as E

ROW_BITS = E.ROW_BITS


class FamilyIndex:
    """Per-row family codes on the device. fwd_fam: 0 = T, 1 + i = E of the i-th entrapment seed (sorted);
    side: -1 forward, j = decoy set j; partners: (J, n) pair partner per decoy set (self where none)."""

    def __init__(self, fwd_fam: np.ndarray, side: np.ndarray, partners: list, n_fwd_fam: int, dev, seeds=None):
        self.J = len(partners)
        self.F = int(n_fwd_fam)
        if self.F * (self.J + 1) > 256 or self.F * self.J * 2 > 256:
            raise ValueError("at most 256 K1 families / retention groups")
        self.seeds = seeds or {}
        self.dev = dev
        self.fwd_fam = torch.as_tensor(np.asarray(fwd_fam, np.int64), device=dev)
        self.side = torch.as_tensor(np.asarray(side, np.int64), device=dev)
        self.partners = torch.stack([torch.as_tensor(np.asarray(p, np.int64)) for p in partners]).to(dev)
        self.kfam = self.fwd_fam * (self.J + 1) + (self.side + 1)
        self.kfam_u8 = None
        self.searches = {}

    @staticmethod
    def codes_from_labels(lab: dict):
        fam = np.asarray(lab["family"])
        seed = np.asarray(lab["entrap_seed"])
        is_e = (fam % 2) == 1
        eseeds = sorted(set(seed[is_e].tolist()))
        fwd_fam = np.zeros(len(fam), np.int64)
        for i, s in enumerate(eseeds):
            fwd_fam[is_e & (seed == s)] = 1 + i
        side = np.asarray(lab["decoy_set"], np.int64)
        return fwd_fam, side, eseeds

    @classmethod
    def from_l1(cls, l1, gi, labels: dict | None = None):
        from ..index.build import read_labels
        lab = labels if labels is not None else read_labels(l1.path)
        if not np.array_equal(np.asarray(lab["pf_id"]), l1.pf_id):
            raise ValueError("labels / peptidoforms row order differ")
        fwd_fam, side, eseeds = cls.codes_from_labels(lab)
        return cls(fwd_fam, side, l1.partners, 1 + len(eseeds), gi.dev,
                   {"entrap_seeds": eseeds, "decoy_seeds": l1.meta.get("params", {}).get("decoy_seeds")})

    def group(self, f: int, j: int, decoy: bool) -> int:
        return (f * self.J + j) * 2 + int(decoy)

    def replicate_groups(self, j: int, s_index: int) -> list:
        """Groups of replicate (decoy set j, entrapment family 1 + s_index)."""
        return [self.group(f, j, d) for f in (0, 1 + s_index) for d in (False, True)]


def _dev_i64(a, dev) -> torch.Tensor:
    return a.long() if isinstance(a, torch.Tensor) else torch.as_tensor(np.asarray(a, np.int64), device=dev)


def _lex_order(*keys):
    """Permutation sorting by keys[0] (major) .. keys[-1] (minor), ascending; ties keep their input order."""
    o = torch.argsort(keys[-1], stable=True)
    for k in reversed(keys[:-1]):
        o = o[torch.argsort(k[o], stable=True)]
    return o


def _segment_rank(sorted_keys: torch.Tensor) -> torch.Tensor:
    """Position of every element inside its run of equal keys (keys sorted)."""
    first = torch.searchsorted(sorted_keys, sorted_keys, side="left")
    return torch.arange(len(sorted_keys), device=sorted_keys.device) - first


def select_family(o_h, o_r, o_c, d_spec, hyp, fi: FamilyIndex, p: E.SearchParams, dev):
    """Per-family top-K, closure over all decoy sets, canonical hypothesis. Returns candidates sorted by
    (spectrum, row): (h, r, k1 count (-1: closure only), shortlisted, spectrum, (spectrum << ROW_BITS) | row)."""
    h = o_h.long()
    r = o_r.long()
    c = o_c.long()
    spec = d_spec[h].long()
    grp = spec * 256 + fi.kfam[r]
    rank_key = ((32767 - c.clamp(max=32767)) << ROW_BITS) | ((1 << ROW_BITS) - 1 - r)   # count desc, row desc
    o = _lex_order(grp, rank_key, h)
    keep = o[_segment_rank(grp[o]) < p.topk]
    h, r, c = h[keep], r[keep], c[keep]
    hs = [h] + [h] * fi.J
    rs = [r] + [fi.partners[j][r] for j in range(fi.J)]
    cs = [c] + [torch.full_like(c, -1)] * fi.J
    ts = [torch.ones_like(r)] + [torch.zeros_like(r)] * fi.J
    h, r, c, t = torch.cat(hs), torch.cat(rs), torch.cat(cs), torch.cat(ts)
    key = (d_spec[h].long() << ROW_BITS) | r
    uk, inv = torch.unique(key, sorted=True, return_inverse=True)
    n = len(uk)
    hsrc = torch.full((n,), 1 << 62, dtype=torch.int64, device=dev).scatter_reduce_(0, inv, h, reduce="amin")
    cc = torch.full((n,), -1, dtype=torch.int64, device=dev).scatter_reduce_(0, inv, c, reduce="amax")
    tk = torch.zeros(n, dtype=torch.int64, device=dev).scatter_reduce_(0, inv, t, reduce="amax")
    us = uk >> ROW_BITS
    ur = uk & ((1 << ROW_BITS) - 1)
    hspec, hlo, hhi = (_dev_i64(hyp[k], dev) for k in ("spec", "lo", "hi"))
    h0 = torch.searchsorted(hspec, us, side="left")
    h1 = torch.searchsorted(hspec, us, side="right")
    hcan = torch.full((n,), -1, dtype=torch.int64, device=dev)
    kmax = int((h1 - h0).max().item()) if n else 0
    for k in range(kmax - 1, -1, -1):            # the lowest hypothesis whose window holds the row wins
        hk = h0 + k
        ok = hk < h1
        hk = torch.where(ok, hk, torch.zeros_like(hk))
        ok &= (hlo[hk] <= ur) & (ur < hhi[hk])
        hcan = torch.where(ok, hk, hcan)
    hcan = torch.where(hcan >= 0, hcan, hsrc)
    return hcan, ur, cc, tk.bool(), us, uk


def select_family_join(ds: "E.DeviceSearch", fi: FamilyIndex, n: int):
    """select_family on the device (vh_family_select) from a DeviceSearch whose front and join K1 counts have
    run: the same candidates as select_family over the appended K1 list, in (spectrum, row) order."""
    import ctypes as C

    from . import _lib as K
    L = K.lib()
    dev = ds.gi.dev
    j, hy = ds.join, ds.hyp
    if fi.kfam_u8 is None:
        fi.kfam_u8 = fi.kfam.to(torch.uint8).contiguous()
    out_off = torch.empty(n + 1, dtype=torch.int64, device=dev)
    ws = torch.empty(max(int(L.vh_family_select_workspace_bytes(n)), 1), dtype=torch.uint8, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    args = (C.byref(ds.vp), n, K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]), K.ptr(j["h_slot"]),
            j["slot_cap"], K.ptr(j["counts"]), K.ptr(fi.kfam_u8), ds.gi.n, K.ptr(fi.partners), fi.J)
    K.check(L.vh_family_select(*args, 0, K.ptr(out_off), None, None, None, None, K.ptr(status), K.ptr(ws),
                               ws.numel(), K.stream()), "family_select (offsets)")
    tot = int(out_off[n].item())
    o_row = torch.empty(max(tot, 1), dtype=torch.int32, device=dev)
    o_hyp = torch.empty_like(o_row)
    o_cnt = torch.empty_like(o_row)
    o_top = torch.empty(max(tot, 1), dtype=torch.uint8, device=dev)
    K.check(L.vh_family_select(*args, 1, K.ptr(out_off), K.ptr(o_row), K.ptr(o_hyp), K.ptr(o_cnt), K.ptr(o_top),
                               K.ptr(status), K.ptr(ws), ws.numel(), K.stream()), "family_select")
    if int(status.item()):
        raise RuntimeError("family_select: a spectrum exceeded the per-spectrum limits (64 hypotheses, 256 rows "
                           "over the K1 threshold, 1,024 shortlist + partner entries) or a window overflowed")
    us = torch.repeat_interleave(torch.arange(n, device=dev), out_off[1:] - out_off[:-1])
    ur = o_row[:tot].long()
    return o_hyp[:tot].long(), ur, o_cnt[:tot].long(), o_top[:tot].bool(), us, (us << ROW_BITS) | ur


def search_family(sp: E.Spectra, gi: E.GpuIndex, l1, fi: FamilyIndex, p: E.SearchParams, select: str = "kernel"):
    """Closed search with family top-2 retention. Returns (table, stats); one table row per (spectrum, retention
    group, grank in {1, 2}). select: "kernel" (join K1 counts + vh_family_select) or "torch" (append-form K1 +
    device sorts; the same candidates)."""
    if select == "kernel":
        return _search_family_kernel(sp, gi, l1, fi, p)
    if select != "torch":
        raise ValueError(select)
    st = E.Stats()
    dev = gi.dev
    with E.Timer(st, "k0"):
        peaks = E.k0_preprocess(sp, p, dev, st, gi)
    with E.Timer(st, "hyp"):              # vh_hypotheses: the host rule (host.py:64-91) on the device
        hd = E.device_hypotheses(gi, peaks["raw"], peaks, p)
        n_h = int(hd["hyp_off"][len(sp)].item())
        hyp = {k: hd[k][:n_h] for k in ("spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz")}
        hyp["hyp_off"] = hd["hyp_off"][:len(sp) + 1]
    with E.Timer(st, "k1"):
        o_h, o_r, o_c, scored, matched, d_spec, d_maxfz = E.run_k1(gi, hyp, peaks, p, st)
    with E.Timer(st, "topk"):
        h, r, c, intop, spec, ukey = select_family(o_h, o_r, o_c, d_spec, hyp, fi, p, dev)
    with E.Timer(st, "k2"):
        k2 = E.run_k2(gi, h, r, d_spec, d_maxfz, peaks, p)
    with E.Timer(st, "features"):
        tbl = family_table(sp, l1, fi, hyp, h, r, intop, spec, ukey, k2, p)
    st.counts.update(n_spectra=len(sp), n_searched=int(np.asarray(peaks["valid"]).sum()), n_hypotheses=n_h,
                     n_rescored=int(len(r)), n_rows=tbl.num_rows)
    return tbl, st


def _search_family_kernel(sp, gi, l1, fi: FamilyIndex, p):
    st = E.Stats()
    dev = gi.dev
    n = len(sp)
    with E.Timer(st, "upload"):
        raw = E.upload_spectra(sp, dev, st)
    ds = fi.searches.get((id(gi), n))      # preallocated buffers, reused by searches of the same size
    if ds is None:
        ds = E.DeviceSearch(gi, p, n, k1_impl="join")
        fi.searches = {(id(gi), n): ds}
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
        h, r, c, intop, spec, ukey = select_family_join(ds, fi, n)
    n_h = int(ds.hyp["hyp_off"][n].item())
    hyp = {k: ds.hyp[k][:n_h] for k in ("spec", "z", "iso", "lo", "hi", "maxfz")}
    with E.Timer(st, "k2"):
        k2 = E.run_k2(gi, h, r, hyp["spec"], hyp["maxfz"], ds.peaks, p)
    with E.Timer(st, "features"):
        tbl = family_table(sp, l1, fi, hyp, h, r, intop, spec, ukey, k2, p)
    st.counts.update(n_spectra=n, n_hypotheses=n_h, n_rescored=int(len(r)), n_rows=tbl.num_rows)
    return tbl, st


def family_table(sp, l1, fi: FamilyIndex, hyp, h, r, intop, spec, ukey, k2, p) -> pa.Table:
    dev = r.device
    ok = (k2["matched_b"].long() + k2["matched_y"].long()) >= p.min_matched_peaks
    side = fi.side[r]
    ff = fi.fwd_fam[r]
    idx_l, g_l = [], []
    for j in range(fi.J):
        # admissible in decoy context j: shortlisted, or the set-j partner of a shortlisted row
        pk = (spec << ROW_BITS) | fi.partners[j][r]
        pos = torch.searchsorted(ukey, pk).clamp(max=max(len(ukey) - 1, 0))
        adm = ok & (intop | ((ukey[pos] == pk) & intop[pos]))
        for m, d in ((adm & (side < 0), 0), (adm & (side == j), 1)):
            ii = torch.nonzero(m).squeeze(1)
            idx_l.append(ii)
            g_l.append((ff[ii] * fi.J + j) * 2 + d)
    idx = torch.cat(idx_l)
    g = torch.cat(g_l)
    s = spec[idx]
    o = _lex_order(s * 256 + g, -k2["hyperscore"][idx], r[idx])
    idx, g, s = idx[o], g[o], s[o]
    rank = _segment_rank(s * 256 + g)
    keep = rank < 2
    idx, g, s, rank = idx[keep], g[keep], s[keep], rank[keep]
    hh = h[idx]
    d = E._to_host({"spec": s, "row": r[idx], "fgroup": g.to(torch.int16), "grank": (rank + 1).to(torch.int8),
                    "hs": k2["hyperscore"][idx], "z": _dev_i64(hyp["z"], dev)[hh].to(torch.int8),
                    "iso": _dev_i64(hyp["iso"], dev)[hh].to(torch.int8), "mb": k2["matched_b"][idx],
                    "my": k2["matched_y"][idx]})
    sn, rows = d["spec"], d["row"]
    return pa.table({
        "run": E.string_column(sp.run, len(sn)),
        "spec_idx": sn.astype(np.int64), "scan": sp.scan[sn],
        "fgroup": d["fgroup"], "grank": d["grank"],
        "pf_row": rows.astype(np.int64), "pf_id": l1.pf_id[rows],
        "hyperscore": d["hs"], "charge": d["z"], "isotope_error": d["iso"],
        "matched_b": d["mb"], "matched_y": d["my"],
    })


def reconstruct(tbl: pa.Table, groups, n: int = 2) -> dict:
    """A replicate's competition from a family table: per spectrum the top-n over `groups` by (hyperscore desc,
    row asc). Returns numpy arrays sorted by (spec_idx, rank)."""
    g = tbl.column("fgroup").to_numpy()
    m = np.isin(g, list(groups))
    s = tbl.column("spec_idx").to_numpy()[m]
    hs = tbl.column("hyperscore").to_numpy()[m]
    row = tbl.column("pf_row").to_numpy()[m]
    pid = tbl.column("pf_id").to_numpy()[m]
    o = np.lexsort((row, -hs, s))
    s, hs, row, pid = s[o], hs[o], row[o], pid[o]
    rank = np.arange(len(s)) - np.searchsorted(s, s, side="left")
    k = rank < n
    return {"spec_idx": s[k], "rank": rank[k] + 1, "hyperscore": hs[k], "pf_id": pid[k], "pf_row": row[k]}


class SubIndex:
    """The rows of one replicate (order kept, so mass order and row tie-breaks are kept) with the pair partner
    of decoy set j renumbered: the index of a separate search of that replicate."""

    def __init__(self, l1, keep: np.ndarray, j: int):
        rows = np.flatnonzero(keep)
        remap = np.full(len(l1), -1, np.int64)
        remap[rows] = np.arange(len(rows))
        self.rows = rows
        self.path = getattr(l1, "path", None)
        self.meta = l1.meta
        for k in ("pf_id", "mass32", "seq_off", "length", "mask", "n_var", "missed"):
            setattr(self, k, getattr(l1, k)[rows])
        self.residues = l1.residues
        pj = remap[l1.partners[j][rows]]
        if (pj < 0).any():
            raise ValueError("a partner lies outside the replicate's rows")
        self.partner = pj
        self.partners = [pj]

    def __len__(self):
        return len(self.rows)
