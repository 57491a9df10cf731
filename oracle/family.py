"""CPU ORACLE -- test infrastructure, not product code.

Numpy restatement of the family top-2 retention of the reference's vulcan/gpu/family.py, the checker of the
device kernels vh_family_select / vh_family_wide* / vh_family_top2. Only tests/ import it.

shortlist  family.py:90-136 (select_family): over every precursor-window row of every hypothesis with K1 count
           >= threshold (counts in the layout of oracle.k1_counts), the best `topk` of each (spectrum, K1 family)
           by (count desc, row desc) -- counts clamped at 32,767 in the key like engine.select_topk
           (engine.py:291-296); identical keys from two hypotheses keep the lower hypothesis first --; each kept
           row adds its pair partner in every decoy set (count -1, not shortlisted); one entry per (spectrum,
           row): lowest source hypothesis, highest count, shortlisted if any source was; the canonical
           hypothesis is the spectrum's lowest hypothesis whose window holds the row, else the lowest source.
top2       family.py:160-203 (family_table): an entry with matched_b + matched_y >= min_matched is admissible in
           decoy context j when it is shortlisted or its set-j partner is a shortlisted entry of the spectrum;
           forward rows join group (fwd_fam * J + j) * 2, rows of decoy set j group (fwd_fam * J + j) * 2 + 1;
           per (spectrum, group) the best two by (hyperscore desc, row asc).
"""
from __future__ import annotations

import numpy as np

ROW_BITS = 28


def _window_entries(hyp: dict, k1: dict):
    """(hypothesis, row, count) of every precursor-window row, in the oracle's window order."""
    lo, hi = np.asarray(hyp["lo"], np.int64), np.asarray(hyp["hi"], np.int64)
    w = np.maximum(hi - lo, 0)
    hh = np.repeat(np.arange(len(lo)), w)
    rr = np.repeat(lo, w) + (np.arange(int(w.sum())) - np.repeat(np.asarray(k1["win_off"][:-1]), w))
    return hh, rr, np.asarray(k1["counts"], np.int64)


def shortlist(hyp: dict, k1: dict, kfam: np.ndarray, partners: np.ndarray, topk: int, thr: int = 1) -> dict:
    """Per spectrum, entries in row order: {"off" (n_spec + 1), "row", "hyp" (canonical, global index),
    "cnt" (-1: partner only), "top" (shortlisted)}."""
    n_spec = len(hyp["hyp_off"]) - 1
    spec_of_h = np.asarray(hyp["spec"], np.int64)
    hh, rr, cc = _window_entries(hyp, k1)
    m = cc >= thr
    hh, rr, cc = hh[m], rr[m], np.minimum(cc[m], 32767)
    ss = spec_of_h[hh]
    kf = np.asarray(kfam, np.int64)[rr]
    o = np.lexsort((hh, -rr, -cc, kf, ss))           # (spectrum, family), count desc, row desc, hypothesis asc
    ss, kf, hh, rr, cc = ss[o], kf[o], hh[o], rr[o], cc[o]
    grp = ss * 256 + kf
    rank = np.arange(len(grp)) - np.searchsorted(grp, grp, side="left")
    keep = rank < topk
    ss, hh, rr, cc = ss[keep], hh[keep], rr[keep], cc[keep]
    J = len(partners)
    e_s = np.concatenate([ss] * (J + 1))
    e_h = np.concatenate([hh] * (J + 1))
    e_r = np.concatenate([rr] + [np.asarray(partners[j], np.int64)[rr] for j in range(J)])
    e_c = np.concatenate([cc] + [np.full(len(cc), -1, np.int64)] * J)
    e_t = np.concatenate([np.ones(len(cc), bool)] + [np.zeros(len(cc), bool)] * J)
    key = (e_s << ROW_BITS) | e_r
    uk, inv = np.unique(key, return_inverse=True)
    n = len(uk)
    hsrc = np.full(n, np.iinfo(np.int64).max)
    np.minimum.at(hsrc, inv, e_h)
    cmax = np.full(n, -1, np.int64)
    np.maximum.at(cmax, inv, e_c)
    top = np.zeros(n, bool)
    np.logical_or.at(top, inv, e_t)
    us, ur = uk >> ROW_BITS, uk & ((1 << ROW_BITS) - 1)
    hoff = np.asarray(hyp["hyp_off"], np.int64)
    lo, hi = np.asarray(hyp["lo"], np.int64), np.asarray(hyp["hi"], np.int64)
    hcan = hsrc.copy()
    for i in range(n):
        for h in range(hoff[us[i]], hoff[us[i] + 1]):
            if lo[h] <= ur[i] < hi[h]:
                hcan[i] = h
                break
    off = np.zeros(n_spec + 1, np.int64)
    np.cumsum(np.bincount(us, minlength=n_spec)[:n_spec], out=off[1:])
    return {"off": off, "row": ur, "hyp": hcan, "cnt": cmax, "top": top}


def top2(sl: dict, hs: np.ndarray, mb: np.ndarray, my: np.ndarray, fwd_fam: np.ndarray, side: np.ndarray,
         partners: np.ndarray, min_matched: int) -> dict:
    """Top-2 per (spectrum, group) over a shortlist `sl` with K2 results per entry. Returns {"spec", "group",
    "rank" (1 or 2), "idx" (entry index)} sorted by (spectrum, group, rank)."""
    off = sl["off"]
    n_spec = len(off) - 1
    rows, top = sl["row"], sl["top"]
    spec = np.repeat(np.arange(n_spec), np.diff(off))
    ok = (np.asarray(mb, np.int64) + np.asarray(my, np.int64)) >= min_matched
    key = (spec << ROW_BITS) | rows
    sd = np.asarray(side, np.int64)[rows]
    ff = np.asarray(fwd_fam, np.int64)[rows]
    J = len(partners)
    idx_l, g_l = [], []
    for j in range(J):
        pk = (spec << ROW_BITS) | np.asarray(partners[j], np.int64)[rows]
        pos = np.minimum(np.searchsorted(key, pk), max(len(key) - 1, 0))
        p_in = (key[pos] == pk) & top[pos] if len(key) else np.zeros(0, bool)
        adm = ok & (top | p_in)
        for m, d in ((adm & (sd < 0), 0), (adm & (sd == j), 1)):
            ii = np.flatnonzero(m)
            idx_l.append(ii)
            g_l.append((ff[ii] * J + j) * 2 + d)
    idx = np.concatenate(idx_l) if idx_l else np.zeros(0, np.int64)
    g = np.concatenate(g_l) if g_l else np.zeros(0, np.int64)
    s = spec[idx]
    o = np.lexsort((rows[idx], -np.asarray(hs)[idx], g, s))
    idx, g, s = idx[o], g[o], s[o]
    gk = s * 65536 + g
    rank = np.arange(len(gk)) - np.searchsorted(gk, gk, side="left")
    k = rank < 2
    return {"spec": s[k], "group": g[k], "rank": rank[k] + 1, "idx": idx[k]}
