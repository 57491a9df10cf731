"""Pins of the family-retention oracle (oracle/family.py) on the CPU, without the GPU:
  * the reference's defining property (tests/gpu/test_family.py:45-67 of the reference): every (decoy set x
    entrapment seed) replicate rebuilt from one search of the multi-replicate table equals a separate search
    of that replicate's rows -- here both searches are the oracle's;
  * with top-K at least the window size every row over the K1 threshold is shortlisted, and on a plain
    target-decoy replicate the family rank 1 is the plain search's rank 1 (oracle.search);
  * a hand-built shortlist case checked entry by entry (top-K per family, closure, dedupe, canonical hypothesis)."""
import numpy as np
import pytest

from conftest import STATIC


def _O():
    from oracle import oracle as O
    return O


@pytest.fixture(scope="module")
def multi_cpu():
    from vulcan.index import build as B
    from vulcan.index import synth
    ip = synth.WORKLOADS[1].index_params()
    l1, lab = B.build_multi_decoy_tables(synth.proteome(200, 13),
                                         {"entrap_seeds": [2001, 2002], "decoy_seeds": [1001, 1002], **ip})
    sp, _, _ = synth.spectra(l1, lab, 1, n=400, seed=9, index_params=ip)
    return l1, lab, sp


def _codes(lab):
    fam, seed = np.asarray(lab["family"]), np.asarray(lab["entrap_seed"])
    ent = (fam % 2) == 1
    seeds = sorted(set(seed[ent].tolist()))
    fwd = np.zeros(len(fam), np.int64)
    for i, s in enumerate(seeds):
        fwd[ent & (seed == s)] = 1 + i
    return fwd, np.asarray(lab["decoy_set"], np.int64), len(seeds) + 1


def _family_search(l1, partners, fwd, side, sp, p):
    from oracle import family as OF
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum
    from vulcan.index import synth
    O = _O()
    vp = make_params(p, STATIC, dict(synth.WORKLOADS[1].var_mods))
    hix = O.HostIndex(l1)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    J = len(partners)
    sl = OF.shortlist(hyp, k1, fwd * (J + 1) + side + 1, partners, p.topk, p.k1_threshold)
    off = sl["off"]
    k2 = O.k2(vp, hix, pk, sl["row"], sl["hyp"], hyp["maxfz"], grp_start=off[:-1], grp_count=np.diff(off))
    t2 = OF.top2(sl, k2["hyperscore"], k2["matched_b"], k2["matched_y"], fwd, side, partners, p.min_matched_peaks)
    i = t2["idx"]
    return {"spec": t2["spec"], "group": t2["group"], "row": sl["row"][i], "hs": k2["hyperscore"][i],
            "pf_id": np.asarray(l1.pf_id)[sl["row"][i]]}, sl, hyp, k1


def _reconstruct(t, groups, n=2):
    m = np.isin(t["group"], groups)
    s, hs, row, pid = t["spec"][m], t["hs"][m], t["row"][m], t["pf_id"][m]
    o = np.lexsort((row, -hs, s))
    s, hs, pid = s[o], hs[o], pid[o]
    rank = np.arange(len(s)) - np.searchsorted(s, s, side="left")
    k = rank < n
    return s[k], rank[k], hs[k], pid[k]


def test_replicates_rebuilt_from_one_search_equal_separate_searches(multi_cpu):
    from vulcan.index import synth
    l1, lab, sp = multi_cpu
    p = synth.WORKLOADS[1].search_params()
    fwd, side, F = _codes(lab)
    partners = np.stack(l1.partners)
    full, _, _, _ = _family_search(l1, partners, fwd, side, sp, p)
    J = len(partners)
    for j in range(J):
        for si in range(F - 1):
            keep = np.isin(fwd, [0, 1 + si]) & np.isin(side, [-1, j])
            rows = np.flatnonzero(keep)
            new_of = np.full(len(l1), -1, np.int64)
            new_of[rows] = np.arange(len(rows))
            from vulcan.index.build import L1Index
            sub = L1Index(arrays={k: np.asarray(getattr(l1, k))[rows] for k in ("pf_id", "mass32", "seq_off",
                                                                              "length", "mask", "n_var", "missed")}
                          | {"partner": new_of[partners[j][rows]], "residues": l1.residues})
            sfwd = np.where(fwd[rows] == 0, 0, 1)
            sside = np.where(side[rows] >= 0, 0, -1)
            sep, _, _, _ = _family_search(sub, np.stack([sub.partner]), sfwd, sside, sp, p)
            groups = [(f * J + j) * 2 + d for f in (0, 1 + si) for d in (0, 1)]
            a = _reconstruct(sep, [0, 1, 2, 3])
            b = _reconstruct(full, groups)
            assert len(a[0]) > 200, (j, si)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (j, si)


def test_wide_topk_family_rank1_is_the_plain_rank1():
    """T + one reversed decoy set, one K1 family per side, top-K beyond every window: the family shortlist holds
    every row over the threshold and its partner, so its rank 1 (groups 0 and 1) is oracle.search's rank 1 with
    the same top-K."""
    from vulcan.gpu._abi import make_params
    from vulcan.index import synth
    O = _O()
    l1, lab = synth.build_index(1, n_proteins=300)
    sp, _, _ = synth.spectra(l1, lab, 1, n=500)
    p = synth.WORKLOADS[1].search_params()
    p.topk = 5000
    side = np.where(np.asarray(lab["is_decoy"]), 0, -1)
    t, sl, hyp, k1 = _family_search(l1, np.stack([l1.partner]), np.zeros(len(l1), np.int64), side, sp, p)
    w = np.maximum(hyp["hi"] - hyp["lo"], 0)
    assert w.max() < p.topk
    assert sl["top"].sum() == int((k1["counts"] >= 1).sum()) - _dup_rows(hyp, k1)
    s, rank, hs, pid = _reconstruct(t, [0, 1], n=1)
    ref = O.search(make_params(p, STATIC, dict(synth.WORKLOADS[1].var_mods)), O.HostIndex(l1), sp)
    got = dict(zip(s.tolist(), pid.tolist()))
    want = {i: int(l1.pf_id[r]) for i, r in enumerate(ref["row"]) if r >= 0}
    assert got == want and len(want) > 300


def _dup_rows(hyp, k1):
    """(spectrum, row) pairs over the threshold in more than one window, counted beyond their first."""
    lo, hi = hyp["lo"], hyp["hi"]
    w = np.maximum(hi - lo, 0)
    hh = np.repeat(np.arange(len(lo)), w)
    rr = np.repeat(lo, w) + (np.arange(int(w.sum())) - np.repeat(k1["win_off"][:-1], w))
    m = k1["counts"] >= 1
    key = (hyp["spec"][hh[m]].astype(np.int64) << 28) | rr[m]
    return len(key) - len(np.unique(key))


def test_shortlist_hand_built_case():
    """Two spectra, three hypotheses, windows of 4 rows, K1 families alternating by row, top-2, one decoy set
    (partner = row ^ 1): every entry written out by hand."""
    from oracle import family as OF
    hyp = {"hyp_off": np.array([0, 2, 3]), "spec": np.array([0, 0, 1]), "lo": np.array([10, 12, 20]),
           "hi": np.array([14, 16, 24])}
    counts = np.array([5, 0, 5, 7,   3, 7, 1, 2,   4, 4, 4, 9], np.int64)     # window order
    k1 = {"win_off": np.array([0, 4, 8, 12]), "counts": counts}
    kfam = np.arange(32) % 2                                                      # even rows family 0, odd 1
    partners = np.array([np.arange(32) ^ 1])
    sl = OF.shortlist(hyp, k1, kfam, partners, topk=2, thr=1)
    # spectrum 0 entries (hyp, row, count): (0,10,5) (0,12,5) (0,13,7) (1,12,3) (1,13,7) (1,14,1) (1,15,2)
    #   family 0 (even rows): (12,5,h0) (10,5,h0) (12,3,h1) (14,1,h1) -> count desc, row desc: (12,5) (10,5)
    #   family 1 (odd rows): (13,7,h0) (13,7,h1) (15,2) -> top-2: (13,7,h0) (13,7,h1)
    #   kept rows 12, 10, 13 (twice); partners 13, 11, 12
    # spectrum 1: rows 20..23 counts 4 4 4 9: family 0 rows 20, 22 (4, 4) -> both; family 1: 23 (9), 21 (4)
    assert sl["off"].tolist() == [0, 4, 8]
    assert sl["row"].tolist() == [10, 11, 12, 13, 20, 21, 22, 23]
    assert sl["cnt"].tolist() == [5, -1, 5, 7, 4, 4, 4, 9]
    assert sl["top"].tolist() == [True, False, True, True, True, True, True, True]
    # canonical: lowest hypothesis whose window holds the row (row 11 is in hypothesis 0's window [10, 14))
    assert sl["hyp"].tolist() == [0, 0, 0, 0, 2, 2, 2, 2]
