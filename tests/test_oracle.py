"""Pins the CPU oracle (oracle/vh_oracle.cpp) before it is trusted as the checker.

Each check is chosen so that a plausible restatement mistake fails it: closed-form known-answer spectra built
from f64 residue masses (an independent route to the fragment ladder), the Stirling closed form of lnfact and its
lgamma bound, brute-force Python orderings on tiny inputs, the reference's own invariants (K1 index == brute
force, pair closure symmetry, label blindness) and the reference's decoy known answer."""
import math

import numpy as np
import pytest

from conftest import STATIC


def _O():
    from oracle import oracle as O
    return O


def _vp(p, static=STATIC, var=None):
    from vulcan.gpu._abi import make_params
    return make_params(p, static, var or {})


def _spectra(peaks, pmz, z):
    from vulcan.gpu.host import Spectra
    n = len(peaks)
    mz = np.concatenate([np.asarray(m, np.float32) for m, _ in peaks]) if n else np.zeros(0, np.float32)
    it = np.concatenate([np.asarray(i, np.float32) for _, i in peaks]) if n else np.zeros(0, np.float32)
    off = np.zeros(n + 1, np.int64)
    np.cumsum([len(m) for m, _ in peaks], out=off[1:])
    return Spectra("t", np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.zeros(n),
                   np.asarray(pmz, np.float64), np.asarray(z, np.int8), mz, it, off)


def _tiny_index(peps, static=STATIC, var=None, max_var=0):
    from vulcan.index.build import build_tables
    prm = {"enzyme": {"cleave_at": "KR", "restrict": "P", "c_terminal": True, "missed_cleavages": 0, "min_len": 5,
                      "max_len": 40},
           "static_mods": static, "variable_mods": var or {}, "max_variable_mods": max_var,
           "peptide_min_mass": 100.0, "peptide_max_mass": 10000.0}
    return build_tables(["".join(peps)], prm)


def _row_of(l1, seq):
    """Row of a sequence in a table built by the I/L-collapsing digest."""
    key = seq.replace("I", "L")
    return [j for j in range(len(l1)) if l1.sequence(j) == key][0]


def _f64_ladder(seq, static=STATIC):
    from vulcan.index._core import H2O_F64, RESIDUE_F64
    rm = np.array([RESIDUE_F64[ord(c) - 65] + static.get(c, 0.0) for c in seq])
    b = np.cumsum(rm)[:-1]
    M = H2O_F64 + rm.sum()
    return b, M - b, M


# ------------------------------------------------------------------------------------------------ constants / ids
def test_decoy_known_answer_and_xxh64():
    import xxhash

    from vulcan.index import _core
    from vulcan.index.decoys import hash_shuffle_perm, pseudo_reverse
    assert pseudo_reverse("PEPTIDEK") == "EDITPEPK"            # tests/index/test_decoys.py:30 of the reference
    q = hash_shuffle_perm("PEPTIDEK", 2001, 0)
    assert q[-1] == 7 and sorted(q.tolist()) == list(range(8))
    for s in (b"", b"A", b"PEPTIDEK", b"M[+15.9949]PEPTIDEK" * 3):
        for seed in (0, 1, 2 ** 40):
            assert _core.xxh64(s, seed) == xxhash.xxh64(s, seed=seed).intdigest()


def test_mass32_close_to_f64_literature_mass():
    """Sage f32 mass of PEPTIDEK vs its monoisotopic mass 927.4549 (f64 residue sum)."""
    l1, _ = _tiny_index(["PEPTIDEK"])
    _, _, M = _f64_ladder("PEPTIDEK")
    assert abs(M - 927.45493) < 1e-3
    i = _row_of(l1, "PEPTIDEK")
    assert abs(float(l1.mass32[i]) - M) < 2e-3


def test_lnfact_closed_form_and_bound():
    O = _O()
    assert O.lnfact(0) == 1.0                                  # Sage quirk (kernels.cu:247)
    for n in range(1, 300):
        x = float(n)
        want = x * math.log(x) - x + 0.5 * math.log(x) + 0.5 * math.log(2 * math.pi * x)
        assert abs(O.lnfact(n) - want) <= 1e-12 * max(1.0, abs(want))
        # Sage's form equals (n+1) ln n - n + ln(2 pi)/2 = ln n! + ln(n)/2 - 1/(12n) + O(n^-3) (Stirling series),
        # so the gap to lgamma lies in (0, 1/(12n)): a dropped or wrong term breaks this bound
        gap = math.lgamma(n + 1) + 0.5 * math.log(n) - O.lnfact(n)
        assert 0.0 < gap < 1.0 / (12 * n) + 1e-12


# ------------------------------------------------------------------------------------------------ K0
def _k0_python(mz, it, fmin, fmax, kmax):
    P32 = np.float32(1.0072764)
    idx = [i for i in range(len(mz)) if not np.isnan(mz[i]) and fmin <= mz[i] <= fmax]
    mass = {i: np.float32(mz[i] - P32) for i in idx}

    def ikey(v):   # NaN intensities first in descending order
        return (0, 0.0) if np.isnan(v) else (1, -float(v))
    sel = sorted(idx, key=lambda i: (ikey(it[i]), -float(mass[i]), i))[:kmax]
    rank = {i: r for r, i in enumerate(sel)}
    fin = sorted(sel, key=lambda i: (float(mass[i]), rank[i]))
    return np.array([mass[i] for i in fin], np.float32), np.array([it[i] for i in fin], np.float32)


def test_k0_matches_python_brute_force():
    from vulcan.gpu.host import SearchParams
    rng = np.random.default_rng(3)
    peaks = []
    for n in (0, 1, 5, 40, 200, 700):
        mz = rng.uniform(100, 2100, n).astype(np.float32)
        it = rng.integers(1, 6, n).astype(np.float32)        # many intensity ties
        if n >= 40:
            mz[:10] = mz[10:20]                               # mass ties
            it[3] = np.nan
            mz[5] = np.nan
            mz[6] = 150.0
            mz[7] = 2000.0
        peaks.append((mz, it))
    sp = _spectra(peaks, [500.0] * len(peaks), [2] * len(peaks))
    p = SearchParams(max_peaks=150, min_peaks=3)
    out = _O().k0(_vp(p), sp)
    for s, (mz, it) in enumerate(peaks):
        m_ref, i_ref = _k0_python(mz, it, 150.0, 2000.0, 150)
        a, b = out["off"][s], out["off"][s + 1]
        assert np.array_equal(out["mass"][a:b], m_ref)
        assert np.array_equal(out["int"][a:b], i_ref, equal_nan=True)
        assert out["valid"][s] == (len(m_ref) >= 3)


# ------------------------------------------------------------------------------------------------ hypotheses
def test_hypotheses_window_boundaries_and_order():
    """searchsorted left on the low bound, right on the high bound: rows exactly at either bound are inside."""
    from vulcan.gpu._abi import max_hyp_per_spectrum
    from vulcan.gpu.host import PROTON, NEUTRON, SearchParams, make_hypotheses
    p = SearchParams(precursor_charge=(2, 3), isotope_errors=(0, 1))
    pmz = np.array([600.3, 712.9, np.nan, 800.0])
    z = np.array([2, 0, 2, 3], np.int8)
    sp = _spectra([(np.zeros(0), np.zeros(0))] * 4, pmz, z)
    valid = np.array([1, 1, 1, 0], bool)
    M = np.float32(np.float32(np.float32(600.3) - PROTON) * np.float32(2))
    lo = np.float32(M + (M * np.float32(-10.0)) / np.float32(1e6))
    hi = np.float32(M + (M * np.float32(10.0)) / np.float32(1e6))
    mass32 = np.sort(np.array([lo, hi, np.nextafter(lo, -np.inf, dtype=np.float32),
                               np.nextafter(hi, np.inf, dtype=np.float32), 500.0, 3000.0], np.float32))
    O = _O()
    h = O.hypotheses(_vp(p), sp, valid, mass32, max_hyp_per_spectrum(p))
    assert h["spec"].tolist() == [0, 0, 1, 1, 1, 1]
    assert h["z"].tolist() == [2, 2, 2, 2, 3, 3] and h["iso"].tolist() == [0, 1, 0, 1, 0, 1]
    assert h["hyp_off"].tolist() == [0, 2, 6, 6, 6]
    assert mass32[h["lo"][0]] == lo and mass32[h["hi"][0] - 1] == hi and h["hi"][0] - h["lo"][0] == 2
    assert h["M"][1] == np.float32(M - np.float32(NEUTRON))
    ref = make_hypotheses(sp, valid, mass32, p)                # the host restatement agrees field by field
    for k in ("spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz"):
        assert np.array_equal(h[k], ref[k]), k


# ------------------------------------------------------------------------------------------------ K1 / K2 KATs
def _exact_spectrum(seq, inten=10.0, zf=(1,)):
    b, y, M = _f64_ladder(seq)
    P = 1.0072764
    mz = np.concatenate([(np.concatenate([b, y]) + z * P) / z for z in zf])
    o = np.argsort(mz)
    return mz[o], np.full(len(mz), inten), (M + 2 * P) / 2


def test_k2_known_answer_exact_ladder():
    """A spectrum holding exactly the b/y 1+ ions of PEPTIDEK: all 7 b and 7 y match, longest ladders = L-2 (the
    run bookkeeping never counts ion index 0), hyperscore = ln((7I+1)^2) + 2 lnfact(7)."""
    from vulcan.gpu.host import SearchParams
    O = _O()
    l1, _ = _tiny_index(["PEPTIDEK"])
    r = _row_of(l1, "PEPTIDEK")
    mz, it, pmz = _exact_spectrum("PEPTIDEK")
    sp = _spectra([(mz, it)], [pmz], [2])
    p = SearchParams(fragment_unit="da", fragment_ppm=(-0.02, 0.02), min_peaks=1, fragment_min_mz=50.0)
    vp = _vp(p)
    pk = O.k0(vp, sp)
    hix = O.HostIndex(l1)
    k2 = O.k2(vp, hix, pk, [r], [0], np.array([2], np.int8), grp_count=[1], grp_start=[0])
    assert k2["matched_b"][0] == 7 and k2["matched_y"][0] == 7
    assert k2["longest_b"][0] == 6 and k2["longest_y"][0] == 6
    assert k2["summed_b"][0] == 70.0 and k2["summed_y"][0] == 70.0
    want = math.log(71.0 * 71.0) + 2 * (7 * math.log(7) - 7 + 0.5 * math.log(7) + 0.5 * math.log(2 * math.pi * 7))
    assert abs(k2["hyperscore"][0] - want) < 1e-12
    assert abs(k2["ppm"][0]) < 1.0                                  # f32 vs f64 ladder: sub-ppm


def test_k2_tie_goes_to_the_later_peak_and_ppm_weighting():
    """Two equally intense peaks inside one window: the later (heavier) one is matched (Sage's `>=`), and the
    intensity-weighted ppm uses its mass."""
    from vulcan.gpu.host import SearchParams
    O = _O()
    l1, _ = _tiny_index(["PEPTIDEK"])
    r = _row_of(l1, "PEPTIDEK")
    b, y, M = _f64_ladder("PEPTIDEK")
    P = 1.0072764
    f = b[3] + P
    mz = np.array([f - 0.01, f + 0.01])
    sp = _spectra([(mz, np.array([5.0, 5.0]))], [(M + 2 * P) / 2], [2])
    p = SearchParams(fragment_unit="da", fragment_ppm=(-0.02, 0.02), min_peaks=1, fragment_min_mz=50.0)
    vp = _vp(p)
    pk = O.k0(vp, sp)
    k2 = O.k2(vp, O.HostIndex(l1), pk, [r], [0], np.array([2], np.int8), grp_count=[1], grp_start=[0])
    assert k2["matched_b"][0] == 1 and k2["matched_y"][0] == 0
    th = np.float32(l1.mass32[r])  # unused; the b ion comes from the f32 ladder
    del th
    m_later = np.float32(np.float32(f + 0.01) - np.float32(P))
    # ppm = I*|mz - m|*2e6/(mz + m) / (sb + sy): the later peak is 0.01 above the ion
    ppm = float(k2["ppm"][0])
    assert abs(ppm - 0.01 * 2e6 / (2 * float(m_later))) < 0.05 * ppm


def test_k1_known_answer_count():
    """Exact b/y 1+ peaks of PEPTIDEK: count = number of indexed ions (ion index > 2, in [fmin, fmax])."""
    from vulcan.gpu._abi import max_hyp_per_spectrum
    from vulcan.gpu.host import SearchParams
    O = _O()
    l1, _ = _tiny_index(["PEPTIDEK"])
    mz, it, pmz = _exact_spectrum("PEPTIDEK")
    sp = _spectra([(mz, it)], [pmz], [2])
    p = SearchParams(fragment_unit="da", fragment_ppm=(-0.02, 0.02), min_peaks=1, isotope_errors=(0, 0))
    vp = _vp(p)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p))
    k1 = O.k1_counts(vp, O.HostIndex(l1), pk, hyp)
    b, y, _ = _f64_ladder("PEPTIDEK")
    L = 8
    want = sum(1 for k in range(L - 1) if k + 1 > 2 and 150 <= b[k] <= 2000) + \
        sum(1 for k in range(L - 1) if L - 1 - k > 2 and 150 <= y[k] <= 2000)
    h, r, c = O.k1_append(hyp, k1)
    got = {int(rr): int(cc) for rr, cc in zip(r, c)}
    assert got[_row_of(l1, "PEPTIDEK")] == want


def test_k1_fast_equals_brute_force(proteome_small):
    """The reference's own K1 invariant (test_k1_k2_small.py:49-65): counts of every window pair agree."""
    from vulcan.gpu._abi import max_hyp_per_spectrum
    O = _O()
    d = proteome_small
    vp = _vp(d["p"], d["static"], d["var"])
    pk = O.k0(vp, d["sp"])
    hyp = O.hypotheses(vp, d["sp"], pk["valid"], d["l1"].mass32, max_hyp_per_spectrum(d["p"]))
    hix = O.HostIndex(d["l1"])
    k1 = O.k1_counts(vp, hix, pk, hyp)
    hh, rr, cc = O.k1_append(hyp, k1, 0)
    sel = np.random.default_rng(1).choice(len(hh), min(len(hh), 30000), replace=False)
    assert np.array_equal(O.k1_brute(vp, hix, hh[sel], rr[sel], hyp, pk), cc[sel])
    nz = k1["counts"] > 0
    assert nz.sum() > 1000 and (hyp["maxfz"] == 3).any()
    scored = np.add.reduceat(nz.astype(np.int64), k1["win_off"][:-1]) if len(nz) else []
    w = hyp["hi"] - hyp["lo"]
    assert np.array_equal(np.where(w > 0, scored, 0), k1["scored"].astype(np.int64))


# ------------------------------------------------------------------------------------------------ top-K
def test_topk_matches_python_brute_force_and_closure_is_symmetric(small_plain):
    from vulcan.gpu._abi import max_hyp_per_spectrum
    O = _O()
    d = small_plain
    p = d["p"]
    vp = _vp(p)
    vp.topk = 3
    pk = O.k0(vp, d["sp"])
    hyp = O.hypotheses(vp, d["sp"], pk["valid"], d["l1"].mass32, max_hyp_per_spectrum(p))
    hix = O.HostIndex(d["l1"])
    k1 = O.k1_counts(vp, hix, pk, hyp)
    tk = O.topk(vp, hix, hyp, k1, 6)
    hh, rr, cc = O.k1_append(hyp, k1, 1)
    partner = d["l1"].partner
    for s in range(200):
        ent = [(int(c), int(r), int(h)) for h, r, c in zip(hh, rr, cc) if hyp["spec"][h] == s]
        ent.sort(key=lambda e: (-min(e[0], 32767), -e[1], e[2]))
        sel = ent[:3]
        pool = [(r, h, c) for c, r, h in sel] + [(int(partner[r]), h, -1) for c, r, h in sel]
        best = {}
        for r, h, c in pool:
            if r not in best:
                best[r] = [h, c]
            else:
                best[r] = [min(best[r][0], h), max(best[r][1], c)]
        want = sorted((r, v[0], v[1]) for r, v in best.items())
        n = tk["n_cand"][s]
        got = list(zip(tk["row"][s * 6:s * 6 + n].tolist(), tk["hyp"][s * 6:s * 6 + n].tolist(),
                       tk["cnt"][s * 6:s * 6 + n].tolist()))
        assert got == want, s
        rows = set(r for r, _, _ in got)
        assert all(int(partner[r]) in rows for r in rows)


def test_rank_matches_python_sort():
    O = _O()
    from vulcan.gpu.host import SearchParams
    vp = _vp(SearchParams())
    rng = np.random.default_rng(2)
    cnt = np.array([0, 1, 5, 9, 3])
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    n = cnt.sum()
    row = rng.integers(0, 50, n).astype(np.int32)
    hs = rng.integers(0, 4, n).astype(np.float64)
    mb = rng.integers(0, 5, n).astype(np.int16)
    my = rng.integers(0, 5, n).astype(np.int16)
    sel, nxt = O.rank(vp, start, cnt, row, hs, mb, my, 3)
    for g in range(len(cnt)):
        idx = [i for i in range(start[g], start[g] + cnt[g]) if mb[i] + my[i] >= 4]
        idx.sort(key=lambda i: (-hs[i], row[i], i))
        for k in range(3):
            assert sel[g * 3 + k] == (idx[k] if k < len(idx) else -1)
            assert nxt[g * 3 + k] == (hs[idx[k + 1]] if k + 1 < len(idx) else 0.0)


# ------------------------------------------------------------------------------------------------ end to end
def test_search_recovers_truth_and_matches_components(small_plain):
    O = _O()
    d = small_plain
    vp = _vp(d["p"])
    hix = O.HostIndex(d["l1"])
    s = O.search(vp, hix, d["sp"])
    sig = d["truth"] >= 0
    assert np.mean(s["row"][sig] == d["truth"][sig]) > 0.9       # test_k1_k2_small.py:66-69 bar
    assert s["stats"]["n_rescored"] == int(s["n_cand"].sum())


def test_label_blindness_role_swap_and_permutation():
    """G0 (reference test_k1_k2_small.py:72-84): swapping target/decoy roles and permuting the pair list give a
    bit-identical index side and identical search results."""
    from vulcan.index import synth
    from vulcan.index.build import build_tables
    O = _O()
    w = synth.WORKLOADS[1]
    prots = synth.proteome(300, 77)
    a, la = build_tables(prots, w.index_params())
    b, lb = build_tables(prots, w.index_params(), shuffle_rows_seed=3, swap_labels=True)
    for k in ("pf_id", "mass32", "seq_off", "length", "mask", "partner", "residues", "n_var", "missed"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.array_equal(la["is_decoy"], lb["is_decoy"])      # the true roles are kept in the labels
    sp, _, _ = synth.spectra(a, la, 1, n=300, seed=9)
    vp = _vp(w.search_params(), STATIC, w.var_mods)
    ra, rb = O.search(vp, O.HostIndex(a), sp), O.search(vp, O.HostIndex(b), sp)
    for k in ("row", "hs", "dn", "n_cand"):
        assert np.array_equal(ra[k], rb[k])


def test_builder_invariants(small_plain):
    l1 = small_plain["l1"]
    assert np.all(np.diff(l1.mass32) >= 0)
    same = np.diff(l1.mass32) == 0
    assert np.all(np.diff(l1.pf_id.astype(np.float64))[same] > 0) or not same.any() or \
        np.all(l1.pf_id[1:][same] > l1.pf_id[:-1][same])
    assert np.array_equal(l1.partner[l1.partner], np.arange(len(l1)))
    assert len(np.unique(l1.pf_id)) == len(l1)
    seqs = [l1.sequence(i) for i in range(0, len(l1), 997)]
    assert all("I" not in s for s in seqs)
    i = 5
    assert sorted(l1.sequence(i)) == sorted(l1.sequence(int(l1.partner[i])))


# ------------------------------------------------------------------------------------------------ search modes
def test_mode_hypotheses_match_host_restatement(proteome_small):
    """Closed, mass-offset and open hypotheses (ptm/search.py:48-77): the C++ oracle and the host numpy
    restatement agree field by field, and every spectrum's hypotheses keep the order closed, offsets, open."""
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum, set_modes
    from vulcan.ptm.search import ModeParams, make_mode_hypotheses
    O = _O()
    d = proteome_small
    p = d["p"]
    mp = ModeParams(offsets=(15.994915, 42.010565), open_da=(-150.0, 500.0))
    vp = set_modes(make_params(p, d["static"], d["var"]), mp.offsets, mp.open_da)
    pk = O.k0(vp, d["sp"])
    h = O.hypotheses(vp, d["sp"], pk["valid"], d["l1"].mass32, max_hyp_per_spectrum(p, 2, True))
    ref = make_mode_hypotheses(d["sp"], pk["valid"].astype(bool), d["l1"].mass32, p, mp)
    for k in ("spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz", "mode"):
        assert np.array_equal(h[k], ref[k]), k
    assert np.array_equal(h["off"], ref["off"], equal_nan=True)
    s = h["spec"]
    for i in range(len(s) - 1):
        if s[i] == s[i + 1]:
            assert h["mode"][i] <= h["mode"][i + 1]
    open_ = h["mode"] == 2
    assert open_.any() and np.all(h["hi"][open_] - h["lo"][open_] > 100)


def test_k2_shift_localizes_the_delta():
    """PEPTIDEK with +79.966331 on residue 3 (T): unshifted b0-b2 and y3-y6, shifted b3-b6 and y0-y2 are all
    present. Unshifted scoring misses half of them; shifted scoring places the delta at residue 3 and matches
    all 7 b and 7 y ions, so hyperscore = ln(71^2) + 2 lnfact(7)."""
    from vulcan.gpu.host import SearchParams
    O = _O()
    l1, _ = _tiny_index(["PEPTIDEK"])
    r = _row_of(l1, "PEPTIDEK")
    b, y, M = _f64_ladder("PEPTIDEK")
    d = 79.966331
    P = 1.0072764
    bs = b + np.where(np.arange(7) >= 3, d, 0.0)
    ys = y + np.where(np.arange(7) < 3, d, 0.0)
    mz = np.sort(np.concatenate([bs, ys]) + P)
    sp = _spectra([(mz, np.full(len(mz), 10.0))], [(M + d + 2 * P) / 2], [2])
    p = SearchParams(fragment_unit="da", fragment_ppm=(-0.02, 0.02), min_peaks=1, fragment_min_mz=50.0)
    vp = _vp(p)
    pk = O.k0(vp, sp)
    hix = O.HostIndex(l1)
    shift = np.float32(np.float32(M + d) - l1.mass32[r])
    closed = O.k2(vp, hix, pk, [r], [0], np.array([2], np.int8), grp_count=[1], grp_start=[0])
    sh = O.k2(vp, hix, pk, [r], [0], np.array([2], np.int8), grp_count=[1], grp_start=[0], shift=[shift])
    assert closed["matched_b"][0] + closed["matched_y"][0] < 14
    assert sh["matched_b"][0] == 7 and sh["matched_y"][0] == 7
    assert sh["summed_b"][0] == 70.0 and sh["summed_y"][0] == 70.0
    want = math.log(71.0 * 71.0) + 2 * (7 * math.log(7) - 7 + 0.5 * math.log(7) + 0.5 * math.log(2 * math.pi * 7))
    assert abs(sh["hyperscore"][0] - want) < 1e-12
    zero = O.k2(vp, hix, pk, [r], [0], np.array([2], np.int8), grp_count=[1], grp_start=[0], shift=[0.0])
    for k in closed:                                   # shift 0 is the closed path exactly
        assert np.array_equal(zero[k], closed[k], equal_nan=True), k


def test_open_search_oracle_finds_shifted_spectra():
    """The oracle's mode pipeline (hypotheses -> K1 counts -> top-K -> shifted K2 -> rank 1) finds spectra whose
    precursor carries an unexplained delta through offset / open hypotheses (> 50% of them at rank 1,
    the reference's bar in test_ptm_modes.py:114-117)."""
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum, set_modes
    from vulcan.index import synth
    from vulcan.index.build import build_tables
    O = _O()
    w1 = synth.WORKLOADS[1]
    l1, lab = build_tables(synth.proteome(200, 104), w1.index_params())
    sp, truth, meta = synth.spectra(l1, lab, 3, n=150, seed=34, index_params=w1.index_params())
    p = w1.search_params()
    p.topk = 20
    vp = set_modes(make_params(p, STATIC, w1.var_mods), (15.994915,), (-150.0, 500.0))
    hix = O.HostIndex(l1)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p, 1, True))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    stride = 2 * p.topk
    tk = O.topk(vp, hix, hyp, k1, stride)
    n = len(sp)
    m = (np.arange(stride)[None, :] < tk["n_cand"][:, None]).ravel()
    rows, hyps = tk["row"][m], tk["hyp"][m]
    shift = np.where(hyp["mode"][hyps] == 0, np.float32(0),
                     (hyp["M"][hyps] - l1.mass32[rows]).astype(np.float32)).astype(np.float32)
    start = np.concatenate([[0], np.cumsum(tk["n_cand"])[:-1]])
    k2 = O.k2(vp, hix, pk, rows, hyps, hyp["maxfz"], grp_start=start, grp_count=tk["n_cand"], shift=shift)
    sel, _ = O.rank(vp, start, tk["n_cand"], rows, k2["hyperscore"], k2["matched_b"], k2["matched_y"], 1)
    shifted = np.flatnonzero(np.abs(meta["shift"]) > 1.0)
    got = [hyp["mode"][hyps[sel[s]]] > 0 for s in shifted if sel[s] >= 0]
    assert len(got) > 20 and np.mean(got) > 0.5
    hit = [rows[sel[s]] == truth[s] for s in shifted if sel[s] >= 0]
    assert np.mean(hit) > 0.5


def test_spectrum_classes_bins():
    """Class = charge bin (2, 3, 4, 5+) x kept-peak bin (< 64, 64-149, >= 150), kept = peaks inside
    [fmin, fmax] capped at max_peaks, as K0 keeps them."""
    import numpy as np
    from vulcan.gpu.host import SearchParams, Spectra
    from vulcan.gpu.schedule import spectrum_classes
    p = SearchParams(max_peaks=150)
    counts = [10, 63, 64, 149, 150, 400]
    charges = [2, 3, 4, 5, 6, 0]
    mz, off = [], [0]
    for c in counts:
        mz += [500.0] * c + [100.0, 2500.0, float("nan")]     # three peaks K0 drops
        off.append(len(mz))
    n = len(counts)
    sp = Spectra("t", np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.zeros(n),
                 np.full(n, 500.0), np.array(charges, np.int8), np.array(mz, np.float32),
                 np.ones(len(mz), np.float32), np.array(off, np.int64))
    cls = spectrum_classes(sp, p)
    assert cls.tolist() == [0 * 3 + 0, 1 * 3 + 0, 2 * 3 + 1, 3 * 3 + 1, 3 * 3 + 2, 0 * 3 + 2]
