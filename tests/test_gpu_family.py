"""Family top-2 retention (vulcan/gpu/family.py) on a small multi-replicate table: every (decoy set x entrapment
seed) replicate rebuilt from one search equals a separate search over that replicate's rows (the reference's
check, tests/gpu/test_family.py:45-67): same spectra, ranks, peptidoforms and bit-equal hyperscores."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

from conftest import STATIC  # noqa: E402


@pytest.fixture(scope="module")
def multi():
    from vulcan.index import build as B
    from vulcan.index import synth
    ip = synth.WORKLOADS[1].index_params()
    l1, lab = B.build_multi_decoy_tables(synth.proteome(300, 11),
                                         {"entrap_seeds": [2001, 2002], "decoy_seeds": [1001, 1002], **ip})
    sp, _, _ = synth.spectra(l1, lab, 1, n=600, seed=7, index_params=ip)
    return l1, lab, sp


def test_family_top2_equals_separate_searches(multi):
    from vulcan.gpu import engine as E
    from vulcan.gpu import family as F
    from vulcan.index import synth
    l1, lab, sp = multi
    print(f"[gpu] family: {len(l1)} rows, {len(sp)} spectra", flush=True)
    p = synth.WORKLOADS[1].search_params()
    var = dict(synth.WORKLOADS[1].var_mods)
    gi = E.GpuIndex(l1, p, STATIC, var)
    fi = F.FamilyIndex.from_l1(l1, gi, lab)
    tbl, st = F.search_family(sp, gi, l1, fi, p)
    assert tbl.num_rows > 0
    ff, side = fi.fwd_fam.cpu().numpy(), fi.side.cpu().numpy()
    for j in range(2):
        for si in range(2):
            keep = np.isin(ff, [0, 1 + si]) & np.isin(side, [-1, j])
            sub = F.SubIndex(l1, keep, j)
            gs = E.GpuIndex(sub, p, STATIC, var)
            fs = F.FamilyIndex(ff[keep], np.where(side[keep] >= 0, 0, -1), [sub.partner], fi.F, gs.dev)
            ts, _ = F.search_family(sp, gs, sub, fs, p)
            a = F.reconstruct(ts, sorted(set(ts.column("fgroup").to_numpy().tolist())))
            b = F.reconstruct(tbl, fi.replicate_groups(j, si))
            assert len(a["spec_idx"]) > 300, (j, si)
            for k in ("spec_idx", "rank", "pf_id"):
                assert np.array_equal(a[k], b[k]), (j, si, k)
            assert np.array_equal(a["hyperscore"].view(np.uint64), b["hyperscore"].view(np.uint64))


def test_family_rank1_is_the_plain_search_rank1_on_a_single_decoy_replicate(multi):
    """On one replicate's rows with the family rule reduced to one K1 family per side, the top-1 over the
    replicate's groups is the plain search's rank-1 whenever the plain search's candidate lists cover it."""
    from vulcan.gpu import engine as E
    from vulcan.gpu import family as F
    from vulcan.index import synth
    l1, lab, sp = multi
    p = synth.WORKLOADS[1].search_params()
    var = dict(synth.WORKLOADS[1].var_mods)
    gi = E.GpuIndex(l1, p, STATIC, var)
    fi = F.FamilyIndex.from_l1(l1, gi, lab)
    ff, side = fi.fwd_fam.cpu().numpy(), fi.side.cpu().numpy()
    keep = (ff == 0) & np.isin(side, [-1, 0])             # T and dec_0(T): a plain target-decoy table
    sub = F.SubIndex(l1, keep, 0)
    gs = E.GpuIndex(sub, p, STATIC, var)
    fs = F.FamilyIndex(np.zeros(len(sub), np.int64), np.where(side[keep] >= 0, 0, -1), [sub.partner], 1, gs.dev)
    ts, _ = F.search_family(sp, gs, sub, fs, p)
    fam1 = F.reconstruct(ts, [0, 1], n=1)
    plain, _, _ = E.search(sp, gs, sub, p, report=1)
    ps = plain.column("spec_idx").to_numpy()
    common = np.intersect1d(fam1["spec_idx"], ps)
    assert len(common) > 0.9 * len(ps)
    fa = dict(zip(fam1["spec_idx"].tolist(), fam1["pf_row"].tolist()))
    pa_ = dict(zip(ps.tolist(), plain.column("pf_row").to_numpy().tolist()))
    same = np.mean([fa[s] == pa_[s] for s in common.tolist()])
    assert same > 0.97, same


def _oracle_shortlist(l1, fi_np, sp, p, var):
    """The oracle's K0 / hypotheses / per-row K1 counts and its family shortlist (oracle/family.py)."""
    from oracle import family as OF
    from oracle import oracle as O
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum
    vp = make_params(p, STATIC, var)
    hix = O.HostIndex(l1)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    fwd, side, partners = fi_np
    kfam = fwd * (len(partners) + 1) + side + 1
    sl = OF.shortlist(hyp, k1, kfam, partners, p.topk, p.k1_threshold)
    return vp, hix, pk, hyp, sl


@pytest.mark.parametrize("case", ["default", "wide"])
def test_family_shortlist_and_table_equal_oracle(multi, case):
    """vh_family_select (and, for spectra beyond one wave's LDS, the wide path) == the oracle's shortlist over the
    per-row K1 counts (rows, canonical hypotheses, counts, shortlisted flags per spectrum); the top-2 table ==
    the oracle's top-2 over the oracle's K2 of that shortlist. "wide": the reference's defaults top-K 500 and
    K1 threshold 1 with a 300 ppm precursor window and 0.5 Da fragments, so about a quarter of the spectra have
    more than 256 rows over the threshold."""
    from oracle import family as OF
    from oracle import oracle as O
    from vulcan.gpu import engine as E
    from vulcan.gpu import family as F
    from vulcan.index import synth
    l1, lab, sp = multi
    p = synth.WORKLOADS[1].search_params()
    if case == "wide":
        p.topk, p.k1_threshold, p.precursor_ppm, p.fragment_ppm = 500, 1, (-300.0, 300.0), (-0.5, 0.5)
    var = dict(synth.WORKLOADS[1].var_mods)
    gi = E.GpuIndex(l1, p, STATIC, var)
    fi = F.FamilyIndex.from_l1(l1, gi, lab)
    fwd, side, _ = F.family_codes(lab)
    partners = np.stack(l1.partners)
    vp, hix, pk, hyp, want = _oracle_shortlist(l1, (fwd, side, partners), sp, p, var)
    n = len(sp)
    print(f"[gpu] family vs oracle ({case}): {len(want['row'])} shortlist entries", flush=True)
    ds = E.DeviceSearch(gi, p, n, k1_impl="join")
    ds.front(E.upload_spectra(sp, gi.dev), n, E._marker(None))
    ds.join_front(n, E._marker(None))
    got = F.select_family_join(ds, fi, n)
    if case == "wide":
        assert got.n_wide > 50
    else:
        assert got.n_wide == 0
    assert np.array_equal(got.off.cpu().numpy(), want["off"])
    for name, dv in (("row", got.row), ("hyp", got.hyp), ("cnt", got.cnt)):
        assert np.array_equal(dv.cpu().numpy().astype(np.int64), want[name]), name
    assert np.array_equal(got.top.cpu().numpy().astype(bool), want["top"])
    # the table: oracle K2 over the oracle shortlist, oracle top-2
    off = want["off"]
    k2 = O.k2(vp, hix, pk, want["row"], want["hyp"], hyp["maxfz"], grp_start=off[:-1], grp_count=np.diff(off))
    t2 = OF.top2(want, k2["hyperscore"], k2["matched_b"], k2["matched_y"], fwd, side, partners, p.min_matched_peaks)
    tbl, st = F.search_family(sp, gi, l1, fi, p)
    assert st.counts["n_wide"] == got.n_wide
    assert len(t2["idx"]) == tbl.num_rows > 0
    assert np.array_equal(tbl.column("spec_idx").to_numpy(), t2["spec"])
    assert np.array_equal(tbl.column("fgroup").to_numpy().astype(np.int64), t2["group"])
    assert np.array_equal(tbl.column("grank").to_numpy().astype(np.int64), t2["rank"])
    assert np.array_equal(tbl.column("pf_row").to_numpy(), want["row"][t2["idx"]])
    assert np.allclose(tbl.column("hyperscore").to_numpy(), k2["hyperscore"][t2["idx"]], rtol=1e-12, atol=0)
    assert np.array_equal(tbl.column("charge").to_numpy(), hyp["z"][want["hyp"][t2["idx"]]])


def test_family_search_parameters_change_between_calls(multi):
    """Two family searches on the same index and batch size with different parameters (top-K, precursor
    tolerance) each equal a search with fresh buffers: the step buffers are not reused across parameters."""
    from vulcan.gpu import engine as E
    from vulcan.gpu import family as F
    from vulcan.index import synth
    l1, lab, sp = multi
    var = dict(synth.WORKLOADS[1].var_mods)
    p1 = synth.WORKLOADS[1].search_params()
    p2 = type(p1)(**{**p1.to_dict(), "topk": 7, "precursor_ppm": (-5.0, 5.0)})
    gi = E.GpuIndex(l1, p1, STATIC, var)
    fi = F.FamilyIndex.from_l1(l1, gi, lab)
    a1, _ = F.search_family(sp, gi, l1, fi, p1)
    a2, _ = F.search_family(sp, gi, l1, fi, p2)
    for p, a in ((p1, a1), (p2, a2)):
        fresh = F.FamilyIndex.from_l1(l1, gi, lab)
        b, _ = F.search_family(sp, gi, l1, fresh, p)
        assert a.equals(b)
    assert not a1.equals(a2)
