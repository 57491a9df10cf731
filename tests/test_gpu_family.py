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


def test_family_select_kernel_equals_device_sorts(multi):
    """vh_family_select on the join K1's window counts == select_family over the appended K1 list (torch
    device sorts): the same (spectrum, row) entries, hypotheses, counts and shortlist flags; and the two search
    paths give the same table."""
    from vulcan.gpu import engine as E
    from vulcan.gpu import family as F
    from vulcan.index import synth
    l1, lab, sp = multi
    p = synth.WORKLOADS[1].search_params()
    var = dict(synth.WORKLOADS[1].var_mods)
    gi = E.GpuIndex(l1, p, STATIC, var)
    fi = F.FamilyIndex.from_l1(l1, gi, lab)
    n = len(sp)
    print("[gpu] family select kernel vs sorts", flush=True)
    st = E.Stats()
    peaks = E.k0_preprocess(sp, p, gi.dev, st, gi)
    hd = E.device_hypotheses(gi, peaks["raw"], peaks, p)
    n_h = int(hd["hyp_off"][n].item())
    hyp = {k: hd[k][:n_h] for k in ("spec", "z", "iso", "pm", "M", "lo", "hi", "maxfz")}
    hyp["hyp_off"] = hd["hyp_off"][:n + 1]
    o_h, o_r, o_c, _, _, d_spec, _ = E.run_k1(gi, hyp, peaks, p, st)
    want = F.select_family(o_h, o_r, o_c, d_spec, hyp, fi, p, gi.dev)
    ds = E.DeviceSearch(gi, p, n, k1_impl="join")
    ds.front(E.upload_spectra(sp, gi.dev), n, E._marker(None))
    ds.join_front(n, E._marker(None))
    got = F.select_family_join(ds, fi, n)
    assert len(got[0]) == len(want[0]) > 0
    for a, b, name in zip(got, want, ("hyp", "row", "count", "shortlisted", "spectrum", "key")):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()), name
    t1, _ = F.search_family(sp, gi, l1, fi, p, select="kernel")
    t2, _ = F.search_family(sp, gi, l1, fi, p, select="torch")
    assert t1.num_rows == t2.num_rows > 0
    for name in t2.column_names:
        assert np.array_equal(t1.column(name).to_numpy(), t2.column(name).to_numpy()), name
