"""HIP path vs the oracle on the semantics of the phospho-scale prefilter workload (STY phospho <= 3 and Met-ox <= 2
per peptidoform, at most 5 variable mods, index top-200 + closure, charges 2-6, 20-500 peaks, per-run
calibration offsets) on a 2,000-protein table, and the recall judge: exhaustive K2 over every window row against
the oracle and against the shortlist (reference engine.py:500-523, scripts/m2/e2_correctness.py:139-181)."""
import numpy as np
import pytest

from conftest import STATIC

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)

HS_RTOL = 1e-12


def _np(t):
    return t.detach().cpu().numpy()


def _O():
    from oracle import oracle as O
    return O


def _take(sp, n):
    from vulcan.gpu.host import Spectra
    off = sp.offsets[: n + 1]
    return Spectra(sp.run, sp.scan[:n], sp.index[:n], sp.rt[:n], sp.precursor_mz[:n], sp.charge[:n],
                   sp.mz[: off[-1]], sp.intensity[: off[-1]], off.copy())


def _flat(ds, n):
    nc = _np(ds.n_cand)[:n]
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    return nc, m


@pytest.fixture(scope="module")
def phospho():
    """2,000 proteins of the phospho workload (3.7M peptidoforms), 3,000 spectra of its generator, its search
    parameters (index top-200, chunk 1024, 320 ppm bins, tile 1024); the oracle's K0 / hypotheses / K1 counts."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum
    from vulcan.index import synth
    print("[gpu] building the phospho table (2,000 proteins)", flush=True)
    w = synth.WORKLOADS[4]
    l1, lab = synth.build_index(4, n_proteins=2000)
    sp, truth, _ = synth.spectra(l1, lab, 4, n=3000)
    p = w.search_params()
    gi = E.GpuIndex(l1, p, STATIC, w.var_mods)
    vp = make_params(p, STATIC, w.var_mods)
    O = _O()
    hix = O.HostIndex(l1)
    pk = O.k0(vp, sp)
    hyp = O.hypotheses(vp, sp, pk["valid"], l1.mass32, max_hyp_per_spectrum(p))
    k1 = O.k1_counts(vp, hix, pk, hyp)
    print(f"[gpu] phospho: {len(l1)} rows, {len(hyp['spec'])} hypotheses, {int(k1['win_off'][-1])} window rows",
          flush=True)
    return {"w": w, "l1": l1, "lab": lab, "sp": sp, "truth": truth, "p": p, "gi": gi, "vp": vp, "hix": hix,
            "pk": pk, "hyp": hyp, "k1": k1, "raw": E.upload_spectra(sp, gi.dev), "n": len(sp)}


def test_phospho_table_has_the_capped_multi_residue_forms(phospho):
    """The table carries phospho on S, T and Y together with Met-ox, never more than 3 phospho, 2 Met-ox or
    5 modifications on one peptidoform (the caps the workload sets)."""
    l1 = phospho["l1"]
    rng = np.random.default_rng(0)
    rows = rng.choice(len(l1), 20000, replace=False)
    seen = {"S": 0, "T": 0, "Y": 0, "M": 0}
    for r in rows:
        L = int(l1.length[r])
        seq = bytes(l1.residues[l1.seq_off[r]:l1.seq_off[r] + L]).decode()
        bits = [(int(l1.mask[r]) >> k) & 1 for k in range(L)]
        mods = [seq[k] for k in range(L) if bits[k]]
        assert all(c in "STYM" for c in mods)
        n_p, n_m = sum(c in "STY" for c in mods), mods.count("M")
        assert n_p <= 3 and n_m <= 2 and n_p + n_m <= 5 and n_p + n_m == int(l1.n_var[r])
        for c in set(mods):
            seen[c] += 1
    assert min(seen.values()) > 100
    assert int(l1.n_var.max()) == 5


@pytest.mark.parametrize("impl", ["join", "tile"])
def test_phospho_search_equals_oracle(phospho, impl):
    """Every join K1 slot count, the top-200 + closure lists, every K2 column and rank 1 equal the oracle's."""
    from vulcan.gpu import engine as E
    e = phospho
    O = _O()
    n = e["n"]
    ds = E.DeviceSearch(e["gi"], e["p"], n, k1_impl=impl)
    ds.run(e["raw"])
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    if impl == "join":
        tot = int(e["k1"]["win_off"][-1])
        assert np.array_equal(_np(ds.join["counts"])[:tot], e["k1"]["counts"])
    tk = O.topk(e["vp"], e["hix"], e["hyp"], e["k1"], ds.stride)
    nc, m = _flat(ds, n)
    assert np.array_equal(nc, tk["n_cand"])
    # the top-200 cut is exercised: spectra whose shortlist was truncated at K before closure
    assert (tk["n_cand"] >= e["p"].topk).sum() > 300
    rows = _np(ds.c_row)[:n * ds.stride][m]
    hyps = _np(ds.c_hyp)[:n * ds.stride][m]
    assert np.array_equal(rows, tk["row"][m]) and np.array_equal(hyps, tk["hyp"][m])
    assert np.array_equal(_np(ds.c_cnt)[:n * ds.stride][m], tk["cnt"][m])
    nh = len(e["hyp"]["spec"])
    assert np.array_equal(_np(ds.scored)[:nh].astype(np.uint32), e["k1"]["scored"])
    assert np.array_equal(_np(ds.matched)[:nh].astype(np.uint64), e["k1"]["matched"])
    start = np.concatenate([[0], np.cumsum(nc)[:-1]])
    ref = O.k2(e["vp"], e["hix"], e["pk"], rows, hyps, e["hyp"]["maxfz"], grp_start=start, grp_count=nc)
    for k in ("matched_b", "matched_y", "longest_b", "longest_y", "summed_b", "summed_y"):
        assert np.array_equal(_np(ds.k2[k])[:n * ds.stride][m], ref[k]), k
    assert np.array_equal(_np(ds.k2["ppm"])[:n * ds.stride][m], ref["ppm"], equal_nan=True)
    assert np.allclose(_np(ds.k2["hyperscore"])[:n * ds.stride][m], ref["hyperscore"], rtol=HS_RTOL, atol=0)
    full = O.search(e["vp"], e["hix"], e["sp"])
    assert np.array_equal(ds.rank1_rows(n), full["row"])
    sel = _np(ds.sel)[:n]
    hs = _np(ds.k2["hyperscore"])[np.maximum(sel, 0)]
    ok = sel >= 0
    assert np.allclose(hs[ok], full["hs"][ok], rtol=HS_RTOL, atol=0)
    truth = e["truth"]
    sig = truth >= 0
    print(f"[gpu] phospho {impl}: truth recovery {np.mean(full['row'][sig] == truth[sig]):.3f}", flush=True)


@pytest.mark.parametrize("mode", [1, 2])
def test_join_counts_of_spectra_beyond_256_peaks_equal_oracle(phospho, mode):
    """max_peaks 512 on the phospho spectra (20-500 raw peaks): the count kernel holds a spectrum's first 256
    peaks in registers and loads the rest on demand; window counts (mode 1) and the per-ion feature counts
    (mode 2, the learned path's K1) equal the oracle's on 1,000 spectra, a share of them beyond 256 kept peaks."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum
    e = phospho
    O = _O()
    p = type(e["p"])(**{**e["p"].to_dict(), "max_peaks": 512})
    vp = make_params(p, STATIC, e["w"].var_mods)
    sub = _take(e["sp"], 1000)
    n = len(sub)
    pk = O.k0(vp, sub)
    kept = np.diff(pk["off"])
    assert (kept > 256).sum() > 20
    hyp = O.hypotheses(vp, sub, pk["valid"], e["l1"].mass32, max_hyp_per_spectrum(p))
    k1 = O.k1_counts(vp, e["hix"], pk, hyp)
    ds = E.DeviceSearch(e["gi"], p, n, k1_impl="join", join_features=mode == 2)
    mark = E._marker(None)
    ds.front(E.upload_spectra(sub, e["gi"].dev), n, mark)     # K0, hypotheses, order
    ds.join_front(n, mark)                                    # pairs + block counts
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    tot = int(k1["win_off"][-1])
    got = _np(ds.join["counts"])[:tot]
    if mode == 2:
        got = (got.astype(np.uint64) >> np.uint64(43)).astype(np.uint32)
    assert np.array_equal(got, k1["counts"])


@pytest.fixture(scope="module")
def phospho_learned(phospho):
    """A prefilter network trained on 1,500 separately seeded spectra of the phospho workload (labels: the exact
    rescorer's hyperscore minus the spectrum best), then the learned and plain steps on the 3,000 test spectra."""
    from vulcan.gpu import engine as E
    from vulcan.index import synth
    from vulcan.rescore import prefilter as PF
    e = phospho
    tr, _, _ = synth.spectra(e["l1"], e["lab"], 4, n=1500, seed=904)
    pp = PF.PrefilterParams(index_topk=200, steps=400, batch=4096)
    x, y, _, _ = PF.training_data(e["gi"], e["p"], E.upload_spectra(tr, e["gi"].dev), pp)
    print(f"[gpu] phospho prefilter: training on {x.shape[0]} candidates", flush=True)
    model = PF.train(PF.PrefilterModel.init(0), x, y, pp, log_every=100)
    n = e["n"]
    ls = PF.LearnedSearch(e["gi"], e["p"], n, model, pp)
    ls.run(e["raw"])
    plain = E.DeviceSearch(e["gi"], e["p"], n)
    plain.run(e["raw"])
    torch.cuda.synchronize()
    assert int(ls.status.item()) == 0 and int(plain.status.item()) == 0
    return {"pp": pp, "ls": ls, "plain": plain, "model": model}


def test_phospho_learned_index_lists_equal_oracle(phospho, phospho_learned):
    """The learned path's index top-200 (no closure) is the oracle's."""
    import ctypes as C
    from vulcan.gpu._abi import VhParams
    e, t = phospho, phospho_learned
    fl = t["ls"].lists
    vp = VhParams()
    C.memmove(C.byref(vp), C.byref(e["vp"]), C.sizeof(VhParams))
    vp.topk, vp.pair_closure = 200, 0
    tk = _O().topk(vp, e["hix"], e["hyp"], e["k1"], 200)
    n = e["n"]
    nc = _np(fl.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(200)[None, :] < nc[:, None]).ravel()
    for dv, k in ((fl.row, "row"), (fl.hyp, "hyp"), (fl.cnt, "cnt")):
        assert np.array_equal(_np(dv)[:n * 200][m], tk[k][m]), k


def test_phospho_learned_rescoring_equals_oracle(phospho, phospho_learned):
    """The kept candidates are a subset of the plain top-200 + closure lists; their K2 outputs and rank 1 equal
    the oracle's on the same candidates; the learned rank-1 never outscores the plain rank-1 (its candidates are
    a subset), and equal scores mean the same row."""
    e, t = phospho, phospho_learned
    O = _O()
    ls, plain = t["ls"], t["plain"]
    n = e["n"]
    nc, m = _flat(ls, n)
    assert nc.max() <= 2 * t["pp"].keep
    lrow = _np(ls.c_row)[:n * ls.stride].reshape(n, ls.stride)
    prow = _np(plain.c_row)[:n * plain.stride].reshape(n, plain.stride)
    pnc = _np(plain.n_cand)[:n]
    for s in range(n):
        assert set(lrow[s, :nc[s]].tolist()) <= set(prow[s, :pnc[s]].tolist()), s
    rows = _np(ls.c_row)[:n * ls.stride][m]
    hyps = _np(ls.c_hyp)[:n * ls.stride][m]
    start = np.concatenate([[0], np.cumsum(nc)[:-1]])
    ref = O.k2(e["vp"], e["hix"], e["pk"], rows, hyps, e["hyp"]["maxfz"], grp_start=start, grp_count=nc)
    for k in ("matched_b", "matched_y", "longest_b", "longest_y", "summed_b", "summed_y"):
        assert np.array_equal(_np(ls.k2[k])[:n * ls.stride][m], ref[k]), k
    assert np.allclose(_np(ls.k2["hyperscore"])[:n * ls.stride][m], ref["hyperscore"], rtol=HS_RTOL, atol=0)
    sel, _ = O.rank(e["vp"], start, nc, rows, ref["hyperscore"], ref["matched_b"], ref["matched_y"], 1)
    flat_of = np.flatnonzero(m)
    assert np.array_equal(_np(ls.sel)[:n], np.where(sel >= 0, flat_of[np.maximum(sel, 0)], -1))
    r_l, r_p = ls.rank1_rows(n), plain.rank1_rows(n)
    hs_l = _np(ls.k2["hyperscore"])[np.maximum(_np(ls.sel)[:n], 0)]
    hs_p = _np(plain.k2["hyperscore"])[np.maximum(_np(plain.sel)[:n], 0)]
    both = (r_l >= 0) & (r_p >= 0)
    assert np.all(hs_l[both] <= hs_p[both])
    assert np.all(r_l[both & (hs_l == hs_p)] == r_p[both & (hs_l == hs_p)])
    from vulcan.rescore import prefilter as PF
    ag = PF.agreement(r_p, r_l)
    print(f"[gpu] phospho learned: top-1 agreement {ag:.4f}, {nc.mean():.1f} rescored per spectrum "
          f"(plain {pnc.mean():.1f})", flush=True)
    assert ag > 0.95


# ------------------------------------------------------------------------------------------------ recall judge
def test_exhaustive_rescoring_is_the_recall_judge(proteome_small):
    """search_exhaustive scores every window row of every hypothesis: its K2 equals the oracle's on those rows,
    its rank 1 equals the oracle's ranking of them, no shortlist rank-1 outscores it, and the shortlist (top-50 +
    closure) recovers most exhaustive rank-1 rows."""
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    d = proteome_small
    O = _O()
    p = type(d["p"])(**d["p"].to_dict())
    gi = E.GpuIndex(d["l1"], p, d["static"], d["var"])
    sub = np.arange(0, len(d["sp"]), 3)
    tbl, _, dbg = E.search_exhaustive(d["sp"], gi, d["l1"], p, spec_subset=sub)
    vp = make_params(p, d["static"], d["var"])
    hix = O.HostIndex(d["l1"])
    pk = O.k0(vp, d["sp"])
    h, r = _np(dbg["h"]), _np(dbg["r"])
    hyp = dbg["hyp"]
    spec = hyp["spec"][h]
    assert np.all(np.diff(spec) >= 0) and len(r) > 10000
    us, cnt = np.unique(spec, return_counts=True)
    start = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    ref = O.k2(vp, hix, pk, r, h, hyp["maxfz"], grp_spec=us, grp_start=start, grp_count=cnt)
    k2 = {k: _np(v) for k, v in dbg["k2"].items()}
    for k in ("matched_b", "matched_y", "longest_b", "longest_y", "summed_b", "summed_y"):
        assert np.array_equal(k2[k], ref[k]), k
    assert np.allclose(k2["hyperscore"], ref["hyperscore"], rtol=HS_RTOL, atol=0)
    sel, _ = O.rank(vp, start, cnt, r, ref["hyperscore"], ref["matched_b"], ref["matched_y"], 1)
    ex_row = {int(us[g]): int(r[i]) for g, i in enumerate(sel) if i >= 0}
    ex_hs = {int(us[g]): float(ref["hyperscore"][i]) for g, i in enumerate(sel) if i >= 0}
    got = dict(zip(tbl.column("spec_idx").to_pylist(), tbl.column("pf_row").to_pylist()))
    assert got == ex_row
    short = O.search(vp, hix, d["sp"])
    window = set(((spec.astype(np.int64) << 28) | r).tolist())
    hit = both = 0
    for s, row in ex_row.items():
        if short["row"][s] >= 0:
            both += 1
            # a closure partner can sit just outside every precursor window (its f32 mass is summed in another
            # residue order); every other shortlisted row is one of the exhaustive candidates
            if (s << 28) | int(short["row"][s]) in window:
                assert short["hs"][s] <= ex_hs[s] * (1 + HS_RTOL)
            hit += int(short["row"][s] == row)
    recall = hit / max(len(ex_row), 1)
    print(f"[gpu] shortlist (top-{p.topk} + closure) recall of the exhaustive rank-1: {recall:.4f} "
          f"({hit} of {len(ex_row)} spectra; {len(r)} window rows rescored)", flush=True)
    assert both > 0.8 * len(ex_row) and recall > 0.9
