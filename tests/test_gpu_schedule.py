"""Uneven batching on the GPU: persistent (work-counter) launches, work-ordered spectra, per-class schedules and
the hipGraph-captured step give exactly the candidate lists, scores and rank-1 rows of the default eager step
(results are per spectrum and must not depend on the schedule); graph replays are checked one by one with eager
work in between."""
import numpy as np
import pytest

from conftest import STATIC

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def uneven(proteome_small):
    """uneven-workload-shaped spectra (16 runs, 20-500 peaks, charges 2-6, calibration offsets) on the proteome-small table."""
    from vulcan.gpu import engine as E
    from vulcan.index import synth
    d = proteome_small
    sp, truth, _ = synth.spectra(d["l1"], d["labels"], 2, n=4000)
    p = synth.WORKLOADS[2].search_params()
    gi = E.GpuIndex(d["l1"], p, STATIC, d["var"])
    raw = E.upload_spectra(sp, gi.dev)
    base = E.DeviceSearch(gi, p, len(sp))
    base.run(raw)
    torch.cuda.synchronize()
    print("[gpu] uneven batch ready", flush=True)
    return {"gi": gi, "p": p, "sp": sp, "raw": raw, "base": base, "n": len(sp)}


def _lists(ds, n):
    nc = _np(ds.n_cand)[:n]
    rows = _np(ds.c_row)[:n * ds.stride].reshape(n, ds.stride)
    hs = _np(ds.k2["hyperscore"])[:n * ds.stride].reshape(n, ds.stride)
    return nc, [tuple(rows[s, :nc[s]]) for s in range(n)], [tuple(hs[s, :nc[s]]) for s in range(n)]


def _same(a, b, n):
    na, ra, ha = _lists(a, n)
    nb, rb, hb = _lists(b, n)
    assert np.array_equal(na, nb)
    assert ra == rb
    assert ha == hb
    assert np.array_equal(a.rank1_rows(n), b.rank1_rows(n))


@pytest.mark.parametrize("band_log2", [28, 12])
def test_work_order_is_a_permutation_sorted_by_band_and_work(uneven, band_log2):
    """vh_spectrum_order_work alone (before any kernel consumes its output): a permutation of the spectra, bands
    ascending, estimated work descending inside a band (estimate recomputed on the host)."""
    import ctypes as C
    from vulcan.gpu import _lib as K
    u = uneven
    ds, n = u["base"], u["n"]
    order = torch.full((n,), -1, dtype=torch.int32, device=u["gi"].dev)
    hy, pk = ds.hyp, ds.peaks
    K.check(K.lib().vh_spectrum_order_work(n, K.ptr(hy["hyp_off"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]),
                                           K.ptr(hy["maxfz"]), K.ptr(pk["off"]), band_log2, K.ptr(order),
                                           K.ptr(ds.order_ws), ds.order_ws.numel(), K.stream()), "order_work")
    o = _np(order)
    assert np.array_equal(np.sort(o), np.arange(n))
    ho = _np(hy["hyp_off"])[:n + 1]
    lo, hi, mfz = _np(hy["lo"]), _np(hy["hi"]), _np(hy["maxfz"]).astype(np.int64)
    npk = np.diff(_np(pk["off"])[:n + 1])
    est = np.zeros(n, np.int64)
    band = np.full(n, 0xFFFFF, np.int64)
    for s in range(n):
        h = np.arange(ho[s], ho[s + 1])
        if len(h):
            est[s] = min(npk[s] * np.maximum(mfz[h] - 1, 1).sum() + (hi[h] - lo[h]).sum() // 4, 0xFFFFF)
            band[s] = min(lo[h[0]] >> band_log2, 0xFFFFF)
    want = np.lexsort((np.arange(n), -est, band))
    assert np.array_equal(o, want)


@pytest.mark.parametrize("mode", [{"persistent": True}, {"order_mode": "work"},
                                  {"persistent": True, "order_mode": "work", "band_log2": 12}])
def test_persistent_and_work_order_equal_default(uneven, mode):
    from vulcan.gpu import engine as E
    u = uneven
    ds = E.DeviceSearch(u["gi"], u["p"], u["n"], **mode)
    ds.run(u["raw"])
    torch.cuda.synchronize()
    assert int(ds.status.item()) == 0
    _same(ds, u["base"], u["n"])
    if mode.get("order_mode") == "work":
        o = _np(ds.order)[:u["n"]]
        assert np.array_equal(np.sort(o), np.arange(u["n"]))


@pytest.mark.parametrize("k1_impl", ["join", "tile"])
def test_class_schedule_graph_replays_equal_eager(uneven, k1_impl):
    from vulcan.gpu import engine as E
    from vulcan.gpu import schedule as S
    u = uneven
    rng = np.random.default_rng(3)
    cands = S.DEFAULT_CANDIDATES
    sched = S.ClassSchedule(u["gi"], u["p"], u["sp"], k1_impl=k1_impl)
    assert len(sched.classes) >= 6
    sched.set_table({c: cands[int(rng.integers(len(cands)))] for c, _, _ in sched.classes})
    sched.run(u["raw"])
    torch.cuda.synchronize()
    _same(sched, u["base"], u["n"])
    g = S.GraphStep(sched, u["raw"])
    other = E.DeviceSearch(u["gi"], u["p"], u["n"], persistent=True)
    for i in range(3):
        sched.c_row.fill_(-7)                       # the replay must rewrite every list
        sched.sel.fill_(-7)
        g.replay()
        torch.cuda.synchronize()
        _same(sched, u["base"], u["n"])
        other.run(u["raw"])                         # eager work between replays
        x = torch.randn(1 << 20, device=u["gi"].dev).sum().item()
        assert np.isfinite(x)
        torch.cuda.synchronize()
    _same(other, u["base"], u["n"])


def test_autotune_picks_a_configuration_for_every_class(uneven):
    from vulcan.gpu import schedule as S
    u = uneven
    sched = S.ClassSchedule(u["gi"], u["p"], u["sp"])
    cands = [{"tile": 512, "xcd_group": 32, "persistent": False}, {"tile": 1024, "xcd_group": 0, "persistent": True}]
    res = S.autotune(sched, u["raw"], candidates=cands, reps=1)
    assert set(res["table"]) == {c for c, _, _ in sched.classes}
    assert all(res["table"][c] in cands for c in res["table"])
    sched.run(u["raw"])
    torch.cuda.synchronize()
    _same(sched, u["base"], u["n"])


def test_tune_step_picks_a_configuration_with_identical_results(uneven):
    from vulcan.gpu import schedule as S
    u = uneven
    cands = [{"kind": "step", "persistent": True, "order_mode": "work", "xcd_group": 0}, {"kind": "classes"}]
    obj, tried = S.tune_step(u["gi"], u["p"], u["sp"], u["raw"], candidates=cands, reps=1)
    assert [c for c, _ in tried] == cands and all(t > 0 for _, t in tried)
    obj.run(u["raw"])
    torch.cuda.synchronize()
    assert int(obj.status.item()) == 0
    _same(obj, u["base"], u["n"])


@pytest.mark.parametrize("kind", ["classes", "persistent_work"])
def test_pipeline_of_tuned_steps_equals_default(uneven, kind):
    """Slots built by step_factory from a tuned step (a ClassSchedule with its table; a persistent, work-ordered
    DeviceSearch), several steps in flight with no synchronisation between submissions: every slot's lists,
    scores and rank 1 equal the default eager step's."""
    from vulcan.gpu import engine as E
    from vulcan.gpu import schedule as S
    u = uneven
    if kind == "classes":
        obj = S.ClassSchedule(u["gi"], u["p"], u["sp"], table={0: {"xcd_group": 0, "persistent": False}})
    else:
        obj = E.DeviceSearch(u["gi"], u["p"], u["n"], persistent=True, order_mode="work")
    pipe = E.SearchPipeline(u["gi"], u["p"], u["n"], depth=3, make=S.step_factory(u["gi"], u["sp"], obj))
    assert all(type(s) is type(obj) for s in pipe.slots)
    if kind == "classes":
        assert all(s.table == obj.table for s in pipe.slots)
    for _ in range(7):
        pipe.run(u["raw"])
    pipe.synchronize()
    for s in pipe.slots:
        assert int(s.status.item()) == 0
        _same(s, u["base"], u["n"])


def test_uneven_batch_equals_oracle(uneven, proteome_small):
    """The uneven batch (charges 2-6, 20-500 peaks, per-run calibration offsets) against the oracle: the default
    step's candidate lists, hyperscores and rank 1, and the rank 1 of the tuned class schedule replayed as a
    hipGraph."""
    from oracle import oracle as O
    from vulcan.gpu import schedule as S
    from vulcan.gpu._abi import make_params, max_hyp_per_spectrum
    u = uneven
    d = proteome_small
    n, ds = u["n"], u["base"]
    vp = make_params(u["p"], STATIC, d["var"])
    hix = O.HostIndex(d["l1"])
    pk = O.k0(vp, u["sp"])
    hyp = O.hypotheses(vp, u["sp"], pk["valid"], d["l1"].mass32, max_hyp_per_spectrum(u["p"]))
    assert len(np.unique(np.asarray(u["sp"].charge))) == 5
    k1 = O.k1_counts(vp, hix, pk, hyp)
    tk = O.topk(vp, hix, hyp, k1, ds.stride)
    nc = _np(ds.n_cand)[:n]
    assert np.array_equal(nc, tk["n_cand"])
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    rows, hyps = _np(ds.c_row)[:n * ds.stride][m], _np(ds.c_hyp)[:n * ds.stride][m]
    assert np.array_equal(rows, tk["row"][m]) and np.array_equal(hyps, tk["hyp"][m])
    start = np.concatenate([[0], np.cumsum(nc)[:-1]])
    ref = O.k2(vp, hix, pk, rows, hyps, hyp["maxfz"], grp_start=start, grp_count=nc)
    assert np.allclose(_np(ds.k2["hyperscore"])[:n * ds.stride][m], ref["hyperscore"], rtol=1e-12, atol=0)
    assert np.array_equal(_np(ds.k2["matched_y"])[:n * ds.stride][m], ref["matched_y"])
    full = O.search(vp, hix, u["sp"])
    assert np.array_equal(ds.rank1_rows(n), full["row"])
    sched = S.ClassSchedule(u["gi"], u["p"], u["sp"])
    g = S.GraphStep(sched, u["raw"])
    g.replay()
    torch.cuda.synchronize()
    assert np.array_equal(sched.rank1_rows(n), full["row"])
