"""Multi-run feed (store read + pinned double-buffered upload overlapped with the previous run's search) vs one
search per run: the same rank-1 rows and the same PSM tables, runs of different sizes through one preallocated
DeviceSearch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("needs a ROCm GPU", allow_module_level=True)


@pytest.fixture(scope="module")
def runs(small_plain, tmp_path_factory):
    from vulcan.gpu import engine as E
    from vulcan.index import synth
    d = small_plain
    out = tmp_path_factory.mktemp("stores")
    paths, sps = [], []
    for i, n in enumerate((700, 1500, 400)):
        sp, _, _ = synth.spectra(d["l1"], d["labels"], 0, n=n, seed=50 + i)
        sp.run = f"run{i}"
        path = str(out / f"run{i}.vspec.arrow")
        E.write_store(sp, path)
        paths.append(path)
        sps.append(E.load_store(path))
    return d, paths, sps


def test_feed_equals_per_run_search(runs, tmp_path):
    from vulcan.gpu import engine as E
    from vulcan.gpu.feed import RunFeed
    d, paths, sps = runs
    print("[gpu] feed: 3 runs (700 / 1500 / 400 spectra)", flush=True)
    gi = E.GpuIndex(d["l1"], d["p"], d["static"], d["var"])
    feed = RunFeed(gi, d["p"], max_spectra=1500, max_peaks=max(int(s.offsets[-1]) for s in sps))
    # capacities far below the first run's pairs / slots, marked fitted: the kernels bound their writes and
    # flag the overflow, the feed refits and searches the run again
    feed.ds._join_alloc(feed.ds.join, 64, 16)
    feed.ds.join["fitted"] = True
    res = feed.run(paths, l1=d["l1"], out_dir=str(tmp_path))
    assert 1 <= feed.refits <= 2          # the 700-spectrum run, perhaps the 1,500 one
    torch.cuda.synchronize()
    assert [r["run"] for r in res] == ["run0", "run1", "run2"]
    for r, sp in zip(res, sps):
        tbl, _, dbg = E.search(sp, gi, d["l1"], d["p"])
        rows = dbg["device"].rank1_rows(len(sp))
        assert r["n_spectra"] == len(sp)
        assert np.array_equal(r["rows"], rows), r["run"]
        assert r["table"].num_rows == tbl.num_rows and r["table"].num_rows > 0
        assert r["table"].schema == tbl.schema
        for name in tbl.column_names:
            a, b = r["table"].column(name).to_numpy(), tbl.column(name).to_numpy()
            if a.dtype.kind == "f":
                assert np.array_equal(a, b, equal_nan=True), (r["run"], name)
            else:
                assert np.array_equal(a, b), (r["run"], name)
        import pyarrow.parquet as pq
        assert pq.read_table(r["path"]).num_rows == tbl.num_rows


def test_feed_equals_oracle_search_per_store(runs):
    """The feed's rank-1 rows and its PSM tables' rank-1 rows and hyperscores equal the oracle's whole search
    (or_search: K0, hypotheses, per-row K1, top-K + closure, K2, rank) on each loaded store's spectra."""
    from oracle import oracle as O
    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    from vulcan.gpu.feed import RunFeed
    d, paths, sps = runs
    print("[gpu] feed vs the oracle's search: 3 runs", flush=True)
    gi = E.GpuIndex(d["l1"], d["p"], d["static"], d["var"])
    feed = RunFeed(gi, d["p"], max_spectra=1500, max_peaks=max(int(s.offsets[-1]) for s in sps))
    res = feed.run(paths, l1=d["l1"])
    torch.cuda.synchronize()
    vp = make_params(d["p"], d["static"], d["var"])
    hix = O.HostIndex(d["l1"])
    for r, sp in zip(res, sps):
        ref = O.search(vp, hix, sp)
        assert np.array_equal(r["rows"], ref["row"].astype(r["rows"].dtype)), r["run"]
        t = r["table"]
        one = t.column("rank").to_numpy() == 1
        s, row, hs = (t.column(k).to_numpy()[one] for k in ("spec_idx", "pf_row", "hyperscore"))
        want = np.flatnonzero(ref["row"] >= 0)
        assert np.array_equal(s, want) and np.array_equal(row, ref["row"][want]), r["run"]
        assert np.allclose(hs, ref["hs"][want], rtol=1e-12, atol=0), r["run"]
        assert len(want) > 0.5 * len(sp)


def test_feed_rejects_oversized_run(runs):
    from vulcan.gpu import engine as E
    from vulcan.gpu.feed import RunFeed
    d, paths, sps = runs
    gi = E.GpuIndex(d["l1"], d["p"], d["static"], d["var"])
    feed = RunFeed(gi, d["p"], max_spectra=800, max_peaks=10 ** 6)
    with pytest.raises(ValueError, match="exceeds the feed"):
        feed.run(paths[:2])


def test_search_many_writes_per_run_tables(runs, tmp_path):
    """batch.search_many (reference batch.py:21-41 result keys) writes the tables one search per run gives."""
    import pyarrow.parquet as pq
    from vulcan.gpu import batch
    from vulcan.gpu import engine as E
    d, paths, sps = runs
    print("[gpu] search_many: 3 runs", flush=True)
    gi = E.GpuIndex(d["l1"], d["p"], d["static"], d["var"])
    r = batch.search_many(paths, gi, d["l1"], d["p"], str(tmp_path / "out"), io_threads=2)
    assert r["n_spectra"] == sum(len(s) for s in sps) and [x["run"] for x in r["runs"]] == ["run0", "run1", "run2"]
    assert all(x["gpu_ms"] > 0 and x["h2d_bytes"] > 0 for x in r["runs"])
    for sp in sps:
        tbl, _, _ = E.search(sp, gi, d["l1"], d["p"])
        got = pq.read_table(str(tmp_path / "out" / f"{sp.run}.psms.parquet"))
        assert got.num_rows == tbl.num_rows
        for name in ("spec_idx", "pf_row", "rank", "hyperscore", "matched_peaks"):
            assert np.array_equal(got.column(name).to_numpy(), tbl.column(name).to_numpy()), (sp.run, name)


def test_run_and_batch_clis_write_the_same_tables(runs, tmp_path):
    """python -m vulcan.gpu.run and python -m vulcan.gpu.batch (reference run.py / batch.py CLIs) over an index
    written to disk: the same PSM files, and equal to one engine.search per run."""
    import json

    import pyarrow.parquet as pq
    from vulcan.gpu import batch, run
    from vulcan.gpu import engine as E
    from vulcan.index.build import L1Index
    d, paths, sps = runs
    idx = str(tmp_path / "idx")
    d["l1"].save(idx, d["labels"])
    l1 = L1Index(idx)
    params = json.dumps({k: list(v) if isinstance(v, tuple) else v for k, v in d["p"].to_dict().items()})
    var = json.dumps(d["var"])
    print("[gpu] run / batch CLIs", flush=True)
    assert run.main(["--index", idx, "--stores", *paths, "--out", str(tmp_path / "r"), "--params", params,
                     "--var", var]) == 0
    assert batch.main(["--index", idx, "--stores", *paths, "--out", str(tmp_path / "b"), "--params", params]) == 0
    assert json.load(open(tmp_path / "r" / "stats.json"))["passes"][0][0]["run"] == "run0"
    gi = E.GpuIndex(l1, d["p"], d["static"], d["var"])
    for sp in sps:
        want, _, _ = E.search(sp, gi, l1, d["p"])
        # batch.main binds the reference's fixed mods (C static, M-ox variable); this table has no variable
        # sites, so its results equal those of the empty variable set
        for out in ("r", "b"):
            a = pq.read_table(str(tmp_path / out / f"{sp.run}.psms.parquet"))
            assert a.num_rows == want.num_rows, out
            for name in ("spec_idx", "pf_row", "pf_id", "hyperscore", "poisson"):
                assert np.array_equal(a.column(name).to_numpy(), want.column(name).to_numpy(), equal_nan=True), \
                    (out, name)


def test_empty_runs(runs, tmp_path):
    """An empty run (no spectra) searched on its own and in the middle of a feed: no PSMs, and the runs around
    it are unaffected."""
    from vulcan.gpu import engine as E
    from vulcan.gpu.feed import RunFeed
    from vulcan.gpu.host import Spectra
    d, paths, sps = runs
    print("[gpu] empty runs", flush=True)
    empty = Spectra("empty", np.zeros(0, np.uint32), np.zeros(0, np.uint32), np.zeros(0), np.zeros(0),
                    np.zeros(0, np.int8), np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(1, np.int64))
    gi = E.GpuIndex(d["l1"], d["p"], d["static"], d["var"])
    tbl, st, _ = E.search(empty, gi, d["l1"], d["p"])
    assert tbl.num_rows == 0 and st.counts["n_spectra"] == 0
    ep = str(tmp_path / "empty.vspec.arrow")
    E.write_store(empty, ep)
    feed = RunFeed(gi, d["p"], max_spectra=1500, max_peaks=max(int(s.offsets[-1]) for s in sps))
    res = feed.run([paths[0], ep, paths[2]], l1=d["l1"])
    assert [r["n_spectra"] for r in res] == [len(sps[0]), 0, len(sps[2])]
    assert res[1]["table"].num_rows == 0 and len(res[1]["rows"]) == 0
    for r, sp in ((res[0], sps[0]), (res[2], sps[2])):
        want, _, dbg = E.search(sp, gi, d["l1"], d["p"])
        assert np.array_equal(r["rows"], dbg["device"].rank1_rows(len(sp)))
        assert r["table"].num_rows == want.num_rows
