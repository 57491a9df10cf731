"""Run store (`<run>.vspec.arrow`) writer / reader and load_store (CPU only): format invariants of the store
specification (spectrum_id, index, sidecar counts, chunked offsets, MS1 filtering, unknown charges)."""
import json

import numpy as np
import pyarrow as pa

from vulcan.io import store as S


def _run(n, seed):
    rng = np.random.default_rng(seed)
    npk = rng.integers(0, 40, n)
    off = np.concatenate([[0], np.cumsum(npk)]).astype(np.int64)
    mz = np.sort(rng.uniform(150, 2000, off[-1])).astype(np.float32)
    it = rng.lognormal(0, 1, off[-1]).astype(np.float32)
    return off, mz, it


def test_roundtrip_and_invariants(tmp_path):
    n = 2 * S.BATCH_SPECTRA + 37                       # three record batches
    off, mz, it = _run(n, 1)
    level = np.where(np.arange(n) % 5 == 0, 1, 2).astype(np.uint8)
    z = np.random.default_rng(2).integers(0, 5, n).astype(np.int8)   # 0 = unknown (stored null)
    pmz = np.where(level == 2, np.random.default_rng(3).uniform(300, 1500, n), np.nan)
    path = str(tmp_path / ("r1" + S.FILE_SUFFIX))
    w = S.StoreWriter(path, {"run_name": "r1"})
    w.write({"offsets": off, "mz": mz, "intensity": it, "ms_level": level, "rt_min": np.arange(n) * 0.01,
             "precursor_mz": pmz, "precursor_charge": z, "scan": np.arange(n) + 100})
    side = w.close()
    assert side == json.load(open(path + ".json"))
    assert side["n_spectra"] == n and side["n_peaks"] == int(off[-1]) and side["n_ms1"] == int((level == 1).sum())
    t = S.read_table(path)
    assert t.num_rows == n and len(t.column("mz").chunks) == 3
    meta = S.run_meta(path)
    assert meta["run_name"] == "r1" and meta["run_id"] == S.default_run_id("r1") and meta["store_version"] == 1
    idx = t.column("index").to_numpy()
    assert np.array_equal(idx, np.arange(n))
    sid = t.column("spectrum_id").to_numpy()
    assert np.array_equal(sid, (np.uint64(meta["run_id"]) << np.uint64(32)) | idx.astype(np.uint64))
    pk = S.peaks_numpy(t)
    assert np.array_equal(pk.offsets, off) and np.array_equal(pk.mz, mz) and np.array_equal(pk.intensity, it)
    ch = t.column("precursor_charge")
    assert ch.null_count == int(((z == 0) | (level == 1)).sum())   # unknown charge, and every MS1 row
    assert t.column("precursor_mz").null_count == int((level == 1).sum())
    assert not any(p.name.endswith(".partial") for p in tmp_path.iterdir())
    # sidecar counts equal the counts recomputed from the file
    lv = t.column("ms_level").to_numpy()
    assert side["n_ms2"] == int((lv == 2).sum()) and side["n_by_level"] == {"1": int((lv == 1).sum()),
                                                                              "2": int((lv == 2).sum())}
    assert side["n_peaks"] == sum(len(c.values) for c in t.column("mz").chunks)


def test_ms1_rows_have_null_precursor_and_activation(tmp_path):
    n = 50
    off, mz, it = _run(n, 8)
    level = np.where(np.arange(n) % 4 == 0, 1, 2).astype(np.uint8)
    path = str(tmp_path / "r5.vspec.arrow")
    w = S.StoreWriter(path, {"run_name": "r5", "run_id": 7})
    w.write({"offsets": off, "mz": mz, "intensity": it, "ms_level": level, "rt_min": np.arange(n) * 0.1,
             "precursor_mz": np.full(n, 500.0), "precursor_charge": np.full(n, 2, np.int8),
             "activation": np.full(n, 2, np.uint8)})
    w.close()
    t = S.read_table(path)
    for f in ("precursor_mz", "precursor_charge", "activation"):
        nulls = t.column(f).is_null().to_numpy(zero_copy_only=False)
        assert np.array_equal(nulls, level == 1), f
    sid = t.column("spectrum_id").to_numpy()
    assert np.array_equal(sid, (np.uint64(7) << np.uint64(32)) | np.arange(n, dtype=np.uint64))
    assert S.run_meta(path)["run_id"] == 7 and 0 <= S.default_run_id("abc") < 2 ** 32
    assert S.open_store(path).num_record_batches == 1
    assert S.counts(path) == (n, int(off[-1]))
    import os
    os.remove(path + ".json")
    assert S.counts(path) == (n, int(off[-1]))        # recomputed from the file without the sidecar


def test_load_store_keeps_ms2(tmp_path):
    from vulcan.gpu import engine as E
    from vulcan.gpu.host import Spectra
    n = 300
    off, mz, it = _run(n, 4)
    level = np.where(np.arange(n) % 3 == 0, 1, 2).astype(np.uint8)
    z = np.random.default_rng(5).integers(0, 5, n).astype(np.int8)
    pmz = np.random.default_rng(6).uniform(300, 1500, n)
    path = str(tmp_path / "r2.vspec.arrow")
    w = S.StoreWriter(path, {"run_name": "r2"})
    w.write({"offsets": off, "mz": mz, "intensity": it, "ms_level": level, "rt_min": np.arange(n) * 0.5,
             "precursor_mz": pmz, "precursor_charge": z, "scan": np.arange(n) + 7})
    w.close()
    sp = E.load_store(path)
    keep = np.flatnonzero(level == 2)
    assert sp.run == "r2" and len(sp) == len(keep)
    assert np.array_equal(sp.scan, keep + 7) and np.array_equal(sp.index, keep)
    assert np.array_equal(sp.charge, z[keep]) and np.array_equal(sp.precursor_mz, pmz[keep])
    for j, i in enumerate(keep):
        assert np.array_equal(sp.mz[sp.offsets[j]:sp.offsets[j + 1]], mz[off[i]:off[i + 1]])
    # write_store is load_store's inverse on MS2 runs
    sp2 = Spectra("r3", sp.scan, sp.index, sp.rt, sp.precursor_mz, sp.charge, sp.mz, sp.intensity, sp.offsets)
    E.write_store(sp2, str(tmp_path / "r3.vspec.arrow"))
    back = E.load_store(str(tmp_path / "r3.vspec.arrow"))
    for k in ("scan", "rt", "precursor_mz", "charge", "mz", "intensity", "offsets"):
        assert np.array_equal(getattr(back, k), getattr(sp2, k)), k
    assert isinstance(S.read_table(str(tmp_path / "r3.vspec.arrow"), ["scan"]), pa.Table)


def test_load_store_into_buffers_and_threads(tmp_path):
    """load_store into preallocated buffers (the feed's pinned slots) and with parallel batch copies equals the
    plain load; a run too large for the buffers is refused before anything is written past them."""
    from vulcan.gpu import engine as E
    n = 3 * S.BATCH_SPECTRA + 5
    off, mz, it = _run(n, 7)
    level = np.where(np.arange(n) % 7 == 3, 1, 2).astype(np.uint8)
    level[S.BATCH_SPECTRA:2 * S.BATCH_SPECTRA] = 2      # one batch without MS1 rows
    path = str(tmp_path / "r4.vspec.arrow")
    w = S.StoreWriter(path, {"run_name": "r4"})
    w.write({"offsets": off, "mz": mz, "intensity": it, "ms_level": level, "rt_min": np.arange(n) * 0.1,
             "precursor_mz": np.full(n, 500.0), "precursor_charge": np.full(n, 2, np.int8)})
    w.close()
    ref = E.load_store(path, threads=1)
    big = {"mz": np.full(len(mz), -1, np.float32), "int": np.full(len(mz), -1, np.float32),
           "off": np.full(n + 1, -1, np.int64), "pmz": np.zeros(n), "z": np.zeros(n, np.int8)}
    got = E.load_store(path, out=big, threads=3)
    for k in ("scan", "index", "rt", "precursor_mz", "charge", "mz", "intensity", "offsets"):
        assert np.array_equal(getattr(got, k), getattr(ref, k)), k
    assert np.shares_memory(got.mz, big["mz"]) and np.shares_memory(got.offsets, big["off"])
    small = {k: v[:len(v) // 2] for k, v in big.items()}
    import pytest
    with pytest.raises(ValueError, match="do not fit"):
        E.load_store(path, out=small)


def test_empty_and_ms1_only_stores(tmp_path):
    from vulcan.gpu import engine as E
    path = str(tmp_path / "empty.vspec.arrow")
    w = S.StoreWriter(path, {"run_name": "empty"})
    w.write({"offsets": np.zeros(1, np.int64), "mz": np.zeros(0, np.float32), "intensity": np.zeros(0, np.float32),
             "rt_min": np.zeros(0)})
    side = w.close()
    assert side["n_spectra"] == 0 and side["n_peaks"] == 0 and S.counts(path) == (0, 0)
    sp = E.load_store(path)
    assert len(sp) == 0 and len(sp.offsets) == 1 and len(sp.mz) == 0
    off, mz, it = _run(20, 9)
    path1 = str(tmp_path / "ms1.vspec.arrow")
    w = S.StoreWriter(path1, {"run_name": "ms1"})
    w.write({"offsets": off, "mz": mz, "intensity": it, "ms_level": np.ones(20, np.uint8), "rt_min": np.arange(20.0)})
    w.close()
    sp1 = E.load_store(path1)
    assert len(sp1) == 0 and sp1.offsets.tolist() == [0] and len(sp1.mz) == 0


def test_reads_a_store_written_by_the_reference_writer():
    """tests/golden/ref_run.vspec.arrow was written by the reference's own StoreWriter (full schema, two record
    batches, MS1 rows, null charges; tests/golden/make_store_fixture.py): this reader and load_store give back
    the arrays it was given, and the columns this build writes have the reference's names and types."""
    import os

    from vulcan.gpu import engine as E
    g = os.path.join(os.path.dirname(__file__), "golden")
    path = os.path.join(g, "ref_run.vspec.arrow")
    want = json.load(open(os.path.join(g, "ref_run_inputs.json")))
    t = S.read_table(path)
    assert t.num_rows == len(want["scan"]) and len(t.column("mz").chunks) == 2
    for f in S.FIELDS:
        assert t.schema.field(f.name).type == f.type, f.name
    meta = S.run_meta(path)
    assert meta["run_name"] == "ref_run" and meta["run_id"] == S.default_run_id("ref_run")
    assert np.array_equal(t.column("spectrum_id").to_numpy(),
                          (np.uint64(meta["run_id"]) << np.uint64(32)) | np.arange(t.num_rows, dtype=np.uint64))
    pk = S.peaks_numpy(t)
    assert np.array_equal(np.diff(pk.offsets), want["peak_counts"])
    assert np.array_equal(pk.mz, np.array(want["mz"], np.float32))
    assert np.array_equal(pk.intensity, np.array(want["intensity"], np.float32))
    assert S.counts(path) == (len(want["scan"]), len(want["mz"]))
    sp = E.load_store(path)
    lv = np.array(want["ms_level"])
    keep = np.flatnonzero(lv == 2)
    assert len(sp) == len(keep) and np.array_equal(sp.scan, np.array(want["scan"])[keep])
    assert np.array_equal(sp.rt, np.array(want["rt_min"])[keep])
    assert np.array_equal(sp.precursor_mz, np.array(want["precursor_mz"], np.float64)[keep])
    z = np.where(np.array(want["precursor_charge_null"]), 0, np.array(want["precursor_charge"]))[keep]
    assert np.array_equal(sp.charge, z.astype(np.int8))
    cnt = np.array(want["peak_counts"])
    starts = np.concatenate([[0], np.cumsum(cnt)])
    for j, i in enumerate(keep):
        assert np.array_equal(sp.mz[sp.offsets[j]:sp.offsets[j + 1]],
                              np.array(want["mz"], np.float32)[starts[i]:starts[i + 1]])
