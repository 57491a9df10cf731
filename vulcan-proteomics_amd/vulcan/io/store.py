"""Spectral store of one run (`<run>.vspec.arrow`): the search path's input format (vulcan/io/STORE_SPEC.md of the
reference, format version 1; reader / writer API of vulcan/io/store.py:62-208).

An Arrow IPC file, one row per spectrum in acquisition order, peaks in list<f32> columns whose offsets are chunked
per record batch (at most BATCH_SPECTRA spectra per batch), run metadata as JSON under the schema-metadata key
"vulcan", final counts in a `<file>.json` sidecar, written as `<file>.partial` and renamed on close. The column
names and types are the format's (STORE_SPEC.md section 2); this module writes the identity, timing, precursor,
acquisition and peak columns the search and the PSM table use, and reads any store that has them.
"""
from __future__ import annotations

import hashlib
import json
import os
from dataclasses import dataclass

import numpy as np
import pyarrow as pa

STORE_VERSION = 1
FILE_SUFFIX = ".vspec.arrow"
BATCH_SPECTRA = 4096
META_KEY = b"vulcan"

FIELDS = [
    pa.field("spectrum_id", pa.uint64(), nullable=False),
    pa.field("run_id", pa.uint32(), nullable=False),
    pa.field("index", pa.uint32(), nullable=False),
    pa.field("scan", pa.uint32(), nullable=False),
    pa.field("ms_level", pa.uint8(), nullable=False),
    pa.field("rt_min", pa.float64(), nullable=False),
    pa.field("precursor_mz", pa.float64()),
    pa.field("precursor_charge", pa.int8()),
    pa.field("activation", pa.uint8()),
    pa.field("analyzer", pa.uint8(), nullable=False),
    pa.field("polarity", pa.int8(), nullable=False),
    pa.field("tic", pa.float32()),
    pa.field("peak_source", pa.uint8(), nullable=False),
    pa.field("mz", pa.list_(pa.float32()), nullable=False),
    pa.field("intensity", pa.list_(pa.float32()), nullable=False),
]


def default_run_id(run_name: str) -> int:
    """First 4 bytes of sha1(run_name), little-endian (the format's default before a catalog assigns one)."""
    return int.from_bytes(hashlib.sha1(run_name.encode()).digest()[:4], "little")


def schema(run_meta: dict) -> pa.Schema:
    return pa.schema(FIELDS, metadata={META_KEY: json.dumps(run_meta).encode()})


class StoreWriter:
    """Batched writer: write(cols) appends spectra (column arrays; `mz` / `intensity` flat with `offsets`),
    close() renames the file into place and writes the sidecar counts."""

    def __init__(self, path: str, run_meta: dict):
        self.path = path
        meta = {"store_version": STORE_VERSION, "peak_order": "mz_ascending", **run_meta}
        meta.setdefault("run_id", default_run_id(meta["run_name"]))
        self.meta = meta
        self._schema = schema(meta)
        self._sink = pa.OSFile(path + ".partial", "wb")
        self._w = pa.ipc.new_file(self._sink, self._schema)
        self.n = 0
        self.n_peaks = 0
        self.by_level: dict = {}

    def write(self, cols: dict) -> None:
        off = np.asarray(cols["offsets"], np.int64)
        n = len(off) - 1
        for a in range(0, n, BATCH_SPECTRA):
            b = min(n, a + BATCH_SPECTRA)
            self._batch({k: v[a:b] for k, v in cols.items() if k not in ("offsets", "mz", "intensity")},
                        cols["mz"][off[a]:off[b]], cols["intensity"][off[a]:off[b]], off[a:b + 1] - off[a])

    def _batch(self, c: dict, mz, inten, off) -> None:
        n = len(off) - 1
        idx = np.arange(self.n, self.n + n, dtype=np.uint32)
        rid = np.uint64(self.meta["run_id"])
        level = np.asarray(c.get("ms_level", np.full(n, 2)), np.uint8)

        ms1 = level == 1

        def opt(name, typ):
            """Nullable column: null where not finite (f) / not positive (charge), and on MS1 rows for the
            precursor and activation fields (the format's invariant 3)."""
            v = c.get(name)
            if v is None:
                return pa.nulls(n, typ)
            v = np.asarray(v)
            mask = ~np.isfinite(v) if v.dtype.kind == "f" else np.zeros(n, bool)
            if name == "precursor_charge":
                mask = v <= 0
            if name in ("precursor_mz", "precursor_charge", "activation"):
                mask = mask | ms1
            return pa.array(v, typ, mask=mask)

        off32 = pa.array(np.asarray(off, np.int32))
        arrays = [
            pa.array((rid << np.uint64(32)) | idx.astype(np.uint64), pa.uint64()),
            pa.array(np.full(n, self.meta["run_id"], np.uint32)), pa.array(idx),
            pa.array(np.asarray(c.get("scan", idx + 1), np.uint32)), pa.array(level),
            pa.array(np.asarray(c["rt_min"], np.float64)),
            opt("precursor_mz", pa.float64()), opt("precursor_charge", pa.int8()), opt("activation", pa.uint8()),
            pa.array(np.asarray(c.get("analyzer", np.ones(n)), np.uint8)),
            pa.array(np.asarray(c.get("polarity", np.ones(n)), np.int8)), opt("tic", pa.float32()),
            pa.array(np.asarray(c.get("peak_source", np.zeros(n)), np.uint8)),
            pa.ListArray.from_arrays(off32, pa.array(np.asarray(mz, np.float32))),
            pa.ListArray.from_arrays(off32, pa.array(np.asarray(inten, np.float32))),
        ]
        self._w.write_batch(pa.record_batch(arrays, schema=self._schema))
        self.n += n
        self.n_peaks += int(off[-1])
        for lv, k in zip(*np.unique(level, return_counts=True)):
            self.by_level[str(int(lv))] = self.by_level.get(str(int(lv)), 0) + int(k)

    def close(self, extra_meta: dict | None = None) -> dict:
        self._w.close()
        self._sink.close()
        os.replace(self.path + ".partial", self.path)
        side = {"n_spectra": self.n, "n_peaks": self.n_peaks, "n_by_level": self.by_level,
                "n_ms1": self.by_level.get("1", 0), "n_ms2": self.by_level.get("2", 0), **(extra_meta or {})}
        with open(self.path + ".json", "w") as fh:
            json.dump(side, fh, indent=1)
        return side


def open_store(path: str) -> pa.ipc.RecordBatchFileReader:
    """Memory-mapped under /tmp (local disk), a buffered file elsewhere (the format's I/O policy)."""
    src = pa.memory_map(path, "r") if os.path.abspath(path).startswith("/tmp/") else pa.OSFile(path, "rb")
    return pa.ipc.open_file(src)


def read_table(path: str, columns=None) -> pa.Table:
    t = open_store(path).read_all()
    return t.select(columns) if columns is not None else t


def counts(path: str) -> tuple:
    """(spectra, peaks) of a store: from its sidecar when present, else recomputed from the file."""
    try:
        with open(path + ".json") as fh:
            side = json.load(fh)
        return int(side["n_spectra"]), int(side["n_peaks"])
    except (OSError, KeyError, ValueError):
        t = read_table(path, ["mz"])
        return t.num_rows, sum(int(c.offsets[-1].as_py() - c.offsets[0].as_py()) for c in t.column("mz").chunks)


def run_meta(path: str) -> dict:
    md = open_store(path).schema.metadata or {}
    return json.loads(md.get(META_KEY, b"{}"))


@dataclass
class Peaks:
    mz: np.ndarray          # f32, flat
    intensity: np.ndarray   # f32, flat
    offsets: np.ndarray     # i64, n + 1


def _chunk_plan(table: pa.Table, keep: np.ndarray | None):
    """Per record batch: (list offsets, kept rows mask or None, kept peaks) of the `mz` column."""
    plan, row = [], 0
    for cm in table.column("mz").chunks:
        o = cm.offsets.to_numpy().astype(np.int64)
        k = len(cm)
        m = None if keep is None else keep[row:row + k]
        if m is not None and m.all():
            m = None
        cnt = np.diff(o)
        plan.append((o, m, int(o[-1] - o[0]) if m is None else int(cnt[m].sum())))
        row += k
    return plan


def peaks_numpy(table: pa.Table, keep: np.ndarray | None = None, out: dict | None = None,
                threads: int = 1) -> Peaks:
    """The list columns flattened (rows where `keep` is set, all rows without it), straight from the record
    batches into one buffer per column with int64 offsets: one copy of the peak bytes, no filtered table.
    out: optional preallocated {"mz", "int" (f32), "off" (i64)} arrays that are large enough (e.g. views of
    pinned host memory); the returned arrays are views of them. threads > 1 copies record batches in parallel
    (numpy releases the GIL for the copies)."""
    mz_c, it_c = table.column("mz"), table.column("intensity")
    plan = _chunk_plan(table, keep)
    rows = [len(o) - 1 if m is None else int(np.count_nonzero(m)) for o, m, _ in plan]
    n, npk = sum(rows), sum(c for _, _, c in plan)
    if out is None:
        out = {"mz": np.empty(npk, np.float32), "int": np.empty(npk, np.float32), "off": np.empty(n + 1, np.int64)}
    elif len(out["mz"]) < npk or len(out["int"]) < npk or len(out["off"]) < n + 1:
        raise ValueError(f"{n} spectra / {npk} peaks do not fit the output buffers")
    mz_o, it_o, off_o = out["mz"][:npk], out["int"][:npk], out["off"][:n + 1]
    off_o[0] = 0
    row0 = np.concatenate([[0], np.cumsum(rows)]).astype(np.int64)
    base0 = np.concatenate([[0], np.cumsum([c for _, _, c in plan])]).astype(np.int64)

    def one(i):
        o, m, c = plan[i]
        k, row, base = rows[i], int(row0[i]), int(base0[i])
        if k == 0:
            return
        cm, ci = mz_c.chunks[i], it_c.chunks[i]
        vm = cm.values.to_numpy(zero_copy_only=False)[o[0]:o[-1]]
        vi = ci.values.to_numpy(zero_copy_only=False)[o[0]:o[-1]]
        cnt = np.diff(o)
        if m is None:
            mz_o[base:base + c] = vm
            it_o[base:base + c] = vi
        else:
            pm = np.repeat(m, cnt)
            np.compress(pm, vm, out=mz_o[base:base + c])
            np.compress(pm, vi, out=it_o[base:base + c])
            cnt = cnt[m]
        np.cumsum(cnt, out=off_o[row + 1:row + k + 1])
        off_o[row + 1:row + k + 1] += base

    if threads > 1 and len(plan) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(min(threads, len(plan))) as ex:
            list(ex.map(one, range(len(plan))))
    else:
        for i in range(len(plan)):
            one(i)
    return Peaks(mz_o, it_o, off_o)
