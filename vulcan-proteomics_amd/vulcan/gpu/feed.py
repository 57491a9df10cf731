"""Multi-run feed: search a list of run stores back to back with the host side of run i+1 (store read, copy into
pinned host memory, host -> device copy) overlapped with the search of run i.

The reference overlaps store loading with the search of the previous run in a thread pool (vulcan/gpu/batch.py:
21-41, cohort.py:512-790) and uploads each run with synchronous copies. Here two pinned host slots and two device
slots alternate: a host thread reads the next store's record batches straight into its pinned slot (one copy of
the peak bytes), the copy to the device runs on a copy stream (default priority) and the search stream waits for
it with an event; the device slot a copy overwrites is released by an event recorded after the search that last
read it, and the pinned slot by the event that ends its previous copy. One DeviceSearch
(preallocated for the largest run) serves every run.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from .engine import DeviceSearch, GpuIndex, device_table, load_store, write_psms
from .host import SearchParams, Spectra

_FIELDS = (("mz", torch.float32), ("int", torch.float32), ("off", torch.int64), ("pmz", torch.float64),
           ("z", torch.int8))


class RunFeed:
    """Searches stores of up to max_spectra spectra and max_peaks raw peaks each."""

    def __init__(self, gi: GpuIndex, p: SearchParams, max_spectra: int, max_peaks: int, report: int = 1,
                 threads: int = 4):
        self.gi, self.p, self.threads = gi, p, threads
        self.max_spectra, self.max_peaks = max_spectra, max_peaks
        dev = gi.dev
        self.ds = DeviceSearch(gi, p, max_spectra, report=report)
        size = {"mz": max_peaks, "int": max_peaks, "off": max_spectra + 1, "pmz": max_spectra, "z": max_spectra}
        self.host = [{k: torch.empty(size[k], dtype=t, pin_memory=True) for k, t in _FIELDS} for _ in range(2)]
        self.dev = [{k: torch.empty(size[k], dtype=t, device=dev) for k, t in _FIELDS} for _ in range(2)]
        self.copy_stream = torch.cuda.Stream(device=dev)
        self.copied = [torch.cuda.Event() for _ in range(2)]
        self.released = [torch.cuda.Event() for _ in range(2)]
        self.staged_n = [0, 0]
        self.staged_bytes = [0, 0]
        self.refits = 0
        self.ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def _load(self, path: str, slot: int) -> Spectra:
        """Read a store straight into pinned slot `slot` (host thread; waits for the slot's previous copy)."""
        self.copied[slot].synchronize()
        h = self.host[slot]
        try:
            return load_store(path, out={k: h[k].numpy() for k, _ in _FIELDS}, threads=self.threads)
        except ValueError as e:
            raise ValueError(f"{path}: exceeds the feed's {self.max_spectra} spectra / {self.max_peaks} peaks "
                             f"({e})") from None

    def _stage(self, slot: int, sp: Spectra) -> None:
        """Pinned slot -> device slot, asynchronous on the copy stream, after the search that last read it."""
        n, npk = len(sp), int(sp.offsets[-1])
        h, d = self.host[slot], self.dev[slot]
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(self.released[slot])
            for k, m in (("mz", npk), ("int", npk), ("off", n + 1), ("pmz", n), ("z", n)):
                d[k][:m].copy_(h[k][:m], non_blocking=True)
            self.copied[slot].record(self.copy_stream)
        self.staged_n[slot] = n
        self.staged_bytes[slot] = 8 * npk + 8 * (n + 1) + 9 * n

    def _raw(self, slot: int) -> dict:
        d = self.dev[slot]
        return {"mz": d["mz"], "int": d["int"], "off": d["off"], "pmz": d["pmz"], "z": d["z"], "n": self.staged_n[slot]}

    def _load_stage(self, path: str, slot: int) -> Spectra:
        with torch.cuda.device(self.gi.dev):
            sp = self._load(path, slot)
            self._stage(slot, sp)
        return sp

    def _search(self, slot: int) -> int:
        self.ds.status.zero_()
        self.ds.run(self._raw(slot))
        self.released[slot].record(torch.cuda.current_stream(self.gi.dev))
        return int(self.ds.status.item())

    def run(self, paths, l1=None, out_dir: str | None = None, report: int | None = None, sink=None) -> list:
        """Search every store of `paths` in order. Returns per run {"run", "n_spectra", "rows" (rank-1 row per
        spectrum, -1 = none), "search_s" (host wall from the wait for its copy to its results), "gpu_ms" (the
        search on the device), "h2d_bytes"} plus "table" (the PSM table) when l1 is given; with out_dir the
        table is also written as <run>.psms.parquet. sink: optional callable given each run's result as soon
        as it is ready (it may take the table away, e.g. to write it in another thread).

        Three stages overlap: the host thread reads run i+1 into its pinned slot and enqueues its copy while
        the search of run i runs. The join K1's pair / slot capacities are fitted on the first run; a later run
        that exceeds them is flagged by the pairs kernel (status 8, writes bounded) and searched again after a
        refit, so no run waits on a host read of its totals."""
        out = []
        if not paths:
            return out
        cur = torch.cuda.current_stream(self.gi.dev)
        with ThreadPoolExecutor(1) as ex:
            fut = ex.submit(self._load_stage, paths[0], 0)
            for i in range(len(paths)):
                slot = i % 2
                sp = fut.result()
                fut = ex.submit(self._load_stage, paths[i + 1], 1 - slot) if i + 1 < len(paths) else None
                t0 = time.perf_counter()
                cur.wait_event(self.copied[slot])
                self.ev[0].record(cur)
                stt = self._search(slot)
                if stt & 8 and self.ds.k1_impl == "join":   # this run outgrew the fitted capacities
                    self.ds.join["fitted"] = False
                    self.refits += 1
                    stt = self._search(slot)
                self.ev[1].record(cur)
                if stt & 9:
                    raise RuntimeError(f"run {sp.run}: search status {stt}")
                n = len(sp)
                res = {"run": sp.run, "n_spectra": n, "h2d_bytes": self.staged_bytes[slot]}
                if l1 is not None:
                    tbl, _, _ = device_table(self.ds, sp, self.gi, l1, self.p, report or self.p.report_psms,
                                             host_hyp=False)
                    res["table"] = tbl
                    if out_dir is not None:
                        res["path"] = write_psms(tbl, out_dir, sp.run)
                res["rows"] = self.ds.rank1_rows(n)
                res["gpu_ms"] = self.ev[0].elapsed_time(self.ev[1])
                res["search_s"] = time.perf_counter() - t0
                if sink is not None:
                    sink(res)
                out.append(res)
        return out
