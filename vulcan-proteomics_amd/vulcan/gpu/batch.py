"""Many run stores through one resident index (the reference's vulcan/gpu/batch.py surface: `search_many` and
its command line).

    python -m vulcan.gpu.batch --index DIR --stores a.vspec.arrow ... --out OUTDIR [--io-threads 2]

Built on RunFeed: the host reads store i+1 straight into pinned memory and its PCIe copy runs on a copy stream
while run i is searched on the device; writer threads write `<out>/<run>.psms.parquet` while later runs are
searched. One DeviceSearch, sized for the largest store, serves every run.
"""
from __future__ import annotations

import json
import os
from concurrent.futures import ThreadPoolExecutor

from ._cli import Clock, Session, cpu_core_s, store_parser, write_json


class _TableWriter:
    """Writes each finished run's PSM table on a thread pool and keeps the run's stats record."""

    def __init__(self, out_dir: str, threads: int):
        self.out_dir = out_dir
        self.pool = ThreadPoolExecutor(max(threads, 1))
        self.pending, self.records = [], []

    def __call__(self, res: dict):
        import pyarrow.parquet as pq
        path = os.path.join(self.out_dir, f"{res['run']}.psms.parquet")
        self.pending.append(self.pool.submit(pq.write_table, res.pop("table"), path))
        self.records.append({"run": res["run"], "n_spectra": res["n_spectra"], "search_wall_s": res["search_s"],
                             "gpu_ms": res["gpu_ms"], "h2d_bytes": res["h2d_bytes"]})

    def close(self):
        try:
            for f in self.pending:
                f.result()
        finally:
            self.pool.shutdown()


def search_many(stores, gi, l1, p, out_dir, io_threads=2) -> dict:
    """Search `stores` in order and write one PSM table per run. Returns the reference's result keys:
    {"wall_s", "cpu_core_s", "runs": [{"run", "n_spectra", "search_wall_s", "gpu_ms", "h2d_bytes"}],
    "n_spectra"}."""
    from ..io import store
    from .feed import RunFeed
    os.makedirs(out_dir, exist_ok=True)
    clk = Clock()
    writer = _TableWriter(out_dir, io_threads)
    try:
        if stores:
            spectra, peaks = zip(*(store.counts(s) for s in stores))
            feed = RunFeed(gi, p, max(spectra), max(peaks), threads=max(io_threads, 1))
            feed.run(list(stores), l1=l1, sink=writer)
    finally:
        writer.close()
    return {"wall_s": clk.wall(), "cpu_core_s": clk.cpu(), "runs": writer.records,
            "n_spectra": sum(r["n_spectra"] for r in writer.records)}


def main(argv=None) -> int:
    ap = store_parser("Search run stores through one resident index, store reads overlapped with searches")
    ap.add_argument("--io-threads", type=int, default=2, help="host threads for store reads and table writes")
    a = ap.parse_args(argv)
    ses = Session.open(a)
    setup_s = ses.setup_s()
    r = search_many(a.stores, ses.gi, ses.l1, ses.params, a.out, a.io_threads)
    r["setup_s"] = setup_s
    r["process_wall_s"] = ses.setup_s()
    r["process_core_s"] = cpu_core_s()
    r["spectra_per_s_steady"] = r["n_spectra"] / r["wall_s"] if r["wall_s"] > 0 else 0.0
    write_json(os.path.join(a.out, "batch_stats.json"), r)
    print(json.dumps({k: v for k, v in r.items() if k != "runs"}, default=float))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
