"""Multi-run search over run stores (reference vulcan/gpu/batch.py:21-41 `search_many`, :44-73 CLI).

    python -m vulcan.gpu.batch --index DIR --stores a.vspec.arrow ... --out OUTDIR [--io-threads 2]

Built on RunFeed: the host reads store i+1 straight into pinned memory and its PCIe copy runs on a copy stream
while run i is searched on the device; the PSM tables are written to `<out>/<run>.psms.parquet` by writer
threads while later runs are searched. One DeviceSearch, preallocated for the largest store, serves every run.
"""
from __future__ import annotations

import argparse
import json
import os
import resource
import time
from concurrent.futures import ThreadPoolExecutor


def _core_s():
    r = resource.getrusage(resource.RUSAGE_SELF)
    return r.ru_utime + r.ru_stime


def search_many(stores, gi, l1, p, out_dir, io_threads=2):
    """Search `stores` in order and write one PSM table per run. Returns {"wall_s", "cpu_core_s", "runs": [{"run",
    "n_spectra", "search_wall_s", "gpu_ms", "h2d_bytes"}], "n_spectra"} (the reference's result keys)."""
    import pyarrow.parquet as pq

    from ..io import store
    from .feed import RunFeed
    os.makedirs(out_dir, exist_ok=True)
    t0, c0 = time.perf_counter(), _core_s()
    if not stores:
        return {"wall_s": 0.0, "cpu_core_s": 0.0, "runs": [], "n_spectra": 0}
    sizes = [store.counts(s) for s in stores]
    feed = RunFeed(gi, p, max(n for n, _ in sizes), max(k for _, k in sizes), threads=max(io_threads, 1))
    stats, writes = [], []
    with ThreadPoolExecutor(max(io_threads, 1)) as ex:
        def sink(res):
            tbl = res.pop("table")
            writes.append(ex.submit(pq.write_table, tbl, os.path.join(out_dir, f"{res['run']}.psms.parquet")))
            stats.append({"run": res["run"], "n_spectra": res["n_spectra"], "search_wall_s": res["search_s"],
                          "gpu_ms": res["gpu_ms"], "h2d_bytes": res["h2d_bytes"]})

        feed.run(stores, l1=l1, sink=sink)
        for w in writes:
            w.result()
    return {"wall_s": time.perf_counter() - t0, "cpu_core_s": _core_s() - c0, "runs": stats,
            "n_spectra": sum(s["n_spectra"] for s in stats)}


def main(argv=None):
    ap = argparse.ArgumentParser(description="Search run stores with one resident index")
    ap.add_argument("--index", required=True)
    ap.add_argument("--stores", nargs="+", required=True)
    ap.add_argument("--out", required=True)
    ap.add_argument("--io-threads", type=int, default=2)
    ap.add_argument("--params", default="{}")
    a = ap.parse_args(argv)
    t_start = time.perf_counter()
    import torch

    from ..index.build import L1Index
    from . import engine as E
    p = E.SearchParams(**json.loads(a.params))
    torch.zeros(1, device="cuda")
    l1 = L1Index(a.index)
    gi = E.GpuIndex(l1, p, {"C": 57.021464}, {"M": 15.9949})
    t_setup = time.perf_counter() - t_start
    r = search_many(a.stores, gi, l1, p, a.out, a.io_threads)
    r.update(setup_s=t_setup, process_wall_s=time.perf_counter() - t_start, process_core_s=_core_s(),
             spectra_per_s_steady=r["n_spectra"] / r["wall_s"] if r["wall_s"] else 0.0)
    with open(os.path.join(a.out, "batch_stats.json"), "w") as fh:
        json.dump(r, fh, indent=1)
    print(json.dumps({k: v for k, v in r.items() if k != "runs"}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
