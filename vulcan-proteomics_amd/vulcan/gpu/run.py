"""Command line: closed search of run stores against an L1 index on disk, one run after another (the
reference's vulcan/gpu/run.py command line and output files).

    python -m vulcan.gpu.run --index DIR --stores a.vspec.arrow [...] --out OUTDIR [--repeat 2] [--params JSON]

Output: OUTDIR/<run>.psms.parquet per store (engine.search, top `report_psms` per spectrum; written on the first
pass only) and OUTDIR/stats.json: the setup timings, then one list of per-run records per pass (store load,
search phases, table write, H2D bytes, CPU core-seconds, counts). Passes after the first measure the warm
steady state. vulcan.gpu.batch overlaps store reads, uploads and searches of many runs instead.
"""
from __future__ import annotations

import json
import os
import time

from ._cli import Clock, Session, cpu_core_s, store_parser, write_json


def _search_store(path: str, ses: Session, out_dir: str | None) -> dict:
    """Load one store, search it, optionally write its table; the run's stats record."""
    import pyarrow.parquet as pq

    from . import engine as E
    clk = Clock()
    sp = E.load_store(path)
    t_load = clk.wall()
    tbl, st, _ = E.search(sp, ses.gi, ses.l1, ses.params)
    t_w = time.perf_counter()
    if out_dir is not None:
        pq.write_table(tbl, os.path.join(out_dir, f"{sp.run}.psms.parquet"))
    total = st.wall_s["total"]
    return {"run": sp.run, "load_store_s": t_load, "write_s": time.perf_counter() - t_w, "wall_s_total": clk.wall(),
            "cpu_core_s": clk.cpu(), "search_wall_s": st.wall_s, "gpu_ms": st.gpu_ms,
            "gpu_s": sum(st.gpu_ms.values()) / 1e3, "h2d_bytes": st.h2d_bytes, "counts": st.counts,
            "spectra_per_s_search": st.counts["n_spectra"] / total if total > 0 else 0.0}


def main(argv=None) -> int:
    ap = store_parser("Closed search of run stores against an L1 index, one run after another")
    ap.add_argument("--repeat", type=int, default=1, help="search passes over the stores (tables from pass 1)")
    a = ap.parse_args(argv)
    ses = Session.open(a)
    os.makedirs(a.out, exist_ok=True)
    passes = [[_search_store(s, ses, a.out if k == 0 else None) for s in a.stores] for k in range(max(a.repeat, 1))]
    import torch
    stats = {"params": ses.params.to_dict(), "index": os.path.abspath(a.index), "index_meta": ses.l1.meta,
             "setup": ses.timings, "passes": passes, "process_wall_s": ses.setup_s(),
             "process_cpu_core_s": cpu_core_s(), "gpu_max_mem_bytes": torch.cuda.max_memory_allocated()}
    write_json(os.path.join(a.out, "stats.json"), stats)
    print(json.dumps({"process_wall_s": stats["process_wall_s"], "setup": ses.timings,
                      "pass_walls": [[r["wall_s_total"] for r in ps] for ps in passes]}, default=float))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
