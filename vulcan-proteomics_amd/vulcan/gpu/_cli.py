"""Shared plumbing of the store-search command lines (vulcan.gpu.run, vulcan.gpu.batch): the common options,
a process clock (wall seconds and CPU core-seconds since a mark) and the resident-index setup with its timings.

The timing keys written by the commands are the ones the reference's run.py / batch.py write to stats.json /
batch_stats.json (reference vulcan/gpu/run.py:70-79, batch.py:63-68), so tools that read those files keep
working; new keys are only added beside them.
"""
from __future__ import annotations

import argparse
import json
import resource
import time
from dataclasses import dataclass, field

PROCESS_T0 = time.perf_counter()     # module import time ~ process start for `python -m vulcan.gpu.*`

DEFAULT_STATIC = '{"C": 57.021464}'
DEFAULT_VAR = '{"M": 15.9949}'


def cpu_core_s() -> float:
    """CPU time (user + system) of this process so far, in core-seconds."""
    u = resource.getrusage(resource.RUSAGE_SELF)
    return u.ru_utime + u.ru_stime


class Clock:
    """Wall and CPU time since construction (or the last `restart`)."""

    def __init__(self):
        self.restart()

    def restart(self):
        self.w0, self.c0 = time.perf_counter(), cpu_core_s()

    def wall(self) -> float:
        return time.perf_counter() - self.w0

    def cpu(self) -> float:
        return cpu_core_s() - self.c0


def store_parser(description: str) -> argparse.ArgumentParser:
    """Options common to both commands: index directory, stores, output directory, SearchParams overrides and
    the static / variable modification tables of the index."""
    ap = argparse.ArgumentParser(description=description)
    ap.add_argument("--index", required=True, help="L1 index directory (vulcan.index.build layout)")
    ap.add_argument("--stores", nargs="+", required=True, help="run stores (<run>.vspec.arrow)")
    ap.add_argument("--out", required=True, help="output directory (<run>.psms.parquet + stats)")
    ap.add_argument("--params", default="{}", help="JSON overrides of SearchParams")
    ap.add_argument("--static", default=DEFAULT_STATIC, help="JSON static modifications of the index")
    ap.add_argument("--var", default=DEFAULT_VAR, help="JSON variable modifications of the index")
    return ap


def search_params(text: str):
    """SearchParams from a JSON object of field overrides (lists become the tuples the dataclass holds)."""
    from .host import SearchParams
    over = {k: tuple(v) if isinstance(v, list) else v for k, v in json.loads(text).items()}
    return SearchParams(**over)


@dataclass
class Session:
    """A loaded L1 table and its resident GpuIndex, with the setup timings."""
    params: object
    l1: object
    gi: object
    timings: dict = field(default_factory=dict)

    @classmethod
    def open(cls, args) -> "Session":
        t_import = time.perf_counter() - PROCESS_T0
        c_import = cpu_core_s()
        import torch

        from ..index.build import L1Index
        from .engine import GpuIndex
        p = search_params(args.params)
        clk = Clock()
        torch.cuda.init()
        torch.empty(1, device="cuda").zero_()           # first device allocation: runtime + context up
        t_dev = clk.wall()
        clk.restart()
        l1 = L1Index(args.index)
        t_l1 = clk.wall()
        clk.restart()
        gi = GpuIndex(l1, p, json.loads(args.static), json.loads(args.var))
        t_gi = clk.wall()
        tm = {"python_import_s": t_import, "cuda_init_s": t_dev, "l1_load_s": t_l1, "gpu_index_build_s": t_gi,
              "gpu_index_bytes": gi.device_bytes, "gpu_index_h2d_bytes": gi.h2d, "n_postings": gi.n_postings,
              "n_bins": gi.n_bins, "n_slices": gi.n_slices, "cpu_core_s_import": c_import,
              "cpu_core_s_setup_total": cpu_core_s()}
        return cls(p, l1, gi, tm)

    def setup_s(self) -> float:
        return time.perf_counter() - PROCESS_T0


def write_json(path: str, obj) -> None:
    with open(path, "w") as fh:
        json.dump(obj, fh, indent=1, default=float)
