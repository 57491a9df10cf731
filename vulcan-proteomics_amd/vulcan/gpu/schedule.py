"""Uneven batches: spectra of many runs packed into one search step, per-class launch schedules chosen by an
autotuner, and the whole step captured as one HIP graph.

No reference counterpart: the reference searches one run at a time with a fixed tile (work items of <= 1024
rows, vulcan/gpu/engine.py:213-229) and a host round trip per phase; its multi-run feed (cohort.py:512-790)
overlaps runs on several streams. Here:

* classes: every spectrum falls in a (precursor charge, kept-peak count) class -- charge 2, 3, 4, 5+ x kept
  peaks < 64, 64-149, >= 150 -- computed on the host from the raw input (the same inputs K0 uses), so class sizes
  are known when the step is built;
* packing: spectra are ordered by (class, precursor mass) on the host; each class is one contiguous range of
  the order array, so K1 and K2 of a class are launched over `order[start:start+count]` with that class's
  parameters, and precursor-mass locality is kept inside a class;
* schedule: per class a K1 LDS tile (rows per pass; tile K1 only), the XCD grouping of the block order, and grid
  vs persistent launch (resident waves pulling spectra from an atomic counter, vh_k1_topk / vh_k1j_topk / vh_k2
  `queue`); with the bucket-join K1 the block-major counting runs once for the whole batch and the per-class
  launches are its per-spectrum select and K2;
* autotune: every candidate configuration is applied to all classes at once, the step runs, and each class
  keeps the configuration under which its own K1 + K2 launches were fastest (HIP events per launch);
* graph: `GraphStep` captures K0 .. rank of the packed batch into one hipGraph (no allocation, no host
  synchronisation and no memset node inside: every counter is cleared by a kernel), replayed per step;
* tune_step: the whole-step configuration (grid / persistent launches, processing order, XCD grouping, or the
  per-class schedule) is chosen by timing each candidate on the batch; the winner is the one captured.

Results never depend on the schedule: candidate lists and scores are per spectrum.
"""
from __future__ import annotations

import ctypes as C
import itertools
import json
import time

import numpy as np
import torch

from . import _lib as K
from .engine import DeviceSearch, GpuIndex, _k2_call, _marker
from .host import SearchParams, Spectra

Z_BINS = (2, 3, 4, 5)          # charge 2, 3, 4, >= 5 (unknown charge counts as 2)
P_BINS = (64, 150)             # kept peaks < 64, 64-149, >= 150 (the default max_peaks)
N_CLASSES = len(Z_BINS) * (len(P_BINS) + 1)

DEFAULT_CANDIDATES = [
    {"tile": t, "xcd_group": g, "persistent": q}
    for t, g, q in itertools.product((256, 512, 1024, 2048), (0, 32), (False, True))
]
# with the bucket-join K1 the counts are block-major for the whole batch; per class only the select and K2
# launches are scheduled (the LDS tile does not apply)
JOIN_CANDIDATES = [{"tile": 512, "xcd_group": g, "persistent": q} for g, q in itertools.product((0, 32), (False, True))]


def spectrum_classes(sp: Spectra, p: SearchParams) -> np.ndarray:
    """Class id per spectrum: charge bin * 3 + kept-peak bin (kept = peaks in [fmin, fmax], capped at
    max_peaks, as K0 keeps them)."""
    mz = sp.mz
    ok = (mz >= p.fragment_min_mz) & (mz <= p.fragment_max_mz) & ~np.isnan(mz)
    cs = np.zeros(len(ok) + 1, np.int64)
    np.cumsum(ok, out=cs[1:])
    kept = np.minimum(cs[sp.offsets[1:]] - cs[sp.offsets[:-1]], p.max_peaks)
    z = np.clip(np.asarray(sp.charge, np.int64), 2, Z_BINS[-1])
    zb = np.searchsorted(np.array(Z_BINS), z, side="left")
    pb = np.searchsorted(np.array(P_BINS), kept, side="right")
    return (zb * (len(P_BINS) + 1) + pb).astype(np.int32)


class ClassSchedule(DeviceSearch):
    """DeviceSearch over a packed multi-run batch whose K1 / K2 are launched per spectrum class with per-class
    parameters (`table`: class -> {"tile", "xcd_group", "persistent"}; missing classes use the defaults)."""

    def __init__(self, gi: GpuIndex, p: SearchParams, sp: Spectra, table: dict | None = None, report: int = 1,
                 k1_impl: str | None = None):
        super().__init__(gi, p, len(sp), report=report, persistent=True, order_mode="fixed", k1_impl=k1_impl)
        cls = spectrum_classes(sp, p)
        o = np.lexsort((np.asarray(sp.precursor_mz), cls))
        self.order.copy_(torch.from_numpy(o.astype(np.int32)).to(gi.dev))
        counts = np.bincount(cls, minlength=N_CLASSES)
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
        self.classes = [(c, int(starts[c]), int(counts[c])) for c in range(N_CLASSES) if counts[c]]
        self.set_table(table or {})

    def set_table(self, table: dict):
        self.table = {c: {**{"tile": self.p.tile, "xcd_group": self.p.xcd_group, "persistent": True},
                          **table.get(c, table.get(str(c), {}))} for c, _, _ in self.classes}
        self.vps = {}
        for c, _, _ in self.classes:
            q = SearchParams(**self.p.to_dict())
            q.tile, q.xcd_group = int(self.table[c]["tile"]), int(self.table[c]["xcd_group"])
            self.vps[c] = self.gi.params_for(q)

    def candidate_stage(self, n: int, mark):
        L = K.lib()
        gi, pk, hy = self.gi, self.peaks, self.hyp
        if self.k1_impl == "join":
            self.join_front(n, mark)
            j = self.join
            for c, s0, cnt in self.classes:
                vp = self.vps[c]
                q = self.queue if self.table[c]["persistent"] else None
                mark(f"k1_c{c}")
                K.check(L.vh_k1j_topk(C.byref(vp), C.byref(gi.view), cnt, K.ptr(self.order[s0:s0 + cnt]),
                                      K.ptr(hy["hyp_off"]), K.ptr(hy["mode"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]),
                                      K.ptr(j["h_slot"]), j["slot_cap"], K.ptr(j["counts"]), self.stride,
                                      K.ptr(self.n_cand), K.ptr(self.c_row), K.ptr(self.c_hyp), K.ptr(self.c_cnt),
                                      K.ptr(self.scored), K.ptr(self.matched), K.ptr(self.status), K.ptr(q),
                                      K.stream()), "k1j_topk")
                mark(f"k1_c{c}")
            return
        for c, s0, cnt in self.classes:
            vp = self.vps[c]
            q = self.queue if self.table[c]["persistent"] else None
            mark(f"k1_c{c}")
            K.check(L.vh_k1_topk(C.byref(vp), C.byref(gi.tile_view()), cnt, K.ptr(self.order[s0:s0 + cnt]),
                                 K.ptr(pk["off"]), K.ptr(pk["mass"]), K.ptr(hy["hyp_off"]), K.ptr(hy["maxfz"]),
                                 K.ptr(hy["mode"]), K.ptr(hy["lo"]), K.ptr(hy["hi"]), self.stride,
                                 K.ptr(self.n_cand), K.ptr(self.c_row), K.ptr(self.c_hyp), K.ptr(self.c_cnt),
                                 K.ptr(self.scored), K.ptr(self.matched), K.ptr(self.status), K.ptr(q),
                                 K.stream()), "k1_topk")
            mark(f"k1_c{c}")

    def back(self, n: int, mark):
        L = K.lib()
        gi, pk, hy = self.gi, self.peaks, self.hyp
        for c, s0, cnt in self.classes:
            q = self.queue if self.table[c]["persistent"] else None
            mark(f"k2_c{c}")
            _k2_call(gi, self.vps[c], cnt, None, None, self.n_cand, self.stride, self.c_row, self.c_hyp,
                     hy["maxfz"], pk, self.k2, order=self.order[s0:s0 + cnt], queue=q)
            mark(f"k2_c{c}")
        mark("rank")
        K.check(L.vh_rank(C.byref(self.vp), n, None, K.ptr(self.n_cand), self.stride, K.ptr(self.c_row),
                          K.ptr(self.k2["hyperscore"]), K.ptr(self.k2["matched_b"]), K.ptr(self.k2["matched_y"]),
                          self.report, K.ptr(self.sel), K.ptr(self.nxt), K.stream()), "rank")
        mark("rank")
        return {"sel": self.sel[:n * self.report], "next": self.nxt[:n * self.report], "n_cand": self.n_cand[:n]}

    def class_ms(self, ev: dict) -> dict:
        """Per class: mean K1 + K2 milliseconds from the events of run(raw, ev)."""
        out = {}
        for c, _, _ in self.classes:
            t = 0.0
            for k in (f"k1_c{c}", f"k2_c{c}"):
                lst = ev.get(k, [])
                pairs = list(zip(lst[0::2], lst[1::2]))
                t += float(np.mean([a.elapsed_time(b) for a, b in pairs])) if pairs else 0.0
            out[c] = t
        return out


def autotune(sched: ClassSchedule, raw: dict, candidates=None, reps: int = 2) -> dict:
    """Per class, the candidate configuration with the fastest K1 + K2 launches. Returns {"table": {class:
    config}, "class_ms": {class: ms}, "tried": n, "seconds": s}; the schedule is left set to the table."""
    cands = candidates or (JOIN_CANDIDATES if sched.k1_impl == "join" else DEFAULT_CANDIDATES)
    t0 = time.perf_counter()
    best = {c: (float("inf"), None) for c, _, _ in sched.classes}
    for cfg in cands:
        sched.set_table({c: cfg for c, _, _ in sched.classes})
        sched.run(raw)                          # warm-up of this configuration
        ev: dict = {}
        for _ in range(reps):
            sched.run(raw, ev)
        torch.cuda.synchronize()
        for c, t in sched.class_ms(ev).items():
            if t < best[c][0]:
                best[c] = (t, dict(cfg))
    table = {c: cfg for c, (t, cfg) in best.items()}
    sched.set_table(table)
    return {"table": table, "class_ms": {c: t for c, (t, _) in best.items()}, "tried": len(cands),
            "seconds": time.perf_counter() - t0}


# whole-step configurations tried by tune_step: launch form (grid / persistent), processing order (precursor mass /
# mass band then estimated work) and XCD grouping of the block order; "classes" = the per-class schedule
STEP_CANDIDATES = [
    {"kind": "step", "persistent": False, "order_mode": "mass", "xcd_group": 32},
    {"kind": "step", "persistent": False, "order_mode": "mass", "xcd_group": 0},
    {"kind": "step", "persistent": True, "order_mode": "mass", "xcd_group": 32},
    {"kind": "step", "persistent": True, "order_mode": "work", "xcd_group": 32},
    {"kind": "classes"},
]


def tune_step(gi: GpuIndex, p: SearchParams, sp: Spectra, raw: dict, candidates=None, reps: int = 3):
    """Time every candidate configuration of the whole search step on this batch (eager, after a warm-up run)
    and return (the fastest step object, [(config, ms)]); a "classes" candidate is a ClassSchedule autotuned per
    class first. Results never depend on the configuration."""
    n = len(sp)
    out = []
    best = (float("inf"), None)
    for cfg in candidates or STEP_CANDIDATES:
        if cfg["kind"] == "classes":
            obj = ClassSchedule(gi, p, sp)
            autotune(obj, raw, reps=1)
        else:
            q = SearchParams(**p.to_dict())
            q.xcd_group = int(cfg["xcd_group"])
            obj = DeviceSearch(gi, q, n, persistent=bool(cfg["persistent"]), order_mode=cfg["order_mode"])
        obj.run(raw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            obj.run(raw)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        out.append((dict(cfg), ms))
        if ms < best[0]:
            best = (ms, obj)
        else:
            del obj
    return best[1], out


def step_factory(gi: GpuIndex, sp: Spectra, obj):
    """A callable that builds a fresh step object configured like `obj` (a tune_step result: a ClassSchedule
    with its table, or a DeviceSearch with its launch form, order and XCD grouping), for the slots of an
    engine.SearchPipeline."""
    if isinstance(obj, ClassSchedule):
        table = dict(obj.table)
        return lambda: ClassSchedule(gi, obj.p, sp, table=table, report=obj.report, k1_impl=obj.k1_impl)
    q = SearchParams(**obj.p.to_dict())
    persistent = obj.queue is not None
    return lambda: DeviceSearch(gi, q, len(sp), report=obj.report, persistent=persistent,
                                order_mode=obj.order_mode, band_log2=obj.band_log2, k1_impl=obj.k1_impl)


def save_table(path: str, table: dict):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in table.items()}, f, indent=1)


def load_table(path: str) -> dict:
    with open(path) as f:
        return {int(k): v for k, v in json.load(f).items()}


class GraphStep:
    """One search step (obj.run(raw) of a DeviceSearch / ClassSchedule / LearnedSearch) captured into a hipGraph;
    replay() launches the whole step. The captured step reads raw and writes obj's buffers: both must stay
    allocated (they are) and raw must keep its shapes and addresses."""

    def __init__(self, obj, raw: dict, warmup: int = 2):
        self.obj, self.raw = obj, raw
        for _ in range(warmup):              # first-call set-up (LDS attributes) happens outside the capture
            obj.run(raw)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            obj.run(raw)
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()
