#!/usr/bin/env python3
"""Benchmark of the plain HIP closed search (spectra/s) on one MI355X.

Workload (default): the proteome-scale plain search -- 20,000 synthetic proteins, 1 missed cleavage, Met-ox <= 2,
reversed decoys (~3.1M peptidoforms), 100,000 synthetic spectra (charges 2/3/4, b/y 2+ for z >= 3), precursor
+-10 ppm, fragment +-0.02 Th, top-K 50 + pair closure. A step = one pass of the search hot path over the whole
batch with the raw spectra already in HBM: K0 peak view -> hypotheses -> K1 (index) + top-K + closure -> K2
exact rescoring -> rank 1. The index build is timed separately and not part of a step. The timed loop keeps
`--depth` steps (default 3) in flight, each slot on its own stream (engine.SearchPipeline): `value` is the
throughput of K such steps between two device synchronisations; `serial_ms_per_step` is one step at a time.

Prints ONE JSON line (rank 0) with the metric, per-kernel HIP-event times, the roofline of the dominant kernel
(algorithmic bytes per launch / its average duration) and the CPU baseline (the C++ oracle's search, OpenMP, on a
bounded sample of the same workload). The default line also carries:
  * recall_vs_exhaustive: the shortlist's rank 1 against K2 over every window row on the first 5,000 spectra;
  * prefilter: the metric's plain vs learned comparison on the phospho-scale workload (36M peptidoforms, 200k
    spectra, index top-200): plain, learned (adaptive keep), learned literal top-16, the K1-count rule at 16 and at
    the learned path's rescoring budget -- spectra/s, top-1 agreement with plain, rescored candidates per
    spectrum, the MLP's MFMA roofline and a bounded-sample CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "vulcan-proteomics_amd"))
sys.path.insert(0, ROOT)

METRIC = "spectra/s (index+rescore) plain vs learned prefilter at ≥99.5% top-1 agreement"
HBM_PEAK_GBS = 8000.0
STATIC = {"C": 57.021464}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _stages(ev):
    import numpy as np
    out = {}
    for k, lst in ev.items():
        pairs = list(zip(lst[0::2], lst[1::2]))
        out[k] = float(np.mean([s.elapsed_time(e) for s, e in pairs]))
    return out


def time_with_features(ds, raw, n, steps):
    """The step followed by the PSM feature columns of rank 1 on the device (vh_psm_features over the step's
    candidate lists, dense; the host table assembly is left out): (seconds per step, feature ms per step)."""
    import torch

    from vulcan.gpu import engine as E
    hy = ds.hyp

    def feats():
        return E.psm_features(ds.gi, n, None, ds.sel[:n], ds.nxt[:n], 1, ds.c_hyp, ds.c_row, ds.c_cnt, ds.k2,
                              hy["hyp_off"], hy["z"], hy["iso"], hy["pm"], ds.scored, ds.matched,
                              ds.peaks["tic"])
    ds.run(raw)
    feats()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ds.run(raw)
        feats()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(torch.cuda.current_stream())
    for _ in range(steps):
        feats()
    b.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    return dt, a.elapsed_time(b) / steps


def time_steps(obj, raw, steps, warmup):
    """Warm up, time `steps` runs (wall clock between synchronisations, no instrumentation in the timed loop),
    then run `steps` more with per-stage HIP events for the stage breakdown."""
    import torch
    for _ in range(warmup):
        obj.run(raw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        obj.run(raw)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    ev: dict = {}
    for _ in range(steps):
        obj.run(raw, ev)
    torch.cuda.synchronize()
    return dt, _stages(ev)


def time_pipeline(pipe, raw, steps, warmup):
    """Wall seconds for `steps` steps of a SearchPipeline over one resident batch, bracketed by device
    synchronisations (every slot warmed up first: the first run of a slot fits its pair / slot capacities)."""
    import torch
    for _ in range(max(warmup, pipe.depth)):
        pipe.run(raw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        pipe.run(raw)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


MLP_FLOP_PER_CANDIDATE = 2 * (8 * 256 + 256 * 256 + 256)   # 135,680
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _take(sp, n):
    from vulcan.gpu.host import Spectra
    off = sp.offsets[: n + 1]
    return Spectra(sp.run, sp.scan[:n], sp.index[:n], sp.rt[:n], sp.precursor_mz[:n], sp.charge[:n],
                   sp.mz[: off[-1]], sp.intensity[: off[-1]], off.copy())


def prefilter_leg(a, w, l1, lab, gi, p, raw, n, sp=None):
    """Plain vs learned prefilter on the same spectra. Plain: index top-K (200) + closure -> K2. Learned: index
    top-K -> MLP -> the best keep_min by score, then every candidate within `margin` of the best up to `keep`, +
    closure -> K2 (the default rule); also the literal top-16 by score. Ablations at the same index top-K: the
    K1-count rule at K = 16 (the configuration's literal budget) and at the learned path's mean kept count. The
    network is trained first on separately seeded spectra of the same workload (labels from the plain K2)."""
    import numpy as np
    import torch

    from vulcan.gpu import engine as E
    from vulcan.index import synth
    from vulcan.rescore import prefilter as PF
    pp = PF.PrefilterParams(index_topk=p.topk, keep=a.keep, keep_min=a.keep_min, margin=a.margin,
                            steps=a.train_steps)
    t0 = time.perf_counter()
    sp_tr, _, _ = synth.spectra(l1, lab, w.key, n=a.n_train, seed=900 + w.key)
    raw_tr = E.upload_spectra(sp_tr, gi.dev)
    x, y, _, _ = PF.training_data(gi, p, raw_tr, pp)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    log(f"prefilter training set: {x.shape[0]} candidates of {len(sp_tr)} spectra ({t_data:.1f} s)")
    model = PF.train(PF.PrefilterModel.init(pp.seed), x, y, pp, log_every=max(pp.steps // 10, 1))
    losses = [v for t, v in model.history if isinstance(t, int)]
    train_s = [v for t, v in model.history if t == "train_s"][0]
    log(f"trained {pp.steps} steps in {train_s:.1f} s, loss {losses[0]:.3f} -> {losses[-1]:.3f}")
    del raw_tr, x, y
    q = lambda k: type(p)(**{**p.to_dict(), "topk": k, "pair_closure": True})  # noqa: E731
    pp16 = PF.PrefilterParams(**{**pp.__dict__, "keep": 16, "keep_min": 16, "margin": 0.0})
    out, rows = {}, {}

    def measure(name, make):
        """`a.depth` steps in flight (the value), then one slot alone (the serial step and its stage times)."""
        pipe = E.SearchPipeline(gi, p, n, depth=a.depth, make=make)
        dt = time_pipeline(pipe, raw, a.steps, a.warmup) / a.steps
        obj = pipe.slots[0]
        dts, st = time_steps(obj, raw, a.steps, a.warmup)
        for d in pipe.slots:
            if int(d.status.item()) != 0:
                raise RuntimeError(f"{name}: k1 status {int(d.status.item())}")
        rows[name] = obj.rank1_rows(n)
        same = all(np.array_equal(d.rank1_rows(n), rows[name]) for d in pipe.slots[1:])
        nc = obj.n_cand[:n].cpu().numpy()
        out[name] = {"spectra_per_s": n / dt, "ms_per_step": dt * 1e3, "serial_spectra_per_s": n / dts,
                     "serial_ms_per_step": dts * 1e3, "slots_rank1_equal": same, "stage_ms": st,
                     "rescored_per_spectrum": float(nc.mean())}
        log(f"{name}: {n / dt:,.0f} spectra/s ({dt * 1e3:.2f} ms/step; one at a time {dts * 1e3:.2f} ms), "
            f"{nc.mean():.1f} rescored per spectrum")
        return obj, st

    plain, _ = measure("plain", lambda: E.DeviceSearch(gi, q(pp.index_topk), n))
    plain_vp = plain.vp
    del plain
    ls, st = measure("learned", lambda: PF.LearnedSearch(gi, p, n, model, pp))
    scored = int(ls.lists.n_cand[:n].sum().item())
    fl = MLP_FLOP_PER_CANDIDATE * scored / (st["mlp"] * 1e-3) / 1e12
    out["learned"]["mlp_roofline"] = {"bound": "mfma", "kernel": "prefilter_score_kernel", "achieved": fl,
                                      "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                      "frac": fl / MFMA_BF16_PEAK_TFLOPS, "candidates_scored": scored,
                                      "flop_per_candidate": MLP_FLOP_PER_CANDIDATE}
    del ls
    measure("learned_top16", lambda: PF.LearnedSearch(gi, p, n, model, pp16))
    measure("count_rule_16", lambda: E.DeviceSearch(gi, q(16), n))
    k_eq = max(1, int(round(out["learned"]["rescored_per_spectrum"] / 2)))
    measure("count_rule_budget", lambda: E.DeviceSearch(gi, q(k_eq), n))
    out["count_rule_budget"]["topk"] = k_eq
    ref = rows["plain"]
    ag = {k: PF.agreement(ref, v) for k, v in rows.items() if k != "plain"}
    for k, v in ag.items():
        out[k]["top1_agreement"] = v
    res = {"index_topk": pp.index_topk, "keep_max": pp.keep, "keep_min": pp.keep_min, "margin": pp.margin,
           "pair_closure": True, "top1_agreement": ag["learned"], "speedup_vs_plain":
           out["learned"]["spectra_per_s"] / out["plain"]["spectra_per_s"],
           "serial_speedup_vs_plain": out["learned"]["serial_spectra_per_s"] / out["plain"]["serial_spectra_per_s"],
           "launch": f"{a.depth} steps in flight on {a.depth} streams",
           "meets_agreement_bar": ag["learned"] >= 0.995,
           "train": {"spectra": a.n_train, "steps": pp.steps, "batch": pp.batch, "train_s": train_s,
                     "loss_first": losses[0], "loss_last": losses[-1], "labels": "K2 hyperscore - spectrum best"},
           **out}
    res["cpu_baseline"] = None
    if sp is not None and not a.no_cpu:
        log("prefilter workload cpu baseline (plain top-200 + closure) ...")
        cb, cres, ncpu = cpu_baseline(plain_vp, l1, sp, a.cpu_seconds / 2)
        res["cpu_baseline"] = cb
        res["plain_rank1_equal_cpu_sample"] = bool(np.array_equal(ref[:ncpu], cres["row"][:ncpu]))
    return res


def prefilter_block(a):
    """The phospho-scale prefilter workload (index top-200, STY phospho <= 3 + Met-ox <= 2, 200k spectra of
    charges 2-6 / 20-500 peaks) built and compared as prefilter_leg does, for the default line."""
    import torch

    from vulcan.gpu import engine as E
    from vulcan.index import synth
    w = synth.WORKLOADS[4]
    t0 = time.perf_counter()
    l1, lab = synth.build_index(4)
    sp, _, _ = synth.spectra(l1, lab, 4, n=a.prefilter_spectra)
    log(f"prefilter workload: {len(l1)} peptidoforms, {len(sp)} spectra ({time.perf_counter() - t0:.1f} s)")
    p = w.search_params()
    gi = E.GpuIndex(l1, p, STATIC, w.var_mods)
    raw = E.upload_spectra(sp, gi.dev)
    res = prefilter_leg(a, w, l1, lab, gi, p, raw, len(sp), sp)
    res["workload"] = w.name
    res["n_spectra"] = len(sp)
    res["n_peptidoforms"] = len(l1)
    res["n_postings"] = gi.n_postings
    del gi, raw
    torch.cuda.empty_cache()
    return res


def recall_vs_exhaustive(gi, l1, p, sp, rows, n_sub: int):
    """Shortlist recall: over the first n_sub spectra, the fraction whose exhaustive rank 1 (K2 over every
    window row, search_exhaustive) is the shortlist's rank 1 (`rows`)."""
    import numpy as np

    from vulcan.gpu import engine as E
    sub = _take(sp, min(n_sub, len(sp)))
    t0 = time.perf_counter()
    tbl, _, dbg = E.search_exhaustive(sub, gi, l1, p)
    ex = dict(zip(tbl.column("spec_idx").to_pylist(), tbl.column("pf_row").to_pylist()))
    hit = sum(int(rows[s] == r) for s, r in ex.items())
    return {"spectra": len(sub), "with_exhaustive_psm": len(ex), "window_rows_rescored": int(len(dbg["r"])),
            "recall": hit / max(len(ex), 1), "seconds": time.perf_counter() - t0}


def uneven_leg(a, w, l1, gi, p, sp, raw, n):
    """Packed multi-run batch: default eager step (grid launches, mass order) vs the tuned step (the fastest of
    grid / persistent launches, mass / work order, XCD grouping and the per-class schedule, timed on this batch)
    replayed as one hipGraph."""
    import numpy as np
    import torch

    from vulcan.gpu import engine as E
    from vulcan.gpu import schedule as S
    base = E.DeviceSearch(gi, p, n)
    dt_base, st_base = time_steps(base, raw, a.steps, a.warmup)
    rows_base = base.rank1_rows(n)
    log(f"eager default: {n / dt_base:,.0f} spectra/s ({dt_base * 1e3:.2f} ms)")
    del base
    t0 = time.perf_counter()
    sched, tried = S.tune_step(gi, p, sp, raw)
    tune_s = time.perf_counter() - t0
    cfg = min(tried, key=lambda t: t[1])[0]
    log(f"tuned the step over {len(tried)} configurations in {tune_s:.1f} s: {cfg}")
    ev: dict = {}
    sched.run(raw)
    for _ in range(3):
        sched.run(raw, ev)
    torch.cuda.synchronize()
    stage = _stages(ev)
    classes = []
    if isinstance(sched, S.ClassSchedule):
        cms = sched.class_ms(ev)
        classes = [{"class": c, "charge_bin": S.Z_BINS[c // (len(S.P_BINS) + 1)], "peak_bin": c % (len(S.P_BINS) + 1),
                    "spectra": cnt, "config": sched.table[c], "k1_k2_ms": cms[c]} for c, _, cnt in sched.classes]
    k1_ms = sum(v for k, v in stage.items() if k.startswith("k1"))
    k2_ms = sum(v for k, v in stage.items() if k.startswith("k2"))
    join_ms = {k: stage[k] for k in ("k1j_pairs", "k1j_count") if k in stage}
    g = S.GraphStep(sched, raw)
    for _ in range(a.warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / a.steps
    rows = sched.rank1_rows(n)
    same = bool(np.array_equal(rows, rows_base))
    log(f"graph schedule: {n / dt:,.0f} spectra/s ({dt * 1e3:.2f} ms), rank-1 equal to eager default: {same}")
    pipe = E.SearchPipeline(gi, p, n, depth=a.depth, make=S.step_factory(gi, sp, sched))
    dtp = time_pipeline(pipe, raw, a.steps, a.warmup) / a.steps
    pipe_same = all(np.array_equal(d.rank1_rows(n), rows_base) for d in pipe.slots)
    del pipe
    log(f"tuned step, {a.depth} in flight: {n / dtp:,.0f} spectra/s ({dtp * 1e3:.2f} ms), rank-1 equal: {pipe_same}")
    return {"pipeline_ms_per_step": dtp * 1e3, "pipeline_spectra_per_s": n / dtp, "pipeline_depth": a.depth,
            "pipeline_rank1_equal_default": pipe_same,
            "graph_ms_per_step": dt * 1e3, "graph_spectra_per_s": n / dt, "eager_default_ms": dt_base * 1e3,
            "eager_default_spectra_per_s": n / dt_base, "eager_default_stage_ms": st_base,
            "rank1_equal_default": same, "tune_s": tune_s, "configs": [{"config": c, "eager_ms": t} for c, t in tried],
            "chosen": cfg, "stage_ms": stage, "k1_ms_sum": k1_ms, "k2_ms_sum": k2_ms, "k1_impl": sched.k1_impl,
            "join_stage_ms": join_ms, "classes": classes}, sched


def algorithmic_bytes(ds, gi, n, l1):
    """Per-launch algorithmic HBM bytes of K1 (fused top-K) and K2 on this batch.
    K1: sum_h [4*matched_h + 8*peaks_s(h)*(maxfz_h - 1)] + 4*peaks_s + 12*n_cand_s + 29 per hypothesis
        (postings that must be read, a bin-bound pair per (peak, z) lookup, peak masses, the candidate list,
        hypothesis window/charge in, scored/matched out).
    K2: sum_c (21 + L_c) + 8*peaks_s + 37 per candidate (seq_off 8 + len 1 + mask 8 + mass32 4 + residues L,
        peak mass+intensity, candidate row/hypothesis/maxfz in, 28-byte result out)."""
    import numpy as np
    hyp_off = ds.hyp["hyp_off"][: n + 1].cpu().numpy()
    nh = int(hyp_off[-1])
    pk_off = ds.peaks["off"][: n + 1].cpu().numpy()
    npk = np.diff(pk_off)
    hspec = ds.hyp["spec"][:nh].cpu().numpy()
    mfz = ds.hyp["maxfz"][:nh].cpu().numpy().astype(np.int64)
    matched = ds.matched[:nh].cpu().numpy().astype(np.int64)
    nc = ds.n_cand[:n].cpu().numpy().astype(np.int64)
    k1 = int((4 * matched + 8 * npk[hspec] * (mfz - 1)).sum() + 4 * npk.sum() + 12 * nc.sum() + 29 * nh)
    m = (np.arange(ds.stride)[None, :] < nc[:, None]).ravel()
    rows = ds.c_row[: n * ds.stride].cpu().numpy()[m]
    k2 = int((21 + l1.length[rows].astype(np.int64)).sum() + 8 * npk[nc > 0].sum() + 37 * len(rows))
    return {"k1_topk": k1, "k2": k2, "n_hyp": nh, "n_rescored": int(len(rows)), "matched": int(matched.sum())}


def algorithmic_bytes_join(ds, gi, n):
    """Per-launch algorithmic HBM bytes of the bucket-join K1 stages on this batch.
    k1j_count: 6 B per posting of every block with at least one pair (key + row, read once) + 8 B block bounds,
        per pair its 32-byte record (peaks, slot, window, peak count, maxfz) and the spectrum's peak masses
        (+ intensities in feature mode), and 4 B (8 B feature mode) per window-row slot written;
    k1j_select: the slots read back, 12 B per candidate written, per hypothesis window/slot/mode in (25 B) and
        scored/matched out (12 B)."""
    import numpy as np
    j = ds.join
    b = gi.block_arrays()
    bp = j["blk_pairs"].cpu().numpy().view(np.uint32).astype(np.int64)
    off = b["off"].cpu().numpy().view(np.uint32).astype(np.int64)
    per_blk = np.diff(bp)
    touched = per_blk > 0
    n_pairs = int(bp[-1])
    hyp_off = ds.hyp["hyp_off"][: n + 1].cpu().numpy()
    nh = int(hyp_off[-1])
    rec = j["pairs"][:n_pairs].cpu().numpy()
    np_pair = rec[:, 3] & 0xFFFFFFFF                                              # peak count field
    slots = int(j["h_slot"][nh].item())
    sw = 8 if j["feat"] else 4
    pk_b = 8 if j["feat"] else 4
    count = int(6 * np.diff(off)[touched].sum() + 8 * touched.sum() + 32 * n_pairs + pk_b * np_pair.sum()
                + sw * slots)
    nc = ds.n_cand[:n].cpu().numpy().astype(np.int64)
    select = int(sw * slots + 12 * nc.sum() + 37 * nh)
    return {"k1j_count": count, "k1j_select": select, "n_pairs": n_pairs, "n_slots": slots,
            "blocks_touched": int(touched.sum())}


def cpu_baseline(vp, l1, sp, target_s: float):
    """The oracle's whole-search (OpenMP over spectra) on a bounded prefix of the same spectra."""
    import numpy as np

    from oracle import oracle as O
    from vulcan.gpu.host import Spectra
    hix = O.HostIndex(l1)

    def take(n):
        off = sp.offsets[: n + 1]
        return Spectra(sp.run, sp.scan[:n], sp.index[:n], sp.rt[:n], sp.precursor_mz[:n], sp.charge[:n],
                       sp.mz[: off[-1]], sp.intensity[: off[-1]], off.copy())
    n0 = min(2000, len(sp))
    t0 = time.perf_counter()
    O.search(vp, hix, take(n0))
    dt = time.perf_counter() - t0
    n = int(min(len(sp), max(n0, n0 * target_s / max(dt, 1e-6))))
    sub = take(n)
    t0 = time.perf_counter()
    res = O.search(vp, hix, sub)
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "spectra/s", "cores": int(O.lib().or_max_threads()), "kind": "port",
            "sample": f"first {n} of {len(sp)} spectra of the same workload, {dt:.1f} s, OpenMP C++ oracle "
                      f"(K0, hypotheses, per-row K1, top-K + closure, K2, rank 1)",
            "cpu_model": _cpu_model()}, res, n


def cpu_baseline_modes(vp, l1, sp, maxh: int, target_s: float):
    """The oracle's component functions in pipeline order for the open / mass-offset search (K0, mode hypotheses,
    per-row K1 counts of every window row, top-K + closure, shifted K2, rank 1), OpenMP inside each, on a
    bounded prefix of the same spectra."""
    import numpy as np

    from oracle import oracle as O
    hix = O.HostIndex(l1)

    def run(sub):
        pk = O.k0(vp, sub)
        hyp = O.hypotheses(vp, sub, pk["valid"], l1.mass32, maxh)
        k1 = O.k1_counts(vp, hix, pk, hyp)
        stride = 2 * vp.topk if vp.pair_closure else vp.topk
        tk = O.topk(vp, hix, hyp, k1, stride)
        nc = tk["n_cand"]
        m = (np.arange(stride)[None, :] < nc[:, None]).ravel()
        rows, hyps = tk["row"][m], tk["hyp"][m]
        shift = np.where(hyp["mode"][hyps] == 0, np.float32(0),
                         (hyp["M"][hyps] - l1.mass32[rows]).astype(np.float32)).astype(np.float32)
        start = np.concatenate([[0], np.cumsum(nc)[:-1]]).astype(np.int64)
        k2 = O.k2(vp, hix, pk, rows, hyps, hyp["maxfz"], grp_start=start, grp_count=nc, shift=shift)
        sel, _ = O.rank(vp, start, nc, rows, k2["hyperscore"], k2["matched_b"], k2["matched_y"], 1)
        return np.where(sel >= 0, rows[np.maximum(sel, 0)], -1), int(k1["win_off"][-1])
    n0 = min(20, len(sp))
    t0 = time.perf_counter()
    run(_take(sp, n0))
    dt = time.perf_counter() - t0
    n = int(min(len(sp), max(n0, n0 * target_s / max(dt, 1e-6))))
    t0 = time.perf_counter()
    rows, n_win = run(_take(sp, n))
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "spectra/s", "cores": int(O.lib().or_max_threads()), "kind": "port",
            "sample": f"first {n} of {len(sp)} spectra of the same workload, {dt:.1f} s, {n_win} window rows; "
                      f"OpenMP C++ oracle functions in pipeline order (K0, mode hypotheses, per-row K1, top-K + "
                      f"closure, shifted K2, rank 1)", "cpu_model": _cpu_model()}, rows, n


def gbm_leg(a):
    """Gradient-boosted tree scorer (vh_gbm_decision via GpuHGB): rows/s on `--gbm-rows` rows of 24 f64 features
    resident in HBM, a 300-tree binary HistGradientBoosting model (at most 31 leaves per tree) fitted on 50k
    synthetic rows; sklearn's decision_function on the host cores as the CPU baseline (a bounded sample)."""
    import numpy as np
    import torch
    from sklearn.ensemble import HistGradientBoostingClassifier

    from vulcan.gpu import _lib as K
    from vulcan.rescore.fast import GpuHGB
    nf, iters = 24, 300
    rng = np.random.default_rng(0)
    X = rng.normal(size=(50000, nf))
    X[rng.random(X.shape) < 0.03] = np.nan
    xz = np.nan_to_num(X)
    y = (xz[:, 0] * xz[:, 1] + np.sin(xz[:, 2] * 3) + rng.normal(size=len(X)) * 0.3 > 0).astype(int)
    m = HistGradientBoostingClassifier(max_iter=iters, max_leaf_nodes=31, early_stopping=False,
                                       random_state=0).fit(X, y)
    g = GpuHGB(m)
    n = a.gbm_rows
    Xt = rng.normal(size=(n, nf))
    Xt[rng.random(Xt.shape) < 0.03] = np.nan
    Xd = torch.from_numpy(Xt).cuda()
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    for _ in range(a.warmup):
        g.decision_device(Xd, out)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(st)
    for _ in range(a.steps):
        g.decision_device(Xd, out)
    e1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / a.steps
    g.check()
    chk = min(n, 100_000)
    equal = bool(np.array_equal(out[:chk].cpu().numpy(), m.decision_function(Xt[:chk])))
    bytes_per_launch = n * (nf * 8 + 8)                  # feature rows read + one f64 score written
    ach = bytes_per_launch / (ms * 1e-3) / 1e9
    workload = f"gbm-{iters}-trees-{nf}-features-{n}-rows"
    tpath = os.path.join(ROOT, "profiles", f"traffic_gbm_kernel_{workload}.json")
    traffic = json.load(open(tpath)).get("hbm_bytes_per_launch") if os.path.exists(tpath) else None
    res = {"metric": "rows/s (gradient-boosted tree scorer, f64, bit-identical to sklearn)", "value": n / (ms * 1e-3),
           "unit": "rows/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": workload, "n_rows": n, "n_features": nf,
                      "n_trees": g.T, "n_nodes": g.M, "node_chunks": g.n_chunks, "max_depth": g.depth,
                      "parallelism": "single-gpu"},
           "kernel_ms": ms, "tree_walks_per_s": n * g.T / (ms * 1e-3), "equal_sklearn_first_rows": equal,
           "roofline": {"bound": "hbm", "kernel": "gbm_kernel", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "algorithmic_bytes_per_launch": bytes_per_launch,
                        "note": "walks are bound by dependent LDS reads (node, then feature), not by HBM"},
           "cpu_baseline": None}
    if not a.no_cpu:
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        ns = min(n, 1_000_000)
        t0 = time.perf_counter()
        m.decision_function(Xt[:ns])
        tc = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": ns / tc, "unit": "rows/s", "cores": cores, "kind": "reference",
                               "sample": f"sklearn {__import__('sklearn').__version__} decision_function (the call "
                                         f"the reference makes on the CPU) on the first {ns} rows"}
    print(json.dumps(res), flush=True)


def family_leg(a):
    """Family top-2 retention (vulcan.gpu.family.search_family) over a multi-replicate table (T, two entrapment
    families, two shuffled decoy sets) built from `--family-proteins` synthetic proteins, on `--n-spectra`
    (default 50,000) synthetic spectra; a step = one whole family search (device K0 -> hypotheses -> K1 ->
    per-family shortlist + closure -> K2 -> top-2 table on the host)."""
    import numpy as np
    import torch

    from vulcan.gpu import engine as E
    from vulcan.gpu import family as F
    from vulcan.index import build as B
    from vulcan.index import synth
    ip = synth.WORKLOADS[1].index_params()
    t0 = time.perf_counter()
    l1, lab = B.build_multi_decoy_tables(synth.proteome(a.family_proteins, 11),
                                         {"entrap_seeds": [2001, 2002], "decoy_seeds": [1001, 1002], **ip})
    log(f"multi-replicate table: {len(l1)} rows ({time.perf_counter() - t0:.1f} s)")
    n = a.n_spectra or 50_000
    sp, _, _ = synth.spectra(l1, lab, 1, n=n, seed=7, index_params=ip)
    p = synth.WORKLOADS[1].search_params()
    gi = E.GpuIndex(l1, p, STATIC, dict(synth.WORKLOADS[1].var_mods))
    fi = F.FamilyIndex.from_l1(l1, gi, lab)
    for _ in range(max(a.warmup, 1)):
        F.search_family(sp, gi, l1, fi, p)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stage = {}
    for _ in range(a.steps):
        tbl, st = F.search_family(sp, gi, l1, fi, p)
        for k, v in st.gpu_ms.items():
            stage[k] = stage.get(k, 0.0) + v / a.steps
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / a.steps
    res = {"metric": "spectra/s (family top-2 retention search)", "value": n / wall, "unit": "spectra/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": wall * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"family-top2-{a.family_proteins}-proteins-{n}-spectra", "n_spectra": n,
                      "n_peptidoforms": len(l1), "decoy_sets": 2, "entrapment_seeds": 2, "topk": p.topk,
                      "parallelism": "single-gpu"},
           "stage_ms": stage, "n_table_rows": tbl.num_rows, "counts": st.counts,
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(res), flush=True)


def big_index_leg(a):
    """An index of more postings than u32 offsets address: the phospho-scale workload's table semantics (STY
    phospho <= 3 + Met-ox <= 2, index top-200 + closure) on `--big-proteins` synthetic proteins (default 100,000:
    about 180M peptidoforms and 4.7G postings), held as row-range slices of at most max_slice_entries postings
    each. A step = one plain search (join K1 over every slice) of `--n-spectra` (default 20,000) spectra; rank-1
    is checked against the oracle's whole search on a bounded sample."""
    import numpy as np
    import torch

    from vulcan.gpu import engine as E
    from vulcan.gpu._abi import make_params
    from vulcan.index import synth
    w = synth.WORKLOADS[4]
    t0 = time.perf_counter()
    l1, lab = synth.build_index(4, n_proteins=a.big_proteins, progress=lambda m: log(f"table {m}"))
    table_s = time.perf_counter() - t0
    log(f"table: {len(l1)} peptidoforms ({table_s:.1f} s)")
    n = a.n_spectra or 20_000
    sp, truth, _ = synth.spectra(l1, lab, 4, n=n, seed=17)
    p = w.search_params()
    t0 = time.perf_counter()
    gi = E.GpuIndex(l1, p, STATIC, dict(w.var_mods))
    log(f"table on the device, {gi.n_postings} postings ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    sl = gi.block_slices()
    build_s = time.perf_counter() - t0
    log(f"index: {len(sl)} row-range slices of {[b['n_postings'] for b in sl]} postings ({build_s:.1f} s)")
    ds = E.DeviceSearch(gi, p, n, k1_impl="join")
    raw = E.upload_spectra(sp, gi.dev)
    dt, stage_ms = time_steps(ds, raw, a.steps, a.warmup)
    if int(ds.status.item()) != 0:
        raise RuntimeError(f"status {int(ds.status.item())}")
    rows = ds.rank1_rows(n)
    sig = truth >= 0
    out = {"metric": "spectra/s (index+rescore), index beyond 2^32 postings", "value": n / dt, "unit": "spectra/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"phospho-semantics-{a.big_proteins}-proteins-{n}-spectra", "n_spectra": n,
                      "n_peptidoforms": len(l1), "n_postings": gi.n_postings, "n_slices": len(sl),
                      "slice_postings": [b["n_postings"] for b in sl], "max_slice_entries": p.max_slice_entries,
                      "topk": p.topk, "pair_closure": True, "parallelism": "single-gpu"},
           "stage_ms": stage_ms, "table_build_s": table_s, "index_build_s": build_s,
           "device_bytes": gi.device_bytes, "top1_truth_recovery": float(np.mean(rows[sig] == truth[sig])),
           "roofline": None, "cpu_baseline": None}
    if not a.no_cpu:
        log("cpu baseline ...")
        cb, res, ncpu = cpu_baseline(make_params(p, STATIC, dict(w.var_mods)), l1, sp, a.cpu_seconds)
        out["cpu_baseline"] = cb
        out["rank1_equal_cpu_sample"] = bool(np.array_equal(rows[:ncpu], res["row"][:ncpu]))
    print(json.dumps(out), flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=1,
                    help="index into vulcan.index.synth.WORKLOADS (1 = the proteome-scale plain search)")
    ap.add_argument("--n-spectra", type=int, default=None)
    ap.add_argument("--depth", type=int, default=3,
                    help="closed / open search legs: steps in flight at once, one stream each (1: one at a time)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the rank-1 comparison with the CPU sample")
    ap.add_argument("--chunk", type=int, default=None, help="override rows per index bucket")
    ap.add_argument("--bin-ppm", type=float, default=None, help="override directory bin width")
    ap.add_argument("--tile", type=int, default=None, help="override K1 LDS tile")
    ap.add_argument("--k1", choices=("auto", "tile", "join", "wide"), default=None,
                    help="K1 implementation (default: the workload's search parameters, auto)")
    ap.add_argument("--block-q", type=int, default=None, help="join K1: log2 postings per index block")
    ap.add_argument("--block-rows", type=int, default=None, help="join K1: log2 max rows per index block")
    ap.add_argument("--prefilter", action="store_true",
                    help="plain vs learned prefilter comparison (default for the phospho-prefilter workload)")
    ap.add_argument("--keep", type=int, default=64, help="prefilter: at most this many candidates kept before closure")
    ap.add_argument("--keep-min", type=int, default=8, help="prefilter: always keep the best this many by score")
    ap.add_argument("--margin", type=float, default=3.0,
                    help="prefilter: also keep candidates scoring within this margin of the best")
    ap.add_argument("--no-prefilter", action="store_true",
                    help="default workload: leave out the plain vs learned comparison on the phospho workload")
    ap.add_argument("--prefilter-spectra", type=int, default=200_000,
                    help="spectra of the phospho workload in the default line's prefilter comparison")
    ap.add_argument("--recall-spectra", type=int, default=5000,
                    help="spectra whose shortlist rank 1 is checked against exhaustive rescoring (0: skip)")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--n-train", type=int, default=20000, help="prefilter: training spectra (own seed)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="only run the u32 copy kernel on this many MiB (HBM counter calibration)")
    ap.add_argument("--gbm", action="store_true", help="gradient-boosted tree scorer leg (rows/s) instead")
    ap.add_argument("--gbm-rows", type=int, default=1_000_000)
    ap.add_argument("--family", action="store_true", help="family top-2 retention leg instead")
    ap.add_argument("--family-proteins", type=int, default=2000)
    ap.add_argument("--big-index", action="store_true",
                    help="search an index beyond 2^32 postings (row-range slices) instead")
    ap.add_argument("--big-proteins", type=int, default=100_000)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {a.gpus} is not supported (the multi-GPU sharding "
              f"design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from vulcan.gpu import _lib as K
    from vulcan.gpu import engine as E
    from vulcan.index import synth
    dev = torch.device("cuda:0")
    if a.calib_copy_mib:
        n = a.calib_copy_mib * 1024 * 1024 // 4
        src = torch.ones(n, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            K.check(K.lib().vh_copy_u32(K.ptr(src), K.ptr(dst), n, K.stream()), "copy")
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes_each_way": n * 4, "launches": 5}))
        return
    if a.gbm:
        gbm_leg(a)
        return
    if a.family:
        family_leg(a)
        return
    if a.big_index:
        big_index_leg(a)
        return
    w = synth.WORKLOADS[a.workload]
    t0 = time.perf_counter()
    l1, lab = synth.build_index(a.workload)
    log(f"table: {len(l1)} peptidoforms ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    sp, truth, _ = synth.spectra(l1, lab, a.workload, n=a.n_spectra)
    n = len(sp)
    log(f"spectra: {n} ({time.perf_counter() - t0:.1f} s)")
    p = w.search_params()
    modes = {}
    if w.open_da is not None:
        from vulcan.ptm.search import ModeParams, mode_search_params
        mp = ModeParams(open_da=w.open_da)
        p = mode_search_params(p, mp)
        modes = {"open_da": mp.open_da, "k1_threshold_open": mp.k1_threshold_open}
    for k, v in (("chunk_size", a.chunk), ("bin_ppm", a.bin_ppm), ("tile", a.tile), ("k1_impl", a.k1),
                 ("block_postings_log2", a.block_q), ("block_rows_log2", a.block_rows)):
        if v is not None:
            setattr(p, k, v)
    torch.zeros(1, device=dev)   # HIP runtime and library initialisation stay out of index_build_s
    K.lib()
    torch.cuda.synchronize()
    gi = E.GpuIndex(l1, p, STATIC, w.var_mods)
    log(f"index: {gi.n_postings} postings, {gi.n_buckets} buckets x {gi.n_bins} bins, build {gi.build_s:.3f} s")
    raw = E.upload_spectra(sp, dev)
    if w.leg == "uneven":
        res, sched = uneven_leg(a, w, l1, gi, p, sp, raw, n)
        ab = algorithmic_bytes(sched, gi, n, l1)
        if sched.k1_impl == "join":
            ab.update(algorithmic_bytes_join(sched, gi, n))
            rk, rb, rms = "k1j_count", ab["k1j_count"], res["join_stage_ms"]["k1j_count"]
        else:
            rk, rb, rms = "k1_topk (sum over classes)", ab["k1_topk"], res["k1_ms_sum"]
        ach = rb / (rms * 1e-3) / 1e9
        out = {"metric": METRIC, "value": res["pipeline_spectra_per_s"], "unit": "spectra/s", "n_gpus": 1,
               "steps": a.steps, "warmup": a.warmup, "ms_per_step": res["pipeline_ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": w.name, "n_spectra": n, "n_runs": w.n_runs, "n_peptidoforms": len(l1),
                          "n_postings": gi.n_postings, "topk": p.topk, "pair_closure": True,
                          "precursor_ppm": list(p.precursor_ppm), "fragment_da": list(p.fragment_ppm),
                          "max_fragment_charge": p.max_fragment_charge, "precursor_charge": list(p.precursor_charge),
                          "parallelism": "single-gpu",
                          "launch": f"the autotuned step, {a.depth} steps in flight on {a.depth} streams"},
               "roofline": {"bound": "hbm", "kernel": rk, "achieved": ach,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                            "algorithmic_bytes_per_launch": rb},
               "uneven": res, "counts": ab, "cpu_baseline": None}
        if not a.no_cpu:
            log("cpu baseline ...")
            cb, cres, ncpu = cpu_baseline(sched.vp, l1, sp, a.cpu_seconds)
            out["cpu_baseline"] = cb
            out["rank1_equal_cpu_sample"] = bool(np.array_equal(sched.rank1_rows(n)[:ncpu], cres["row"][:ncpu]))
        print(json.dumps(out), flush=True)
        return
    if a.prefilter or w.leg == "prefilter":
        res = prefilter_leg(a, w, l1, lab, gi, p, raw, n, sp)
        L = res["learned"]
        st = L["stage_ms"]
        mr = L["mlp_roofline"]
        out = {"metric": METRIC, "value": L["spectra_per_s"], "unit": "spectra/s", "n_gpus": 1, "steps": a.steps,
               "warmup": a.warmup, "ms_per_step": L["ms_per_step"], "higher_is_better": True, "scaling": "weak",
               "vs_baseline": None, "dtype": "f32+bf16", "data": "synthetic",
               "config": {"workload": w.name, "n_spectra": n, "n_peptidoforms": len(l1),
                          "n_postings": gi.n_postings, "index_topk": res["index_topk"], "keep_min": res["keep_min"],
                          "keep_max": res["keep_max"], "keep_margin": res["margin"],
                          "pair_closure": True, "precursor_ppm": list(p.precursor_ppm),
                          "fragment_da": list(p.fragment_ppm), "max_fragment_charge": p.max_fragment_charge,
                          "parallelism": "single-gpu", "launch": f"{a.depth} steps in flight on {a.depth} streams"},
               "stage_ms": st, "index_build_s": gi.build_s,
               "roofline": dict(mr, kernel="prefilter_score", traffic=None),
               "prefilter": res, "cpu_baseline": res.pop("cpu_baseline", None)}
        print(json.dumps(out), flush=True)
        return
    pipe = E.SearchPipeline(gi, p, n, depth=a.depth, **modes)
    wall = time_pipeline(pipe, raw, a.steps, a.warmup)
    log(f"timed {a.steps} steps, {a.depth} in flight: {wall:.3f} s")
    ds = pipe.slots[0]
    dt_serial, stage_ms = time_steps(ds, raw, a.steps, a.warmup)   # one step at a time; stages from a second pass
    log(f"one step at a time: {dt_serial * 1e3:.3f} ms per step")
    for d in pipe.slots:
        if int(d.status.item()) != 0:
            raise RuntimeError(f"k1_topk status {int(d.status.item())}")
    sel = ds.sel[:n].cpu().numpy()
    rows = ds.c_row[: n * ds.stride].cpu().numpy()
    best = np.where(sel >= 0, rows[np.maximum(sel, 0)], -1)
    slots_equal = all(np.array_equal(d.rank1_rows(n), best) for d in pipe.slots[1:])
    sig = truth >= 0
    recovery = float(np.mean(best[sig] == truth[sig]))
    ab = algorithmic_bytes(ds, gi, n, l1)
    kernels = ("k1_topk", "k2")
    if ds.k1_impl == "wide":
        ab["k1w_topk"] = ab["k1_topk"]
        kernels = ("k1w_topk", "k2")
    if ds.k1_impl == "join":
        ab.update(algorithmic_bytes_join(ds, gi, n))
        stage_ms["k1"] = sum(stage_ms[k] for k in ("k1j_pairs", "k1j_count", "k1j_select"))
        kernels = ("k1j_count", "k1j_select", "k2")
    dom = max(kernels, key=lambda k: stage_ms[k])
    geometry = {"chunk_size": 1 << gi.bucket_log2, "bin_shift": gi.bin_shift, "n_bins": gi.n_bins, "tile": p.tile,
                "xcd_group": p.xcd_group, "k1_impl": ds.k1_impl}
    if ds.k1_impl == "join":
        bv = gi.blocks()
        geometry.update(block_postings_log2=p.block_postings_log2, block_rows_log2=p.block_rows_log2,
                        n_blocks=bv.n_blocks, max_block_rows=bv.max_rows, max_block_postings=bv.max_postings)
    achieved = ab[dom] / (stage_ms[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_{w.name}.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        traffic = tj.get("hbm_bytes_per_launch")
    value = n * a.steps / wall
    out = {"metric": METRIC, "value": value, "unit": "spectra/s", "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": wall / a.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": w.name, "n_spectra": n, "n_peptidoforms": len(l1), "n_postings": gi.n_postings,
                      "topk": p.topk, "pair_closure": True, "precursor_ppm": list(p.precursor_ppm),
                      "fragment_da": list(p.fragment_ppm), "max_fragment_charge": p.max_fragment_charge,
                      "parallelism": "single-gpu", "launch": f"{a.depth} steps in flight on {a.depth} streams"},
           "serial_ms_per_step": dt_serial * 1e3, "serial_spectra_per_s": n / dt_serial,
           "slots_rank1_equal": slots_equal,
           "stage_ms": stage_ms, "index_build_s": gi.build_s, "index_geometry": geometry,
           "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                        "algorithmic_bytes_per_launch": ab[dom]},
           "counts": ab, "top1_truth_recovery": recovery}
    if ds.report == 1:
        dtf, fms = time_with_features(ds, raw, n, a.steps)
        out["with_psm_features"] = {"spectra_per_s": n / dtf, "ms_per_step": dtf * 1e3, "feature_ms": fms,
                                    "note": "the step + rank-1 PSM feature columns on the device "
                                            "(vh_psm_features); host table assembly not included"}
    if modes:
        out["config"]["open_da"] = list(modes["open_da"])
    out["cpu_baseline"] = None
    if not a.no_cpu:
        log("cpu baseline ...")
        if modes:
            from vulcan.gpu._abi import max_hyp_per_spectrum
            cb, crow, ncpu = cpu_baseline_modes(ds.vp, l1, sp, max_hyp_per_spectrum(p, 0, True), a.cpu_seconds)
        else:
            cb, res, ncpu = cpu_baseline(ds.vp, l1, sp, a.cpu_seconds)
            crow = res["row"]
        out["cpu_baseline"] = cb
        if not a.no_check:
            out["rank1_equal_cpu_sample"] = bool(np.array_equal(best[:ncpu], crow[:ncpu]))
    if not modes and a.recall_spectra > 0:
        log("shortlist recall vs exhaustive rescoring ...")
        out["recall_vs_exhaustive"] = recall_vs_exhaustive(gi, l1, p, sp, best, a.recall_spectra)
        log(f"recall {out['recall_vs_exhaustive']['recall']:.4f}")
    if a.workload == 1 and not a.no_prefilter:
        del ds, gi, raw
        torch.cuda.empty_cache()
        out["prefilter"] = prefilter_block(a)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
